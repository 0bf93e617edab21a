"""GPU parity: the HIP kernels through the C ABI against the CPU oracle and
the reference's golden outputs.  Integer work (tables, V records) is
compared bit for bit; the FP64 step too (the library is compiled without FP
contraction and follows the reference's evaluation order)."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, TINY_CASES, load_tiny

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def random_mask(dims, p, seed):
    return (np.random.default_rng(seed).random(dims) < p).astype(np.uint8)


def say(*a):
    print("[case]", *a, flush=True)


# ---------------------------------------------------------------- index build
@pytest.mark.parametrize("name", TINY_CASES)
def test_preprocess_matches_reference_fixture(plb, name):
    d = load_tiny(name)
    for np_ in d["ranks"]:
        say("preprocess", name, np_)
        V, start = plb.preprocess(d["flag"], int(np_))
        x, y, z, t, pid, nid = O.semantic_fields(V)
        assert np.array_equal(start, d[f"np{np_}_start"])
        for k, v in dict(x=x, y=y, z=z, type=t, pid=pid, id=nid).items():
            assert np.array_equal(v, d[f"np{np_}_V_{k}"].astype(np.int64)), (name, np_, k)
        # fields the reference leaves undefined are written as zero
        solid = V["nb"]["node_type"] == O.SOLID
        assert not V["nb"]["node_id"][solid].any() and not V["nb"]["node_pid"][solid].any()


@pytest.mark.parametrize("dims,np_,kind", [((64, 48, 40), 4, "spheres"), ((96, 64, 64), 3, "noise"),
                                           ((128, 128, 128), 8, "spheres")])
def test_preprocess_bitexact_vs_oracle(plb, dims, np_, kind):
    say("preprocess", dims, np_, kind)
    flag = (plb.mask_spheres(dims, 4.0, 8.0, 0.3, 3) if kind == "spheres"
            else plb.mask_noise(dims, 2.0, 0.3, 4))
    V, start = plb.preprocess(flag, np_)
    Vo, so = O.build_V(flag, np_)
    assert np.array_equal(start, so)
    for a, b in zip(O.semantic_fields(V), O.semantic_fields(Vo)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("mode", [O.BB_REFCOMPAT, O.BB_HALFWAY])
@pytest.mark.parametrize("np_", [1, 2, 4])
def test_table_from_V_and_flag_bitexact(plb, mode, np_):
    flag = random_mask((16, 11, 13), 0.55, 40 + np_)
    say("table", mode, np_)
    V, start = O.build_V(flag, np_)
    want = O.pull_table(V, start, mode)
    with plb.Lattice.from_V(V, start, flag.shape, bb_mode=mode) as lat:
        assert np.array_equal(lat.read_table(), want)
    with plb.Lattice.from_flag(flag, np_, bb_mode=mode) as lat:
        assert np.array_equal(lat.read_table(), want)
    for layout, order in [(plb.LAYOUT_AOSOA, plb.ORDER_REFERENCE), (plb.LAYOUT_SOA, plb.ORDER_BRICK),
                          (plb.LAYOUT_AOSOA, plb.ORDER_BRICK), (plb.LAYOUT_AOSOA, plb.ORDER_PENCIL)]:
        for index in (plb.INDEX_I32, plb.INDEX_TILE8):
            with plb.Lattice.from_flag(flag, np_, bb_mode=mode, layout=layout, ordering=order,
                                       brick=(4, 2, 3), index_kind=index) as lat:
                assert np.array_equal(lat.read_table(), want), (layout, order, index)


def test_table_large_bitexact(plb):
    """Wide rows (x-wraps span more than a byte of ids) exercise the TILE8
    int32 fallback."""
    flag = plb.mask_spheres((64, 160, 640), 6.0, 12.0, 0.3, 11)
    say("table large")
    V, start = O.build_V(flag, 4)
    for mode in (O.BB_REFCOMPAT, O.BB_HALFWAY):
        want = O.pull_table(V, start, mode)
        with plb.Lattice.from_flag(flag, 4, bb_mode=mode) as lat:
            assert np.array_equal(lat.read_table(), want)
            b32 = lat.bytes_per_update()
        for layout in (0, 1):
            with plb.Lattice.from_flag(flag, 4, bb_mode=mode, index_kind=plb.INDEX_TILE8,
                                       layout=layout) as lat:
                assert np.array_equal(lat.read_table(), want)
                b8 = lat.bytes_per_update()
        assert b32 == 2 * 19 * 8 + 72 and 2 * 19 * 8 + 18 < b8 < 2 * 19 * 8 + 24


# ---------------------------------------------------------------- FP64 step parity
@pytest.mark.parametrize("name", TINY_CASES)
def test_fp64_step_matches_reference_fixture_bitwise(plb, name):
    d = load_tiny(name)
    tau, gx, steps = float(d["tau"]), float(d["gx"]), int(d["steps"])
    for np_ in d["ranks"]:
        say("fp64 fixture", name, np_)
        V, start = O.build_V(d["flag"], int(np_))
        with plb.Lattice.from_V(V, start, d["flag"].shape, tau=tau, gx=gx) as lat:
            lat.step(0, steps)
            assert np.array_equal(lat.read_populations(), d[f"np{np_}_f0"]), (name, np_)
            assert np.array_equal(lat.read_macros(), d[f"np{np_}_macros"]), (name, np_)


def test_fp64_config1_matches_reference(plb):
    """Config 1 (64^3 noise, tau 1.0, g 1e-6): 1000 steps on 1 rank and 200
    steps on 4 ranks, SHA-256 of populations and macros as the reference wrote
    them."""
    meta = json.load(open(os.path.join(GOLDEN, "config1.json")))
    m = meta["mask"]
    flag = plb.mask_noise(tuple(m["dims"]), m["sigma"], m["porosity"], m["seed"])
    assert sha(flag) == m["sha256"]
    for np_ in (1, 4):
        g = meta[f"np{np_}"]
        say("config1", np_, g["steps"])
        with plb.Lattice.from_flag(flag, np_, tau=1.0, gx=1e-6) as lat:
            lat.step(0, g["steps"])
            f = lat.read_populations()
            mac = lat.read_macros()
        idx = np.array(g["sample_idx"])
        assert np.array_equal(f[idx], np.array(g["sample_f0"]))
        assert sha(f) == g["f0_sha256"]
        assert sha(mac) == g["macros_sha256"]


@pytest.mark.parametrize("mode", [O.BB_REFCOMPAT, O.BB_HALFWAY])
def test_fp64_layouts_orders_blocks_identical(plb, mode):
    flag = plb.mask_spheres((40, 36, 44), 3.0, 6.0, 0.35, 2)
    V, start = O.build_V(flag, 2)
    tab = O.pull_table(V, start, mode)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 60, 0.7, 1e-4)
    variants = [dict(), dict(layout=1), dict(ordering=1), dict(layout=1, ordering=1, brick=(8, 8, 8)),
                dict(block_size=64), dict(block_size=1024), dict(ordering=1, brick=(40, 1, 1)),
                dict(index_kind=1), dict(index_kind=1, layout=1, block_size=512),
                dict(index_kind=1, layout=1, ordering=1, brick=(8, 8, 8)),
                dict(index_kind=1, layout=1, ordering=2, brick=(44, 4, 1)),
                dict(ordering=2, brick=(1000, 7, 3)),
                dict(nodes_per_thread=2), dict(nodes_per_thread=2, index_kind=1, layout=1, block_size=128),
                dict(nodes_per_thread=2, index_kind=1, layout=1, ordering=2, brick=(44, 8, 1), block_size=512),
                dict(waves_per_simd=4), dict(index_kind=1, layout=1, ordering=2, brick=(44, 4, 1), waves_per_simd=5),
                dict(waves_per_simd=3, block_size=128, nodes_per_thread=2), dict(waves_per_simd=1, block_size=1024),
                dict(waves_per_simd=7, block_size=64, index_kind=1, layout=1)]
    for kw in variants:
        say("variant", mode, kw)
        with plb.Lattice.from_flag(flag, 2, tau=0.7, gx=1e-4, bb_mode=mode, **kw) as lat:
            lat.step(0, 25)
            lat.step(25, 35)
            assert np.array_equal(lat.read_populations(), f0), kw
            assert np.array_equal(lat.read_macros(), mac), kw


def test_fp64_halfway_vs_oracle_push(plb):
    flag = random_mask((20, 18, 16), 0.6, 77)
    V, start = O.build_V(flag, 1)
    f0, f1 = O.init_equilibrium(len(V))
    mac = O.push_run(V, start, f0, f1, 50, 0.9, 5e-4, mode=O.BB_HALFWAY)
    with plb.Lattice.from_V(V, start, flag.shape, tau=0.9, gx=5e-4, bb_mode=plb.BB_HALFWAY) as lat:
        lat.step(0, 50)
        assert np.array_equal(lat.read_populations(), f0)
        assert np.array_equal(lat.read_macros(), mac)


def test_write_read_populations_roundtrip_and_resume(plb):
    """Checkpoint/resume: a state written back continues bit-identically."""
    flag = random_mask((12, 12, 12), 0.6, 5)
    with plb.Lattice.from_flag(flag, tau=0.8, gx=1e-4) as a, plb.Lattice.from_flag(flag, tau=0.8, gx=1e-4) as b:
        a.step(0, 30)
        snap = a.read_populations()
        a.step(30, 20)
        b.write_populations(snap)
        assert np.array_equal(b.read_populations(), snap)
        b.step(30, 20)
        assert np.array_equal(a.read_populations(), b.read_populations())


# ---------------------------------------------------------------- reductions, guard
def test_reduce_vs_oracle(plb):
    flag = plb.mask_spheres((48, 48, 48), 4.0, 8.0, 0.3, 6)
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start, O.BB_HALFWAY)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 100, 0.8, 1e-5)
    want = O.reduce(f0, mac[1], flag.shape, 0.8, 1e-5)
    with plb.Lattice.from_flag(flag, tau=0.8, gx=1e-5, bb_mode=plb.BB_HALFWAY) as lat:
        lat.step(0, 100)
        r = lat.reduce()
    for k in ("mass", "sum_ux", "mean_ux", "q", "K"):
        assert r[k] == pytest.approx(want[k], rel=1e-12), k


def test_density_guard_matches_reference(plb):
    """The reference exits with -6 at the first node whose rho leaves
    [0.5, 1.2] (porelb.c:338-342); the fixture holds its message."""
    d = dict(np.load(os.path.join(GOLDEN, "fault.npz")))
    flag, tau, gx = d["flag"], float(d["tau"]), float(d["gx"])
    msg = str(d["message"])
    assert int(d["returncode"]) == 250  # exit(-6) & 0xff
    V, start = O.build_V(flag, 1)
    f0, f1 = O.init_equilibrium(len(V))
    with pytest.raises(O.Fault) as ei:
        O.push_run(V, start, f0, f1, 200, tau, gx)
    with plb.Lattice.from_V(V, start, flag.shape, tau=tau, gx=gx) as lat:
        assert lat.step_rc(0, 200) == plb.ERR_DENSITY
        ft = lat.last_fault()
        assert lat.step_rc(200, 5) == plb.ERR_DENSITY  # state stays frozen
        f_after = lat.read_populations()
    assert ft["tripped"] == 1 and ft["step"] == ei.value.step and ft["node"] == ei.value.node
    text = f"{ft['pid']} >>> at {ft['local_id']} ({ft['x']}, {ft['y']}, {ft['z']}),  Macro denstity  = {ft['rho']:g} , < 0.5 !"
    assert text == msg
    # the frozen state is the output of the faulting step
    g0, g1 = O.init_equilibrium(len(V))
    if ei.value.step > 0:
        O.push_run(V, start, g0, g1, ei.value.step, tau, gx)
    with pytest.raises(O.Fault):
        O.push_run(V, start, g0, g1, 1, tau, gx, it0=ei.value.step)
    assert np.array_equal(f_after, g0)


# ---------------------------------------------------------------- reduced precision
@pytest.mark.parametrize("prec,rnd,tol_u,tol_k", [(1, 0, 2e-5, 1e-5), (2, 0, 2e-2, 1e-2), (2, 1, 5e-3, 1e-4)])
@pytest.mark.parametrize("bb,kw", [(1, dict()), (0, dict(index_kind=1, layout=1, ordering=2, brick=(48, 8, 1)))])
def test_shifted_storage_vs_fp64(plb, prec, rnd, tol_u, tol_k, bb, kw):
    """FP32/FP16 shifted storage against FP64 with the same bounce-back; the
    reference bounce order includes constant entries (TILE8 code 254).  FP16
    rounded to nearest drifts (1e-2); stochastic rounding keeps K within the
    metric's 1e-4."""
    flag = plb.mask_noise((48, 48, 48), 2.0, 0.35, 12)
    res = {}
    for p in (0, prec):
        say("precision", p, rnd, bb, kw)
        with plb.Lattice.from_flag(flag, tau=1.0, gx=1e-6, precision=p, bb_mode=bb, fp16_rounding=rnd,
                                   **kw) as lat:
            lat.step(0, 2000)
            res[p] = (lat.read_macros(), lat.reduce())
    u64, u = res[0][0][1:], res[prec][0][1:]
    err_u = np.linalg.norm(u - u64) / np.linalg.norm(u64)
    err_k = abs(res[prec][1]["K"] / res[0][1]["K"] - 1)
    say("err_u", err_u, "err_k", err_k)
    assert err_u < tol_u and err_k < tol_k
    assert abs(res[prec][1]["mass"] / res[0][1]["mass"] - 1) < 1e-6


def test_halfway_mass_conserved_at_scale(plb):
    """Size-independent property at a 256^3 sphere pack: half-way bounce-back
    conserves mass to rounding (FP64)."""
    flag = plb.mask_spheres((256, 256, 256), 6.0, 12.0, 0.3, 1)
    with plb.Lattice.from_flag(flag, tau=0.8, gx=1e-6, bb_mode=plb.BB_HALFWAY) as lat:
        m0 = lat.reduce()["mass"]
        lat.step(0, 200)
        m1 = lat.reduce()["mass"]
        n = lat.N
    assert abs(m0 - n) < 1e-9 * n
    assert abs(m1 - m0) < 1e-11 * n
