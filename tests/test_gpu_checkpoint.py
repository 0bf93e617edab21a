"""Checkpoint/resume (the empty save_state/recovery_state stubs of
porelb.c:367-373) anchored to the oracle, and the state API behind it.

- porelb.x resumed from a checkpoint writes the part files, and its next
  checkpoint holds the populations, of the oracle's straight push run of
  porelb.c over the same total steps: several ranks, a resume across a save
  step, FP64 bit for bit.
- FP32 and FP16 storage resume bit for bit from the stored elements
  (plb_read_state / plb_write_state), in the library and through porelb.x;
  a checkpoint of another precision is refused.
- A state written back after a density-guard trip continues like the oracle.
- Two V records on one voxel are refused when the library reorders nodes."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "porelb_amd", "bin")
HDR = 40  # checkpoint header bytes: magic, precision, elem, scale, step, nodes, rho0


def say(*a):
    print("[case]", *a, flush=True)


def run(args, cwd, env=None):
    e = dict(os.environ)
    e.update(env or {})
    say(" ".join(args), env)
    return subprocess.run(args, cwd=cwd, capture_output=True, text=True, env=e, timeout=120)


def read_parts(cwd, num_p, it):
    parts = [np.fromfile(os.path.join(cwd, f"part_{p:03d}_{it:07d}.bin")).reshape(4, -1) for p in range(num_p)]
    return np.concatenate(parts, axis=1)


def read_ckpt(path, n, dtype):
    raw = open(path, "rb").read()
    hdr = np.frombuffer(raw[:16], dtype="<i4")
    step, nodes = np.frombuffer(raw[16:32], dtype="<i8")
    assert hdr[0] == 0x32424C50 and nodes == n
    return dict(precision=int(hdr[1]), elem=int(hdr[2]), scale=int(hdr[3]), step=int(step),
                rho0=float(np.frombuffer(raw[32:40], dtype="<f8")[0]),
                state=np.frombuffer(raw[HDR:], dtype=dtype).reshape(n, 19))


@pytest.mark.parametrize("num_p", [1, 3])
def test_cli_resume_equals_oracle_straight_run(tmp_path, num_p):
    """23 steps with a checkpoint, then a restart to 47: the part files at
    10, 20 (first run) and 30, 40, 47 (resumed run, across save steps) and
    the populations of the final checkpoint equal the oracle's straight
    push run of 47 steps (porelb.c:426-437, num_p emulated ranks)."""
    flag = (np.random.default_rng(60 + num_p).random((12, 10, 14)) < 0.62).astype(np.uint8)
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), str(num_p), str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    tau, gx = 0.85, 2e-4
    env = {"PLB_TAU": repr(tau), "PLB_GX": repr(gx), "PLB_PRINT": "0", "PLB_SAVE": "10"}
    r = run([os.path.join(BIN, "porelb.x"), "23"], tmp_path, dict(env, PLB_CHECKPOINT="ck1.bin"))
    assert r.returncode == 0, r.stderr
    r = run([os.path.join(BIN, "porelb.x"), "47"], tmp_path, dict(env, PLB_RESTART="ck1.bin",
                                                                     PLB_CHECKPOINT="ck2.bin"))
    assert r.returncode == 0, r.stderr
    V, start = O.build_V(flag, num_p)
    f0, f1 = O.init_equilibrium(len(V))
    # A save at step t inside the loop holds the state after step t, i.e.
    # after t + 1 steps (porelb.c:436); the final save after the loop holds
    # the state after it_max steps (porelb.c:440).
    it = 0
    for name, done in ((10, 11), (20, 21), (23, 23), (30, 31), (40, 41), (47, 47)):
        mac = O.push_run(V, start, f0, f1, done - it, tau, gx, it0=it)
        it = done
        assert np.array_equal(read_parts(tmp_path, num_p, name), mac), name
        if name == 23:
            ck = read_ckpt(tmp_path / "ck1.bin", len(V), np.float64)
            assert ck["step"] == 23 and ck["precision"] == 0 and ck["elem"] == 8
            assert np.array_equal(ck["state"], f0)
    ck = read_ckpt(tmp_path / "ck2.bin", len(V), np.float64)
    assert ck["step"] == 47
    assert np.array_equal(ck["state"], f0)


@pytest.mark.parametrize("prec", [0, 1, 2])
def test_state_resume_is_bitwise(plb, prec):
    """A lattice stepped straight and one resumed from the stored elements
    of step 30 hold identical states after step 55, in every precision."""
    flag = plb.mask_spheres((24, 20, 28), 2.0, 4.0, 0.45, 8)
    kw = dict(tau=0.7, gx=2e-5, precision=prec, bb_mode=1, index_kind=1, layout=1, ordering=2, brick=(28, 4, 1))
    with plb.Lattice.from_flag(flag, **kw) as a, plb.Lattice.from_flag(flag, **kw) as b:
        a.step(0, 30)
        snap = a.read_state()
        assert snap.dtype == {0: np.float64, 1: np.float32, 2: np.uint16}[prec]
        a.step(30, 25)
        b.write_state(snap)
        assert np.array_equal(b.read_state(), snap)
        b.step(30, 25)
        assert np.array_equal(a.read_state(), b.read_state())
        assert np.array_equal(a.read_macros(), b.read_macros())
        info = a.state_info()
        assert info["elem_bytes"] == {0: 8, 1: 4, 2: 2}[prec] and info["n_nodes"] == a.N
        assert (info["fp16_scale_log2"] > 0) == (prec == 2)
        # the FP64 view of the state is the stored element un-shifted
        f = a.read_populations()
        s = a.read_state()
        if prec == 0:
            assert np.array_equal(f, s)
        elif prec == 1:
            assert np.array_equal(f, O.TP[None, :] + s.astype(np.float64))


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_cli_resume_bitwise_reduced_precision(tmp_path, prec):
    """porelb.x with FP32/FP16 storage: 30 steps + checkpoint + restart to 50
    writes the same part files as 50 straight steps; a restart into another
    precision is refused with -5."""
    flag = (np.random.default_rng(70).random((14, 12, 16)) < 0.62).astype(np.uint8)
    nz, ny, nx = flag.shape
    flag.tofile(tmp_path / "flag.bin")
    assert run([os.path.join(BIN, "preprocess.x"), "2", str(nx), str(ny), str(nz)], tmp_path).returncode == 0
    env = {"PLB_TAU": "0.8", "PLB_GX": "1e-4", "PLB_PRINT": "0", "PLB_SAVE": "1000",
           "PLB_PRECISION": prec, "PLB_BB": "halfway"}
    straight = tmp_path / "straight"
    straight.mkdir()
    os.link(tmp_path / "V.bin", straight / "V.bin")
    assert run([os.path.join(BIN, "porelb.x"), "50"], straight, env).returncode == 0
    r = run([os.path.join(BIN, "porelb.x"), "30"], tmp_path, dict(env, PLB_CHECKPOINT="ck.bin"))
    assert r.returncode == 0, r.stderr
    ck = read_ckpt(tmp_path / "ck.bin", int(np.count_nonzero(flag)), np.float32 if prec == "fp32" else np.uint16)
    assert ck["precision"] == {"fp32": 1, "fp16": 2}[prec] and ck["step"] == 30
    r = run([os.path.join(BIN, "porelb.x"), "50"], tmp_path, dict(env, PLB_RESTART="ck.bin"))
    assert r.returncode == 0, r.stderr
    assert np.array_equal(read_parts(straight, 2, 50), read_parts(tmp_path, 2, 50))
    other = "fp64" if prec == "fp32" else "fp32"
    r = run([os.path.join(BIN, "porelb.x"), "50"], tmp_path, dict(env, PLB_RESTART="ck.bin", PLB_PRECISION=other))
    assert r.returncode == (-5) & 0xFF and "holds precision" in r.stderr, (r.returncode, r.stderr)


@pytest.mark.parametrize("api", ["populations", "state"])
def test_restore_after_density_trip_continues_like_oracle(plb, api):
    """The guard trips at step 9 (fixture mask, g_x 0.25); writing back the
    oracle's state after 5 steps clears it, and steps 5..7 then equal the
    oracle's (porelb.c:338-342 stops the reference at the trip)."""
    d = dict(np.load(os.path.join(GOLDEN, "fault.npz")))
    flag, tau, gx = d["flag"], float(d["tau"]), 0.25
    V, start = O.build_V(flag, 1)
    f0, f1 = O.init_equilibrium(len(V))
    O.push_run(V, start, f0, f1, 5, tau, gx)
    snap = f0.copy()
    mac = O.push_run(V, start, f0, f1, 3, tau, gx, it0=5)
    with plb.Lattice.from_V(V, start, flag.shape, tau=tau, gx=gx) as lat:
        assert lat.step_rc(0, 200) == plb.ERR_DENSITY
        assert lat.last_fault()["step"] == 9
        if api == "populations":
            lat.write_populations(snap)
        else:
            lat.write_state(snap)
        assert lat.last_fault()["tripped"] == 0
        assert lat.step_rc(5, 3) == 0
        assert np.array_equal(lat.read_populations(), f0)
        assert np.array_equal(lat.read_macros(), mac)


@pytest.mark.parametrize("ordering", [1, 2])
def test_duplicate_voxel_records_refused(plb, ordering):
    """Two V records on one voxel would share an internal id in brick or
    pencil order: plb_create refuses them (-3); the V order accepts them
    as the reference does."""
    flag = (np.random.default_rng(80).random((6, 7, 8)) < 0.7).astype(np.uint8)
    V, start = O.build_V(flag, 1)
    for k in ("x", "y", "z"):
        V[5][k] = V[6][k]
    with pytest.raises(plb.PlbError) as ei:
        plb.Lattice.from_V(V, start, flag.shape, ordering=ordering, brick=(8, 2, 2))
    assert ei.value.code == -3
    with plb.Lattice.from_V(V, start, flag.shape) as lat:
        lat.step(0, 2)


@pytest.mark.parametrize("ordering,brick", [(0, (16, 4, 4)), (1, (8, 4, 2)), (2, (28, 4, 1))])
def test_fp16_stochastic_rounding_independent_of_node_order(plb, ordering, brick):
    """The stochastic-rounding stream is keyed by the node's V-order id, so FP16
    storage gives the same state in every node order and layout, and a batch
    sample the same as its own lattice."""
    flag = plb.mask_spheres((24, 20, 28), 2.0, 4.0, 0.45, 8)
    kw = dict(tau=0.7, gx=2e-5, precision=2, bb_mode=1, fp16_rounding=1, fp16_scale_log2=20)
    with plb.Lattice.from_flag(flag, **kw) as ref:
        ref.step(0, 40)
        want = ref.read_state()
    with plb.Lattice.from_flag(flag, index_kind=1, layout=1, ordering=ordering, brick=brick, **kw) as lat:
        lat.step(0, 40)
        assert np.array_equal(lat.read_state(), want)
    other = plb.mask_noise((20, 18, 22), 2.0, 0.4, 3)
    b = plb.Batch([other, flag], [0.9, 0.7], [1e-5, 2e-5], precision=2, bb_mode=1, fp16_rounding=1,
                  fp16_scale_log2=20, index_kind=1, layout=1, ordering=ordering, brick=brick)
    try:
        b.run(40, 40, 0.0)
        with plb.Lattice.from_flag(flag, **kw) as ref:
            ref.step(0, 40)
            assert np.array_equal(b.read_macros(1), ref.read_macros())
    finally:
        b.close()
