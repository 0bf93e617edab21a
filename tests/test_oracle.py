"""The CPU oracle pinned against the reference's own outputs (golden fixtures
written by the reference programs built from their sources), against the
reference binaries themselves where they are present, and against closed
forms and brute force."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, TINY_CASES, load_tiny


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def vfields(O, V):
    x, y, z, t, pid, nid = O.semantic_fields(V)
    return dict(x=x, y=y, z=z, type=t, pid=pid, id=nid)


@pytest.mark.parametrize("name", TINY_CASES)
def test_build_V_matches_reference_fixture(oracle_mod, name):
    O = oracle_mod
    d = load_tiny(name)
    for np_ in d["ranks"]:
        V, start = O.build_V(d["flag"], int(np_))
        got = vfields(O, V)
        for k in ["x", "y", "z", "type", "pid", "id"]:
            assert np.array_equal(got[k], d[f"np{np_}_V_{k}"].astype(np.int64)), (name, np_, k)
        assert np.array_equal(start, d[f"np{np_}_start"])


@pytest.mark.parametrize("name", TINY_CASES)
def test_push_and_pull_match_reference_fixture_bitwise(oracle_mod, name):
    O = oracle_mod
    d = load_tiny(name)
    tau, gx, steps = float(d["tau"]), float(d["gx"]), int(d["steps"])
    for np_ in d["ranks"]:
        V, start = O.build_V(d["flag"], int(np_))
        f0, f1 = O.init_equilibrium(len(V))
        mac = O.push_run(V, start, f0, f1, steps, tau, gx)
        assert np.array_equal(f0, d[f"np{np_}_f0"]), (name, np_)
        assert np.array_equal(mac, d[f"np{np_}_macros"]), (name, np_)
        tab = O.pull_table(V, start, O.BB_REFCOMPAT)
        g0, _ = O.init_equilibrium(len(V))
        mac2 = O.pull_run(tab, g0, steps, tau, gx)
        assert np.array_equal(g0, f0) and np.array_equal(mac2, mac), (name, np_)


def test_config1_np4_matches_reference(oracle_mod, plb):
    """64^3 noise mask, 4 ranks, 200 steps: SHA-256 of f0 and macros."""
    O = oracle_mod
    meta = json.load(open(os.path.join(GOLDEN, "config1.json")))
    m = meta["mask"]
    flag = plb.mask_noise(tuple(m["dims"]), m["sigma"], m["porosity"], m["seed"])
    assert sha(flag) == m["sha256"]
    V, start = O.build_V(flag, 4)
    tab = O.pull_table(V, start, O.BB_REFCOMPAT)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, meta["np4"]["steps"], 1.0, 1e-6)
    assert sha(f0) == meta["np4"]["f0_sha256"]
    assert sha(mac) == meta["np4"]["macros_sha256"]


@pytest.mark.slow
def test_config1_np1_1000_steps_matches_reference(oracle_mod, plb):
    """Config 1 exactly: 64^3, tau 1.0, g 1e-6, 1000 steps, 1 rank."""
    O = oracle_mod
    meta = json.load(open(os.path.join(GOLDEN, "config1.json")))
    m = meta["mask"]
    flag = plb.mask_noise(tuple(m["dims"]), m["sigma"], m["porosity"], m["seed"])
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start, O.BB_REFCOMPAT)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, meta["np1"]["steps"], 1.0, 1e-6)
    assert sha(f0) == meta["np1"]["f0_sha256"]
    assert sha(mac) == meta["np1"]["macros_sha256"]


def random_mask(dims, p, seed):
    return (np.random.default_rng(seed).random(dims) < p).astype(np.uint8)


@pytest.mark.parametrize("seed", [11, 12, 13])
@pytest.mark.parametrize("np_", [1, 2, 3, 4])
def test_pull_table_equals_push_refcompat(oracle_mod, seed, np_):
    """The pull-table rule reproduces the sequential push + bounce + receive
    order of porelb.c:285-312 bitwise, for any rank count."""
    O = oracle_mod
    flag = random_mask((12, 7, 9), 0.55, seed)
    V, start = O.build_V(flag, np_)
    f0, f1 = O.init_equilibrium(len(V))
    m1 = O.push_run(V, start, f0, f1, 25, 0.7, 2e-3)
    tab = O.pull_table(V, start, O.BB_REFCOMPAT)
    g0, _ = O.init_equilibrium(len(V))
    m2 = O.pull_run(tab, g0, 25, 0.7, 2e-3)
    assert np.array_equal(f0, g0) and np.array_equal(m1, m2)


@pytest.mark.parametrize("np_", [1, 3])
def test_pull_table_equals_push_halfway(oracle_mod, np_):
    O = oracle_mod
    flag = random_mask((12, 10, 8), 0.6, 5)
    V, start = O.build_V(flag, np_)
    f0, f1 = O.init_equilibrium(len(V))
    m1 = O.push_run(V, start, f0, f1, 30, 0.8, 1e-3, mode=O.BB_HALFWAY)
    tab = O.pull_table(V, start, O.BB_HALFWAY)
    g0, _ = O.init_equilibrium(len(V))
    m2 = O.pull_run(tab, g0, 30, 0.8, 1e-3)
    assert np.array_equal(f0, g0) and np.array_equal(m1, m2)


def test_halfway_conserves_mass_refcompat_does_not(oracle_mod):
    O = oracle_mod
    flag = random_mask((10, 9, 8), 0.6, 21)
    V, start = O.build_V(flag, 1)
    N = len(V)
    res = {}
    for mode in (O.BB_HALFWAY, O.BB_REFCOMPAT):
        tab = O.pull_table(V, start, mode)
        f0, _ = O.init_equilibrium(N)
        O.pull_run(tab, f0, 100, 0.8, 1e-3)
        res[mode] = abs(f0.sum() - N) / N
    assert res[O.BB_HALFWAY] < 1e-13
    assert res[O.BB_REFCOMPAT] > 1e-9


@pytest.mark.parametrize("tau", [0.6, 0.8, 1.0, 1.7])
def test_all_fluid_momentum_closed_form(oracle_mod, tau):
    """Uniform periodic fluid: collision with the reference's forcing term adds
    (1-omega/2)F at it=0 and F at every later step (sum_j e_jx S_j feq_j =
    (1-omega/2)F rho, and the BGK part relaxes toward rho(u) = m + F/2), so
    m_n = (1-omega/2)F + (n-1)F and Ux(it) = m_it + F/2."""
    O = oracle_mod
    flag = np.ones((4, 3, 5), dtype=np.uint8)
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start, O.BB_REFCOMPAT)
    F, om = 1e-5, 1.0 / tau
    for n in (1, 2, 7, 40):
        f0, _ = O.init_equilibrium(len(V))
        mac = O.pull_run(tab, f0, n, tau, F)
        m = f0 @ O.E[:, 0].astype(float)
        expect = (1 - om / 2) * F + (n - 1) * F
        assert np.allclose(m, expect, rtol=1e-10, atol=0)
        assert np.allclose(f0.sum(axis=1), 1.0, rtol=0, atol=1e-14)
        if n >= 2:
            assert np.allclose(mac[1], (1 - om / 2) * F + (n - 2) * F + F / 2, rtol=1e-10, atol=0)


def test_build_V_invariants_brute_force(oracle_mod):
    """coords(nb) = (x+e_j) mod n, types from the mask, rank rule, local ids,
    V_sub_start = prefix of rank sizes, reverse links."""
    O = oracle_mod
    for seed, dims, np_ in [(1, (8, 5, 6), 2), (2, (9, 4, 7), 3), (3, (4, 6, 5), 4), (4, (3, 3, 3), 1)]:
        flag = random_mask(dims, 0.6, seed)
        nz, ny, nx = dims
        V, start = O.build_V(flag, np_)
        slab = nz // np_
        pores = [(z, y, x) for z in range(nz) for y in range(ny) for x in range(nx) if flag[z, y, x]]
        by_rank = {p: [c for c in pores if c[0] // slab == p] for p in range(np_)}
        assert list(start) == list(np.cumsum([0] + [len(by_rank[p]) for p in range(np_)])[:-1])
        gid = {}
        for p in range(np_):
            for k, c in enumerate(by_rank[p]):
                gid[c] = (p, k, start[p] + k)
        for c, (p, k, g) in gid.items():
            rec = V[g]
            assert (rec["z"], rec["y"], rec["x"]) == c
            for j in range(1, 19):
                nb = rec["nb"][j - 1]
                cz, cy, cx = ((c[0] + O.E[j, 2]) % nz, (c[1] + O.E[j, 1]) % ny, (c[2] + O.E[j, 0]) % nx)
                if not flag[cz, cy, cx]:
                    assert nb["node_type"] == O.SOLID
                    continue
                q, kq, gq = gid[(cz, cy, cx)]
                assert nb["node_id"] == kq and nb["node_pid"] == q
                assert nb["node_type"] == (O.FLUID if q == p else O.EXTERN)
                back = V[gq]["nb"][O.RE[j] - 1]
                assert back["node_id"] == k and back["node_pid"] == p


def test_build_V_rejects_bad_input(oracle_mod):
    O = oracle_mod
    flag = random_mask((7, 4, 4), 0.5, 1)
    with pytest.raises(ValueError):
        O.build_V(flag, 2)  # 7 % 2: pid = z/3 reaches 2 (preprocess.c:107)
    bad = flag.copy()
    bad[0, 0, 0] = 2
    with pytest.raises(ValueError):
        O.build_V(bad, 1)


def test_getK_bcc_physics(oracle_mod, plb):
    O = oracle_mod
    # Dilute limit: with d* -> 1, K1 -> 1/(6 pi a n), the Stokes drag of
    # isolated spheres at number density n = 2/L^3 (two spheres per BCC cell).
    L = 100.0
    k = O.getK_bcc(1e-4, int(L) + 1)
    stokes = 1.0 / (6 * np.pi * k["a"] * (2 / L**3))
    assert abs(k["K1"] / stokes - 1) < 1e-3
    # Touching spheres at chi = 1: corner-to-centre distance sqrt(3)L/2 = 2a.
    k1 = O.getK_bcc(1.0, 65)
    assert abs(2 * k1["a"] - np.sqrt(3) * 64 / 2) < 1e-12
    # Porosity formula vs the voxelised BCC mask (within voxelisation error).
    for chi in (0.3, 0.6, 0.8):
        n = 64
        m = plb.mask_bcc(n, chi)
        kk = O.getK_bcc(chi, n + 1)
        assert abs(m.mean() - kk["eps"]) < 0.01
    # K1 decreases monotonically with sphere size.
    Ks = [O.getK_bcc(c, 65)["K1"] for c in np.linspace(0.1, 0.9, 9)]
    assert all(a > b for a, b in zip(Ks, Ks[1:]))


def test_reduce_definition(oracle_mod):
    O = oracle_mod
    flag = random_mask((6, 6, 6), 0.6, 9)
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 10, 0.8, 1e-4)
    r = O.reduce(f0, mac[1], flag.shape, 0.8, 1e-4)
    assert abs(r["mass"] - f0.sum()) < 1e-12 * len(V)
    assert r["sum_ux"] == pytest.approx(mac[1].sum(), rel=1e-14)
    assert r["q"] == pytest.approx(mac[1].sum() / 216, rel=1e-14)
    assert r["K"] == pytest.approx((0.8 - 0.5) / 3 * r["q"] / 1e-4, rel=1e-14)


def poiseuille_flag(ny, nx=4, nz=3):
    """Slit between two solid planes y = 0 and y = ny - 1, periodic in x and z."""
    flag = np.ones((nz, ny, nx), dtype=np.uint8)
    flag[:, 0, :] = 0
    flag[:, ny - 1, :] = 0
    return flag


def poiseuille_profile(y, ny, tau, g):
    """Force-driven slit flow with the half-way walls at y = 1/2 and y = ny - 3/2."""
    nu = (tau - 0.5) / 3.0
    return g / (2.0 * nu) * (y - 0.5) * (ny - 1.5 - y)


@pytest.mark.parametrize("tau,tol", [(0.5 + np.sqrt(3.0 / 16.0), 1e-9), (1.0, 3e-3)])
def test_poiseuille_halfway(oracle_mod, tau, tol):
    """Slit Poiseuille flow with half-way bounce-back, a physics known answer:
    at tau = 1/2 + sqrt(3/16) BGK with half-way walls reproduces the
    parabola exactly (He, Zou, Luo & Dembo, J. Stat. Phys. 87, 1997); other
    tau leave the known tau-dependent wall slip, here 1.4e-3."""
    O = oracle_mod
    ny, g = 20, 1e-6
    flag = poiseuille_flag(ny)
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start, O.BB_HALFWAY)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 30000, tau, g)
    ua = poiseuille_profile(V["y"].astype(float), ny, tau, g)
    assert np.linalg.norm(mac[1] - ua) / np.linalg.norm(ua) < tol
    assert np.abs(mac[2:]).max() < 1e-8 * ua.max()  # no cross flow beyond rounding
