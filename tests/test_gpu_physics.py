"""Physics known answers through the HIP path.

- Slit Poiseuille flow with half-way bounce-back: exact (to convergence) at
  tau = 1/2 + sqrt(3/16) for BGK with half-way walls (He, Zou, Luo & Dembo,
  J. Stat. Phys. 87, 1997), for every index encoding, layout and order, and
  within FP32 rounding for the shifted FP32 storage.
- Permeability of a body-centred cubic sphere array against the
  Sangani-Acrivos series that getK.c tabulates (getK.c:5-72): a few percent at
  32 voxels per cell."""
import numpy as np
import pytest

import oracle as O
from test_oracle import poiseuille_flag, poiseuille_profile

pytestmark = pytest.mark.gpu
MAGIC_TAU = 0.5 + np.sqrt(3.0 / 16.0)


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("kw,tol", [
    (dict(), 1e-9),
    (dict(index_kind=1, layout=1, ordering=2, brick=(4, 8, 1)), 1e-9),
    (dict(index_kind=2, layout=1, ordering=2, brick=(4, 8, 1)), 1e-9),
    (dict(index_kind=1, layout=1, graph_steps=64), 1e-9),
    (dict(precision=1, index_kind=1, layout=1, ordering=2, brick=(4, 8, 1)), 1e-4),
])
def test_poiseuille_exact_at_magic_tau(plb, kw, tol):
    ny, g = 20, 1e-6
    flag = poiseuille_flag(ny)
    say("poiseuille", kw)
    with plb.Lattice.from_flag(flag, 1, tau=MAGIC_TAU, gx=g, bb_mode=plb.BB_HALFWAY, **kw) as lat:
        lat.step(0, 30000)
        mac = lat.read_macros()  # V order
    V, _ = O.build_V(flag, 1)
    ua = poiseuille_profile(V["y"].astype(float), ny, MAGIC_TAU, g)
    err = np.linalg.norm(mac[1] - ua) / np.linalg.norm(ua)
    say("poiseuille rel L2 error", err)
    assert err < tol


@pytest.mark.parametrize("chi", [0.5, 0.8])
def test_bcc_permeability_vs_sangani_acrivos(plb, chi):
    """K = nu q / g_x of a BCC array (two spheres per cubic cell of 32 voxels,
    radius sqrt(3) L chi / 4) against K1 of getK.c; FP32 storage agrees with
    FP64 far inside that margin."""
    n, tau, g = 32, 1.0, 1e-6
    flag = plb.mask_bcc(n, chi)
    k1 = O.getK_bcc(chi, n + 1)["K1"]  # getK.c takes the resolution nx and uses L = nx - 1
    Ks = {}
    for prec in (0, 1):
        say("bcc", chi, prec)
        with plb.Lattice.from_flag(flag, 1, tau=tau, gx=g, precision=prec, bb_mode=plb.BB_HALFWAY,
                                   index_kind=1, layout=1, ordering=2, brick=(n, 8, 1)) as lat:
            lat.step(0, 8000)
            Ks[prec] = lat.reduce()["K"]
    say("bcc K", Ks, "K1", k1)
    assert abs(Ks[0] / k1 - 1.0) < 0.05
    assert abs(Ks[1] / Ks[0] - 1.0) < 1e-4


@pytest.mark.parametrize("prec,tol_mass,tol_mom", [(0, 1e-13, 1e-12), (1, 1e-7, 1e-6)])
@pytest.mark.parametrize("tau", [0.55, 0.8, 1.7])
def test_collision_conserves_mass_and_momentum(plb, prec, tol_mass, tol_mom, tau):
    """All-fluid periodic box, g_x = 0, random positive populations: streaming
    moves populations and the BGK collision relaxes to feq, whose zeroth and
    first moments are rho and rho*u (sum_j w_j = 1, sum_j w_j e_j = 0,
    sum_j w_j e_ja e_jb = delta_ab / 3, third moments zero).  Total mass and
    all three momentum components stay at their initial values to rounding
    (FP64) or FP32 rounding of the deviations (FP32 storage; observed 3e-9
    relative in mass and 3e-8 absolute in momentum over 40 steps)."""
    flag = np.ones((6, 7, 8), dtype=np.uint8)
    ex, ey, ez = (np.array(O.E)[:, k] for k in range(3))
    rng = np.random.default_rng(int(tau * 100))
    with plb.Lattice.from_flag(flag, 1, tau=tau, gx=0.0, precision=prec, bb_mode=plb.BB_HALFWAY,
                               index_kind=1, layout=1) as lat:
        f = np.array(O.TP)[None, :] * (1.0 + 0.05 * rng.standard_normal((lat.N, 19)))
        lat.write_populations(f)
        f = lat.read_populations()  # as stored (FP32 storage rounds the deviations)
        m0 = np.array([f.sum(), (f * ex).sum(), (f * ey).sum(), (f * ez).sum()])
        lat.step(1, 40)  # it > 0: no half-force shift (and g_x = 0 anyway)
        g = lat.read_populations()
        m1 = np.array([g.sum(), (g * ex).sum(), (g * ey).sum(), (g * ez).sum()])
    say("moments before", m0, "after", m1)
    assert abs(m1[0] / m0[0] - 1) < tol_mass
    assert np.all(np.abs(m1[1:] - m0[1:]) < tol_mom)


def test_run_to_convergence_rule_and_state(plb):
    """plb_run_to_convergence stops at the first check where <u_x> moved by at
    most tol * max(|<u_x>|, 1e-3 g_x) since the previous check, and its state
    equals plain stepping to the same step bit for bit; tol = 0 runs all steps."""
    flag = plb.mask_spheres((24, 20, 28), 3.0, 5.0, 0.4, 11)
    kw = dict(tau=0.9, gx=1e-5, bb_mode=plb.BB_HALFWAY, index_kind=1, layout=1)
    tol, check = 1e-5, 50
    with plb.Lattice.from_flag(flag, **kw) as lat:
        n, r = lat.run_to_convergence(0, 20000, check, tol)
        mac = lat.read_macros()
    say("converged after", n, "steps, <u_x>", r["mean_ux"])
    assert 0 < n < 20000 and n % check == 0
    means = []
    with plb.Lattice.from_flag(flag, **kw) as lat:
        for s in range(0, n, check):
            lat.step(s, check)
            means.append(lat.reduce()["mean_ux"])
        assert np.array_equal(lat.read_macros(), mac)
        assert lat.reduce() == r
    rule = [abs(b - a) <= tol * max(abs(b), 1e-3 * kw["gx"]) for a, b in zip(means, means[1:])]
    assert rule[-1] and not any(rule[:-1])
    with plb.Lattice.from_flag(flag, **kw) as lat:
        n0, _ = lat.run_to_convergence(0, 130, check, 0.0)
        assert n0 == 130
