"""Full-size properties at the benchmark's 1024^3 sphere pack (214.7 M pore
nodes), where the CPU oracle cannot follow: results that must not depend on
the index encoding, layout or node order, and mass conservation of half-way
bounce-back in FP32 storage."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DIMS, RMIN, RMAX, PHI, SEED = (1024, 1024, 1024), 6.0, 12.0, 0.20, 4
PENCIL = dict(index_kind=1, layout=1, ordering=2, brick=(1024, 128, 1))


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def flag(plb):
    say("mask 1024^3")
    return plb.mask_spheres(DIMS, RMIN, RMAX, PHI, SEED)


def test_fp64_reference_order_independent_of_index_and_order(plb, flag):
    """FP64 with the reference's bounce order, 20 steps: the int32 table in
    V order and TILE8 in pencils give bit-identical rho, u at every node (read
    back in V order).  The FP64 sums agree to rounding only: the reduction is
    deterministic but adds nodes in storage order, which pencils permute."""
    macs, sums = [], []
    for kw in (dict(), PENCIL):
        say("fp64 refcompat", kw)
        with plb.Lattice.from_flag(flag, 1, tau=0.6, gx=5e-7, **kw) as lat:
            assert lat.N == 214748230
            lat.step(0, 20)
            r = lat.reduce()
            sums.append((r["mass"], r["sum_ux"]))
            macs.append(lat.read_macros())
    assert np.array_equal(macs[0], macs[1])
    assert abs(sums[0][0] / sums[1][0] - 1) < 1e-14
    assert abs(sums[0][1] / sums[1][1] - 1) < 1e-12


def test_fp32_halfway_conserves_mass_and_tracks_fp64(plb, flag):
    """Half-way bounce-back conserves mass; FP32 shifted storage keeps it to
    FP32 rounding over 300 steps and stays within 1e-5 of FP64 in sum u_x."""
    res = {}
    for prec in (0, 1):
        say("halfway", prec)
        with plb.Lattice.from_flag(flag, 1, tau=0.6, gx=5e-7, precision=prec, bb_mode=plb.BB_HALFWAY,
                                   **PENCIL) as lat:
            m0 = lat.reduce()["mass"]
            lat.step(0, 300)
            r = lat.reduce()
            res[prec] = (m0, r["mass"], r["sum_ux"])
    m0, m1, _ = res[0]
    assert abs(m1 / m0 - 1) < 1e-12
    m0, m1, _ = res[1]
    assert abs(m1 / m0 - 1) < 1e-7
    assert abs(res[1][2] / res[0][2] - 1) < 1e-5
