"""Edge cases of the domain: no pores, one pore, one-voxel-thick and
two-voxel domains (neighbours wrap onto the node itself), an all-fluid box,
and the flag-value check of preprocess.c."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def oracle_run(flag, np_, steps, tau, gx, mode=O.BB_REFCOMPAT):
    V, start = O.build_V(flag, np_)
    tab = O.pull_table(V, start, mode)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, steps, tau, gx)
    return V, start, f0, mac


def test_no_pores(plb):
    flag = np.zeros((4, 5, 6), dtype=np.uint8)
    for kw in (dict(index_kind=0), dict(index_kind=1), dict(index_kind=2, bb_mode=1),
               dict(index_kind=1, graph_steps=2)):
        with plb.Lattice.from_flag(flag, **kw) as lat:
            assert lat.N == 0
            lat.step(0, 5)
            assert lat.read_populations().shape == (0, 19)
            r = lat.reduce()
            assert r["mass"] == 0.0 and r["n_nodes"] == 0
    V, start = plb.preprocess(flag, 2)
    assert len(V) == 0 and list(start) == [0, 0]


@pytest.mark.parametrize("dims", [(1, 1, 1), (1, 1, 5), (2, 2, 2), (3, 1, 4), (1, 7, 1)])
@pytest.mark.parametrize("mode", [O.BB_REFCOMPAT, O.BB_HALFWAY])
def test_degenerate_dims_bitwise(plb, dims, mode):
    rng = np.random.default_rng(sum(dims))
    flag = (rng.random(dims) < 0.7).astype(np.uint8)
    flag.flat[0] = 1
    V, start, f0, mac = oracle_run(flag, 1, 20, 0.8, 1e-4, mode)
    variants = [dict(), dict(index_kind=1, layout=1), dict(ordering=2, brick=(dims[2], 1, 1))]
    if mode == O.BB_HALFWAY:
        variants += [dict(index_kind=2), dict(index_kind=2, layout=1, ordering=2, brick=(dims[2], 1, 1))]
    for kw in variants:
        with plb.Lattice.from_flag(flag, bb_mode=mode, tau=0.8, gx=1e-4, **kw) as lat:
            assert np.array_equal(lat.read_table(), O.pull_table(V, start, mode))
            lat.step(0, 20)
            assert np.array_equal(lat.read_populations(), f0), kw
            assert np.array_equal(lat.read_macros(), mac), kw


@pytest.mark.parametrize("np_", [8, 16])
def test_one_plane_per_rank_reference_order(plb, np_):
    """z-slabs one plane thick (nz = num_p): every z-link crosses ranks, so the
    reference's receive-after-stream order (porelb.c:285-312) decides every
    slot with a z component.  FP64 table, populations and macros equal the
    oracle's emulated ranks bit for bit, both from the mask and from V.bin
    records."""
    rng = np.random.default_rng(np_)
    flag = (rng.random((np_, 9, 11)) < 0.6).astype(np.uint8)
    V, start, f0, mac = oracle_run(flag, np_, 25, 0.7, 2e-4)
    assert len(start) == np_ and all(start[k] <= start[k + 1] for k in range(np_ - 1))
    for kw in (dict(), dict(index_kind=1, layout=1, ordering=2, brick=(11, 4, 1), waves_per_simd=5)):
        with plb.Lattice.from_flag(flag, np_, tau=0.7, gx=2e-4, **kw) as lat:
            assert np.array_equal(lat.read_table(), O.pull_table(V, start, O.BB_REFCOMPAT)), kw
            lat.step(0, 25)
            assert np.array_equal(lat.read_populations(), f0), kw
            assert np.array_equal(lat.read_macros(), mac), kw
    with plb.Lattice.from_V(V, start, flag.shape, tau=0.7, gx=2e-4) as lat:
        lat.step(0, 25)
        assert np.array_equal(lat.read_macros(), mac)


def test_single_pore_surrounded_by_solid(plb):
    flag = np.zeros((5, 5, 5), dtype=np.uint8)
    flag[2, 2, 2] = 1
    V, start, f0, mac = oracle_run(flag, 1, 10, 0.8, 1e-3, O.BB_HALFWAY)
    with plb.Lattice.from_flag(flag, bb_mode=plb.BB_HALFWAY, tau=0.8, gx=1e-3) as lat:
        assert (lat.read_table() == -1).all()      # every direction bounces
        lat.step(0, 10)
        assert np.array_equal(lat.read_populations(), f0)
        assert lat.reduce()["mass"] == pytest.approx(1.0, abs=1e-15)


def test_all_fluid_box_momentum_closed_form(plb):
    """Uniform periodic fluid: m_n = (1 - omega/2) F + (n - 1) F for every
    precision (FP32/FP16 to their storage precision)."""
    flag = np.ones((8, 6, 10), dtype=np.uint8)
    tau, F, n = 0.7, 1e-5, 25
    om = 1 / tau
    expect = (1 - om / 2) * F + (n - 1) * F
    for prec, rtol in ((0, 1e-10), (1, 1e-5), (2, 3e-3)):
        with plb.Lattice.from_flag(flag, tau=tau, gx=F, precision=prec, fp16_scale_log2=24) as lat:
            lat.step(0, n)
            f = lat.read_populations()
        m = f @ O.E[:, 0].astype(float)
        assert np.allclose(m, expect, rtol=rtol, atol=0), prec


def test_flag_values_checked(plb):
    flag = np.ones((4, 4, 4), dtype=np.uint8)
    flag[1, 2, 3] = 7
    with pytest.raises(plb.PlbError) as ei:
        plb.Lattice.from_flag(flag)
    assert ei.value.code == -25
    with pytest.raises(plb.PlbError) as ei:
        plb.preprocess(flag, 1)
    assert ei.value.code == -25


def test_bad_V_records_rejected(plb):
    """check_V (porelb.c:115-141): coordinates out of range -> -3, an EXTERN
    neighbour on the node's own rank -> -4."""
    flag = (np.random.default_rng(1).random((6, 5, 4)) < 0.6).astype(np.uint8)
    V, start = O.build_V(flag, 2)
    bad = V.copy()
    bad["x"][3] = 4
    with pytest.raises(plb.PlbError) as ei:
        plb.Lattice.from_V(bad, start, flag.shape)
    assert ei.value.code == -3
    bad = V.copy()
    k = np.argwhere(bad["nb"]["node_type"] == O.EXTERN)[0]
    bad["nb"]["node_pid"][k[0], k[1]] = 0 if start[1] <= k[0] else 1  # point at the own rank
    own = 0 if k[0] < start[1] else 1
    bad["nb"]["node_pid"][k[0], k[1]] = own
    with pytest.raises(plb.PlbError) as ei:
        plb.Lattice.from_V(bad, start, flag.shape)
    assert ei.value.code == -4


def test_read_into_page_locked_buffer(plb):
    """plb_host_alloc buffers receive the same macros as pageable arrays."""
    flag = (np.random.default_rng(8).random((10, 9, 11)) < 0.6).astype(np.uint8)
    with plb.Lattice.from_flag(flag, tau=0.8, gx=1e-4) as lat:
        lat.step(0, 20)
        a = lat.read_macros()
        b = lat.read_macros(out=plb.host_array((4, lat.N)))
        assert np.array_equal(a, b)
