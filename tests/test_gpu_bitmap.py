"""The BITMAP index: no stored pull table; the step kernel derives each
source from the pore bitmap of the node order and the node's voxel position.
Its sources must be exactly the half-way pull table of the oracle (bit for
bit, in V order), and a step with it must be bit-identical to a step with
the stored tables."""
import hashlib

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def random_mask(dims, p, seed):
    return (np.random.default_rng(seed).random(dims) < p).astype(np.uint8)


def say(*a):
    print("[case]", *a, flush=True)


# (dims z,y,x; num_p; ordering; brick) -- x not a multiple of 32, one- and
# two-voxel extents (a node is its own periodic neighbour), pencils that do
# not divide ny, pencils of 1 and 2 rows (no interior rows at all).
TABLE_CASES = [
    ((16, 11, 13), 1, 0, (1, 1, 1)),
    ((16, 11, 13), 2, 0, (1, 1, 1)),
    ((15, 11, 13), 3, 0, (1, 1, 1)),
    ((20, 11, 13), 4, 0, (1, 1, 1)),
    ((16, 11, 13), 1, 2, (13, 4, 1)),
    ((16, 11, 13), 2, 2, (100, 3, 1)),
    ((9, 10, 33), 3, 2, (33, 1, 1)),
    ((9, 10, 33), 1, 2, (40, 2, 1)),
    ((1, 7, 9), 1, 0, (1, 1, 1)),
    ((5, 1, 9), 1, 2, (9, 4, 1)),
    ((6, 7, 1), 2, 2, (1, 3, 1)),
    ((2, 2, 2), 1, 0, (1, 1, 1)),
    ((3, 64, 70), 1, 2, (70, 64, 1)),
]


@pytest.mark.parametrize("dims,np_,order,brick", TABLE_CASES)
def test_bitmap_sources_equal_halfway_table(plb, dims, np_, order, brick):
    flag = random_mask(dims, 0.6, sum(dims) + np_)
    say("bitmap table", dims, np_, order, brick)
    V, start = O.build_V(flag, np_)
    want = O.pull_table(V, start, O.BB_HALFWAY)
    for layout in (plb.LAYOUT_SOA, plb.LAYOUT_AOSOA):
        kw = dict(bb_mode=plb.BB_HALFWAY, layout=layout, ordering=order, brick=brick,
                  index_kind=plb.INDEX_BITMAP)
        with plb.Lattice.from_flag(flag, np_, **kw) as lat:
            assert np.array_equal(lat.read_table(), want), layout
        with plb.Lattice.from_V(V, start, flag.shape, **kw) as lat:
            assert np.array_equal(lat.read_table(), want), layout


def test_bitmap_sources_large_pencils(plb):
    """Mostly interior waves (the per-wave offset path) plus every pencil,
    x and periodic edge of a 96x80x72 sphere pack."""
    flag = plb.mask_spheres((72, 80, 96), 3.0, 7.0, 0.3, 5)
    V, start = O.build_V(flag, 2)
    want = O.pull_table(V, start, O.BB_HALFWAY)
    for order, brick in [(0, (1, 1, 1)), (2, (96, 16, 1)), (2, (96, 24, 1)), (2, (200, 80, 1))]:
        say("bitmap large", order, brick)
        with plb.Lattice.from_flag(flag, 2, bb_mode=plb.BB_HALFWAY, layout=plb.LAYOUT_AOSOA,
                                   ordering=order, brick=brick, index_kind=plb.INDEX_BITMAP) as lat:
            assert np.array_equal(lat.read_table(), want), (order, brick)
            b = lat.bytes_per_update()
            assert 2 * 19 * 8 + 4 < b < 2 * 19 * 8 + 4 + 2.0


@pytest.mark.parametrize("kw", [dict(), dict(layout=1), dict(layout=1, ordering=2, brick=(30, 4, 1)),
                                dict(layout=1, ordering=2, brick=(30, 2, 1), block_size=64),
                                dict(layout=1, ordering=2, brick=(30, 5, 1), nodes_per_thread=2,
                                     block_size=128)])
def test_bitmap_fp64_step_bitwise_vs_oracle(plb, kw):
    flag = plb.mask_spheres((28, 26, 30), 2.0, 5.0, 0.4, 8)
    V, start = O.build_V(flag, 2)
    tab = O.pull_table(V, start, O.BB_HALFWAY)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 40, 0.75, 2e-4)
    say("bitmap fp64", kw)
    with plb.Lattice.from_flag(flag, 2, tau=0.75, gx=2e-4, bb_mode=plb.BB_HALFWAY,
                               index_kind=plb.INDEX_BITMAP, **kw) as lat:
        lat.step(0, 15)
        lat.step(15, 25)
        assert np.array_equal(lat.read_populations(), f0)
        assert np.array_equal(lat.read_macros(), mac)


@pytest.mark.parametrize("prec", [1, 2])
def test_bitmap_shifted_storage_equals_tile8(plb, prec):
    """Same sources, same arithmetic: FP32/FP16 states are bit-identical to
    the TILE8 path after many steps."""
    flag = plb.mask_spheres((48, 40, 64), 3.0, 6.0, 0.3, 3)
    shape = dict(tau=0.6, gx=5e-6, precision=prec, bb_mode=plb.BB_HALFWAY, layout=plb.LAYOUT_AOSOA,
                 ordering=plb.ORDER_PENCIL, brick=(64, 8, 1), fp16_scale_log2=16)
    out = []
    for index in (plb.INDEX_TILE8, plb.INDEX_BITMAP):
        say("bitmap shifted", prec, index)
        with plb.Lattice.from_flag(flag, 1, index_kind=index, **shape) as lat:
            lat.step(0, 200)
            out.append((sha(lat.read_populations()), sha(lat.read_macros()), lat.reduce()["K"]))
    assert out[0] == out[1]


def test_bitmap_at_scale_equals_tile8(plb):
    """256^3 sphere pack (about 5 M pore nodes), FP32, bench-style pencils:
    populations after 30 steps hash-identical to the TILE8 path."""
    flag = plb.mask_spheres((256, 256, 256), 6.0, 12.0, 0.30, 2)
    shape = dict(tau=0.8, gx=1e-6, precision=plb.FP32, bb_mode=plb.BB_HALFWAY, layout=plb.LAYOUT_AOSOA,
                 ordering=plb.ORDER_PENCIL, brick=(256, 64, 1))
    out = []
    for index in (plb.INDEX_TILE8, plb.INDEX_BITMAP):
        say("bitmap scale", index)
        with plb.Lattice.from_flag(flag, 1, index_kind=index, **shape) as lat:
            lat.step(0, 30)
            out.append((sha(lat.read_populations()), lat.reduce()["mass"]))
    assert out[0] == out[1]


def test_bitmap_rejected_configurations(plb):
    flag = random_mask((8, 9, 10), 0.6, 1)
    bad = [dict(bb_mode=plb.BB_REFCOMPAT), dict(bb_mode=plb.BB_HALFWAY, ordering=plb.ORDER_BRICK),
           dict(bb_mode=plb.BB_HALFWAY, ordering=plb.ORDER_PENCIL, brick=(10, 3, 2)),
           dict(bb_mode=plb.BB_HALFWAY, ordering=plb.ORDER_PENCIL, brick=(9, 3, 1))]
    for kw in bad:
        with pytest.raises(plb.PlbError) as e:
            plb.Lattice.from_flag(flag, 1, index_kind=plb.INDEX_BITMAP, **kw)
        assert e.value.code == -20, kw
    with pytest.raises(plb.PlbError) as e:
        plb.Batch([flag], [0.8], [1e-5], index_kind=plb.INDEX_BITMAP, bb_mode=plb.BB_HALFWAY)
    assert e.value.code == -20
