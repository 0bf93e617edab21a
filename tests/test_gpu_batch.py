"""Batched samples: every sample of a batch evolves exactly as the same
sample alone, and per-sample convergence stopping freezes each sample at its
own step count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def masks(plb):
    return [plb.mask_spheres((24, 20, 28), 3.0, 5.0, 0.4, 1),
            plb.mask_noise((32, 16, 24), 2.0, 0.45, 2),
            plb.mask_spheres((17, 23, 19), 2.0, 4.0, 0.35, 3)]


@pytest.mark.parametrize("prec,index,order,bb,waves", [(0, 0, 0, 1, 0), (1, 0, 0, 1, 0), (1, 1, 0, 1, 0),
                                                       (2, 1, 0, 1, 0), (1, 1, 2, 1, 0), (0, 0, 1, 1, 0),
                                                       (0, 1, 2, 0, 0), (1, 1, 2, 0, 0), (1, 1, 2, 1, 5)])
def test_batch_equals_single_runs(plb, prec, index, order, bb, waves):
    fl = masks(plb)
    taus, gxs = [0.8, 1.0, 0.6], [1e-5, 2e-5, 5e-6]
    brick = (8, 4, 2) if order == 1 else (64, 4, 1)
    b = plb.Batch(fl, taus, gxs, precision=prec, bb_mode=bb, index_kind=index, layout=index,
                  fp16_scale_log2=20, ordering=order, brick=brick, waves_per_simd=waves)
    steps, upd, ms = b.run(300, 100, 0.0)   # tol 0: nobody stops early
    assert list(steps) == [300, 300, 300]
    assert upd == 300 * sum(b.nodes)
    red = b.reduce()
    for s, f in enumerate(fl):
        sbrick = (f.shape[2],) + brick[1:] if order == 2 else brick
        with plb.Lattice.from_flag(f, tau=taus[s], gx=gxs[s], precision=prec,
                                   bb_mode=bb, index_kind=index, layout=index,
                                   fp16_scale_log2=20, ordering=order, brick=sbrick) as lat:
            assert lat.N == b.nodes[s]
            lat.step(0, 300)
            r = lat.reduce()
        assert red[s]["sum_ux"] == r["sum_ux"], s
        assert red[s]["mass"] == pytest.approx(r["mass"], rel=1e-14), s


def test_batch_convergence_stops_per_sample(plb):
    fl = masks(plb)
    b = plb.Batch(fl, [0.8, 1.0, 0.6], [1e-5, 2e-5, 5e-6], precision=1, bb_mode=plb.BB_HALFWAY)
    steps, upd, ms = b.run(20000, 50, 1e-6)
    assert (steps < 20000).all() and len(set(steps.tolist())) >= 2
    assert upd == sum(int(s) * n for s, n in zip(steps, b.nodes))
    red = b.reduce()
    for s, f in enumerate(fl):
        with plb.Lattice.from_flag(f, tau=[0.8, 1.0, 0.6][s], gx=[1e-5, 2e-5, 5e-6][s], precision=1,
                                   bb_mode=plb.BB_HALFWAY) as lat:
            lat.step(0, int(steps[s]))
            assert lat.reduce()["sum_ux"] == red[s]["sum_ux"]
            # the sample's macros are those of its own last step, in V order
            assert np.array_equal(b.read_macros(s), lat.read_macros()), s


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_autotune_shape_keeps_results(plb, persistent):
    """plb_batch_autotune, timing either driver: 16 stage-1 batches (2 index
    kinds x 2 layouts x 4 node orders), 2 other block sizes, the winner again
    and 3 wave caps; the shape it picks gives every sample the macros of the
    default shape."""
    fl = masks(plb)
    taus, gxs = [0.8, 1.0, 0.6], [1e-5, 2e-5, 5e-6]
    base = plb.default_config(precision=1, bb_mode=plb.BB_HALFWAY, block_size=256)
    best, res = plb.batch_autotune(fl, taus, gxs, steps=8, cfg=base, persistent=persistent)
    assert len(res) == 16 + 2 + 1 + 3 and all(r["mlups"] > 0 for r in res)
    assert {(r["index_kind"], r["layout"]) for r in res[:16]} == {(0, 0), (0, 1), (1, 0), (1, 1)}
    assert {r["block"] for r in res[16:18]} == {128, 512}
    assert [r["waves"] for r in res[19:]] == [6, 5, 4]
    assert best.nodes_per_thread == 1 and best.precision == 1 and best.bb_mode == plb.BB_HALFWAY
    top = max(res, key=lambda r: r["mlups"])
    assert (top["index_kind"], top["layout"], top["ordering"], top["block"], top["waves"]) == \
        (best.index_kind, best.layout, best.ordering, best.block_size, best.waves_per_simd)
    out = []
    for cfg in (base, best):
        b = plb.Batch(fl, taus, gxs, cfg=cfg)
        try:
            b.run(120, 40, 0.0, persistent=persistent)
            out.append([b.read_macros(s) for s in range(len(fl))])
        finally:
            b.close()
    for s in range(len(fl)):
        assert np.array_equal(out[0][s], out[1][s]), s
