"""Autotuner and copy-ceiling calibration on the GPU."""
import time

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def test_copy_bandwidth_plausible(plb):
    g = plb.copy_bandwidth(1 << 30, 3)
    assert 1000.0 < g < 8000.0  # GB/s, below the 8 TB/s HBM peak


def test_autotune_cache_and_result(plb, tmp_path):
    flag = plb.mask_spheres((48, 40, 64), 3.0, 6.0, 0.3, 8)
    cache = str(tmp_path / "tune.txt")
    base = plb.default_config(tau=0.8, gx=1e-5, precision=plb.FP64, bb_mode=plb.BB_REFCOMPAT)
    best, res = plb.autotune(flag, 1, 5, cache, cfg=base)
    # stage 1: 2 index kinds x 2 layouts x 4 orders; stage 2: 3 blocks x 2 nodes/thread x 4 wave caps
    assert len(res) == 2 * 2 * 4 + 3 * 2 * 4 and all(r["mlups"] > 0 for r in res)
    assert {r["index_kind"] for r in res} == {plb.INDEX_I32, plb.INDEX_TILE8}
    assert {r["waves"] for r in res[16:]} == {0, 6, 5, 4}
    assert max(r["mlups"] for r in res) == pytest.approx(max(r["mlups"] for r in res))
    text = open(cache).read().split()
    assert len(text) == 13
    t0 = time.perf_counter()
    best2, res2 = plb.autotune(flag, 1, 5, cache, cfg=base)
    assert res2 == [] and time.perf_counter() - t0 < 2.0
    for k in ("index_kind", "layout", "ordering", "block_size", "nodes_per_thread", "waves_per_simd"):
        assert getattr(best, k) == getattr(best2, k)
    # the tuned shape changes speed only, never the FP64 result
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start, O.BB_REFCOMPAT)
    f0, _ = O.init_equilibrium(len(V))
    O.pull_run(tab, f0, 30, 0.8, 1e-5)
    with plb.Lattice.from_flag(flag, 1, cfg=best) as lat:
        lat.step(0, 30)
        assert np.array_equal(lat.read_populations(), f0)


def test_autotune_aa_keeps_stream_mode(plb, tmp_path):
    """Tuning a one-buffer AA configuration returns an AA configuration with
    one node per thread (the only form AA runs), cached apart from the pull
    mode, and its result equals the oracle's half-way run."""
    flag = plb.mask_spheres((40, 36, 48), 3.0, 6.0, 0.3, 9)
    cache = str(tmp_path / "tune.txt")
    pull = plb.default_config(tau=0.8, gx=1e-5, precision=plb.FP64, bb_mode=plb.BB_HALFWAY)
    aa = plb.default_config(tau=0.8, gx=1e-5, precision=plb.FP64, bb_mode=plb.BB_HALFWAY,
                            stream_mode=plb.STREAM_AA)
    plb.autotune(flag, 1, 4, cache, cfg=pull)
    best, res = plb.autotune(flag, 1, 4, cache, cfg=aa)
    assert res and all(r["npt"] == 1 for r in res)
    assert best.stream_mode == plb.STREAM_AA and best.nodes_per_thread == 1
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start, O.BB_HALFWAY)
    f0, _ = O.init_equilibrium(len(V))
    O.pull_run(tab, f0, 21, 0.8, 1e-5)
    with plb.Lattice.from_flag(flag, 1, cfg=best) as lat:
        lat.step(0, 21)
        assert np.array_equal(lat.read_populations(), f0)
