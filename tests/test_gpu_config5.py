"""The batching workload at its full size: 64 samples with edge L in
64..192 and porosity in 0.15..0.45, alternately sphere packs and smoothed
noise (seed 100, the benchmark's batch study), tau 1.0, g_x 1e-6.

- FP32 storage, tol 0, 200 steps: every sample's reductions and V-order
  macros from the batch equal its own single-lattice run bit for bit.
- FP64 with the reference bounce order: the three smallest samples, batched,
  equal the oracle's pull restatement of porelb.c bit for bit.
- A convergence run stops every sample at the step count a run of that
  sample alone reaches."""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TAU, GX = 1.0, 1e-6
SHAPE = dict(precision=1, bb_mode=1, index_kind=1, layout=1, ordering=2, brick=(192, 32, 1), block_size=256,
             waves_per_simd=5)


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def flags(plb):
    sys.path.insert(0, REPO)
    import bench

    f = bench.batch_samples(plb, 64, 100)
    say("samples", len(f), "pore nodes", sum(int(x.sum()) for x in f))
    return f


def test_batch_equals_single_runs(plb, flags):
    n = len(flags)
    b = plb.Batch(flags, [TAU] * n, [GX] * n, **SHAPE)
    try:
        steps, upd, _ = b.run(200, 50, 0.0)
        assert (steps == 200).all() and upd == 200 * sum(b.nodes)
        red = b.reduce()
        for s, f in enumerate(flags):
            say("sample", s, f.shape, b.nodes[s])
            with plb.Lattice.from_flag(f, tau=TAU, gx=GX, **{k: v for k, v in SHAPE.items() if k != "brick"},
                                       brick=(f.shape[2], 32, 1)) as lat:
                assert lat.N == b.nodes[s]
                lat.step(0, 200)
                r = lat.reduce()
                for k in ("mass", "sum_ux", "K"):
                    assert red[s][k] == r[k], (s, k)
                assert np.array_equal(b.read_macros(s), lat.read_macros()), s
    finally:
        b.close()


def test_batch_fp64_refcompat_smallest_vs_oracle(plb, flags):
    order = np.argsort([int(f.sum()) for f in flags])[:3]
    sub = [flags[k] for k in order]
    b = plb.Batch(sub, [TAU] * 3, [GX] * 3, precision=0, bb_mode=0, index_kind=1, layout=1, ordering=2,
                  brick=(192, 32, 1), block_size=256)
    try:
        b.run(200, 200, 0.0)
        red = b.reduce()
        for s, f in enumerate(sub):
            say("oracle sample", int(order[s]), f.shape)
            V, start = O.build_V(f, 1)
            tab = O.pull_table(V, start, O.BB_REFCOMPAT)
            f0, _ = O.init_equilibrium(len(V))
            mac = O.pull_run(tab, f0, 200, TAU, GX)
            assert np.array_equal(b.read_macros(s), mac), s
            want = O.reduce(f0, mac[1], f.shape, TAU, GX)
            assert red[s]["mass"] == pytest.approx(want["mass"], rel=1e-12)
            assert red[s]["K"] == pytest.approx(want["K"], rel=1e-12)
    finally:
        b.close()


def test_batch_convergence_steps_equal_sequential(plb, flags):
    n = len(flags)
    b = plb.Batch(flags, [TAU] * n, [GX] * n, **SHAPE)
    try:
        steps, _, _ = b.run(20000, 100, 1e-3)
    finally:
        b.close()
    say("batched steps", steps.min(), steps.max())
    assert steps.max() <= 20000
    for s, f in enumerate(flags):
        one = plb.Batch([f], [TAU], [GX], **SHAPE)
        try:
            st, _, _ = one.run(20000, 100, 1e-3)
        finally:
            one.close()
        assert st[0] == steps[s], s


def test_persistent_launch_equals_per_step_driver(plb, flags):
    """The whole batch as one persistent launch: the same stop step, the same
    reductions and the same V-order macros for every sample as one launch
    per step."""
    n = len(flags)
    res = []
    for persistent in (False, True):
        b = plb.Batch(flags, [TAU] * n, [GX] * n, **SHAPE)
        try:
            steps, upd, ms = b.run(20000, 100, 1e-3, persistent=persistent)
            say("persistent" if persistent else "per-step", f"{ms:.1f} ms", f"{upd / (ms * 1e-3) / 1e6:.0f} MLUPS")
            res.append((steps, upd, b.reduce(), [b.read_macros(s) for s in range(n)]))
        finally:
            b.close()
    (s0, u0, r0, m0), (s1, u1, r1, m1) = res
    assert list(s0) == list(s1) and u0 == u1
    for s in range(n):
        assert r0[s] == r1[s], s
        assert np.array_equal(m0[s], m1[s]), s
