"""Seeded sweep over random batches: 2-5 small random masks per batch, each
with its own tau and g_x, in random valid batch configurations (bounce-back
mode, I32 or TILE8, layout, node order, block size, waves-per-SIMD cap).
Every sample's FP64 macros (read back in V order) equal the oracle's pull
run bit for bit, and its mass, sum u_x and K equal the oracle's reduction to
1e-12 relative (the GPU sums in node order, the oracle sequentially)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N_CASES = 24
STEPS = 23


def say(*a):
    print("[case]", *a, flush=True)


def draw_batch(k):
    rng = np.random.default_rng(7000 + k)
    n = int(rng.integers(2, 6))
    flags, taus, gxs = [], [], []
    for _ in range(n):
        nx, ny, nz = (int(v) for v in rng.integers(3, 21, size=3))
        f = (rng.random((nz, ny, nx)) < rng.uniform(0.4, 0.9)).astype(np.uint8)
        f.flat[int(rng.integers(f.size))] = 1
        flags.append(f)
        taus.append(float(rng.uniform(0.6, 1.4)))
        gxs.append(float(rng.uniform(1e-5, 1e-4)))
    index = int(rng.choice([0, 1]))
    kw = dict(precision=0, bb_mode=int(rng.choice([0, 1])), index_kind=index, layout=int(rng.choice([0, 1])),
              ordering=int(rng.choice([0, 1, 2])), brick=tuple(int(v) for v in rng.integers(1, 9, size=3)),
              block_size=int(rng.choice([64, 128, 256, 512])), waves_per_simd=int(rng.choice([0, 0, 4, 5, 6])))
    return flags, taus, gxs, kw


@pytest.mark.parametrize("k", range(N_CASES))
def test_random_batch_matches_oracle(plb, k):
    flags, taus, gxs, kw = draw_batch(k)
    say(k, [f.shape[::-1] for f in flags], kw)
    b = plb.Batch(flags, taus, gxs, **kw)
    try:
        steps, upd, _ = b.run(STEPS, STEPS, 0.0)  # tol 0: every sample runs all steps
        red = b.reduce()
        nodes = list(b.nodes)
        macs = [b.read_macros(s) for s in range(len(flags))]
    finally:
        b.close()
    assert list(steps) == [STEPS] * len(flags)
    for s, f in enumerate(flags):
        V, start = O.build_V(f, 1)
        assert nodes[s] == len(V)
        tab = O.pull_table(V, start, kw["bb_mode"])
        f0, _ = O.init_equilibrium(len(V))
        mac = O.pull_run(tab, f0, STEPS, taus[s], gxs[s])
        assert np.array_equal(macs[s], mac), s  # V order, bit for bit
        want = O.reduce(f0, mac[1], f.shape, taus[s], gxs[s])
        for key in ("mass", "sum_ux", "K"):
            assert red[s][key] == pytest.approx(want[key], rel=1e-12, abs=1e-300), (s, key)
