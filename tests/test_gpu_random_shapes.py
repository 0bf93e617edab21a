"""Seeded sweep over small random masks and random valid configurations: the
FP64 path (every index encoding, layout, node order, brick, block size,
nodes per thread, graph replay length and waves-per-SIMD cap, on 1-4 z-slab
ranks) must equal
the oracle's pull restatement bit for bit, populations and macros.  Dims go
down to 1 voxel and node counts are ragged against the 64-node tiles.  The
same configurations with FP32 shifted storage stay within FP32 rounding of
the FP64 oracle: |d rho| <= 5e-7 and |d u| <= 1e-5 max|u| + 1e-9 after 17
steps (observed: 6e-8 and 1e-6 max|u|)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N_CASES = 160


def say(*a):
    print("[case]", *a, flush=True)


def draw_case(k):
    rng = np.random.default_rng(9000 + k)
    np_ = int(rng.choice([1, 1, 2, 3, 4]))
    nx, ny = (int(v) for v in rng.integers(1, 25, size=2))
    nz = np_ * int(rng.integers(1, 7))
    flag = (rng.random((nz, ny, nx)) < rng.uniform(0.3, 0.9)).astype(np.uint8)
    flag.flat[int(rng.integers(flag.size))] = 1
    index = int(rng.choice([0, 1, 2]))
    mode = O.BB_HALFWAY if index == 2 else int(rng.choice([O.BB_REFCOMPAT, O.BB_HALFWAY]))
    kw = dict(index_kind=index, layout=int(rng.choice([0, 1])),
              block_size=int(rng.choice([64, 128, 256, 512, 1024])))
    if index == 2:
        kw["ordering"] = int(rng.choice([0, 2]))
        if kw["ordering"] == 2:
            kw["brick"] = (nx + int(rng.integers(0, 3)), int(rng.integers(1, ny + 3)), 1)
    else:
        kw["ordering"] = int(rng.choice([0, 1, 2]))
        kw["brick"] = tuple(int(v) for v in rng.integers(1, 10, size=3))
    if rng.random() < 0.3 and kw["block_size"] <= 512:
        kw["nodes_per_thread"] = 2
    if rng.random() < 0.3:
        kw["graph_steps"] = int(rng.choice([2, 4, 6]))
    if rng.random() < 0.3:
        kw["waves_per_simd"] = int(rng.integers(1, 9))
    tau, gx = float(rng.uniform(0.6, 1.4)), float(rng.uniform(-2e-4, 2e-4))
    return flag, np_, mode, kw, tau, gx


@pytest.mark.parametrize("k", range(N_CASES))
def test_random_shape_fp64_bitwise(plb, k):
    flag, np_, mode, kw, tau, gx = draw_case(k)
    say(k, flag.shape[::-1], "np", np_, "bb", mode, kw, "tau", round(tau, 4), "gx", gx)
    V, start = O.build_V(flag, np_)
    tab = O.pull_table(V, start, mode)
    f0, _ = O.init_equilibrium(len(V))
    steps = 17
    mac = O.pull_run(tab, f0, steps, tau, gx)
    with plb.Lattice.from_flag(flag, np_, tau=tau, gx=gx, bb_mode=mode, **kw) as lat:
        assert lat.N == len(V)
        lat.step(0, 9)
        lat.step(9, steps - 9)
        assert np.array_equal(lat.read_populations(), f0)
        assert np.array_equal(lat.read_macros(), mac)


@pytest.mark.parametrize("k", range(0, N_CASES, 4))
def test_random_shape_fp32_within_rounding(plb, k):
    flag, np_, mode, kw, tau, gx = draw_case(k)
    say(k, "fp32", flag.shape[::-1], "np", np_, "bb", mode, kw)
    V, start = O.build_V(flag, np_)
    tab = O.pull_table(V, start, mode)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 17, tau, gx)
    with plb.Lattice.from_flag(flag, np_, tau=tau, gx=gx, bb_mode=mode, precision=plb.FP32, **kw) as lat:
        lat.step(0, 17)
        got = lat.read_macros()
    umax = np.abs(mac[1:]).max()
    say("max |d rho|", np.abs(got[0] - mac[0]).max(), "max |d u|", np.abs(got[1:] - mac[1:]).max(), "max |u|", umax)
    assert np.abs(got[0] - mac[0]).max() <= 5e-7
    assert np.abs(got[1:] - mac[1:]).max() <= 1e-5 * umax + 1e-9


@pytest.mark.parametrize("k", range(2, N_CASES, 4))
def test_random_shape_fp16_within_rounding(plb, k):
    """FP16 storage (deviations scaled by 2^k, FP32 collision) on the same
    random configurations stays within FP16 rounding of the FP64 oracle after
    17 steps: |d rho| <= 4e-6, |d u| <= 1e-2 max|u| + 1e-9 (observed: 4e-7 and
    8e-4 max|u|)."""
    flag, np_, mode, kw, tau, gx = draw_case(k)
    say(k, "fp16", flag.shape[::-1], "np", np_, "bb", mode, kw)
    V, start = O.build_V(flag, np_)
    tab = O.pull_table(V, start, mode)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 17, tau, gx)
    with plb.Lattice.from_flag(flag, np_, tau=tau, gx=gx, bb_mode=mode, precision=plb.FP16, **kw) as lat:
        lat.step(0, 17)
        got = lat.read_macros()
    umax = np.abs(mac[1:]).max()
    say("fp16 max |d rho|", np.abs(got[0] - mac[0]).max(), "max |d u|", np.abs(got[1:] - mac[1:]).max(),
        "max |u|", umax)
    assert np.abs(got[0] - mac[0]).max() <= 4e-6
    assert np.abs(got[1:] - mac[1:]).max() <= 1e-2 * umax + 1e-9


AA_CASES = [k for k in range(N_CASES) if draw_case(k)[2] == O.BB_HALFWAY]


@pytest.mark.parametrize("k", AA_CASES)
def test_random_shape_aa_fp64_bitwise(plb, k):
    """The half-way cases again with one-buffer AA streaming (one node per
    thread, no graph replays): the oracle's populations and macros bit for
    bit after an odd number of steps, read back through the AA store map."""
    flag, np_, mode, kw, tau, gx = draw_case(k)
    kw = {n: v for n, v in kw.items() if n not in ("nodes_per_thread", "graph_steps")}
    say(k, "aa", flag.shape[::-1], "np", np_, kw)
    V, start = O.build_V(flag, np_)
    tab = O.pull_table(V, start, mode)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 17, tau, gx)
    with plb.Lattice.from_flag(flag, np_, tau=tau, gx=gx, bb_mode=mode, stream_mode=plb.STREAM_AA, **kw) as lat:
        lat.step(0, 8)
        lat.step(8, 9)
        assert np.array_equal(lat.read_populations(), f0)
        assert np.array_equal(lat.read_macros(), mac)
