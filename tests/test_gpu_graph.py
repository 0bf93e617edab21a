"""Graph replays (graph_steps > 0): runs of steps replay one captured HIP
graph whose kernels read the step number from device memory.  Results must be
bit-identical to one launch per step, across repeated replays with other
device work and host reads in between, including a density-guard fault
inside a replay."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def say(*a):
    print("[case]", *a, flush=True)


def test_graph_config1_matches_reference(plb):
    """Config 1 (64^3 noise): 1000 FP64 steps as graph replays of 16 steps,
    issued in uneven calls with reads in between; SHA-256 as the reference."""
    meta = json.load(open(os.path.join(GOLDEN, "config1.json")))
    m = meta["mask"]
    flag = plb.mask_noise(tuple(m["dims"]), m["sigma"], m["porosity"], m["seed"])
    g = meta["np1"]
    say("graph config1")
    with plb.Lattice.from_flag(flag, 1, tau=1.0, gx=1e-6, graph_steps=16) as lat:
        it = 0
        for n in (1, 17, 100, 333, 49, 500):
            lat.step(it, n)
            it += n
            lat.read_macros()
        assert it == g["steps"]
        assert sha(lat.read_populations()) == g["f0_sha256"]
        assert sha(lat.read_macros()) == g["macros_sha256"]


@pytest.mark.parametrize("kw", [
    dict(precision=1, bb_mode=1, layout=1, ordering=2, brick=(48, 8, 1), index_kind=1),
    dict(precision=1, bb_mode=1, layout=1, ordering=2, brick=(48, 8, 1), index_kind=2),
    dict(precision=2, bb_mode=1, layout=1, index_kind=1, fp16_scale_log2=16),
    dict(precision=0, bb_mode=0, layout=1, index_kind=1, nodes_per_thread=2, block_size=256),
])
def test_graph_equals_plain_launches(plb, kw):
    flag = plb.mask_spheres((40, 36, 48), 3.0, 6.0, 0.35, 4)
    out = []
    for G in (0, 2, 6, 64):
        say("graph vs plain", G, kw)
        with plb.Lattice.from_flag(flag, 1, tau=0.7, gx=1e-5, graph_steps=G, **kw) as lat:
            it = 0
            for n in (7, 130, 1, 64, 65):
                lat.step(it, n)
                it += n
                r = lat.reduce()
            out.append((sha(lat.read_populations()), sha(lat.read_macros()), r["mass"]))
    assert all(o == out[0] for o in out)


def test_graph_density_guard(plb):
    """The reference's density-guard case trips at step 3; with replays of 2
    or 6 steps the fault lands inside a replay, and the state, step, node and
    density match the plain path."""
    d = dict(np.load(os.path.join(GOLDEN, "fault.npz")))
    flag, tau, gx = d["flag"], float(d["tau"]), float(d["gx"])
    V, start = O.build_V(flag, 1)
    res = []
    for G in (0, 2, 6):
        say("graph fault", G)
        with plb.Lattice.from_V(V, start, flag.shape, tau=tau, gx=gx, graph_steps=G) as lat:
            assert lat.step_rc(0, 200) == plb.ERR_DENSITY
            ft = lat.last_fault()
            assert lat.step_rc(200, 20) == plb.ERR_DENSITY
            res.append((ft, sha(lat.read_populations())))
    assert res[0] == res[1] == res[2]
    assert res[0][0]["step"] == 3


def test_graph_steps_validated(plb):
    flag = (np.random.default_rng(2).random((6, 6, 6)) < 0.6).astype(np.uint8)
    for G in (-2, 3, 5000):
        with pytest.raises(plb.PlbError) as e:
            plb.Lattice.from_flag(flag, 1, graph_steps=G)
        assert e.value.code == -20
