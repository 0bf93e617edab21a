"""The persistent batch launch (plb_batch_run_persistent) against the
one-launch-per-step driver (plb_batch_run): the same steps, the same
per-sample stop steps, the same reductions and macros, bit for bit; and a
density-guard trip inside the launch is reported like the per-step driver's."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def masks(plb):
    return [plb.mask_spheres((24, 20, 28), 3.0, 5.0, 0.4, 1),
            plb.mask_noise((32, 16, 24), 2.0, 0.45, 2),
            plb.mask_spheres((17, 23, 19), 2.0, 4.0, 0.35, 3),
            plb.mask_noise((40, 36, 44), 2.0, 0.3, 4)]


TAUS, GXS = [0.8, 1.0, 0.6, 0.9], [1e-5, 2e-5, 5e-6, 1e-5]


def both(plb, fl, max_steps, check, tol, **kw):
    out = []
    for persistent in (False, True):
        say("persistent" if persistent else "per-step", kw, max_steps, check, tol)
        b = plb.Batch(fl, TAUS[:len(fl)], GXS[:len(fl)], **kw)
        try:
            steps, upd, ms = b.run(max_steps, check, tol, persistent=persistent)
            out.append(dict(steps=steps, upd=upd, red=b.reduce(),
                            mac=[b.read_macros(s) for s in range(len(fl))]))
        finally:
            b.close()
    return out


@pytest.mark.parametrize("kw", [dict(precision=1, bb_mode=1, index_kind=1, layout=1, ordering=2, brick=(64, 4, 1)),
                                dict(precision=0, bb_mode=0, index_kind=0, layout=0),
                                dict(precision=2, bb_mode=1, index_kind=1, layout=1, fp16_scale_log2=20),
                                dict(precision=1, bb_mode=1, index_kind=1, layout=1, block_size=128,
                                     waves_per_simd=5)])
def test_fixed_steps_equal_per_step_driver(plb, kw):
    a, b = both(plb, masks(plb), 137, 40, 0.0, **kw)
    assert list(a["steps"]) == list(b["steps"]) == [137] * 4
    assert a["upd"] == b["upd"]
    for s in range(4):
        assert a["red"][s] == b["red"][s], s
        assert np.array_equal(a["mac"][s], b["mac"][s]), s


def test_convergence_stops_equal_per_step_driver(plb):
    a, b = both(plb, masks(plb), 20000, 50, 1e-6, precision=1, bb_mode=1, index_kind=1, layout=1)
    say("steps", a["steps"], b["steps"])
    assert list(a["steps"]) == list(b["steps"])
    assert (a["steps"] < 20000).all() and len(set(a["steps"].tolist())) >= 2
    for s in range(4):
        assert a["red"][s] == b["red"][s], s
        assert np.array_equal(a["mac"][s], b["mac"][s]), s


@pytest.mark.parametrize("kw", [dict(), dict(index_kind=1, layout=1, ordering=2, brick=(64, 4, 1))])
def test_density_guard_in_persistent_launch(plb, kw):
    """A density-guard trip in one sample stops both drivers with -6, and
    plb_batch_last_fault names the sample and reports the node, coordinates,
    step and rho that the sample's own lattice reports (porelb.c:338-342)."""
    d = dict(np.load(os.path.join(GOLDEN, "fault.npz")))
    fl = [masks(plb)[0], d["flag"]]
    with plb.Lattice.from_flag(d["flag"], tau=float(d["tau"]), gx=float(d["gx"])) as lat:
        assert lat.step_rc(0, 200) == plb.ERR_DENSITY
        want = lat.last_fault()
    for persistent in (False, True):
        b = plb.Batch(fl, [0.8, float(d["tau"])], [1e-5, float(d["gx"])], precision=0, bb_mode=0, **kw)
        try:
            with pytest.raises(plb.PlbError) as ei:
                b.run(200, 10, 0.0, persistent=persistent)
            assert ei.value.code == plb.ERR_DENSITY
            s, f = b.last_fault()
            assert s == 1
            for k in ("tripped", "step", "node", "local_id", "x", "y", "z", "rho"):
                assert f[k] == want[k], (persistent, k)
        finally:
            b.close()
