"""One-buffer AA streaming (plb_config.stream_mode = PLB_STREAM_AA): every
step stores exactly what the two-buffer pull step stores, in every
precision, index encoding, layout and node order; populations and stored
elements read back in the pull layout after odd and even step counts; a
state written into an AA lattice resumes bit for bit, also when it crosses
between the stream modes; the convergence driver stops at the same step; a
density-guard trip freezes the same state; FP64 equals the oracle's half-way
pull run."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def mask(plb):
    return plb.mask_spheres((26, 22, 30), 2.0, 4.0, 0.45, 11)


CASES = [(0, 0, 0, 0), (0, 1, 1, 2), (1, 0, 0, 0), (1, 1, 1, 2), (1, 2, 1, 2), (1, 1, 0, 1), (2, 1, 1, 2),
         (2, 0, 1, 0), (2, 2, 1, 2)]


@pytest.mark.parametrize("prec,index,layout,order", CASES)
@pytest.mark.parametrize("nsteps", [7, 8])
def test_aa_equals_pull(plb, prec, index, layout, order, nsteps):
    flag = mask(plb)
    say(f"prec {prec} index {index} layout {layout} order {order} steps {nsteps}")
    brick = (8, 4, 4) if order == 1 else (30, 4, 1)
    kw = dict(tau=0.7, gx=2e-5, precision=prec, bb_mode=plb.BB_HALFWAY, index_kind=index, layout=layout,
              ordering=order, brick=brick, fp16_rounding=plb.ROUND_STOCHASTIC, block_size=128)
    with plb.Lattice.from_flag(flag, **kw) as pull, \
            plb.Lattice.from_flag(flag, stream_mode=plb.STREAM_AA, **kw) as aa:
        pull.step(0, nsteps)
        aa.step(0, nsteps)
        assert np.array_equal(aa.read_state(), pull.read_state())
        assert np.array_equal(aa.read_populations(), pull.read_populations())
        assert np.array_equal(aa.read_macros(), pull.read_macros())
        ra, rp = aa.reduce(), pull.reduce()
        assert ra["sum_ux"] == rp["sum_ux"]
        assert ra["mass"] == pytest.approx(rp["mass"], rel=1e-13)
        # the read-backs left the buffer as it was: the runs stay equal
        pull.step(nsteps, 5)
        aa.step(nsteps, 5)
        assert np.array_equal(aa.read_state(), pull.read_state())


@pytest.mark.parametrize("split", [7, 8])
def test_aa_state_resume(plb, split):
    """Stored elements read from an AA lattice after `split` steps and written
    into a fresh AA lattice continue like a straight pull run."""
    flag = mask(plb)
    kw = dict(tau=0.8, gx=1e-5, precision=1, bb_mode=plb.BB_HALFWAY, index_kind=1, layout=1, ordering=2,
              brick=(30, 4, 1))
    with plb.Lattice.from_flag(flag, stream_mode=plb.STREAM_AA, **kw) as a, \
            plb.Lattice.from_flag(flag, stream_mode=plb.STREAM_AA, **kw) as b, \
            plb.Lattice.from_flag(flag, **kw) as ref:
        a.step(0, split)
        snap = a.read_state()
        b.write_state(snap)
        assert np.array_equal(b.read_state(), snap)
        b.step(split, 20 - split)
        ref.step(0, 20)
        assert np.array_equal(b.read_state(), ref.read_state())
        # writing into an AA lattice after an odd number of its own steps
        a.step(split, 3)
        a.write_state(snap)
        a.step(split, 20 - split)
        assert np.array_equal(a.read_state(), ref.read_state())


@pytest.mark.parametrize("prec", [0, 2])
@pytest.mark.parametrize("split", [7, 8])
def test_state_crosses_stream_modes(plb, prec, split):
    """Stored elements are the same in both stream modes: a state read from a
    pull lattice continues in an AA lattice, and one read from an AA lattice
    continues in a pull lattice, each like a straight pull run."""
    flag = mask(plb)
    kw = dict(tau=0.75, gx=3e-5, precision=prec, bb_mode=plb.BB_HALFWAY, index_kind=1, layout=1, ordering=2,
              brick=(30, 4, 1), fp16_rounding=plb.ROUND_STOCHASTIC)
    say(f"prec {prec} split {split}")
    with plb.Lattice.from_flag(flag, **kw) as ref, \
            plb.Lattice.from_flag(flag, **kw) as pull, \
            plb.Lattice.from_flag(flag, stream_mode=plb.STREAM_AA, **kw) as aa:
        ref.step(0, split)
        snap = ref.read_state()
        ref.step(split, 20 - split)
        aa.write_state(snap)
        aa.step(split, 20 - split)
        assert np.array_equal(aa.read_state(), ref.read_state())
        aa.init_equilibrium()
        aa.step(0, split)
        pull.write_state(aa.read_state())
        pull.step(split, 20 - split)
        assert np.array_equal(pull.read_state(), ref.read_state())


@pytest.mark.parametrize("check", [25, 24])
def test_aa_run_to_convergence_equals_pull(plb, check):
    """The convergence driver stops an AA lattice at the same step, with the
    same sums and stored elements, as a pull lattice (check intervals of 25
    and 24 steps)."""
    flag = mask(plb)
    kw = dict(tau=0.9, gx=1e-5, precision=1, bb_mode=plb.BB_HALFWAY, index_kind=1, layout=1, ordering=2,
              brick=(30, 4, 1))
    with plb.Lattice.from_flag(flag, **kw) as pull, \
            plb.Lattice.from_flag(flag, stream_mode=plb.STREAM_AA, **kw) as aa:
        n_p, r_p = pull.run_to_convergence(0, 20000, check, 1e-4)
        n_a, r_a = aa.run_to_convergence(0, 20000, check, 1e-4)
        say(f"check {check}: stopped after {n_p} steps")
        assert 0 < n_p < 20000
        assert n_a == n_p
        assert r_a["sum_ux"] == r_p["sum_ux"]
        assert np.array_equal(aa.read_state(), pull.read_state())
        assert np.array_equal(aa.read_macros(), pull.read_macros())


def test_aa_fp64_equals_oracle(plb):
    flag = mask(plb)
    V, start = O.build_V(flag, 1)
    tab = O.pull_table(V, start, O.BB_HALFWAY)
    f0, _ = O.init_equilibrium(len(V))
    mac = O.pull_run(tab, f0, 25, 0.9, 5e-4)
    with plb.Lattice.from_V(V, start, flag.shape, tau=0.9, gx=5e-4, bb_mode=plb.BB_HALFWAY,
                            stream_mode=plb.STREAM_AA, index_kind=1, layout=1) as lat:
        lat.step(0, 25)
        assert np.array_equal(lat.read_populations(), f0)
        assert np.array_equal(lat.read_macros(), mac)


@pytest.mark.parametrize("gx", [0.25, 0.2])
def test_aa_density_trip_freezes_like_pull(plb, gx):
    """A guard trip (fixture mask, large g_x) leaves the AA lattice holding the
    pull lattice's frozen state, whichever phase the trip falls in."""
    d = dict(np.load(os.path.join(GOLDEN, "fault.npz")))
    flag = d["flag"]
    kw = dict(tau=float(d["tau"]), gx=gx, precision=0, bb_mode=plb.BB_HALFWAY, index_kind=1, layout=1)
    with plb.Lattice.from_flag(flag, **kw) as pull, \
            plb.Lattice.from_flag(flag, stream_mode=plb.STREAM_AA, **kw) as aa:
        rp = pull.step_rc(0, 200)
        ra = aa.step_rc(0, 200)
        assert rp == ra == plb.ERR_DENSITY
        fp, fa = pull.last_fault(), aa.last_fault()
        say("trip at step", fp["step"])
        assert fa == fp
        assert np.array_equal(aa.read_populations(), pull.read_populations())


def test_aa_rejected_combinations(plb):
    flag = mask(plb)
    for kw in (dict(bb_mode=plb.BB_REFCOMPAT), dict(bb_mode=plb.BB_HALFWAY, nodes_per_thread=2),
               dict(bb_mode=plb.BB_HALFWAY, graph_steps=16)):
        with pytest.raises(plb.PlbError):
            plb.Lattice.from_flag(flag, stream_mode=plb.STREAM_AA, **kw)
    with pytest.raises(plb.PlbError):
        plb.Batch([flag], [0.8], [1e-5], bb_mode=plb.BB_HALFWAY, stream_mode=plb.STREAM_AA)
