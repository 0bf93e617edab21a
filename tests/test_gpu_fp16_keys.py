"""FP16 stochastic rounding keys read as runs per 64-node tile: a pencil tile's
V-order keys are one or two runs of consecutive values (a pencil's rows of one
z-plane follow each other in V order too), stored as 16 bytes per tile; tiles
with more than three runs (bricks) load their keys per node.  Every driver
(one and two nodes per thread, AA streaming, batches per step and persistent)
must give the state of the V-order lattice, whose keys are the node indices
themselves, bit for bit; and the runs must carry the pencil tiles, so the fast
path is the one tested."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (nz, ny, nx): rows of about 200, 45 and 13 pores
MASKS = {"long": ((6, 8, 400), 0.5), "medium": ((10, 12, 90), 0.5), "short": ((24, 20, 28), 0.45)}


def say(*a):
    print("[case]", *a, flush=True)


def mask(plb, name):
    dims, phi = MASKS[name]
    return plb.mask_spheres(dims, 2.0, 4.0, phi, 17)


KW = dict(tau=0.7, gx=2e-5, precision=2, bb_mode=1, fp16_rounding=1, fp16_scale_log2=20)


def v_order_state(plb, flag, steps):
    with plb.Lattice.from_flag(flag, **KW) as ref:
        ref.step(0, steps)
        return ref.read_state(), ref.read_macros()


@pytest.mark.parametrize("name", list(MASKS))
@pytest.mark.parametrize("shape", ["pencil", "pencil_npt2", "aa", "soa_pencil", "brick"])
def test_fp16_keys_equal_v_order(plb, name, shape):
    flag = mask(plb, name)
    nx = flag.shape[2]
    say(name, shape)
    want, _ = v_order_state(plb, flag, 30)
    kw = dict(index_kind=1, layout=0 if shape == "soa_pencil" else 1, ordering=2, brick=(nx, 4, 1))
    if shape == "brick":
        kw.update(ordering=1, brick=(8, 4, 2))
    if shape == "pencil_npt2":
        kw["nodes_per_thread"] = 2
    if shape == "aa":
        kw["stream_mode"] = plb.STREAM_AA
    with plb.Lattice.from_flag(flag, **KW, **kw) as lat:
        lat.step(0, 30)
        assert np.array_equal(lat.read_state(), want)


@pytest.mark.parametrize("persistent", [False, True])
def test_fp16_keys_batch_equal_v_order(plb, persistent):
    names = ["long", "medium", "short"]
    flags = [mask(plb, n) for n in names]
    b = plb.Batch(flags, [0.7] * 3, [2e-5] * 3, precision=2, bb_mode=1, fp16_rounding=1, fp16_scale_log2=20,
                  index_kind=1, layout=1, ordering=2, brick=(16, 4, 1))
    try:
        b.run(30, 30, 0.0, persistent=persistent)
        for s, f in enumerate(flags):
            _, mac = v_order_state(plb, f, 30)
            assert np.array_equal(b.read_macros(s), mac), names[s]
    finally:
        b.close()


@pytest.mark.parametrize("name", ["long", "short"])
def test_fp16_key_bytes_follow_the_runs(plb, name):
    """Rounding-key bytes per update (the difference between stochastic and
    round-to-nearest FP16 lattices of one shape): 16 bytes per tile plus 4 per
    node of the tiles that load keys, over N nodes in T = Npad / 64 tiles, so
    between 16 T / N (every tile in runs) and 272 T / N (none).  Pencils put
    nearly every tile in runs; 8 x 4 x 2 bricks (runs of at most 8 voxels)
    leave most tiles loading keys."""
    flag = mask(plb, name)
    got = {}
    for order, brick in (("pencil", (flag.shape[2], 4, 1)), ("brick", (8, 4, 2))):
        kw = dict(KW, index_kind=1, layout=1, ordering=2 if order == "pencil" else 1, brick=brick)
        with plb.Lattice.from_flag(flag, **kw) as sr, \
                plb.Lattice.from_flag(flag, **dict(kw, fp16_rounding=0)) as rn:
            got[order] = sr.bytes_per_update() - rn.bytes_per_update()
    say(name, "key bytes per update", got)
    assert 0.25 <= got["pencil"] < 0.6
    assert 2.0 < got["brick"] <= 4.5
