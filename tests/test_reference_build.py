"""The oracle against the reference programs built from the reference's own
sources (oracle/_ref), on fresh random masks.  Skipped where the reference
build is absent (it is built only where the reference sources exist)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.skipif(not O.have_reference_build(), reason="oracle/_ref not built")


@pytest.mark.parametrize("seed,dims,np_", [(101, (10, 6, 7), 1), (102, (10, 6, 7), 2),
                                            (103, (12, 5, 5), 3), (104, (8, 9, 6), 4)])
def test_oracle_equals_reference_binary(seed, dims, np_):
    flag = (np.random.default_rng(seed).random(dims) < 0.58).astype(np.uint8)
    ref = O.run_reference(flag, np_, 35, 0.75, 3e-4, save_every=1000)
    V, start = O.build_V(flag, np_)
    for a, b in zip(O.semantic_fields(V), O.semantic_fields(ref["vbin"]["V"])):
        assert np.array_equal(a, b)
    f0, f1 = O.init_equilibrium(len(V))
    mac = O.push_run(V, start, f0, f1, 35, 0.75, 3e-4)
    assert np.array_equal(f0, ref["f0"])
    assert np.array_equal(f1, ref["f1"])
    assert np.array_equal(mac, ref["macros"])
