"""plb_read_macros_noalloc -- the macro read-back for callers that cannot
allocate device memory -- against plb_read_macros: two calls in a row on one
lattice (the second smaller), through the C ABI, bit for bit; sizes beyond
the fixed staging array are refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def read_noalloc(plb, lat, first, n):
    out = np.zeros((4, n))
    rc = plb.lib().plb_read_macros_noalloc(lat._h, first, n, *(o.ctypes.data_as(C.c_void_p) for o in out))
    return rc, out


def test_noalloc_read_equals_read_macros(plb):
    flag = plb.mask_spheres((96, 80, 72), 4.0, 8.0, 0.35, 11)
    with plb.Lattice.from_flag(flag, tau=0.8, gx=1e-5, precision=1, bb_mode=1, index_kind=1, layout=1,
                               ordering=2, brick=(72, 8, 1)) as lat:
        lat.step(0, 50)
        want = lat.read_macros()
        n = lat.N
        assert n > 100000
        rc, got = read_noalloc(plb, lat, 0, n)           # whole lattice (below the limit)
        assert rc == 0 and np.array_equal(got, want)
        lat.step(50, 7)
        want = lat.read_macros()
        rc, got = read_noalloc(plb, lat, 1234, 5000)     # second call, smaller, at an offset
        assert rc == 0 and np.array_equal(got, want[:, 1234:6234])
        assert read_noalloc(plb, lat, n - 10, 20)[0] == -20       # past the end
    big = plb.mask_noise((160, 160, 160), 2.0, 0.45, 2)
    with plb.Lattice.from_flag(big, tau=0.8, gx=1e-5) as lat:
        assert lat.N > plb.NOALLOC_MAX_NODES
        lat.step(0, 3)
        assert read_noalloc(plb, lat, 0, plb.NOALLOC_MAX_NODES + 1)[0] == -23   # beyond the array
        rc, got = read_noalloc(plb, lat, 3, plb.NOALLOC_MAX_NODES)
        assert rc == 0 and np.array_equal(got, lat.read_macros()[:, 3:3 + plb.NOALLOC_MAX_NODES])
