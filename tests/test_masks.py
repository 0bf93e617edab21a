"""Synthetic mask generators (host code of libplb; no GPU needed)."""
import numpy as np
import pytest


def test_noise_exact_porosity_and_determinism(plb):
    a = plb.mask_noise((24, 20, 16), 2.0, 0.35, 7)
    b = plb.mask_noise((24, 20, 16), 2.0, 0.35, 7)
    c = plb.mask_noise((24, 20, 16), 2.0, 0.35, 8)
    assert a.dtype == np.uint8 and set(np.unique(a)) <= {0, 1}
    assert int(a.sum()) == round(0.35 * a.size)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_noise_is_smooth_and_periodic(plb):
    """Blurred noise: neighbouring voxels agree far more often than i.i.d.
    voxels would, across the periodic boundary as much as inside."""
    m = plb.mask_noise((32, 32, 32), 3.0, 0.5, 3).astype(int)
    inner = (m[:, :, 1:] == m[:, :, :-1]).mean()
    wrap = (m[:, :, 0] == m[:, :, -1]).mean()
    assert inner > 0.8 and wrap > 0.8
    raw = plb.mask_noise((32, 32, 32), 0.0, 0.5, 3).astype(int)
    assert abs((raw[:, :, 1:] == raw[:, :, :-1]).mean() - 0.5) < 0.05


def test_spheres_stop_at_target(plb):
    for phi in (0.2, 0.3, 0.45):
        m = plb.mask_spheres((48, 40, 32), 3.0, 6.0, phi, 5)
        por = m.mean()
        assert por <= phi
        assert por > phi - 0.02  # one sphere removes at most (4/3)pi 6^3 voxels
    a = plb.mask_spheres((32, 32, 32), 2.0, 4.0, 0.3, 9)
    assert np.array_equal(a, plb.mask_spheres((32, 32, 32), 2.0, 4.0, 0.3, 9))


def test_bcc_geometry(plb):
    m = plb.mask_bcc(32, 0.5)
    assert m[0, 0, 0] == 0 and m[16, 16, 16] == 0  # sphere centres are solid
    assert m[8, 0, 0] == 1                         # between spheres along x
    r = (-np.arange(32)) % 32                      # x -> -x (mod n) on every axis
    assert np.array_equal(m, m[np.ix_(r, r, r)])
    assert np.array_equal(m, m.transpose(1, 2, 0))  # cubic symmetry


def test_invalid_arguments(plb):
    with pytest.raises(plb.PlbError):
        plb.mask_noise((8, 8, 8), 1.0, 1.5, 1)
    with pytest.raises(plb.PlbError):
        plb.mask_spheres((8, 8, 8), 3.0, 2.0, 0.3, 1)
