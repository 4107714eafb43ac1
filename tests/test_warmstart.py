"""Host side of the learned warm start: the cell features (closed forms on a
uniform disk) and the field assembly."""
import math

import numpy as np
import pytest

from mcmax_amd import configs, warmstart


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    return configs.load_config("small", str(tmp_path_factory.mktemp("ws")))


def test_feature_optical_depths_closed_form(small):
    m = small
    dens = np.full((m.nr, m.nt), 1e-18)
    dens[0] = 1e-60
    X = warmstart.features(m, dens=dens).reshape(m.nr - 1, m.nt, 8)
    kx = float(m.w[5, 5] @ warmstart.kext_at(m))
    krho = kx * 1e-18
    # radially: from R_in to the cell's radial midpoint
    Rmid = 0.5 * (m.R[1:-1] + m.R[2:])
    tr = krho * (Rmid - m.R[1]) * warmstart.AU
    np.testing.assert_allclose(10 ** X[:, 7, 2] - 1e-10, tr, rtol=1e-9)
    # vertically: from the pole to the cell's midpoint in theta, at R_av
    th = np.arccos(m.Theta)
    tmid = 0.5 * (th[:-1] + th[1:])
    tv = krho * m.R_av[1:, None] * tmid[None, :]
    np.testing.assert_allclose(10 ** X[..., 3] - 1e-10, tv, rtol=1e-9)


def test_feature_columns(small):
    m = small
    X = warmstart.features(m).reshape(m.nr - 1, m.nt, 8)
    np.testing.assert_allclose(X[:, 0, 0], np.log10(m.R_av[1:] / warmstart.AU))
    np.testing.assert_allclose(X[0, :, 1], m.theta_av)
    assert np.all(X[..., 4] == m.Tstar / 1000.0)
    assert np.allclose(X[..., 5], math.log10(m.Lstar / warmstart.LSUN))
    assert np.all(X[..., 6] == 1.125)
    assert np.allclose(X[..., 7], -4.0, atol=1e-3)
    # the optical depths grow outwards and towards the midplane (near the
    # pole the density underflows and the depths stay at the floor)
    assert np.all(np.diff(X[:, :, 2], axis=0) >= 0) and np.all(np.diff(X[:, -1, 2]) > 0)
    assert np.all(np.diff(X[:, :, 3], axis=1) >= 0) and np.all(np.diff(X[-1, :, 3])[-20:] > 0)


def test_field_from_log_t(small):
    m = small
    y = np.linspace(1.0, 3.0, (m.nr - 1) * m.nt)
    T = warmstart.field_from_logT(m, y)
    assert T.shape == (m.nr, m.nt)
    np.testing.assert_allclose(np.log10(T[1:]).ravel(), y)
    assert warmstart.median_rel(T, T) == 0.0
    assert abs(warmstart.median_rel(1.1 * T, T) - 0.1) < 1e-12
