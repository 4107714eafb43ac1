"""Pins of the oracle's SED ray tracing (TraceMono.f TracePath/TraceFlux):
closed forms for the stellar disk and for optically thin thermal emission."""
import numpy as np
import pytest

import pyoracle
from mcmax_amd import configs


@pytest.fixture(scope="module")
def naked(workdir):
    return configs.load_config("tiny", workdir + "/naked", mdust=1e-14, scattype="NONE")


def test_image_radii_cover_star_and_disk(naked):
    o = pyoracle.Oracle(naked)
    obs = pyoracle.make_obs(nphot=0, nphotstar=0)
    R, R01 = o.image_radii(obs)
    assert np.all(np.diff(R) > 0)
    assert R[0] == 0.0 and R[-1] == pytest.approx(naked.R[-1] * float(np.float32(0.9999)))
    # nstar-1 points inside the stellar disk, plus projected rim radii near the pole
    assert np.sum(R < naked.R[0]) >= 29
    assert naked.R[1] <= R01 <= naked.R[-2]


def test_naked_star_flux_is_fstar(naked):
    """Rays that hit the star carry Fstar/(pi R*^2); the image quadrature
    over the stellar disk returns Fstar (TraceMono.f:486-488, 499-520)."""
    o = pyoracle.Oracle(naked)
    obs = pyoracle.make_obs(nphot=0, nphotstar=0)
    T = np.full((naked.nr, naked.nt), 10.0)
    s = o.sed(T, obs, naked.lam[::3])
    assert np.allclose(s[:, 0], naked.Fstar[::3], rtol=5e-6)
    assert np.all(s[:, 1] == 0.0)


def test_thin_thermal_emission_closed_form(workdir):
    """An optically thin, non-scattering disk adds sum_c B(T_c) kappa_c rho_c V_c
    to the stellar flux (the integral of the source function over the image
    plane), independent of inclination."""
    m = configs.load_config("tiny", workdir + "/thinsed", mdust=1e-12, scattype="NONE", nrad=30, ntheta=16)
    o = pyoracle.Oracle(m)
    T = np.empty((m.nr, m.nt))
    T[:] = (300.0 * (m.R_av / m.R_av[1]) ** -0.5)[:, None]
    lams = m.lam[10:36:5]
    for inc in (10.0, 60.0):
        obs = pyoracle.make_obs(angle_deg=inc, nphot=0, nphotstar=0, nj0=6)
        s = o.sed(T, obs, lams)
        for k, lam0 in enumerate(lams):
            l = int(np.argmin(np.abs(m.lam - lam0)))
            assert m.lam[l] == lam0
            B = np.array([[m.BB[int(T[i, j]), l] * (int(T[i, j]) + 1 - T[i, j]) +
                           m.BB[int(T[i, j]) + 1, l] * (T[i, j] - int(T[i, j])) for j in range(m.nt)]
                          for i in range(m.nr)])
            dust = np.sum(B[1:] * m.Kabs[0, l] * m.dens[1:] * m.V[1:])
            expect = m.Fstar[l] + dust
            assert abs(s[k, 0] / expect - 1) < 0.03, (inc, lam0, s[k, 0] / expect)
