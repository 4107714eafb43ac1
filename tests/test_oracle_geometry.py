"""Cell-crossing geometry of the oracle (Trace2edge/hitR/hitT/hitTsame,
MCRadiation.f:1150-1355) against closed forms and brute force: every traced
segment must lie in the cell it is attributed to (the reference's checkbox,
MCRadiation.f:1361-1387), at every one of many sample points along the ray,
and the segments must add up to the chord to the outer sphere."""
import numpy as np
import pytest

import pyoracle


@pytest.fixture(scope="module")
def orc(small_model):
    return pyoracle.Oracle(small_model)


def locate(m, x, y, z):
    """Brute-force cell of a point: radial and |cos theta| bins."""
    r = np.sqrt(x * x + y * y + z * z)
    i = np.searchsorted(m.R, r) - 1
    ct = np.abs(z) / r
    jj = np.searchsorted(-m.Theta, -ct) - 1
    return i, jj


def test_midplane_radial_ray_closed_form(orc, small_model):
    m = small_model
    z0 = 1e-7 * m.R[0]
    cells, lens, code = orc.trace_ray(m.R[0], 0.0, z0, 1.0, 0.0, 0.0, 0, m.nt - 1, onEdge=1, edgeNr=1)
    assert code == 1  # escaped through the outer sphere
    assert list(cells) == [i * m.nt + (m.nt - 1) for i in range(m.nr)]
    R = np.sqrt(m.R ** 2 - z0 ** 2)
    assert np.allclose(lens, np.diff(R), rtol=1e-12, atol=0)


@pytest.mark.parametrize("seed", range(6))
def test_random_rays_brute_force(orc, small_model, seed):
    m = small_model
    rng = np.random.default_rng(seed)
    for _ in range(40):
        # a random interior point of a random cell, a random direction
        i = int(rng.integers(1, m.nr))
        jj = int(rng.integers(0, m.nt))
        r = m.R[i] + (m.R[i + 1] - m.R[i]) * rng.uniform(0.05, 0.95)
        ct = m.Theta[jj + 1] + (m.Theta[jj] - m.Theta[jj + 1]) * rng.uniform(0.05, 0.95)
        ct *= rng.choice([-1.0, 1.0])
        ph = rng.uniform(0, 2 * np.pi)
        st = np.sqrt(1 - ct * ct)
        p = np.array([r * st * np.cos(ph), r * st * np.sin(ph), r * ct])
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        cells, lens, code = orc.trace_ray(*p, *v, i, jj, onEdge=0, edgeNr=0)
        assert code in (1, 2), code
        # total length: to the outer sphere (code 1) or the star (code 2)
        b = p @ v
        Rend = m.R[-1] if code == 1 else m.R[0]
        disc = b * b - (p @ p - Rend ** 2)
        t_end = -b + np.sqrt(disc) if code == 1 else -b - np.sqrt(disc)
        assert abs(lens.sum() / t_end - 1) < 1e-9
        # every sample point between segment ends lies in the traced cell
        s = np.concatenate([[0.0], np.cumsum(lens)])
        for k, (a, e) in enumerate(zip(s[:-1], s[1:])):
            if e - a < 1e-9 * t_end:
                continue
            ts = a + (e - a) * np.linspace(0.02, 0.98, 9)
            q = p[None, :] + ts[:, None] * v[None, :]
            ii, jjs = locate(m, q[:, 0], q[:, 1], q[:, 2])
            assert np.all(ii * m.nt + jjs == cells[k]), (k, cells[k], ii, jjs)


def test_ray_through_the_pole_axis(orc, small_model):
    """A ray crossing the z axis (the degenerate cone cos^2 = 1) stays in the
    polar cell row j = 0 with the reference's jnext = 0 -> 1 rule."""
    m = small_model
    r = 0.5 * (m.R[50] + m.R[51])
    ct = 0.5 * (m.Theta[0] + m.Theta[1])
    st = np.sqrt(1 - ct * ct)
    p = np.array([r * st, 0.0, r * ct])
    v = np.array([-1.0, 0.0, 0.0])
    cells, lens, code = orc.trace_ray(*p, *v, 50, 0, onEdge=0, edgeNr=0)
    assert code == 1
    s = np.concatenate([[0.0], np.cumsum(lens)])
    for k in range(len(lens)):
        t = 0.5 * (s[k] + s[k + 1])
        q = p + t * v
        i, jj = locate(m, *q)
        assert i * m.nt + jj == cells[k]
