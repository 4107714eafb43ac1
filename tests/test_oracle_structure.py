"""Pins of the oracle's vertical structure (TempAverage, DiskStructure,
derivs, compareCells: DiskStructure.f:25-584; odeint/rkqs/rkck:
Subroutines.f:1933-2085).

The checks are fixed by the physics, not by the oracle's formulas:
isothermal hydrostatic equilibrium is a Gaussian in z (the integrand is
linear in z, so the Cash-Karp step is exact); a temperature linear in z
has a closed-form column (log terms) that an error in the tableau or the
step control would miss; the renormalisation conserves every column's mass;
a wedge is a step function; compareCells of shifted states is a sum of
known chi-square terms."""
import types

import numpy as np
import pytest

import pyoracle
from mcmax_amd import configs

MU = float(np.float32(2.3)) * 1.67262158e-24
G = 6.67300e-8
KB = 1.3806503e-16


def _a(scale, Mstar, R):
    """mu G M* / (scale^2 kb R^3): dy/dz = -a z / T - dlnT/dz"""
    return MU * G * Mstar / (scale ** 2 * KB * R ** 3)


def _column_linear_T(z1, z2, T1, T2, a):
    """y(z2) - y(z1) for T linear in z: ln(T1/T2) - a int z/T dz"""
    s = (T2 - T1) / (z2 - z1)
    if s == 0.0:
        return -a * (z2 * z2 - z1 * z1) / (2.0 * T1)
    integral = ((z1 - T1 / s) * np.log(T2 / T1) + (T2 - T1) / s) / s
    return np.log(T1 / T2) - a * integral


def _f32(a, b):
    return float(np.float32(a) / np.float32(b))


def test_isothermal_step_follows_the_single_precision_tableau():
    """For an integrand linear in z a Runge-Kutta step adds h (W x + S h) times
    the slope, W = sum C_i and S = sum C_i A_i; with exact Cash-Karp weights
    W = 1 and S = 1/2 (exact result).  The reference declares them as default
    REAL constants (Subroutines.f:2052-2060), so W - 1 = -1.5e-8 and
    S - 1/2 = +3.9e-9: the column's ratio to the exact Gaussian lies between
    the per-step ratios (W x + S h) / (x + h/2) at x/h >= z1/(z2 - z1)."""
    C1, C3, C4, C6 = _f32(37, 378), _f32(250, 621), _f32(125, 594), _f32(512, 1771)
    A3, A4, A6 = float(np.float32(0.3)), float(np.float32(0.6)), 0.875
    W, S = C1 + C3 + C4 + C6, C3 * A3 + C4 * A4 + C6 * A6
    assert W - 1.0 < -1e-8
    R, M = 1.5e13, 2e33
    for z1, z2, T in ((1e11, 4e11, 80.0), (0.0, 2e12, 300.0), (3e12, 9e12, 20.0)):
        y, n = pyoracle.vertical_step(0.0, z1, z2, T, T, 1.0, M, R)
        ratio = y / _column_linear_T(z1, z2, T, T, _a(1.0, M, R))
        q = z1 / (z2 - z1)
        lo, hi = sorted((W, (W * q + S) / (q + 0.5)))
        assert lo - 1e-15 <= ratio <= hi + 1e-15, (ratio - 1, lo - 1, hi - 1)
        assert 0 < n < 400


@pytest.mark.parametrize("T1,T2", [(300.0, 30.0), (50.0, 400.0), (1000.0, 999.0)])
def test_linear_temperature_step_closed_form(T1, T2):
    """Within the integrator's tolerance (eps = 1e-6 of |y| + |h dy/dz| per
    step) of the closed form; dropping the dlnT term or the log terms of the
    gravity integral misses by more than 1e-2."""
    R, M = 3e13, 1e33
    a = _a(1.3, M, R)
    z1, z2 = 2e11, 3.5e12
    for y0 in (0.0, -17.5):
        y, n = pyoracle.vertical_step(y0, z1, z2, T1, T2, 1.3, M, R)
        exact = y0 + _column_linear_T(z1, z2, T1, T2, a)
        assert abs(y - exact) < 5e-6 * (abs(y0) + abs(exact - y0) + 0.1), (y, exact)
        assert 0 < n < 60


def test_radiation_pressure_term():
    """beta constant: gravity scales by (1 + beta); the step control is
    relative, so the steps are the same and the column scales exactly"""
    R, M, T = 1.5e13, 2e33, 100.0
    y, n = pyoracle.vertical_step(0.0, 1e11, 1e12, T, T, 1.0, M, R, F1=0.5, F2=0.5)
    y0, n0 = pyoracle.vertical_step(0.0, 1e11, 1e12, T, T, 1.0, M, R)
    assert n == n0 and y == pytest.approx(1.5 * y0, rel=1e-13)
    assert y == pytest.approx(1.5 * _column_linear_T(1e11, 1e12, T, T, _a(1.0, M, R)), rel=3e-8)


def _grid(nr=12, nt=10, ng=2, seed=0):
    m = configs.load_config("tiny", "/tmp/orc_struct_cfg", nrad=nr - 1, ntheta=nt)
    rng = np.random.default_rng(seed)
    q = types.SimpleNamespace(nr=m.nr, nt=m.nt, ng=ng, R_av=m.R_av, theta_av=m.theta_av, V=m.V,
                              Mstar_g=m.Mstar_g)
    w = rng.uniform(0.2, 1.0, size=(m.nr, m.nt, ng))
    w /= w.sum(axis=2, keepdims=True)
    state = dict(dens=m.dens.copy(), dens0=m.dens.copy(), gasdens=m.dens.copy(), w=w, w0=w.copy())
    return q, state


def _gauss_expect(q, i, H2, M0):
    """renormalised Gaussian column and its log-profile y (y = 0 at the
    midplane cell)"""
    z = q.R_av[i] * np.cos(q.theta_av)
    y = -(z * z - z[-1] ** 2) / (2.0 * H2)
    e = np.exp(y)
    return e * M0 / np.sum(e * q.V[i]), y


def _close_in_log(got, expect, y):
    """the single-precision Runge-Kutta weights leave |dy| < 2e-8 |y|
    (the tableau test above) on top of the renormalisation's common factor,
    itself within 2e-8 max|y|; cells below the 1e-50 floor are floored"""
    floor = expect < 1e-50
    assert np.all(got[floor] == 1e-50)
    assert not floor[-1]
    d = np.log(got[~floor]) - np.log(expect[~floor])
    off = d[-1]
    assert abs(off) <= 2e-8 * np.max(np.abs(y)) + 1e-12
    assert np.all(np.abs(d - off) <= 2e-8 * np.abs(y[~floor]) + 1e-12), d - off


def test_isothermal_disk_is_gaussian():
    q, state = _grid()
    T = np.full((q.nr, q.nt), 150.0)
    sh = np.linspace(0.8, 1.4, q.nr)
    new, rho, nsteps = pyoracle.disk_structure(q, state, T, shscale=sh)
    assert nsteps > 0
    for i in range(1, q.nr):
        H2 = KB * T[i, 0] * q.R_av[i] ** 3 * sh[i] ** 2 / (MU * G * q.Mstar_g)
        Mgas = np.sum(state["gasdens"][i] * q.V[i])
        _close_in_log(rho[0, i], *_gauss_expect(q, i, H2, Mgas))
        for s in range(q.ng):
            Ms = np.sum(state["dens0"][i] * q.V[i] * state["w0"][i, :, s])
            _close_in_log(rho[s + 1, i], *_gauss_expect(q, i, H2, Ms))


def test_mass_conservation_and_composition():
    q, state = _grid(seed=4)
    rng = np.random.default_rng(1)
    T = 40.0 + 400.0 * rng.uniform(size=(q.nr, q.nt))
    gsh = np.array([np.full(q.nr, 0.5), np.full(q.nr, 1.0)])
    new, rho, _ = pyoracle.disk_structure(q, state, T, gshscale=gsh, settle=[1, 0])
    V = q.V
    nstrict = 0
    for i in range(1, q.nr):
        # every column keeps its mass (cells floored at 1e-50 aside)
        assert np.sum(new["gasdens"][i] * V[i]) == pytest.approx(np.sum(state["gasdens"][i] * V[i]), rel=1e-12)
        for s in range(q.ng):
            before = np.sum(state["dens0"][i] * V[i] * state["w0"][i, :, s])
            after = np.sum(new["dens0"][i] * V[i] * new["w0"][i, :, s])
            assert after == pytest.approx(before, rel=1e-12)
        # the settled species (half the scale height) is more concentrated;
        # where the column fits in the midplane cell both are that cell
        z = q.R_av[i] * np.cos(q.theta_av)
        c0 = np.sum(rho[1, i] * V[i] * z * z) / np.sum(rho[1, i] * V[i])
        c1 = np.sum(rho[2, i] * V[i] * z * z) / np.sum(rho[2, i] * V[i])
        assert c0 <= c1 * (1.0 + 1e-12)
        nstrict += c0 < c1 * (1.0 - 1e-3)
    assert nstrict >= (q.nr - 1) // 2
    assert np.allclose(new["w0"][1:].sum(axis=2), 1.0, rtol=1e-14)
    assert np.allclose(new["w"][1:].sum(axis=2), 1.0, rtol=1e-14)
    # no dust was lost to the gas phase: dens follows dens0
    assert np.allclose(new["dens"], new["dens0"], rtol=1e-14)
    # the void row is left alone
    for k in ("dens", "dens0", "gasdens", "w", "w0"):
        assert np.array_equal(new[k][0], state[k][0])


def test_gas_phase_fraction_is_kept():
    """Half of species 1 in the gas phase (dens*w = dens0*w0/2) stays so."""
    q, state = _grid(ng=1, seed=2)
    state["dens"] = state["dens0"] * 0.5
    T = np.full((q.nr, q.nt), 100.0)
    new, _, _ = pyoracle.disk_structure(q, state, T)
    sel = new["dens0"][1:] > 1e-40
    assert np.allclose(new["dens"][1:][sel], 0.5 * new["dens0"][1:][sel], rtol=1e-13)


def test_midplane_temperature_option():
    q, state = _grid(ng=1)
    T = np.tile(np.linspace(500.0, 60.0, q.nt), (q.nr, 1))
    a, _, _ = pyoracle.disk_structure(q, state, T, mpstr=True)
    b, _, _ = pyoracle.disk_structure(q, state, np.repeat(T[:, -1:], q.nt, axis=1))
    for k in a:
        assert np.array_equal(a[k], b[k])


def test_wedge():
    q, state = _grid(ng=1)
    T = np.full((q.nr, q.nt), 100.0)
    sh = np.full(q.nr, -0.3)  # wedge of half-opening z/r = 0.3
    new, rho, n = pyoracle.disk_structure(q, state, T, shscale=sh)
    assert n == 0
    inside = 1.0 / np.tan(q.theta_av) < 0.3
    inside[-1] = True
    for i in range(1, q.nr):
        e = np.where(inside, 1.0, np.exp(-60.0))
        M0 = np.sum(state["gasdens"][i] * q.V[i])
        assert np.allclose(rho[0, i], np.maximum(e * M0 / np.sum(e * q.V[i]), 1e-50), rtol=1e-13)


def test_floor_at_1e_50():
    """a cold, thin column puts the upper cells below the 1e-50 floor"""
    q, state = _grid(ng=1)
    T = np.full((q.nr, q.nt), 3.0)
    new, rho, _ = pyoracle.disk_structure(q, state, T, shscale=np.full(q.nr, 0.2))
    assert np.min(rho[:, 1:]) == 1e-50
    assert np.min(new["dens"][1:]) >= 1e-50 and np.min(new["gasdens"][1:]) >= 1e-50


def test_temp_average():
    T = np.array([1.0, 2.0, 2.7, 100.0])
    first = pyoracle.temp_average(T, None, 0.5, first=True)
    assert np.array_equal(first, T)
    t27 = float(np.float32(2.7))
    one = pyoracle.temp_average(T, first, 1.0)
    assert np.array_equal(one, [t27, t27, t27, 100.0])
    half = pyoracle.temp_average(np.full(4, 200.0), first, 0.5)
    assert np.allclose(half, 0.5 * first + 100.0, rtol=1e-15)


def test_compare_cells():
    nr, nt = 6, 5
    rng = np.random.default_rng(5)
    a = {k: rng.uniform(1.0, 2.0, size=(nr, nt)) for k in ("T", "dT", "dens", "ddens", "EJv", "dEJv")}
    assert pyoracle.compare_cells(nr, nt, a, a) == (0.0, 0.0, 0.0)
    b = dict(a)
    k = 3.0
    b["T"] = a["T"] + k * a["dT"] * np.sqrt(2.0)
    b["EJv"] = a["EJv"] + 2.0 * a["dEJv"] * np.sqrt(2.0)
    b["dens"] = a["dens"] + a["ddens"] * np.sqrt(2.0)
    N = (nr - 1) * nt
    eT, eR, eE = pyoracle.compare_cells(nr, nt, a, b)
    assert eT == pytest.approx(k * k * N / (N - 1), rel=1e-12)
    assert eE == pytest.approx(4.0 * N / (N - 1), rel=1e-12)
    assert eR == pytest.approx(np.sqrt(N / (N - 1)), rel=1e-12)
    # cells without an error estimate do not count; the void row never does
    c = dict(b)
    c["dT"] = a["dT"].copy()
    c["dT"][2, :] = 0.0
    c["T"] = a["T"].copy()
    c["T"][0, :] += 100.0
    eT2, _, _ = pyoracle.compare_cells(nr, nt, a, c)
    assert eT2 == 0.0
