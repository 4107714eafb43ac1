"""Pins of the oracle (oracle/mcx_oracle.c) against known answers, closed
forms and invariants -- what makes it trustworthy as the checker.  The
reference ships no tests or golden vectors for this path, so every pin here
is a published known answer, a closed form or a brute-force computation.
"""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle

pyoracle.build()
L = pyoracle.load()


def arr(n):
    return np.zeros(n)


# ---------------------------------------------------------------- RNG ------
def test_ran1_park_miller_check_value():
    """Park & Miller (1988): the minimal-standard LCG from seed 1 reaches
    1043618065 at step 10000.  ran1 warms up 40 LCG steps at initialisation
    (NTAB+8) and takes one step per call."""
    s = pyoracle.Ran1State(idum=-1, iy=0)
    for _ in range(10000 - 40):
        L.orc_ran1(C.byref(s))
    assert s.idum == 1043618065


def test_ran1_range_and_mean():
    s = pyoracle.Ran1State(idum=-42, iy=0)
    x = np.array([L.orc_ran1(C.byref(s)) for _ in range(200000)])
    assert x.min() > 0.0 and x.max() < 1.0
    assert abs(x.mean() - 0.5) < 0.003
    assert abs(x.var() - 1 / 12) < 0.002


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox4x32_10_known_answers(ctr, key, expect):
    """Random123 known-answer vectors for philox4x32-10 (Salmon et al. 2011)."""
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox4x32_10(c, k, o)
    assert tuple(o) == expect


def test_philox_uniform_layout():
    """Draw k of stream (seed, pid) is half k%2 of block k/2 (top 53 bits)."""
    seed, pid = 0x123456789abcdef, 77
    for k in range(6):
        blk = k // 2
        c = (C.c_uint32 * 4)(blk & 0xffffffff, blk >> 32, pid & 0xffffffff, pid >> 32)
        kk = (C.c_uint32 * 2)(seed & 0xffffffff, seed >> 32)
        o = (C.c_uint32 * 4)()
        L.orc_philox4x32_10(c, kk, o)
        v = (o[1] << 32 | o[0]) if k % 2 == 0 else (o[3] << 32 | o[2])
        assert L.orc_philox_uniform(seed, pid, k) == ((v >> 11) + 0.5) * 2.0 ** -53


# ------------------------------------------------------- Planck tables ----
def test_planck_integral_is_sigma_T4_over_pi():
    """Trapezoid of B_nu over a fine grid -> sigma T^4 / pi with the Planck
    routine's own constants."""
    k, c, h = 1.3807e-16, 2.9979e10, 6.6261e-27
    sig = 2 * math.pi ** 5 * k ** 4 / (15 * h ** 3 * c ** 2)
    n = 4000
    lam, nu = arr(n), arr(n)
    L.orc_lam_grid(n, 0.01, 1e5, pyoracle.p(lam), pyoracle.p(nu))
    for T in (30.0, 300.0, 3000.0, 9000.0):
        B = np.array([L.orc_planck(T, l) for l in lam])
        I = L.orc_integrate(pyoracle.p(nu), pyoracle.p(B), n)
        assert abs(I / (sig * T ** 4 / math.pi) - 1) < 2e-4, T


def test_planck_wien_and_rayleigh_jeans_limits():
    k, c, h = 1.3807e-16, 2.9979e10, 6.6261e-27
    T = 1000.0
    lam = 1e4  # micron: h nu / k T = 1.4e-3 -> Rayleigh-Jeans
    nu = c / (lam * 1e-4)
    assert abs(L.orc_planck(T, lam) / (2 * nu ** 2 * k * T / c ** 2) - 1) < 1e-3
    lam = 0.5  # h nu / kT = 28.8 -> Wien
    nu = c / (lam * 1e-4)
    wien = 2 * h * nu ** 3 / c ** 2 * math.exp(-h * nu / (k * T))
    assert abs(L.orc_planck(T, lam) / wien - 1) < 1e-11
    assert L.orc_planck(1.0, 0.1) == 0.0  # x > 1e3 cut-off


def test_star_luminosity_check():
    """Init.f:2633: the trapezoid luminosity of a Planck star agrees with the
    closed form sigma T^4 R^2 (the reference's L/4pi convention)."""
    n = 3000
    lam, nu = arr(n), arr(n)
    L.orc_lam_grid(n, 0.01, 1e5, pyoracle.p(lam), pyoracle.p(nu))
    F, Ls = arr(n), C.c_double()
    R = 2.0 * L.orc_const(3)
    L.orc_star(n, pyoracle.p(lam), pyoracle.p(nu), R, 4000.0, pyoracle.p(F), C.byref(Ls))
    assert abs(Ls.value / L.orc_luminosity(4000.0, R) - 1) < 1e-3


def test_kp_is_linear_in_kappa():
    n = 60
    lam, nu = arr(n), arr(n)
    L.orc_lam_grid(n, 0.1, 1000.0, pyoracle.p(lam), pyoracle.p(nu))
    BB = np.zeros((10001, n))
    BBint = arr(10001)
    L.orc_bb_table(n, pyoracle.p(lam), pyoracle.p(nu), pyoracle.p(BB), pyoracle.p(BBint))
    K = np.full(n, 3.0)
    Kp = arr(10001)
    L.orc_kp_table(n, pyoracle.p(nu), pyoracle.p(BB), pyoracle.p(K), pyoracle.p(Kp))
    assert np.allclose(Kp, 3.0 * BBint, rtol=1e-14, atol=0)


# ---------------------------------------------------- phase function -----
def test_hg_bin_average_against_quadrature():
    from scipy.integrate import quad
    g = 0.5
    hg = lambda t: (1 - g * g) / (1 - 2 * g * math.cos(t) + g * g) ** 1.5 / 2
    for a in (1, 10, 45, 90, 170, 180):
        th = (a - 0.5) * math.pi / 180
        lo, hi = th - 0.5 * math.pi / 180, th + 0.5 * math.pi / 180
        exact = quad(lambda t: hg(t) * math.sin(t), lo, hi)[0] / quad(math.sin, lo, hi)[0]
        assert abs(L.orc_hg(g, th) / exact - 1) < 2e-5, a


def test_mueller_normalisation_and_asymmetry():
    nl = 3
    F = np.zeros((nl, 6, 180))
    th = (np.arange(1, 181) - 0.5) * math.pi / 180
    for l in range(nl):
        F[l, 0] = [L.orc_hg(0.5, t) * 7.0 for t in th]
        F[l, 1] = -0.3 * F[l, 0]
    L.orc_normalize_mueller(nl, pyoracle.p(F))
    s = np.sin(th)
    assert np.allclose((F[:, 0] * s).sum(axis=1), s.sum(), rtol=1e-14)
    gmean = (F[0, 0] * np.cos(th) * s).sum() / (F[0, 0] * s).sum()
    assert abs(gmean - 0.5) < 2e-3
    IF11, IF12 = arr(nl), arr(nl)
    L.orc_mueller_integrals(nl, pyoracle.p(F), pyoracle.p(IF11), pyoracle.p(IF12))
    # normalised F11 integrates to sum sin * pi/180 ~ 2 (int sin = 2)
    assert np.allclose(IF11, math.pi / 180 * s.sum(), rtol=1e-13)
    assert np.allclose(IF12, -0.3 * IF11, rtol=1e-12)


# --------------------------------------------------------------- grids ---
def test_theta_grid_properties():
    for nt in (10, 60, 100):
        Th, tav = arr(nt + 1), arr(nt)
        assert L.orc_theta_grid(nt, pyoracle.p(Th), pyoracle.p(tav)) == nt
        assert Th[0] == 1.0 and Th[-1] == 0.0
        assert np.all(np.diff(Th) < 0)
        assert np.all((np.cos(tav) < Th[:-1]) & (np.cos(tav) > Th[1:]))
        # midpoint rule in cos theta for the centres
        assert np.allclose(np.cos(tav), (Th[:-1] + Th[1:]) / 2, rtol=1e-14)


def test_r_grid_is_logarithmic():
    R, Rav = arr(102), arr(101)
    L.orc_r_grid(100, 0.1, 100.0, 2.0 * L.orc_const(3), pyoracle.p(R), pyoracle.p(Rav))
    assert R[1] == pytest.approx(0.1, rel=1e-15) and R[-1] == pytest.approx(100.0, rel=1e-14)
    q = R[2:] / R[1:-1]
    assert np.allclose(q, 1000.0 ** (1 / 100), rtol=1e-12)
    assert np.allclose(Rav[1:], L.orc_const(1) * np.sqrt(R[1:-1] * R[2:]), rtol=1e-13)


def test_min_density_mass_and_profile():
    nrad, nt = 40, 20
    R, Rav = arr(nrad + 2), arr(nrad + 1)
    AU = L.orc_const(1)
    L.orc_r_grid(nrad, 0.1, 100.0, 2.0 * L.orc_const(3), pyoracle.p(R), pyoracle.p(Rav))
    Th, tav = arr(nt + 1), arr(nt)
    L.orc_theta_grid(nt, pyoracle.p(Th), pyoracle.p(tav))
    nr = nrad + 1
    dens, V = np.zeros((nr, nt)), np.zeros((nr, nt))
    M = 1e-4 * L.orc_const(4)
    L.orc_min_density(nr, nt, pyoracle.p(R), pyoracle.p(Th), pyoracle.p(Rav), pyoracle.p(tav), M, 0.1, 100.0,
                      0.05, 1.125, 1.0, pyoracle.p(dens), pyoracle.p(V))
    assert abs((dens[1:] * V[1:]).sum() / M - 1) < 1e-12
    # closed-form shape: rho ~ r^-(p+shpow) exp(-(z/h)^2), h = 0.05 r^1.125
    r = Rav[1:, None] * np.sin(tav)[None, :] / AU
    z = Rav[1:, None] * np.cos(tav)[None, :] / AU
    shape = r ** (-1.0 - 1.125) * np.exp(-(z / (0.05 * r ** 1.125)) ** 2)
    ok = shape > 1e-250
    ratio = dens[1:][ok] / shape[ok]
    assert np.allclose(ratio, ratio[0], rtol=1e-10)
    assert np.all(dens[1:][~ok] < 1e-200)
    # shell volumes add up to the sphere volume between R[0] and R[-1]
    Vtot = 4 * math.pi / 3 * (R[-1] ** 3 - R[0] ** 3) * AU ** 3
    assert abs(V.sum() / Vtot - 1) < 1e-12


# ---------------------------------------------------- T inversion --------
def _grey_model(kappa=5.0, nlam=200):
    """A one-cell model with constant kappa on a wide grid (for determineT)."""
    lam, nu = arr(nlam), arr(nlam)
    L.orc_lam_grid(nlam, 0.01, 1e5, pyoracle.p(lam), pyoracle.p(nu))
    BB = np.zeros((10001, nlam))
    BBint = arr(10001)
    L.orc_bb_table(nlam, pyoracle.p(lam), pyoracle.p(nu), pyoracle.p(BB), pyoracle.p(BBint))
    Kp = arr(10001)
    K = np.full(nlam, kappa)
    L.orc_kp_table(nlam, pyoracle.p(nu), pyoracle.p(BB), pyoracle.p(K), pyoracle.p(Kp))

    class M:
        pass

    m = M()
    m.keep = dict(Kp=Kp, w=np.ones(4))
    m.om = pyoracle.OrcModel(ng=1, Kp=pyoracle.p(Kp), w=pyoracle.p(m.keep["w"]), nr=2, nt=1)
    return m, Kp


def test_determineT_inverts_the_planck_mean():
    m, Kp = _grey_model()
    rng = np.random.default_rng(1)
    for E in 10 ** rng.uniform(np.log10(Kp[2]), np.log10(Kp[9990]), 300):
        for guess in (0.0, 50.0, 9000.0):
            T = L.orc_determineT(C.byref(m.om), 0, float(E), guess)
            i = int(T)
            assert Kp[i] <= E * (1 + 1e-12) and E <= Kp[i + 1] * (1 + 1e-12)
            # T^4-linear interpolation within the bracket
            f = (T ** 4 - i ** 4) / ((i + 1) ** 4 - i ** 4)
            assert abs(Kp[i] + f * (Kp[i + 1] - Kp[i]) - E) <= 1e-6 * E


def test_determineT_grey_closed_form():
    """For constant kappa, Kp(T) = kappa sigma T^4 / pi, so T(E) is closed form."""
    k, c, h = 1.3807e-16, 2.9979e10, 6.6261e-27
    sig = 2 * math.pi ** 5 * k ** 4 / (15 * h ** 3 * c ** 2)
    m, Kp = _grey_model(kappa=5.0)
    for T0 in (20.0, 150.0, 1234.5, 5000.0):
        E = 5.0 * sig * T0 ** 4 / math.pi
        T = L.orc_determineT(C.byref(m.om), 0, E, 100.0)
        assert abs(T / T0 - 1) < 5e-4, (T0, T)


def test_determineT_limits():
    m, Kp = _grey_model()
    assert L.orc_determineT(C.byref(m.om), 0, 0.5 * Kp[1], 10.0) == 1.0
    assert L.orc_determineT(C.byref(m.om), 0, 2 * Kp[10000], 10.0) == 9999.0
    assert L.orc_determineT(C.byref(m.om), 0, float("nan"), 10.0) == 3.0
