"""Pins of the oracle's samplers against distributions computed here, and
of Lucy's MC error estimate against the seed-to-seed spread.

- scatangle (Radiation.f:381-495): the scattering angle's 1-degree bin is
  drawn with probability proportional to sin(theta_i) (E F11_i + Q F12_i)
  (Radiation.f:436-452) and theta is uniform within its bin; for the
  Henyey-Greenstein grain <cos theta> is g;
- the Bjorkman-Wood reemission (Radiation.f:187-213 with emit,
  Radiation.f:940-1001): the wavelength bin [l, l+1] is drawn with
  probability proportional to the trapezoid 0.5 |nu_l - nu_l+1| (S_l +
  S_l+1) of the difference spectrum S = (B(T1) - B(T0)) KabsTot, the weight
  wl1 uniform within the bin; for a small step the spectrum follows
  dB_nu/dT kappa of the Planck function;
- dT (Main.f:515-528): of the order of the seed-to-seed standard deviation
  of T in well-sampled cells.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))

H, KB, CL = 6.62607015e-27, 1.380649e-16, 2.99792458e10


def chi2_ok(counts, p, nsig=5.0):
    """chi-square of counts against probabilities p (bins with expected
    counts >= 20), within nsig standard deviations of its mean."""
    n = counts.sum()
    e = n * p
    sel = e >= 20
    x2 = np.sum((counts[sel] - e[sel]) ** 2 / e[sel])
    dof = sel.sum() - 1
    return x2, dof, x2 < dof + nsig * np.sqrt(2 * dof)


@pytest.fixture(scope="module")
def hg(workdir):
    import pyoracle
    from mcmax_amd import configs
    m = configs.load_config("tiny", workdir + "/pin_hg")
    return m, pyoracle.Oracle(m)


@pytest.fixture(scope="module")
def dhs1(workdir):
    """one computepart DHS species (from the oracle's ComputePart): a
    forward-peaked F11 with a non-zero F12"""
    import pyoracle
    from make_staged_fixtures import OracleDust
    from mcmax_amd import configs
    ov = {"abun01": 1.0, "computepart01:file": "dhs.dat", "computepart01:amin": 0.05,
          "computepart01:amax": 1000.0, "computepart01:apow": 3.5, "computepart01:fmax": 0.8,
          "computepart01:ngrains": 1, "computepart01:blend": False}
    m = configs.load_config("tiny", workdir + "/pin_dhs1", tr=OracleDust(), **ov)
    return m, pyoracle.Oracle(m)


def _theta_bins(cost):
    th = np.degrees(np.arccos(np.clip(cost, -1.0, 1.0)))
    return np.bincount(np.minimum(th.astype(int), 179), minlength=180), th


def test_scatangle_hg(hg):
    m, o = hg
    n = 200000
    for l in (5, 20):
        cost = o.sample_scatangle(0, l, n, seed=3 + l)
        th = (np.arange(180) + 0.5) * np.pi / 180
        p = np.sin(th) * m.F[0, l, 0]
        p /= p.sum()
        counts, _ = _theta_bins(cost)
        x2, dof, ok = chi2_ok(counts, p)
        assert ok, (l, x2, dof)
        # theta uniform inside its bin: the mean of cos over each bin
        edges = np.arange(181) * np.pi / 180
        Ec = np.sum(p * (np.sin(edges[1:]) - np.sin(edges[:-1])) / (np.pi / 180))
        assert abs(cost.mean() - Ec) < 5 * cost.std() / np.sqrt(n), (cost.mean(), Ec)
        # the binned HG keeps <cos theta> = g = 0.5 of the analytic grain
        assert abs(Ec - 0.5) < 2e-3


def test_scatangle_dhs_forward_peak(dhs1):
    """A DHS grain's forward-peaked F11 (unpolarised light)."""
    m, o = dhs1
    th = (np.arange(180) + 0.5) * np.pi / 180
    for l in (9, 30):
        F11 = m.F[0, l, 0]
        assert F11[0] > 20 * np.median(F11)
        counts, _ = _theta_bins(o.sample_scatangle(0, l, 300000, seed=5 + l))
        p = np.sin(th) * F11
        x2, dof, ok = chi2_ok(counts, p / p.sum())
        assert ok, (l, x2, dof)


def test_scatangle_polarised_azimuth(dhs1):
    """Polarised light (Q != 0): the azimuth's 1-degree bin k has weight
    IF11 E + IF12 (Q cos 2phi_k + U sin 2phi_k) (Radiation.f:386-404), the
    scattering plane tracking the polarisation; unpolarised, phi is
    uniform.  F12 enters through IF12 with its sign."""
    m, o = dhs1
    l = 9
    IF11, IF12 = m.IF11[0, l], m.IF12[0, l]
    assert abs(IF12) > 0.05 * IF11
    n = 300000
    for Q, pol in ((0.0, False), (0.7, True), (-0.7, True)):
        _, ph = o.sample_scatangle(0, l, n, E=1.0, Q=Q, pol=pol, seed=13, with_phi=True)
        phd = np.degrees(np.mod(ph, np.pi))
        counts = np.bincount(np.minimum(phd.astype(int), 179), minlength=180)
        w = IF11 + IF12 * Q * m.cos2phi[1:181]
        x2, dof, ok = chi2_ok(counts, w / w.sum())
        assert ok, (Q, x2, dof)
        if Q != 0.0:
            w2 = IF11 - IF12 * Q * m.cos2phi[1:181]
            assert not chi2_ok(counts, w2 / w2.sum())[2], Q


def _diff_spectrum(m, c, T0, T1):
    """(B(T1) - B(T0)) KabsTot with B linear between the integer-K rows of
    the Planck table; an absorption inside one 1 K step uses that step"""
    Ka = m.w.reshape(-1, m.ng)[c] @ m.Kabs
    B = lambda T: (T - int(T)) * m.BB[int(T) + 1] + (1 - (T - int(T))) * m.BB[int(T)]
    if int(T0) == int(T1):
        return (m.BB[int(T0) + 1] - m.BB[int(T0)]) * Ka
    return (B(T1) - B(T0)) * Ka


def test_reemission_wavelengths(hg):
    m, o = hg
    c = 5 * m.nt + 3
    n = 200000
    nu = m.nu
    for T0, T1 in ((300.2, 300.7), (299.6, 301.3), (40.5, 95.25)):
        il, wl = o.sample_reemission(c, T0, T1, n, seed=7)
        S = _diff_spectrum(m, c, T0, T1)
        p = 0.5 * np.abs(nu[:-1] - nu[1:]) * (S[:-1] + S[1:])
        p /= p.sum()
        counts = np.bincount(il - 1, minlength=m.nlam)[: m.nlam - 1]
        x2, dof, ok = chi2_ok(counts, p)
        assert ok, (T0, T1, x2, dof)
        # wl1 is uniform in the chosen bin (Kolmogorov-Smirnov, 1%)
        u = np.sort(wl)
        d = np.max(np.abs(u - (np.arange(1, n + 1) - 0.5) / n))
        assert d < 1.63 / np.sqrt(n), d


def test_reemission_follows_planck_derivative(hg):
    """A small step T -> T + 0.3 K inside one 1 K step: the reemitted
    spectrum is dB_nu/dT kappa of the Planck function at T (the table's
    one-kelvin difference differs by O(1/T))."""
    m, o = hg
    c = 7 * m.nt + 4
    T = 800.1
    n = 300000
    il, _ = o.sample_reemission(c, T, T + 0.3, n, seed=9)
    x = H * m.nu / (KB * T)
    dB = (2 * H * m.nu ** 3 / CL ** 2) * x / T * np.exp(x) / np.expm1(x) ** 2
    Ka = m.w.reshape(-1, m.ng)[c] @ m.Kabs
    S = dB * Ka
    p = 0.5 * np.abs(m.nu[:-1] - m.nu[1:]) * (S[:-1] + S[1:])
    p /= p.sum()
    counts = np.bincount(il - 1, minlength=m.nlam)[: m.nlam - 1]
    sel = n * p > 2000
    rel = counts[sel] / (n * p[sel]) - 1
    assert np.max(np.abs(rel)) < 0.05, rel


def test_lucy_error_matches_seed_spread(hg):
    """Twelve passes with different seeds: in cells every pass samples well
    the standard deviation of T across the seeds and Lucy's dT agree to a
    factor of two (a dropped sqrt(N) or E^2 for E would be orders off)."""
    m, o = hg
    Ts, dTs, Ni = [], [], []
    for s in range(12):
        r = o.mcradiation(20000, seed=100 + s)
        T, dT, _ = o.temperature(r)
        Ts.append(T[1:])
        dTs.append(dT[1:])
        Ni.append(r["Ni"][1:])
    Ts, dTs, Ni = np.array(Ts), np.array(dTs), np.array(Ni)
    sel = Ni.min(axis=0) > 200
    ratio = Ts.std(axis=0, ddof=1)[sel] / np.median(dTs, axis=0)[sel]
    assert sel.sum() > 50
    assert 0.5 < np.median(ratio) < 2.0, np.median(ratio)
