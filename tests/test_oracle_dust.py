"""Pins of the oracle's dust-opacity restatement (DMiLay, MeerhoffMie,
gauleg2, ComputePart) by independent means: a textbook Mie series built on
scipy's spherical Bessel functions, the coated sphere's Rayleigh limit,
Gauss-Legendre exactness, and the closed-form opacity of one solid
sphere."""
import math

import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn

TH = (np.arange(180) + 0.5) * np.pi / 180


def textbook_mie(m, x, mus):
    """Bohren & Huffman (4.53, 4.56, 4.74) with m = n + ik: Qext, Qsca, S1, S2."""
    nmax = int(x + 4.05 * x ** (1 / 3) + 10)
    n = np.arange(1, nmax + 1)
    jn, yn = spherical_jn(n, x), spherical_yn(n, x)
    jm1, ym1 = spherical_jn(n - 1, x), spherical_yn(n - 1, x)
    psi, dpsi = x * jn, x * jm1 - n * jn
    xi, dxi = x * (jn + 1j * yn), x * (jm1 + 1j * ym1) - n * (jn + 1j * yn)
    mx = m * x
    jmn, jmm1 = spherical_jn(n, mx), spherical_jn(n - 1, mx)
    psim, dpsim = mx * jmn, mx * jmm1 - n * jmn
    a = (m * psim * dpsi - psi * dpsim) / (m * psim * dxi - xi * dpsim)
    b = (psim * dpsi - m * psi * dpsim) / (psim * dxi - m * xi * dpsim)
    qext = 2 / x ** 2 * np.sum((2 * n + 1) * (a + b).real)
    qsca = 2 / x ** 2 * np.sum((2 * n + 1) * (abs(a) ** 2 + abs(b) ** 2))
    f = (2 * n + 1) / (n * (n + 1))
    S1, S2 = [], []
    for mu in mus:
        pi = np.zeros(nmax + 1)
        tau = np.zeros(nmax + 1)
        pi[1], tau[1] = 1.0, mu
        for k in range(2, nmax + 1):
            pi[k] = (2 * k - 1) / (k - 1) * mu * pi[k - 1] - k / (k - 1) * (pi[k - 2] if k > 2 else 0.0)
            tau[k] = k * mu * pi[k] - (k + 1) * pi[k - 1]
        S1.append(np.sum(f * (a * pi[1:] + b * tau[1:])))
        S2.append(np.sum(f * (a * tau[1:] + b * pi[1:])))
    return qext, qsca, np.array(S1), np.array(S2)


CASES = [(1.5, 0.0, 5.0), (1.7, 0.03, 12.0), (1.33, 0.01, 30.0), (2.0, 1.0, 3.0)]


@pytest.mark.parametrize("mr,mi,x", CASES)
def test_homogeneous_sphere_matches_textbook_mie(oracle_lib, mr, mi, x):
    lam = 2 * np.pi
    cs, ce, F11, F12, F33, F34 = oracle_lib.meerhoff_mie(x, lam, mr, mi)
    qe, qs, S1, S2 = textbook_mie(mr + 1j * mi, x, np.cos(TH))
    G = np.pi * x * x
    assert cs / G == pytest.approx(qs, rel=1e-11)
    # the reference truncates its sums once a term drops below 1e-8
    assert ce / G == pytest.approx(qe, rel=1e-7)
    p11 = 2 * (abs(S1) ** 2 + abs(S2) ** 2) / (x * x * qs)
    scale = p11.max()
    assert np.max(np.abs(F11 - p11)) < 2e-7 * scale
    assert np.max(np.abs(F12 - 2 * (abs(S2) ** 2 - abs(S1) ** 2) / (x * x * qs))) < 2e-7 * scale
    assert np.max(np.abs(F33 - 4 * (S1 * S2.conj()).real / (x * x * qs))) < 2e-7 * scale
    # F34 carries the sign of the m = n - ik convention the reference uses
    assert np.max(np.abs(F34 + 4 * (S2 * S1.conj()).imag / (x * x * qs))) < 2e-7 * scale
    if mi > 0:
        assert np.max(np.abs(F34 - 4 * (S2 * S1.conj()).imag / (x * x * qs))) > 1e-3 * scale


@pytest.mark.parametrize("mr,mi,x", CASES)
def test_layered_sphere_with_one_material_is_the_homogeneous_sphere(oracle_lib, mr, mi, x):
    """DMiLay with core and shell of the same material against the textbook
    homogeneous sphere at the REAL inputs it receives."""
    m = complex(np.float32(mr), -np.float32(mi))
    mu = oracle_lib.dhs_mu()
    d = oracle_lib.dmilay(np.float32(0.6 * x), np.float32(x), np.float32(1.0), m, m, mu)
    assert d["err"] == 0
    qe, qs, S1, S2 = textbook_mie(m.conjugate(), float(np.float32(x)), mu.astype(np.float64))
    assert d["qext"] == pytest.approx(qe, rel=2e-6)
    assert d["qsca"] == pytest.approx(qs, rel=2e-6)
    sc = max(np.max(abs(S1) ** 2), np.max(abs(S2) ** 2))
    assert np.max(np.abs(d["m1"][:, 0] - abs(S1) ** 2)) < 1e-5 * sc
    assert np.max(np.abs(d["m2"][:, 0] - abs(S2) ** 2)) < 1e-5 * sc
    # the supplements (k = 1) are the amplitudes at pi - theta
    qe2, qs2, T1, T2 = textbook_mie(m.conjugate(), float(np.float32(x)), -mu.astype(np.float64))
    assert np.max(np.abs(d["m1"][:, 1] - abs(T1) ** 2)) < 1e-5 * sc
    # a vacuum core of the same size is a different particle
    h = oracle_lib.dmilay(np.float32(0.6 * x), np.float32(x), np.float32(1.0), m, 1.0 + 0.0j, mu)
    assert abs(h["qsca"] / d["qsca"] - 1) > 1e-3


@pytest.mark.parametrize("mr,mi,fv", [(1.7, 0.03, 0.5), (1.5, 0.2, 0.8), (3.0, 0.5, 0.1)])
def test_hollow_sphere_rayleigh_limit(oracle_lib, mr, mi, fv):
    """x = 1e-3: Qabs = 4x Im(alpha), Qsca = 8/3 x^4 |alpha|^2 with the coated
    sphere's polarisability (Bohren & Huffman 5.36), vacuum core."""
    m = complex(np.float32(mr), -np.float32(mi))
    a = np.float32(1e-3)
    rc = np.float32(float(a) * fv ** (1 / 3))
    d = oracle_lib.dmilay(rc, a, np.float32(1.0), m)
    f = (float(rc) / float(a)) ** 3
    e1, e2 = 1.0, m.conjugate() ** 2
    alpha = ((e2 - 1) * (e1 + 2 * e2) + f * (e1 - e2) * (1 + 2 * e2)) / \
            ((e2 + 2) * (e1 + 2 * e2) + f * (2 * e2 - 2) * (e1 - e2))
    x = float(a)
    assert d["qext"] - d["qsca"] == pytest.approx(4 * x * alpha.imag, rel=2e-5)
    assert d["qsca"] == pytest.approx(8 / 3 * x ** 4 * abs(alpha) ** 2, rel=2e-5)


def test_gauleg2_integrates_polynomials(oracle_lib):
    for n in (3, 10, 20):
        x, w = oracle_lib.gauleg2(0.01, 0.8, n)
        assert np.all(np.diff(x) > 0) and x[0] > 0.01 and x[-1] < 0.8
        for k in range(0, 2 * n):
            exact = (0.8 ** (k + 1) - 0.01 ** (k + 1)) / (k + 1)
            got = float(np.sum(w.astype(np.float64) * x.astype(np.float64) ** k))
            assert got == pytest.approx(exact, rel=3e-6, abs=1e-9), (n, k)


def test_solid_sphere_opacity_closed_form(oracle_lib):
    """ns = 1, fmax = 0: Kabs = 1e4 Cabs / (rho 4/3 pi r^3) per wavelength
    with Cabs from the textbook series; F11 normalised over the sphere."""
    lam = np.array([0.5, 1.0, 3.0, 10.0])
    e1 = np.full((1, 4), 1.6, np.float32)
    e2 = np.full((1, 4), 0.05, np.float32)
    o = oracle_lib.compute_part(lam, e1, e2, [1.0], [2.5], 0.8, 0.8, 3.5, 0.0, 1, 1)
    r = float(np.float32(10 ** ((np.float32(math.log10(0.8)) * 2) / 2.0)))
    for i, l in enumerate(lam):
        qe, qs, _, _ = textbook_mie(complex(1.6, 0.05), float(np.float32(r)) * 2 * np.pi / l, [1.0])
        mass = float(np.float32(2.5)) * 4 * np.pi * r ** 3 / 3
        assert o["Kabs"][i] == pytest.approx(1e4 * (qe - qs) * np.pi * r * r / mass, rel=1e-5)
        assert o["Ksca"][i] == pytest.approx(1e4 * qs * np.pi * r * r / mass, rel=1e-5)
    norm = (o["F"][:, 0, :] * np.sin(TH)).sum(axis=1) / np.sin(TH).sum()
    assert np.all(np.abs(norm - 1) < 1e-6)
    assert o["rho"] == pytest.approx(2.5, rel=1e-6)
    assert o["ndhs"] == 0 and o["nhom"] == 4


def test_dhs_mixture_properties(oracle_lib):
    """DHS over sizes and fractions: Kext = Kabs + Ksca at REAL precision,
    F11 normalised, |F12|, |F33|, |F34| <= F11 beyond the first bin, a
    non-absorbing material absorbs nothing, and a layered series that needs
    more terms than x|m| falls back to homogeneous spheres."""
    lam = np.logspace(-1, 2, 8)
    e1 = np.full((1, 8), 1.7, np.float32)
    e2 = np.full((1, 8), 0.03, np.float32)
    o = oracle_lib.compute_part(lam, e1, e2, [1.0], [3.0], 0.05, 10.0, 3.5, 0.8, 4, 5)
    assert o["ndhs"] == 8 * 4 * 5 and o["nerr"] == 0
    assert np.allclose(o["Kext"], o["Kabs"] + o["Ksca"], rtol=1e-6)
    norm = (o["F"][:, 0, :] * np.sin(TH)).sum(axis=1) / np.sin(TH).sum()
    assert np.all(np.abs(norm - 1) < 1e-6)
    for q in (1, 3, 4):
        assert np.all(np.abs(o["F"][:, q, 1:]) <= o["F"][:, 0, 1:] * (1 + 1e-5))
    n = oracle_lib.compute_part(lam, e1, np.zeros((1, 8), np.float32), [1.0], [3.0], 0.05, 10.0, 3.5, 0.8, 4, 5)
    assert np.all(n["Kabs"] < 1e-6 * n["Kext"])
    z = oracle_lib.compute_part(lam[:4], np.full((1, 4), 1.02, np.float32), np.full((1, 4), 1e-3, np.float32),
                                [1.0], [3.0], 1.0, 50.0, 3.5, 0.8, 3, 4)
    assert z["nerr"] > 0 and z["nhom"] >= z["nerr"]
