"""Physics invariants of the oracle's transport pass (MCRadiation.f):
energy conservation, the naked-star spectrum, and the optically thin
temperature closed form."""
import numpy as np
import pytest

import pyoracle
from mcmax_amd import configs


def test_energy_conservation(tiny_model):
    """Every packet ends in exactly one escape of energy Lstar/N
    (star hits re-emit, MCRadiation.f:560-563)."""
    r = pyoracle.Oracle(tiny_model).mcradiation(3000, seed=4)
    assert r["nescape"] == 3000
    assert abs(r["Eescape"] / tiny_model.Lstar - 1) < 1e-12
    # every emission is a star emission or a reemission after absorption
    assert r["nemit"] == 3000 + r["nstarhit"] + r["nraar"] + r["nabs"]


def test_reference_rng_mode_runs(tiny_model):
    o = pyoracle.Oracle(tiny_model)
    r = o.mcradiation(500, rng="ran1")
    assert r["nescape"] == 500 and abs(r["Eescape"] / tiny_model.Lstar - 1) < 1e-12


@pytest.fixture(scope="module")
def thin_model(workdir):
    return configs.load_config("tiny", workdir + "/thin", mdust=1e-14, nrad=30, ntheta=12)


def test_naked_star_spectrum(thin_model):
    """With (almost) no dust the MC spectrum is the stellar one: the packet
    counts per wavelength point follow the trapezoid emission probabilities
    (half of each bin to each end point), and escape directions are
    isotropic."""
    m = thin_model
    n = 200000
    r = pyoracle.Oracle(m).mcradiation(n, seed=8)
    assert r["nabs"] == 0
    trap = np.abs(np.diff(m.nu)) * 0.5 * (m.Fstar[:-1] + m.Fstar[1:])
    pb = trap / trap.sum()
    pl = np.zeros(m.nlam)
    pl[:-1] += 0.5 * pb
    pl[1:] += 0.5 * pb
    cnt = r["ispec"].sum(axis=0)
    expect = n * pl
    sel = expect > 20
    assert np.all(np.abs(cnt[sel] - expect[sel]) < 5 * np.sqrt(expect[sel]) + 2)
    per_angle = r["ispec"][1:].sum(axis=1)
    assert np.all(np.abs(per_angle - n / 30) < 5 * np.sqrt(n / 30))
    # spec(l,a) = sum of wangle(a) E wlam(l) (MCRadiation.f:631-645):
    # expectation wangle(a)/30 Lstar wlam(l) p(l), with the interior weights
    # wlam(l) = 1e23 / ((nu(l-1) - nu(l+1))/2)
    wlam = 1e23 / ((m.nu[:-2] - m.nu[2:]) / 2)
    ratio = r["spec"][1:, 1:-1].mean(axis=0) / (m.Lstar * wlam * pl[1:-1])
    good = expect[1:-1] > 2000
    assert np.all(np.abs(ratio[good] - 1) < 0.05)
    # and for a smooth stellar spectrum on a fine grid that is 1e23 Fstar
    # (the stellar column of MCSpec*.dat, MCRadiation.f:752)


def test_optically_thin_mean_intensity(thin_model):
    """Optically thin limit: EJv/V = Lstar <Kabs> <1/r^2> / (4 pi), with
    <Kabs> the emission-weighted opacity of the sampled spectrum and
    <1/r^2> the shell average 3 (R2-R1)/(R2^3-R1^3)."""
    m = thin_model
    n = 200000
    r = pyoracle.Oracle(m).mcradiation(n, seed=9)
    AU = float(np.float32(1.49598e13))
    trap = np.abs(np.diff(m.nu)) * 0.5 * (m.Fstar[:-1] + m.Fstar[1:])
    K = m.Kabs[0]
    kmean = (trap * 0.5 * (K[:-1] + K[1:])).sum() / trap.sum()
    R1, R2 = m.R[:-1] * AU, m.R[1:] * AU
    inv_r2 = 3 * (R2 - R1) / (R2 ** 3 - R1 ** 3)
    expect = m.Lstar * kmean * inv_r2 / (4 * np.pi)
    got = r["EJv"] / m.V
    rows = slice(5, m.nr)  # well outside the stellar radius
    ratio = got[rows] / expect[rows, None]
    assert abs(np.median(ratio) - 1) < 0.02
    assert np.mean(np.abs(ratio - 1) < 0.15) > 0.9
