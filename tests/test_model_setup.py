"""The product's host setup (mcmax_amd.model, restating Init.f) against the
oracle's independent C restatement, plus the particle-file reader and the
keyword parser."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle
from mcmax_amd import keywords, model

L = pyoracle.load()
P = pyoracle.p


def ulps(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = np.maximum(np.abs(a), np.abs(b))
    return np.max(np.where(scale > 0, np.abs(a - b) / (scale * 2.0 ** -52 + 1e-320), 0.0))


def test_wavelength_and_planck_tables(small_model):
    m = small_model
    lam, nu = np.zeros(m.nlam), np.zeros(m.nlam)
    L.orc_lam_grid(m.nlam, 0.1, 1000.0, P(lam), P(nu))
    assert ulps(m.lam, lam) <= 2 and ulps(m.nu, nu) <= 2
    BB, BBint = np.zeros((10001, m.nlam)), np.zeros(10001)
    L.orc_bb_table(m.nlam, P(m.lam), P(m.nu), P(BB), P(BBint))
    assert ulps(m.BB, BB) <= 2
    assert ulps(m.BBint, BBint) <= 4
    Kp = np.zeros(10001)
    L.orc_kp_table(m.nlam, P(m.nu), P(m.BB), P(np.ascontiguousarray(m.Kabs[0])), P(Kp))
    assert ulps(m.Kp[0], Kp) == 0
    F, Ls = np.zeros(m.nlam), C.c_double()
    L.orc_star(m.nlam, P(m.lam), P(m.nu), m.Rstar_cm, m.Tstar, P(F), C.byref(Ls))
    assert ulps(m.Fstar, F) <= 2 and ulps(m.Lstar, Ls.value) <= 4


def test_grids_and_density(small_model):
    m = small_model
    Th, tav = np.zeros(m.nt + 1), np.zeros(m.nt)
    L.orc_theta_grid(m.nt, P(Th), P(tav))
    assert ulps(m.Theta, Th) <= 2 and ulps(m.theta_av, tav) <= 2
    R, Rav = np.zeros(m.nr + 1), np.zeros(m.nr)
    L.orc_r_grid(m.nr - 1, m.Rin, m.Rout, m.Rstar_cm, P(R), P(Rav))
    assert ulps(m.R, R) <= 2 and ulps(m.R_av, Rav) <= 2
    dens, V = np.zeros((m.nr, m.nt)), np.zeros((m.nr, m.nt))
    L.orc_min_density(m.nr, m.nt, P(m.R), P(m.Theta), P(m.R_av), P(m.theta_av), m.Mdust_g, m.Rin, m.Rout,
                      0.05, 1.125, 1.0, P(dens), P(V))
    assert ulps(m.V, V) <= 4
    ok = dens > 1e-300
    # exp(-(z/h)^2) turns a one-ulp libm difference in h into ~x ulps
    AU = float(np.float32(1.49598e13))
    r = m.R_av[1:, None] * np.sin(m.theta_av)[None, :] / AU
    z = m.R_av[1:, None] * np.cos(m.theta_av)[None, :] / AU
    x = np.vstack([np.zeros((1, m.nt)), (z / (0.05 * r ** 1.125)) ** 2])
    rel = np.abs(m.dens - dens) / np.maximum(dens, 1e-300)
    assert np.all(rel[ok] <= (8 + 8 * x[ok]) * 2.0 ** -52)
    assert abs((m.dens[1:] * m.V[1:]).sum() / m.Mdust_g - 1) < 1e-12


def test_mueller_tables(small_model):
    m = small_model
    IF11, IF12 = np.zeros(m.nlam), np.zeros(m.nlam)
    F = np.ascontiguousarray(m.F[0])
    L.orc_mueller_integrals(m.nlam, P(F), P(IF11), P(IF12))
    assert ulps(m.IF11[0], IF11) <= 2 and ulps(m.IF12[0], IF12) <= 2
    c2, s2 = np.zeros(361), np.zeros(361)
    L.orc_phi_tables(P(c2), P(s2))
    assert np.allclose(m.cos2phi, c2, rtol=0, atol=1e-15) and np.allclose(m.sin2phi, s2, rtol=0, atol=1e-15)
    # the analytic grain: HG bin averages, normalised
    th = (np.arange(1, 181) - 0.5) * np.pi / 180
    hg = np.array([L.orc_hg(0.5, t) for t in th])
    s = np.sin(th)
    assert np.allclose(m.F[0, 0, 0], hg * s.sum() / (hg * s).sum(), rtol=1e-13)


def test_read_particle_interpolation_and_extrapolation(tmp_path):
    """readparticle: linear in wavelength between file points, kappa_abs ~
    lam^-2 and kappa_sca ~ lam^-4 beyond the last one (Init.f:4723-4786)."""
    lam_f = np.array([0.5, 1.0, 2.0, 4.0])
    ka = 10.0 + 3.0 * lam_f
    ks = 20.0 - 2.0 * lam_f
    path = tmp_path / "lin.particle"
    with open(path, "w") as f:
        for l, a, s in zip(lam_f, ka, ks):
            f.write(f"{l} {a + s} {a} {s}\n")
            for k in range(6):
                f.write(" ".join(["1.0" if k == 0 else "0.0"] * 180) + "\n")
    grid = np.array([0.5, 0.7, 1.3, 2.0, 3.9, 5.0, 8.0])
    Ka, Ks, Ke, F = model.read_particle(str(path), grid)
    inside = grid <= 4.0
    assert np.allclose(Ka[inside], 10.0 + 3.0 * grid[inside], rtol=1e-14)
    assert np.allclose(Ks[inside], 20.0 - 2.0 * grid[inside], rtol=1e-14)
    last = np.where(inside)[0][-1]
    assert np.allclose(Ka[~inside], Ka[last] * (grid[last] / grid[~inside]) ** 2, rtol=1e-14)
    assert np.allclose(Ks[~inside], Ks[last] * (grid[last] / grid[~inside]) ** 4, rtol=1e-14)
    assert np.allclose(F[:, 0, :], 1.0)


def test_keyword_values_and_rejection(tmp_path):
    kw = keywords.parse_lines(["* a comment", "", "Mdust=1d-4", "TCONTACT=.true.", "nrad=100",
                               "part1='my.particle'", "scattype=FULL", "computepart01:amin=0.05"])
    assert kw["mdust"] == 1e-4 and kw["tcontact"] is True and kw["nrad"] == 100
    assert kw["part1"] == "my.particle" and kw["scattype"] == "FULL"
    assert kw["computepart01:amin"] == 0.05
    with pytest.raises(keywords.KeywordError):
        keywords.parse_lines(["notakeyword=3"])
    path = tmp_path / "x.in"
    keywords.write_file(str(path), {"mdust": 1e-4, "tcontact": False, "part1": "a.particle"})
    back = keywords.parse_file(str(path), overrides=["mdust=2d-4"])
    assert back == {"mdust": 2e-4, "tcontact": False, "part1": "a.particle"}


def test_setup_rejects_unsupported_options(workdir):
    from mcmax_amd import configs
    with pytest.raises(model.SetupError):
        configs.load_config("tiny", os.path.join(workdir, "bad"), gridrefine=True)


def test_radiation_pressure_is_refused(tmp_path):
    """radpress needs the pass's radiation-force tally, which is not built:
    the setup refuses it rather than run without it."""
    from mcmax_amd import configs, model
    with pytest.raises(model.SetupError):
        configs.load_config("tiny", str(tmp_path), radpress=True)
