"""GPU parity of the IMAGE observation (mcx_image) against the oracle: the
ray images I, Q, U, V of TraceFlux traced at the image wavelength, and
MakeImage's spread of those rays onto pixels (TraceMono.f:2759-3020); then
the command line writing an image's FITS maps.

The rays of a non-scattering disk are deterministic and compared to
rounding.  With scattering the packet histories follow the oracle's Philox
streams, and a packet that a one-ulp difference of a device math function
sends across a cell edge differently follows its own history, so the
polarised rays are compared to their noise.  The spread is compared on the
same rays (the device's, handed to the oracle): the same samples, deposits
summed in another order."""
import os

import numpy as np
import pytest

from mcmax_amd import __main__ as cli
from mcmax_amd import configs, fits, transport

pytestmark = pytest.mark.gpu


def _power_law_T(m, T1=400.0):
    T = np.empty((m.nr, m.nt))
    T[:] = (T1 * (m.R_av / m.R_av[1]) ** -0.5)[:, None]
    return T


def _dhs(ngrains):
    return {"abun01": 1.0, "computepart01:file": "dhs.dat", "computepart01:amin": 0.05,
            "computepart01:amax": 1000.0, "computepart01:apow": 3.5, "computepart01:fmax": 0.8,
            "computepart01:ngrains": ngrains, "computepart01:blend": False}


def _both(gpu, oracle_lib, m, T, lam, npix, nint, Rmax, pa=0.0, seed=5, **kw):
    gpu.set_model(m)
    print(f"image: lam {lam:.4g} npix {npix} nint {nint} Rmax {Rmax} pa {pa} {kw}", flush=True)
    g, gR, grays = gpu.image(transport.make_obs(seed=seed, **kw), lam, npix, nint, Rmax, pa_deg=pa, seed=seed,
                             T=T, rays=True)
    o = oracle_lib.Oracle(m)
    oR, orays = o.ray_image(T, oracle_lib.make_obs(**kw), lam, seed=seed)
    return g, gR, grays, oR, orays, o


def test_image_thermal_rays_and_spread(gpu, oracle_lib, workdir):
    """Extinction and thermal emission only: every ray's I to 1e-9 of the
    oracle's, no polarisation; the spread of the device's rays equals the
    oracle's spread of the same rays."""
    m = configs.load_config("tiny", workdir + "/gimg_therm", mdust=1e-6, scattype="NONE", nrad=30, ntheta=16)
    T = _power_law_T(m, 300.0)
    lam = float(m.lam[20])
    g, gR, grays, oR, orays, o = _both(gpu, oracle_lib, m, T, lam, 64, 40, 60.0, angle_deg=60.0, nphot=0,
                                       nphotstar=0, nj0=3, nphi=12)
    assert np.allclose(gR, oR, rtol=1e-14, atol=0.0)
    assert np.allclose(grays[0], orays[0], rtol=1e-9, atol=1e-12 * orays[0].max())
    assert np.all(grays[1:] == 0.0) and np.all(orays[1:] == 0.0)
    ref = o.make_image(gR, grays, m.scat_how, 64, 40, 60.0, 0.0, 5)
    assert np.count_nonzero(g[0]) > 500
    assert np.allclose(g[0], ref[0], rtol=1e-11, atol=0.0)
    # the image holds the SED's flux: the same rays, spread instead of summed
    sed = gpu.sed(transport.make_obs(angle_deg=60.0, nphot=0, nphotstar=0, nj0=3, nphi=12), [lam], T=T)
    assert g[0].sum() == pytest.approx(sed[0, 0], rel=0.05)


@pytest.fixture(scope="module")
def dhs_model(workdir, gpu):
    return configs.load_config("tiny", workdir + "/gimg_dhs", tr=gpu, mdust=1e-5, **_dhs(1))


@pytest.mark.parametrize("pa", [0.0, 30.0])
def test_image_polarised_rays_and_spread(gpu, oracle_lib, dhs_model, pa):
    """Full scattering with DHS grains (F12 != 0): the rays' I, Q, U, V
    against the oracle's, and the polarised spread (log interpolation of
    I - |P| and the two amplitudes, linear where that is not finite) of the
    device's rays against the oracle's spread of the same rays."""
    m = dhs_model
    assert m.scat_how == 2
    T = _power_law_T(m)
    lam = float(m.lam[10])
    g, gR, grays, oR, orays, o = _both(gpu, oracle_lib, m, T, lam, 80, 60, 0.15, pa=pa, angle_deg=40.0,
                                       nphot=400, nphotstar=400, nphi=16)
    assert np.allclose(gR, oR, rtol=1e-14, atol=0.0)
    assert np.abs(grays[2]).max() > 0 and np.abs(orays[2]).max() > 0
    scale = np.abs(orays[0]).max()
    for q in range(4):
        # most rays follow identical histories and agree to rounding
        same = np.isclose(grays[q], orays[q], rtol=1e-8, atol=1e-12 * scale)
        assert same.mean() > 0.8, (q, same.mean())
        # the integrals over the image plane agree to the Monte Carlo noise
        assert abs(grays[q].sum() - orays[q].sum()) <= 2e-2 * np.abs(orays[0]).sum()
    ref = o.make_image(gR, grays, m.scat_how, 80, 60, 0.15, pa, 5)
    assert np.count_nonzero(g[0]) > 1000
    assert np.allclose(g[0], ref[0], rtol=1e-10, atol=0.0)
    big = np.abs(ref[0]).max()
    for q in (1, 2, 3):
        assert np.allclose(g[q], ref[q], rtol=1e-9, atol=1e-12 * big), q
    assert np.abs(g[2]).max() > 1e-4 * big


def test_cli_image_observation(tmp_path):
    """An IMAGE observation through the command line: RImage and the FITS
    maps, the total of AI the ray-traced flux at the same wavelength."""
    inp = configs.write_config("tiny", str(tmp_path / "in"))
    obs = tmp_path / "obs.in"
    obs.write_text("type=IMAGE\ninc=45\nlam=2.0\nnpix=100\nnint=50\nfov=250\nnphot=200\nnphotstar=200\n"
                   "type=SPECTRUM\ninc=45\nlam1=2.0\nlam2=2.0\nnlam=1\nnphot=200\nnphotstar=200\n")
    out = tmp_path / "run"
    print("cli: tiny, 20000 packets, IMAGE + SPECTRUM", flush=True)
    assert cli.main([inp, "20000", str(obs), "-o", str(out)]) == 0
    names = set(os.listdir(out))
    tag = "_i45.0_l00002.00"
    assert f"RImage{tag}.dat" in names
    assert f"ImageAI{tag}_fov00250.0.fits" in names
    hdr, ai = fits.read(str(out / f"ImageAI{tag}_fov00250.0.fits"))[0]
    assert ai.shape == (100, 100) and hdr["NAXIS1"] == 100
    assert hdr["CTYPE1"] == "X [AU]" and hdr["CRVAL1"] == pytest.approx(-125.0)
    assert hdr["CDELT1"] == pytest.approx(2.5)
    spec = np.loadtxt(out / "spectrum45.0.dat", ndmin=2)
    assert ai.sum() == pytest.approx(spec[0, 1], rel=0.05)
    rows = [[float(x) for x in line.split()] for line in open(out / f"RImage{tag}.dat")]
    # per ring: R, flux inside, scaled intensity (+ Q, U fluxes and the
    # polarised intensity with full scattering); the outermost ring's row
    # has no Q, U fluxes
    assert all(r[0] < q[0] for r, q in zip(rows, rows[1:]))
    # the flux inside the outermost ring is the spectrum's flux (rays laid
    # out at the image wavelength rather than at 0.55 micron)
    assert rows[-1][1] == pytest.approx(spec[0, 1], rel=0.05)
    m = configs.load_config("tiny", str(tmp_path / "m"))
    if m.scat_how == 2:
        for add in ("AQ", "AU", "AV", "RQ", "RU", "LP", "CP", "DQ", "DU", "QU", "EX"):
            assert f"Image{add}{tag}_fov00250.0.fits" in names
