"""Pins of the oracle's MakeImage restatement (TraceMono.f:2816-3020) and of
the host-side image maps (scaling, RotateStokes, DQ/DU, file names).

The spread of a ray image onto pixels is pinned by what the geometry fixes:
the total of a constant field is its area, the deposits at +phi and -phi
mirror each other (U and V with opposite signs), a field of constant Stokes
vector keeps its ratios through the polarised log interpolation, an
unpolarised field falls back to linear interpolation, and the log-linear
interpolation between ray centres matches an independent restatement with
numpy.interp at every sample of the same Philox draws."""
import math

import numpy as np
import pytest

import pyoracle
from mcmax_amd import configs, model, output

AU = model.AU


@pytest.fixture(scope="module")
def lib(oracle_lib):
    return oracle_lib.load()


def _rings(n, R0=0.0, R1=50.0):
    return np.linspace(R0, R1, n)


def _draws(lib, seed, pid, nint):
    u = np.array([lib.orc_philox_uniform(seed, pid, k) for k in range(2 * nint)])
    return u[0::2], u[1::2]


def test_constant_field_total_is_its_area(lib):
    """I = const: every sample deposits I w2 R / (2 nint) at +phi and -phi
    with w2 = 2 pi AU^2 dR / nphi, so the image total is I AU^2 2 pi/nphi
    sum_ij dR_i mean_k(R_k), exactly, and pi AU^2 (R_n^2 - R_0^2) on average."""
    nimg, nphi, nint, npix = 7, 5, 64, 40
    R = _rings(nimg, 0.5, 30.0)
    rays = np.zeros((4, nimg, nphi))
    rays[0] = 2.5
    out = pyoracle.make_image(lib, R, rays, 1, npix, nint, 31.0, 0.0, 9)
    expect = 0.0
    for i in range(nimg - 1):
        for j in range(nphi):
            ranR, _ = _draws(lib, 9, i * nphi + j, nint)
            Rk = R[i] + (R[i + 1] - R[i]) * ranR
            expect += 2.5 * 2.0 * math.pi * AU ** 2 * (R[i + 1] - R[i]) / nphi * np.mean(Rk)
    assert out[0].sum() == pytest.approx(expect, rel=1e-12)
    area = 2.5 * math.pi * AU ** 2 * (R[-1] ** 2 - R[0] ** 2)
    assert out[0].sum() == pytest.approx(area, rel=0.05)
    assert np.all(out[1:] == 0.0)


def test_mirror_symmetry(lib):
    """pa = 0: pixel (ix, iy) and (ix, npix+1-iy) receive the same samples;
    I and Q are equal there, U and V opposite, bit for bit."""
    rng = np.random.default_rng(3)
    nimg, nphi = 9, 6
    R = _rings(nimg, 0.0, 20.0)
    I = rng.uniform(1.0, 2.0, (nimg, nphi))
    P = rng.uniform(0.0, 0.8, (nimg, nphi)) * I
    ang = rng.uniform(0.0, np.pi, (nimg, nphi))
    rays = np.stack([I, P * np.cos(2 * ang), P * np.sin(2 * ang), 0.1 * P])
    for npix in (30, 31):
        out = pyoracle.make_image(lib, R, rays, 2, npix, 50, 21.0, 0.0, 4)
        im, imQ, imU, imV = out
        assert np.count_nonzero(im) > npix
        assert np.array_equal(im, im[:, ::-1])
        assert np.array_equal(imQ, imQ[:, ::-1])
        assert np.array_equal(imU, -imU[:, ::-1])
        assert np.array_equal(imV, -imV[:, ::-1])
        assert np.any(imU != 0.0) and np.any(imV != 0.0)


def test_constant_stokes_vector_keeps_its_ratios(lib):
    """Identical rays (I, Q, U, V) with ring spacing above the pixel size take
    the polarised log interpolation: al^2 + ar^2 + (I - |P|) = I,
    al^2 - ar^2 = Q, 2 al ar c = U, 2 al ar s = V, so every pixel holds
    Q/I and, on the +phi half (iy > npix/2), U/I and V/I; the -phi half the
    opposite U and V."""
    nimg, nphi, npix = 6, 4, 20
    R = _rings(nimg, 0.0, 10.0)
    rays = np.zeros((4, nimg, nphi))
    rays[0], rays[1], rays[2], rays[3] = 1.0, 0.3, 0.2, 0.05
    im, imQ, imU, imV = pyoracle.make_image(lib, R, rays, 2, npix, 40, 10.5, 0.0, 2)
    hit = im > 0
    assert hit.sum() > 100
    assert np.allclose(imQ[hit] / im[hit], 0.3, rtol=1e-12)
    up = np.zeros_like(hit)
    up[:, npix // 2:] = True
    assert np.allclose(imU[hit & up] / im[hit & up], 0.2, rtol=1e-12)
    assert np.allclose(imU[hit & ~up] / im[hit & ~up], -0.2, rtol=1e-12)
    assert np.allclose(imV[hit & up] / im[hit & up], 0.05, rtol=1e-12)
    assert np.allclose(imV[hit & ~up] / im[hit & ~up], -0.05, rtol=1e-12)


def test_unpolarised_rays_fall_back_to_linear(lib):
    """Q = U = V = 0 make c = U/(2 al ar) = 0/0; number_invalid sends the
    sample to linear interpolation: I as the log interpolation of the same
    constant, Q, U, V zero, nothing non-finite."""
    nimg, nphi, npix = 6, 4, 20
    R = _rings(nimg, 0.0, 10.0)
    rays = np.zeros((4, nimg, nphi))
    rays[0] = 3.0
    full = pyoracle.make_image(lib, R, rays, 2, npix, 40, 10.5, 0.0, 5)
    plain = pyoracle.make_image(lib, R, rays, 1, npix, 40, 10.5, 0.0, 5)
    assert np.all(np.isfinite(full))
    assert np.allclose(full[0], plain[0], rtol=1e-13, atol=0.0)
    assert np.all(full[1:] == 0.0)


def test_log_linear_interpolation_between_ray_centres(lib):
    """scat_how 1: log10 I interpolated linearly in azimuth between the ray
    centres pi (j - 1/2)/nphi (held at the first and last centre) and
    linearly in R between rings, at the Philox draws of each cell; restated
    with numpy.interp and compared pixel by pixel."""
    rng = np.random.default_rng(8)
    nimg, nphi, nint, npix, Rmax, seed = 5, 6, 12, 16, 9.0, 21
    R = np.array([0.0, 1.5, 3.0, 5.5, 8.5])
    I = 10.0 ** rng.uniform(-3.0, 2.0, (nimg, nphi))
    rays = np.zeros((4, nimg, nphi))
    rays[0] = I
    got = pyoracle.make_image(lib, R, rays, 1, npix, nint, Rmax, 0.0, seed)[0]
    want = np.zeros((npix, npix))
    centres = math.pi * (np.arange(nphi) + 0.5) / nphi
    for i in range(nimg - 1):
        for j in range(nphi):
            ranR, ranPhi = _draws(lib, seed, i * nphi + j, nint)
            for k in range(nint):
                r = R[i] + (R[i + 1] - R[i]) * ranR[k]
                phi = math.pi * (j + ranPhi[k]) / nphi
                a = np.interp(phi, centres, np.log10(I[i]))
                b = np.interp(phi, centres, np.log10(I[i + 1]))
                flux = 10.0 ** ((1.0 - ranR[k]) * a + ranR[k] * b) * (2.0 * math.pi * AU ** 2 *
                                                                      (R[i + 1] - R[i]) / nphi) * r
                for ph in (phi, -phi):
                    ix = int(npix * (-r * math.cos(ph) + Rmax) / (2 * Rmax) + 1)
                    iy = int(npix * (r * math.sin(ph) + Rmax) / (2 * Rmax) + 1)
                    if 1 <= ix <= npix and 1 <= iy <= npix:
                        want[ix - 1, iy - 1] += flux / (2 * nint)
    assert np.count_nonzero(want) > 50
    assert np.allclose(got, want, rtol=1e-10, atol=0.0)


def test_position_angle_turns_the_image(lib):
    """pa = 90 degrees puts at (y, -x) the sample that pa = 0 put at (x, y):
    the map turns by a quarter, b[p, q] = a[n-1-q, p]."""
    rng = np.random.default_rng(1)
    nimg, nphi, npix = 6, 4, 24
    R = _rings(nimg, 0.0, 10.0)
    rays = np.zeros((4, nimg, nphi))
    rays[0] = rng.uniform(1.0, 2.0, (nimg, nphi))
    a = pyoracle.make_image(lib, R, rays, 1, npix, 30, 11.0, 0.0, 3)[0]
    b = pyoracle.make_image(lib, R, rays, 1, npix, 30, 11.0, 90.0, 3)[0]
    assert b.sum() == pytest.approx(a.sum(), rel=1e-12)
    # x -> -R cos(phi + 90) = R sin(phi) = y; y -> R sin(phi + 90) = R cos(phi) = -x
    turned = np.rot90(a, k=-1)
    close = np.isclose(b, turned, rtol=1e-9, atol=0.0)
    assert close.mean() > 0.9


# ---------------------------------------------------------------------------
# host-side maps

def test_rotate_stokes_closed_form():
    """Linear polarisation P at angle chi: turning the reference direction
    from x (S = (1, 1e-5)) to the pixel direction theta gives
    Q' = P cos 2(chi - theta), U' = P sin 2(chi - theta); the position-angle
    rotation (S = (cos -pa, sin -pa) to x) gives Q cos 2pa + U sin 2pa."""
    rng = np.random.default_rng(5)
    chi = rng.uniform(0, np.pi, 200)
    th = rng.uniform(-np.pi, np.pi, 200)
    P, E = 0.4, 2.0
    Q, U = P * np.cos(2 * chi), P * np.sin(2 * chi)
    Qr, Ur = output.rotate_stokes_image(Q, U, E, 1.0, 0.00001, np.cos(th), np.sin(th))
    # S = (1, 1e-5) is not of unit length: cos t = cos theta + 1e-5 sin theta
    # turns the frame by O(1e-5) more than theta; the length P is kept
    assert np.allclose(Qr, P * np.cos(2 * (chi - th)), atol=2e-5)
    assert np.allclose(Ur, P * np.sin(2 * (chi - th)), atol=2e-5)
    assert np.allclose(np.hypot(Qr, Ur), P, rtol=1e-12)
    pa = 30.0
    Qp, Up = output.rotate_stokes_image(Q, U, E, math.cos(-pa * math.pi / 180), math.sin(-pa * math.pi / 180),
                                        1.0, 0.0)
    c2, s2 = math.cos(2 * pa * math.pi / 180), math.sin(2 * pa * math.pi / 180)
    assert np.allclose(Qp, Q * c2 + U * s2, atol=1e-12)
    assert np.allclose(Up, -Q * s2 + U * c2, atol=1e-12)
    q0, u0 = output.rotate_stokes_image(np.zeros(3), np.zeros(3), 1.0, 1.0, 1e-5, 0.6, 0.8)
    assert np.all(q0 == 0.0) and np.all(u0 == 0.0)


def _rayleigh(m):
    th = (np.pi * (np.arange(1, 181) - 0.5)) / 180.0
    c = np.cos(th)
    F = np.zeros_like(m.F)
    F[..., 0, :] = 0.75 * (1 + c ** 2)
    F[..., 1, :] = -0.75 * (1 - c ** 2)
    F[..., 2, :] = F[..., 0, :]
    F[..., 3, :] = 1.5 * c
    F[..., 5, :] = 1.5 * c
    m.F = F
    m.IF11, m.IF12 = model.mueller_integrals(F)
    m.gsym = model.asymmetry(F)
    return m


def test_face_on_polarisation_is_centrosymmetric(workdir, lib):
    """A face-on disk of Rayleigh scatterers: the ray image's polarisation is
    azimuthal (Q ~ -P cos 2phi, U ~ P sin 2phi), and in the maps the
    polarisation turned to each pixel's direction from the centre (DQ) holds
    it while DU, its other component, stays small: a wrong sign of U, or of
    the -phi half's U, would move it into DU."""
    m = _rayleigh(configs.load_config("tiny", workdir + "/img_faceon", mdust=1e-5))
    o = pyoracle.Oracle(m)
    T = np.empty((m.nr, m.nt))
    T[:] = (400.0 * (m.R_av / m.R_av[1]) ** -0.5)[:, None]
    obs = pyoracle.make_obs(angle_deg=1.0, nphot=2000, nphotstar=2000, nphi=16)
    R, rays = o.ray_image(T, obs, m.lam[8], seed=7)
    I, Q, U, _ = rays
    phi = np.pi * (np.arange(16) + 0.5) / 16
    rings = [i for i in range(len(R)) if np.all(I[i] > 0) and np.all(np.hypot(Q[i], U[i]) > 0.2 * I[i])]
    assert len(rings) > 20
    P = np.hypot(Q[rings], U[rings])
    cu = [np.corrcoef(U[i], np.sin(2 * phi))[0, 1] for i in rings]
    cq = [np.corrcoef(Q[i], np.cos(2 * phi))[0, 1] for i in rings]
    assert np.median(cu) > 0.95 and np.median(cq) < -0.95
    assert np.corrcoef((U[rings] / P).ravel(), np.tile(np.sin(2 * phi), len(rings)))[0, 1] > 0.9
    assert np.corrcoef((Q[rings] / P).ravel(), np.tile(np.cos(2 * phi), len(rings)))[0, 1] < -0.9
    npix, Rmax = 64, 0.12
    raw = pyoracle.make_image(lib, R, rays, 2, npix, 200, Rmax, 0.0, 3)
    maps = output.image_maps(m, raw, 1.0, True)
    lp, dq, du = maps["LP"], maps["DQ"], maps["DU"]
    disk = np.isfinite(lp) & (lp > 0.2) & (maps["AI"] > 0)
    assert disk.sum() > 200
    assert np.median(np.abs(dq[disk]) / lp[disk]) > 0.95
    assert np.median(np.abs(du[disk]) / lp[disk]) < 0.1
    # one sign over the disk: the polarisation is tangential everywhere
    assert abs(np.mean(np.sign(dq[disk]))) > 0.95


def test_image_file_names():
    """FITSImage's name (TraceMono.f:3466-3502): inclination recovered from
    radians with the 9.95 carry, wavelength and field in single digits."""
    ang = math.pi * 60.0 / 180.0
    assert output.image_filename("AI", ang, 0.55, 100.0) == "ImageAI_i60.0_l00000.55_fov00200.0.fits"
    assert output.image_filename("DQ", math.pi * 45.0 / 180.0, 1.6, 2.5, "_x") == \
        "ImageDQ_i45.0_l00001.60_fov00005.0_x.fits"
    assert output.image_filename("AI", math.pi * 9.97 / 180.0, 12.0, 49.98) == \
        "ImageAI_i10.0_l00012.00_fov00100.0.fits"
    assert output.rimage_filename(ang, 2.2, "_b") == "RImage_i60.0_l00002.20_b.dat"


def test_image_maps_scaling_and_ratios(tiny_model):
    """AI..CP are the raw images times 1e23 zscale / d^2 and their ratios;
    without full scattering or usepol only AI."""
    m = tiny_model
    rng = np.random.default_rng(2)
    raw = rng.uniform(0.1, 1.0, (4, 8, 8))
    raw[1:] *= 0.3
    d2 = m.distance_cm * m.distance_cm
    maps = output.image_maps(m, raw, 2.0, m.scat_how == 2)
    assert np.allclose(maps["AI"], raw[0] * 1e23 * 2.0 / d2, rtol=1e-15)
    if m.scat_how == 2:
        assert set(maps) == {"AI", "AQ", "AU", "AV", "RQ", "RU", "LP", "CP", "DQ", "DU", "QU", "EX"}
        assert np.allclose(maps["RQ"], raw[1] / raw[0], rtol=1e-14)
        assert np.allclose(maps["LP"], np.hypot(raw[1], raw[2]) / raw[0], rtol=1e-14)
        # every pixel but ((n+1)/2, (n+1)/2), which the reference skips and
        # which keeps the previous map's value divided by I once more
        rest = np.ones((8, 8), dtype=bool)
        rest[3, 3] = False
        assert np.allclose((maps["DQ"] ** 2 + maps["DU"] ** 2)[rest], (maps["LP"] ** 2)[rest], rtol=1e-9)
        assert maps["DQ"][3, 3] == maps["CP"][3, 3] / maps["AI"][3, 3]
        assert maps["DU"][3, 3] == maps["DQ"][3, 3] / maps["AI"][3, 3]
        assert maps["EX"].shape == (4, 8, 8)
    assert set(output.image_maps(m, raw, 2.0, False)) == {"AI"}


def test_observation_file_image(tiny_model, tmp_path):
    from mcmax_amd import observe
    m = tiny_model
    p = tmp_path / "obs.in"
    p.write_text("lam=2.2\nnpix=64\ntype=IMAGE\ninc=30\nfov=120\nfov2=60\nnint=20\n"
                 "type=IMAGE\nlam=1.1\ntype=SPECTRUM\n")
    a, b, c = observe.parse_observation_file(str(p), m)
    assert a["type"] == "IMAGE" and a["lam1"] == 2.2 and a["npix"] == 64 and a["nint"] == 20
    assert a["fov"] == [120.0, 60.0] and a["inc"] == 30.0
    assert b["lam1"] == 1.1 and b["fov"] == [2.0 * float(m.R[-1])] and b["nint"] == 1000
    assert c["type"] == "SPECTRUM" and c["lam1"] == m.lam[0]
    for bad in ("d=8.0", "width=0.1", "psffile='psf.fits'", "opening=10"):
        p.write_text(f"type=IMAGE\n{bad}\n")
        with pytest.raises(observe.ObservationError):
            observe.parse_observation_file(str(p), m)
    p.write_text("type=VISIBILITY\n")
    with pytest.raises(observe.ObservationError):
        observe.parse_observation_file(str(p), m)


def test_rimage_profile_closed_form(tiny_model, tmp_path):
    """RImage's flux column for rays of constant I is the trapezoid sum of
    2 pi R I AU^2 over the rings, exact for that linear integrand:
    1e23 I pi AU^2 (R_i^2 - R_0^2) / d^2 at ring i+1; the intensity column
    is the ray at azimuth nphi/2 times the pixel area scaling."""
    m = tiny_model
    R = np.array([0.0, 0.5, 1.5, 4.0, 9.0])
    nphi = 6
    rays = np.zeros((4, len(R), nphi))
    rays[0] = 2.0
    rays[0, :, nphi // 2 - 1] = 3.0   # the column RImage prints
    p = tmp_path / "r.dat"
    output.write_rimage(str(p), m, R, rays, 1.0, 1.0, 20.0, 100)
    rows = [[float(x) for x in line.split()] for line in open(p)]
    d2 = m.distance_cm ** 2
    scaling = (1e23 / d2) * (20.0 ** 2 * AU ** 2) / 100 ** 2
    assert [r[0] for r in rows] == list(R)
    # the ray at nphi/2 carries 3, the other five 2: mean 13/6 per azimuth;
    # row k holds R_k and the flux inside R_{k+1} (the last row, R_n and the
    # total)
    C = 1e23 * (13.0 / 6.0) * math.pi * AU ** 2 / d2
    for k in range(len(R) - 1):
        assert rows[k][1] == pytest.approx(C * (R[k + 1] ** 2 - R[0] ** 2), rel=1e-13), k
    assert rows[-1][1] == pytest.approx(C * (R[-1] ** 2 - R[0] ** 2), rel=1e-13)
    assert all(r[2] == pytest.approx(3.0 * scaling, rel=1e-15) for r in rows)
