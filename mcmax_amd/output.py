"""Output products of the stage, in the reference's text formats.

- the structure file (outputstruct, InputOutput.f:6-227, written as
  denstemp.dat with DENS, TEMP, COMP, GASDENS, DENS0: Main.f:605; with the
  keyword outputfits its FITS form denstemp.fits.gz, InputOutput.f:230-572),
  and its reader for a restart from temperatures (the Nphot = 0 path,
  Main.f:380-384);
- the ray-traced spectrum of an observation of type SPECTRUM
  (spectrum<inc>.dat, Telescope.f:140-164);
- the Monte Carlo spectra per inclination bin and angle-integrated
  (MCSpec<inc>.dat, MCSpec.dat, MCRadiation.f:733-762);
- an IMAGE observation's maps as FITS files (Image<map>_i.._l.._fov...fits,
  FITSImage and ExPoImage, TraceMono.f:3446-3720) and its radial profile
  (RImage_i.._l...dat, TraceMono.f:3375-3444).

Numbers are written with 17 significant digits in E format, one value per
field as the reference's list-directed writes; any list-directed reader
(including the reference's own) reads them back exactly.
"""
import math
import os

import numpy as np

GAS2DUST = 100.0  # Init.f:276
FORMAT_NUMBER = 5  # InputOutput.f:26


class OutputError(ValueError):
    pass


def _num(x):
    return f"{float(x): .16E}"


def _line(*xs):
    return "  ".join(_num(x) for x in xs) + "\n"


def angle_tag(angle_deg):
    """The inclination in a file name: i1 of angle/10 followed by f3.1 of
    the remainder (Telescope.f:145-147, MCRadiation.f:740-741)."""
    a = float(angle_deg)
    tens = int(a / 10.0)
    rest = a - 10.0 * tens
    t, r = f"{tens:d}", f"{rest:.1f}"
    # a value wider than its field prints as asterisks
    return (t if len(t) <= 1 else "*") + (r.rjust(3) if len(r) <= 3 else "***")


# ---------------------------------------------------------------------------
def write_denstemp(path, m, T, state=None):
    """The structure file of radii 1..nr-1 and all theta cells: DENS, TEMP,
    COMP, GASDENS (times gas2dust), DENS0.  state holds dens, dens0, gasdens
    [nr, nt] and w [nr, nt, ng] (default: the model's own, as set up)."""
    nr, nt, ng = m.nr, m.nt, m.ng
    if state is None:
        state = dict(dens=m.dens, dens0=m.dens, gasdens=m.dens, w=m.w)
    T = np.asarray(T, dtype=np.float64).reshape(nr, nt)
    dens = np.asarray(state["dens"]).reshape(nr, nt)
    dens0 = np.asarray(state["dens0"]).reshape(nr, nt)
    gas = np.asarray(state["gasdens"]).reshape(nr, nt)
    w = np.asarray(state["w"]).reshape(nr, nt, ng)
    with open(path, "w") as f:
        f.write("# Format number\n")
        f.write(f"{FORMAT_NUMBER:6d}\n")
        f.write("# NR, NT, NGRAINS, NGRAINS2\n")
        f.write(f"{nr - 1:6d}{nt:6d}{ng:6d}{1:6d}\n")
        f.write("# Spherical radius grid [cm] (middle of cell)\n")
        for i in range(1, nr):
            f.write(_line(m.R_av[i]))
        f.write("# Theta grid [rad, from pole] (middle of cell)\n")
        for j in range(nt):
            f.write(_line(m.theta_av[j]))
        blocks = (("Density", dens), ("Temperature", T))
        for name, arr in blocks:
            f.write(f"# {name} array (for ir=0,nr-1 do for it=0,nt-1 do ...)\n")
            for i in range(1, nr):
                for j in range(nt):
                    f.write(_line(arr[i, j]))
        f.write("# Composition array (for ir=0,nr-1 do for it=0,nt-1 do ...)\n")
        for i in range(1, nr):
            for j in range(nt):
                f.write(_line(*w[i, j]))
        f.write("# Gas density array (for ir=0,nr-1 do for it=0,nt-1 do ...)\n")
        for i in range(1, nr):
            for j in range(nt):
                f.write(_line(gas[i, j] * GAS2DUST))
        f.write("# Density0 array (for ir=0,nr-1 do for it=0,nt-1 do ...)\n")
        for i in range(1, nr):
            for j in range(nt):
                f.write(_line(dens0[i, j]))


def read_denstemp(path):
    """Read a structure file written by write_denstemp (or by the reference's
    ASCII outputstruct with the same variables), or its FITS form (names
    ending in .fits or .fits.gz).  Returns a dict with R_av [nr-1], theta_av
    [nt], and dens, T, gasdens (divided by gas2dust), dens0 [nr-1, nt] and w
    [nr-1, nt, ng]."""
    if path.endswith(".fits") or path.endswith(".fits.gz"):
        return read_denstemp_fits(path)
    with open(path) as f:
        lines = [l.strip() for l in f if l.strip()]
    pos = 0

    def header(expect):
        nonlocal pos
        if not lines[pos].startswith("#") or expect.lower() not in lines[pos].lower():
            raise OutputError(f"{path}: expected '# {expect}', found {lines[pos]!r}")
        pos += 1

    def values(n, per=1):
        nonlocal pos
        out = []
        while len(out) < n * per:
            out.extend(float(x.replace("D", "E").replace("d", "e")) for x in lines[pos].split())
            pos += 1
        return np.array(out[: n * per])

    header("Format number")
    fmt = int(lines[pos]); pos += 1
    if fmt != FORMAT_NUMBER:
        raise OutputError(f"{path}: format number {fmt}, expected {FORMAT_NUMBER}")
    header("NR, NT")
    nr1, nt, ng, ng2 = (int(x) for x in lines[pos].split()[:4]); pos += 1
    if ng2 != 1:
        raise OutputError(f"{path}: temperature-dependent opacities (NGRAINS2={ng2}) are not supported")
    out = {}
    header("Spherical radius grid")
    out["R_av"] = values(nr1)
    header("Theta grid")
    out["theta_av"] = values(nt)
    header("Density array")
    out["dens"] = values(nr1 * nt).reshape(nr1, nt)
    header("Temperature array")
    out["T"] = values(nr1 * nt).reshape(nr1, nt)
    header("Composition array")
    out["w"] = values(nr1 * nt, ng).reshape(nr1, nt, ng)
    header("Gas density array")
    out["gasdens"] = values(nr1 * nt).reshape(nr1, nt) / GAS2DUST
    header("Density0 array")
    out["dens0"] = values(nr1 * nt).reshape(nr1, nt)
    return out


def write_denstemp_fits(path, m, T, state=None):
    """The structure file as FITS (outputstruct_fits, InputOutput.f:230-572,
    written as denstemp.fits.gz when the keyword outputfits is set): the
    primary HDU holds the grid (R_av/AU and theta_av for every cell, shape
    [2, nt, nr-1]) with the keys Rin, Rout [AU], nR, nTheta, ngrains,
    ngrains2, nlam, nHDU and the names of the extensions, which hold DENS,
    TEMP, COMP ([ngrains2+1, ngrains, nt, nr-1]: w, then wopac = 1),
    GASDENS (times gas2dust) and DENS0 ([nt, nr-1] each).  Keyword names
    are upper case and HDU<k> has no blank (valid FITS keywords)."""
    from . import fits
    from .model import AU
    nr, nt, ng = m.nr, m.nt, m.ng
    if state is None:
        state = dict(dens=m.dens, dens0=m.dens, gasdens=m.dens, w=m.w)
    T = np.asarray(T, dtype=np.float64).reshape(nr, nt)
    cells = lambda a: np.asarray(a, dtype=np.float64).reshape(nr, nt)[1:].T.copy()  # [nt, nr-1]
    grid = np.zeros((2, nt, nr - 1))
    grid[0] = (np.asarray(m.R_av)[1:] / AU)[None, :]
    grid[1] = np.asarray(m.theta_av)[:, None]
    w = np.asarray(state["w"], dtype=np.float64).reshape(nr, nt, ng)
    comp = np.ones((2, ng, nt, nr - 1))
    comp[0] = np.transpose(w[1:], (2, 1, 0))
    names = ("DENS", "TEMP", "COMP", "GASDENS", "DENS0")
    keys = [("RIN", float(m.Rin), "[AU]"), ("ROUT", float(m.Rout), "[AU]"), ("NR", nr - 1), ("NTHETA", nt),
            ("NGRAINS", ng), ("NGRAINS2", 1), ("NLAM", int(m.nlam)), ("NHDU", len(names))]
    keys += [(f"HDU{k}", n) for k, n in enumerate(names, 1)]
    arrays = [grid, cells(state["dens"]), cells(T), comp, cells(state["gasdens"]) * GAS2DUST, cells(state["dens0"])]
    fits.write(path, arrays, keys)


def read_denstemp_fits(path):
    """Read a structure file written by write_denstemp_fits: the same dict
    as read_denstemp (arrays [nr-1, nt])."""
    from . import fits
    from .model import AU
    hdus = fits.read(path)
    hdr, grid = hdus[0]
    names = [hdr.get(f"HDU{k}") for k in range(1, int(hdr.get("NHDU", 0)) + 1)]
    if names[:5] != ["DENS", "TEMP", "COMP", "GASDENS", "DENS0"] or len(hdus) < 6:
        raise OutputError(f"{path}: expected DENS, TEMP, COMP, GASDENS, DENS0 extensions, found {names}")
    out = {"R_av": grid[0, 0, :] * AU, "theta_av": grid[1, :, 0].copy()}
    for k, key in ((1, "dens"), (2, "T"), (4, "gasdens"), (5, "dens0")):
        out[key] = hdus[k][1].T.copy()
    out["gasdens"] = out["gasdens"] / GAS2DUST
    out["w"] = np.transpose(hdus[3][1][0], (2, 1, 0)).copy()
    return out


# ---------------------------------------------------------------------------
def arrinterpol_log(x, xarr, yarr):
    """arrinterpol(x, xarr, yarr, n, 1) (Subroutines.f:2506-2532) as the
    reference computes it: outside the table it returns the abscissa at the
    end, and its scan keeps the last index i with x <= xarr(i), so inside the
    table the power law through the last two points is evaluated."""
    n = len(xarr)
    if x <= xarr[0]:
        return float(xarr[0])
    if x >= xarr[n - 1]:
        return float(xarr[n - 1])
    ix = 1
    for i in range(2, n + 1):
        if x <= xarr[i - 1]:
            ix = i
    x1, x2, y1, y2 = xarr[ix - 2], xarr[ix - 1], yarr[ix - 2], yarr[ix - 1]
    pw = math.log10(y2 / y1) / math.log10(x2 / x1)
    return y1 * (x / x1) ** pw


def ism_extinction(m):
    """ExtISM of the spectra: 1 without foreground extinction (keyword av,
    Reddening in Reddening.f, which this build does not restate)."""
    av = float(getattr(m, "keywords", {}).get("av", 0.0))
    if av != 0.0:
        raise OutputError("interstellar extinction (av != 0) is not supported by this build")
    return 1.0


def write_spectrum(path, m, lam_obs, flux, scatflux, fluxQ=None):
    """spectrum<inc>.dat: lambda, flux, scattered flux, stellar flux [Jy]
    and, with full scattering, the Q flux (Telescope.f:150-163)."""
    ext = ism_extinction(m)
    d2 = m.distance_cm ** 2
    with open(path, "w") as f:
        for k, lam in enumerate(lam_obs):
            fstar1 = arrinterpol_log(float(lam), m.lam, m.Fstar)
            cols = [lam, 1e23 * flux[k] * ext / d2, 1e23 * scatflux[k] * ext / d2, 1e23 * fstar1 / d2]
            if m.scat_how == 2:
                cols.append(1e23 * fluxQ[k] * ext / d2)
            f.write(_line(*cols))


def spectrum_filename(angle_deg, flag="", noangle=False):
    if noangle:
        return f"spectrum{flag}.dat"
    return f"spectrum{angle_tag(angle_deg)}{flag}.dat"


def write_mcspec(outdir, m, tal):
    """MCSpec<inc>.dat for the 30 inclination bins and MCSpec.dat (the
    angle-integrated spectrum), from the pass's tallies (spec already holds
    the reference's 1e23 wangle wlam weights).  The error column is
    spec/sqrt(real(ispec)) (the multiwav=.false. branch); real() is single
    precision."""
    ext = ism_extinction(m)
    d2 = m.distance_cm ** 2
    spec = np.asarray(tal["spec"]).reshape(-1, m.nlam)
    scat = np.asarray(tal["scatspec"]).reshape(-1, m.nlam)
    iscat = np.asarray(tal["iscatspec"]).reshape(-1, m.nlam)
    ispec = np.asarray(tal["ispec"]).reshape(-1, m.nlam)
    nangle = spec.shape[0] - 1
    paths = []
    for j in range(1, nangle + 1):
        ang = 180.0 * math.acos((float(j) - 0.5) / float(nangle)) / math.pi
        p = os.path.join(outdir, f"MCSpec{angle_tag(ang)}.dat")
        with np.errstate(divide="ignore", invalid="ignore"):
            dspec = spec[j] / np.sqrt(ispec[j].astype(np.float32)).astype(np.float64)
        with open(p, "w") as f:
            for i in range(m.nlam):
                f.write(_line(m.lam[i], spec[j, i] * ext / d2, dspec[i] * ext / d2, scat[j, i] * ext / d2)
                        [:-1] + f"  {int(iscat[j, i]):12d}  " + _num(1e23 * m.Fstar[i] / d2) + "\n")
        paths.append(p)
    p = os.path.join(outdir, "MCSpec.dat")
    with open(p, "w") as f:
        for i in range(m.nlam):
            f.write(_line(m.lam[i], spec[0, i] * ext / d2))
    paths.append(p)
    return paths


# ---------------------------------------------------------------------------
# images
def _ifmt(n):
    """Fortran i1: one digit, an asterisk when it does not fit."""
    return str(n) if 0 <= n <= 9 else "*"


def _ffmt(x, w, d):
    """Fortran fw.d: asterisks when the value does not fit."""
    t = f"{x:{w}.{d}f}"
    return t if len(t) <= w else "*" * w


def _image_angle_lam(angle_rad, lam):
    """The _i<inc>_l<lam> part of an image file name (TraceMono.f:3466-3478):
    the inclination recovered from radians, a remainder of 9.95 or more
    carried into the tens (less 9.95), and the wavelength as four single
    digits and f4.2."""
    a = 180.0 * angle_rad / math.pi
    i10 = int(a / 10.0)
    i1 = a - 10.0 * i10
    if i1 >= 9.95:
        i1 = i1 - 9.95
        i10 = i10 + 1
    l = float(lam)
    l4 = int(l / 10000.0)
    l3 = int((l - 10000.0 * l4) / 1000.0)
    l2 = int((l - 1000.0 * l3 - 10000.0 * l4) / 100.0)
    l1 = int((l - 100.0 * l2 - 1000.0 * l3 - 10000.0 * l4) / 10.0)
    l0 = l - 10.0 * l1 - 100.0 * l2 - 1000.0 * l3 - 10000.0 * l4
    return (f"_i{_ifmt(i10)}{_ffmt(i1, 3, 1)}_l{_ifmt(l4)}{_ifmt(l3)}{_ifmt(l2)}{_ifmt(l1)}"
            f"{_ffmt(l0, 4, 2)}")


def image_filename(add, angle_rad, lam, Rmax, flag=""):
    """Image<add>_i<inc>_l<lam>_fov<2 Rmax><flag>.fits (FITSImage,
    TraceMono.f:3466-3502): the field's width as four digits and f3.1,
    a remainder of 9.95 or more carried upwards."""
    r = 2.0 * float(Rmax)
    r4 = int(r / 10000.0)
    r3 = int((r - 10000.0 * r4) / 1000.0)
    r2 = int((r - 1000.0 * r3 - 10000.0 * r4) / 100.0)
    r1 = int((r - 100.0 * r2 - 1000.0 * r3 - 10000.0 * r4) / 10.0)
    r0 = r - 10.0 * r1 - 100.0 * r2 - 1000.0 * r3 - 10000.0 * r4
    if r0 >= 9.95:
        r1 = r1 + 1
        r0 = r0 - 9.95
    if r1 >= 10:
        r2, r1 = r2 + 1, r1 - 10
    if r2 >= 10:
        r3, r2 = r3 + 1, r2 - 10
    if r3 >= 10:
        r4, r3 = r4 + 1, r3 - 10
    return (f"Image{add}{_image_angle_lam(angle_rad, lam)}_fov{_ifmt(r4)}{_ifmt(r3)}{_ifmt(r2)}{_ifmt(r1)}"
            f"{_ffmt(r0, 3, 1)}{flag}.fits")


def rimage_filename(angle_rad, lam, flag=""):
    """RImage_i<inc>_l<lam><flag>.dat (RImage, TraceMono.f:3386-3399)."""
    return f"RImage{_image_angle_lam(angle_rad, lam)}{flag}.dat"


def write_image_fits(path, a, Rmax, rscale, scaletype):
    """One map (numpy [ix-1, iy-1], or a cube [k, ix-1, iy-1]) with the
    reference's coordinate keywords (TraceMono.f:3544-3567): the corner
    pixel's centre 0.5, its value -Rmax rscale, pixel scale 2 Rmax rscale /
    npix, in AU (scaletype 1) or arcsec (scaletype 2).  The reference stores
    array(iy, ix) = im(ix, iy), so NAXIS1 runs over iy: numpy [ix-1, iy-1]."""
    from . import fits
    a = np.asarray(a, dtype=np.float64)
    unit = "AU" if scaletype == 1 else "arcsec"
    pix = 2.0 * Rmax * rscale / float(a.shape[-1])
    keys = [("CTYPE1", f"X [{unit}]", "X-coordinate"), ("CTYPE2", f"Y [{unit}]", "Y-coordinate"),
            ("CRPIX1", 0.5, "Corner of the image"), ("CRPIX2", 0.5, "Corner of the image"),
            ("CRVAL1", -Rmax * rscale, "Center of the image"), ("CRVAL2", -Rmax * rscale, "Center of the image"),
            ("CDELT1", pix, f"Pixel scale (in {unit})"), ("CDELT2", pix, f"Pixel scale (in {unit})")]
    fits.write(path, [a], keys)


def rotate_stokes_image(Q, U, E, Sx, Sy, x, y):
    """RotateStokes (Radiation.f:506-533) for a photon travelling along +z
    with reference direction (Sx, Sy, 0), turned to (x, y, 0): the rotation
    by twice the angle between them, rescaled to keep (Q/E)^2 + (U/E)^2;
    unchanged where that is 0.  Element-wise over arrays."""
    with np.errstate(all="ignore"):
        P = (U / E) ** 2 + (Q / E) ** 2
        cost = Sx * x + Sy * y
        sint = (x - (Sx * 1.0 - 0.0) * cost) / (0.0 - 1.0 * Sy)
        cos2t = cost ** 2 - sint ** 2
        sin2t = (2.0 * sint) * cost
        Qn = Q * cos2t + U * sin2t
        Un = -Q * sin2t + U * cos2t
        P2 = (Un / E) ** 2 + (Qn / E) ** 2
        f = np.sqrt(P / P2)
        keep = (P == 0.0)
        return np.where(keep, Q, f * Qn), np.where(keep, U, f * Un)


def image_maps(m, raw, zscale, usepol, pa_deg=0.0, ptel=0.0, aptel=0.0, iprad=0.0):
    """MakeImage's maps from the raw images raw [4, npix, npix] (I, Q, U, V
    before scaling; TraceMono.f:3014-3237): the scaling to 1e23 zscale /
    distance^2, the rotation of Q, U by the position angle, the telescope's
    own polarisation (ptel, aptel), then AI and, with full scattering and
    usepol, AQ, AU, AV, RQ = Q/I, RU = U/I, LP, CP = V/I, DQ and DU (Q, U in
    the frame of each pixel's direction from the centre, over I), QU =
    sqrt(Q^2 + U^2), the cube EX and, with iprad > 0, CX.  Returns
    {name: array [npix, npix] (or [4, npix, npix])}."""
    d = m.distance_cm
    npix = raw.shape[-1]
    pol = (m.scat_how == 2) and usepol
    with np.errstate(all="ignore"):
        im = raw[0] * 1e23 * zscale / (d * d)
        out = {}
        if not pol:
            return {"AI": im}
        imQ = raw[1] * 1e23 * zscale / (d * d)
        imU = raw[2] * 1e23 * zscale / (d * d)
        imV = raw[3] * 1e23 * zscale / (d * d)
        if pa_deg != 0.0:
            rot = im != 0.0
            Qr, Ur = rotate_stokes_image(imQ, imU, im, math.cos(-pa_deg * math.pi / 180.0),
                                         math.sin(-pa_deg * math.pi / 180.0), 1.0, 0.0)
            imQ, imU = np.where(rot, Qr, imQ), np.where(rot, Ur, imU)
        if ptel != 0.0:
            imQ = imQ + ptel * math.cos(aptel * math.pi / 180.0) * im
            imU = imU + ptel * math.sin(aptel * math.pi / 180.0) * im
        out["AI"] = im
        out["AQ"], out["AU"], out["AV"] = imQ, imU, imV
        out["RQ"] = imQ / im
        out["RU"] = imU / im
        out["LP"] = np.sqrt((imQ / im) ** 2 + (imU / im) ** 2)
        imP = imV / im
        out["CP"] = imP
        # pixel directions from the centre; the pixel ((n+1)/2, (n+1)/2) is
        # skipped and keeps the previous map's value
        c = (npix + 1) // 2
        xi = np.arange(1, npix + 1, dtype=np.float64) - float(npix + 1) / 2.0
        X, Y = np.meshgrid(xi, xi, indexing="ij")
        R = np.sqrt(X ** 2 + Y ** 2)
        X, Y = X / R, Y / R
        skip = np.zeros((npix, npix), dtype=bool)
        skip[c - 1, c - 1] = True
        Qd, Ud = rotate_stokes_image(imQ, imU, im, 1.0, 0.00001, X, Y)
        imP = np.where(skip, imP, Qd) / im
        out["DQ"] = imP
        imP = np.where(skip, imP, Ud) / im
        out["DU"] = imP
        out["QU"] = np.sqrt(imQ ** 2 + imU ** 2)
        out["EX"] = np.stack([im, imQ, imU, np.sqrt(imQ ** 2 + imU ** 2)])
        if iprad > 0.0:
            inside = R <= iprad
            fq, fu, fl = np.sum(imQ[inside]), np.sum(imU[inside]), np.sum(im[inside])
            Qc, Uc = imQ - im * fq / fl, imU - im * fu / fl
            out["CX"] = np.stack([im, Qc, Uc, np.sqrt(Qc ** 2 + Uc ** 2)])
    return out


def write_rimage(path, m, Rimg, rays, rscale, zscale, fov1, npix):
    """The radial profile of the ray image (RImage, TraceMono.f:3375-3444):
    per ring R rscale, the flux inside it (1e23/d^2), the intensity of the
    ray at azimuth nphi/2 times the pixel's area scaling and, with full
    scattering, the Q and U fluxes inside and the polarised intensity (its
    per-azimuth normalisation applied as the reference applies it)."""
    from .model import AU
    d2 = m.distance_cm ** 2
    I, Q, U = rays[0], rays[1], rays[2]
    nr, nphi = I.shape
    scaling = (zscale * 1e23 / d2) * (fov1 ** 2 * AU ** 2) / float(npix ** 2)
    kh = nphi // 2 - 1
    flux = Qflux = Uflux = 0.0
    with open(path, "w") as f:
        for i in range(nr - 1):
            ps = 0.0
            dR = Rimg[i + 1] - Rimg[i]
            for k in range(nphi):
                w1 = 2.0 * math.pi * abs(Rimg[i]) * AU ** 2 / float(nphi)
                w2 = 2.0 * math.pi * abs(Rimg[i + 1]) * AU ** 2 / float(nphi)
                flux = flux + dR * (w1 * I[i, k] + w2 * I[i + 1, k]) / 2.0
                if m.scat_how == 2:
                    Qflux = Qflux + dR * (w1 * Q[i, k] + w2 * Q[i + 1, k]) / 2.0
                    Uflux = Uflux + dR * (w1 * U[i, k] + w2 * U[i + 1, k]) / 2.0
                    ps = ps + dR * math.sqrt(((w1 * U[i, k] + w2 * U[i + 1, k]) / 2.0) ** 2 +
                                             ((w1 * Q[i, k] + w2 * Q[i + 1, k]) / 2.0) ** 2)
                ps = ps / (2.0 * math.pi * AU ** 2 * (Rimg[i + 1] ** 2 - Rimg[i] ** 2))
            cols = [Rimg[i] * rscale, flux * 1e23 / d2, scaling * I[i, kh]]
            if m.scat_how == 2:
                cols += [Qflux * 1e23 / d2, Uflux * 1e23 / d2, scaling * ps]
            f.write(_line(*cols))
        cols = [Rimg[nr - 1] * rscale, flux * 1e23 / d2, scaling * I[nr - 1, kh]]
        if m.scat_how == 2:
            cols.append(scaling * math.sqrt(Q[nr - 1, kh] ** 2 + U[nr - 1, kh] ** 2))
        f.write(_line(*cols))
