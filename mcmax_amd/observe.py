"""Observations of types SPECTRUM and IMAGE: the observation files the
reference reads after the input file on its command line (determineoutput,
Telescope.f:563-1190), the wavelengths a spectrum asks for
(Telescope.f:40-71), the ray-traced SED (Observe, Telescope.f:138-164) and
the ray-traced image with its polarisation maps (Telescope.f:165-181 with
MakeImage, TraceMono.f:2759-3237, and RImage, TraceMono.f:3375-3444).
"""
import math

import numpy as np

from . import keywords as kwmod
from . import output


class ObservationError(ValueError):
    pass


# Telescope.f:590-657 defaults that SPECTRUM and IMAGE use: inc defaults to
# the model's incangle, lam1/lam2 to the wavelength grid's ends, an image's
# wavelength to their mean and its field of view to the disk's diameter
DEFAULTS = dict(nphi=45, nrad=2, nphot=1000, nphotstar=1000, nexits=5, usepol=True, flag="", noangle=False,
                nlam=-1, lamfile="", npix=500, nint=1000, scaletype=1, d=-1.0, width=0.0, psffile="",
                opening=0.0, ptel=0.0, aptel=0.0, texp=-1.0, iprad=0.0)

_KEYS = {"type", "inc", "lam1", "lam2", "nlam", "nphi", "nrad", "nphot", "nphotstar", "nexits", "usepol",
         "flag", "noangle", "lamfile", "lam", "npix", "nint", "scaletype", "d", "width", "psffile", "opening",
         "ptel", "aptel", "texp", "iprad"}
_TEXT = ("flag", "lamfile", "psffile")
TYPES = ("SPECTRUM", "IMAGE")


def _new_observation(kind, default):
    """A new observation of `kind` from the defaults (Telescope.f:680-935):
    SPECTRUM takes lam1/lam2, IMAGE the default wavelength `lam` and the
    field-of-view list."""
    tel = {k: v for k, v in default.items() if k not in ("deflam", "fov")}
    tel["type"] = kind
    if kind == "IMAGE":
        tel["lam1"] = default["deflam"]
        tel["fov"] = list(default["fov"])
    return tel


def _set_fov(fov, key, value):
    """fov sets the first field of view, fovN the N-th, growing the list
    (Telescope.f:981-1003); entries not given keep the value they held."""
    n = 1 if key == "fov" else int(key[3:])
    if n < 1:
        raise ObservationError(f"bad field-of-view keyword {key!r}")
    while len(fov) < n:
        fov.append(fov[-1] if fov else 0.0)
    fov[n - 1] = float(value)


def parse_observation_file(path, m):
    """The observations of one observation file, each a dict.  Keys given
    before the first `type=` line set defaults for the ones after it
    (Telescope.f:935-1020); there `lam` is an image's wavelength, while in
    an observation it sets lam1."""
    default = dict(DEFAULTS, inc=float(m.incangle), lam1=float(m.lam[0]), lam2=float(m.lam[-1]),
                   deflam=(float(m.lam[0]) + float(m.lam[-1])) / 2.0, fov=[float(m.R[-1]) * 2.0])
    obs = []
    with open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            if not line.strip() or line.lstrip().startswith("*"):
                continue
            if "=" not in line:
                raise ObservationError(f"{path}: not a key=value line: {line!r}")
            k, v = line.split("=", 1)
            key = k.strip().lower()
            val = v.strip()
            if len(val) >= 2 and val[0] in "\"'":
                val = val[1:-1]
            if key.startswith("fov") and (key == "fov" or key[3:].isdigit()):
                target = obs[-1].setdefault("fov", list(default["fov"])) if obs else default["fov"]
                _set_fov(target, key, kwmod.parse_value(val))
                continue
            if key not in _KEYS:
                raise ObservationError(f"{path}: observation keyword not supported by this build: {key!r}")
            if key == "type":
                kind = val.strip().upper()
                if kind not in TYPES:
                    raise ObservationError(f"{path}: observation type {val!r} is not supported "
                                           f"(one of {', '.join(TYPES)})")
                obs.append(_new_observation(kind, default))
                continue
            value = val if key in _TEXT else kwmod.parse_value(val)
            if key == "lam":
                key = "lam1" if obs else "deflam"
            (obs[-1] if obs else default)[key] = value
    for tel in obs:
        if tel["type"] == "IMAGE":
            check_image(tel)
    return obs


def check_image(tel):
    """Refuse what an IMAGE observation of this build does not make: the
    telescope PSF (keywords d, width, psffile: Convolution and
    ConvolutionFILE with FFTs), the photon noise that needs a telescope
    diameter, and an opened disk (opendisk)."""
    if float(tel["d"]) > 0.0 or float(tel["width"]) > 0.0 or str(tel["psffile"]).strip():
        raise ObservationError("image: telescope PSF convolution (d, width, psffile) is not supported by this build")
    if float(tel["opening"]) != 0.0:
        raise ObservationError("image: an opened disk (opening != 0) is not supported by this build")
    if int(tel["scaletype"]) not in (1, 2):
        raise ObservationError(f"image: scaletype {tel['scaletype']} (1 or 2)")
    if int(tel["npix"]) < 1 or int(tel["nint"]) < 1 or not all(float(x) > 0.0 for x in tel["fov"]):
        raise ObservationError("image: npix and nint must be positive and every fov > 0")


def observed_wavelengths(tel, lam):
    """lam_obs of one observation (Telescope.f:40-71): the model grid between
    lam1 and lam2 with both ends, a file's list, or nlam log-spaced points
    (the spacing's ratio (i-1)/(nlam-1) in single precision)."""
    nlam = int(tel["nlam"])
    if nlam < 0 and not tel["lamfile"]:
        inner = [float(x) for x in lam if tel["lam1"] < x < tel["lam2"]]
        return np.array([float(tel["lam1"])] + inner + [float(tel["lam2"])])
    if tel["lamfile"]:
        vals = []
        with open(tel["lamfile"]) as f:
            for line in f:
                try:
                    vals.append(float(line.split()[0].replace("d", "e").replace("D", "E")))
                except (ValueError, IndexError):
                    continue
        return np.array(vals)
    nlam = min(nlam, len(lam))
    l1, l2 = float(tel["lam1"]), float(tel["lam2"])
    out = []
    for i in range(1, nlam + 1):
        ratio = float(np.float32(i - 1) / np.float32(nlam - 1)) if nlam > 1 else 0.0
        out.append(10.0 ** (math.log10(l1) + math.log10(l2 / l1) * ratio))
    return np.array(out)


def _make_obs(tel, seed):
    from .transport import make_obs
    return make_obs(angle_deg=float(tel["inc"]), nphi=int(tel["nphi"]), nj0=int(tel["nrad"]),
                    nphot=int(tel["nphot"]), nphotstar=int(tel["nphotstar"]), nexits=int(tel["nexits"]),
                    usepol=bool(tel["usepol"]), seed=seed)


def observe_spectrum(tr, m, tel, outdir, T=None, seed=1):
    """Run one SPECTRUM observation on the device and write its file;
    returns (path, lam_obs, flux array [n, 3])."""
    lam_obs = observed_wavelengths(tel, m.lam)
    res = tr.sed(_make_obs(tel, seed), lam_obs, T=T)
    path = f"{outdir}/{output.spectrum_filename(tel['inc'], tel['flag'], bool(tel['noangle']))}"
    output.write_spectrum(path, m, lam_obs, res[:, 0], res[:, 1], res[:, 2])
    return path, lam_obs, res


def image_scales(m, tel, fov):
    """(Rmax in AU, rscale, zscale) of one field of view (Telescope.f:171-178):
    scaletype 1 keeps AU and Jy per pixel; scaletype 2 reads the field in
    arcsec and gives arcsec and mJy per square arcsec."""
    from .model import PARSEC
    if int(tel["scaletype"]) == 1:
        return float(fov) / 2.0, 1.0, 1.0
    d = m.distance_cm
    ext = output.ism_extinction(m)
    fov_au = float(fov) * d / PARSEC
    zscale = ext * 1e3 * ((int(tel["npix"]) / fov_au) * d / PARSEC) ** 2
    return fov_au / 2.0, PARSEC / d, zscale


def observe_image(tr, m, tel, outdir, T=None, seed=1):
    """Run one IMAGE observation on the device (rays traced at the image
    wavelength, MakeImage per field of view) and write its files: RImage
    for the first field, then the FITS maps of every field.  Returns the
    list of paths written."""
    check_image(tel)
    ob = _make_obs(tel, seed)
    lam0 = float(tel["lam1"])
    pa = float(getattr(m, "keywords", {}).get("posangle", 0.0))
    angle = math.pi * float(tel["inc"]) / 180.0
    paths = []
    for i, fov in enumerate(tel["fov"]):
        Rmax, rscale, zscale = image_scales(m, tel, fov)
        raw, Rimg, rays = tr.image(ob, lam0, int(tel["npix"]), int(tel["nint"]), Rmax, pa_deg=pa, seed=seed,
                                   T=T, rays=True)
        if i == 0:
            path = f"{outdir}/{output.rimage_filename(angle, lam0, tel['flag'])}"
            output.write_rimage(path, m, Rimg, rays, rscale, zscale, 2.0 * Rmax, int(tel["npix"]))
            paths.append(path)
        maps = output.image_maps(m, raw, zscale, bool(tel["usepol"]), pa, float(tel["ptel"]), float(tel["aptel"]),
                                 float(tel["iprad"]))
        for add, a in maps.items():
            path = f"{outdir}/{output.image_filename(add, angle, lam0, Rmax, tel['flag'])}"
            output.write_image_fits(path, a, Rmax, rscale, int(tel["scaletype"]))
            paths.append(path)
    return paths


def observe(tr, m, tel, outdir, T=None, seed=1):
    """Run one observation (SPECTRUM or IMAGE); returns the paths written."""
    if tel["type"] == "IMAGE":
        return observe_image(tr, m, tel, outdir, T=T, seed=seed)
    return [observe_spectrum(tr, m, tel, outdir, T=T, seed=seed)[0]]
