"""Observations of type SPECTRUM: the observation files the reference reads
after the input file on its command line (determineoutput,
Telescope.f:563-1190), the wavelengths they ask for (Telescope.f:40-71) and
the ray-traced SED they produce (Observe, Telescope.f:138-164).
"""
import math

import numpy as np

from . import keywords as kwmod
from . import output


class ObservationError(ValueError):
    pass


# Telescope.f:590-657 defaults that SPECTRUM uses (inc defaults to the
# model's incangle, lam1/lam2 to the wavelength grid's ends)
DEFAULTS = dict(nphi=45, nrad=2, nphot=1000, nphotstar=1000, nexits=5, usepol=True, flag="", noangle=False,
                nlam=-1, lamfile="")

_KEYS = {"type", "inc", "lam1", "lam2", "nlam", "nphi", "nrad", "nphot", "nphotstar", "nexits", "usepol",
         "flag", "noangle", "lamfile", "lam"}


def parse_observation_file(path, m):
    """The SPECTRUM observations of one observation file, each a dict.
    Keys given before the first `type=` line set defaults for the ones after
    it (Telescope.f:935-1020)."""
    default = dict(DEFAULTS, inc=float(m.incangle), lam1=float(m.lam[0]), lam2=float(m.lam[-1]))
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
            if key not in _KEYS:
                raise ObservationError(f"{path}: observation keyword not supported by this build: {key!r}")
            if key == "type":
                if val.upper() != "SPECTRUM":
                    raise ObservationError(f"{path}: observation type {val!r} is not supported (SPECTRUM only)")
                obs.append(dict(default))
                continue
            if key == "lam":
                key = "lam1"
            value = val if key in ("flag", "lamfile") else kwmod.parse_value(val)
            (obs[-1] if obs else default)[key] = value
    return obs


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


def observe(tr, m, tel, outdir, T=None, seed=1):
    """Run one SPECTRUM observation on the device and write its file;
    returns (path, lam_obs, flux array [n, 3])."""
    from .transport import make_obs
    lam_obs = observed_wavelengths(tel, m.lam)
    ob = make_obs(angle_deg=float(tel["inc"]), nphi=int(tel["nphi"]), nj0=int(tel["nrad"]),
                  nphot=int(tel["nphot"]), nphotstar=int(tel["nphotstar"]), nexits=int(tel["nexits"]),
                  usepol=bool(tel["usepol"]), seed=seed)
    res = tr.sed(ob, lam_obs, T=T)
    path = f"{outdir}/{output.spectrum_filename(tel['inc'], tel['flag'], bool(tel['noangle']))}"
    output.write_spectrum(path, m, lam_obs, res[:, 0], res[:, 1], res[:, 2])
    return path, lam_obs, res
