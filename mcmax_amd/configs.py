"""Keyword files for the benchmark and test workloads.

small   -- plain method, CPU-runnable: 4000 K / 2 Rsun / 0.5 Msun star, dust
           mass 1e-4 Msun, Rin 0.1 AU, Rout 100 AU, Sigma ~ r^-1, MIN vertical
           profile exp(-(z/h)^2) with h/r = 0.05 at 1 AU and flaring 1.125
           (the reference's MIN convention, Init.f:3163: sh1AU is the h of
           exp(-(z/h)^2)), 100 r x 60 theta cells, 100 wavelengths 0.1-1000 um,
           one analytic grain (kappa_abs = 1e3 cm^2/g (lam/um)^-1 above 1 um,
           flat below; albedo 0.5; HG g = 0.5), FULL scattering.
production -- the same star; dust mass 1e-3 Msun; 200 r x 100 theta cells,
           200 wavelengths.  Dust: the analytic grain above (the ten-size DHS
           mixture from a constant refractive index is not generated yet).
tiny    -- a 20 x 10 cell, 40-wavelength variant for fast tests.
"""
import os

from . import model as mdl

SMALL = dict(
    startype="PLANCK", tstar=4000.0, rstar=2.0, mstar=0.5, distance=100.0,
    mdust=1e-4, rin=0.1, rout=100.0, denstype="MIN", denspow=1.0, sh1au=0.05, shpow=1.125,
    nrad=100, ntheta=60, gridrefine=False, iter=False, lam1=0.1, lam2=1000.0, nlam=100,
    scattype="FULL", tcontact=True, incangle=45.0,
)
PRODUCTION = dict(SMALL, mdust=1e-3, nrad=200, ntheta=100, nlam=200)
TINY = dict(SMALL, nrad=20, ntheta=10, nlam=40, mdust=1e-4)

CONFIGS = {"small": SMALL, "production": PRODUCTION, "tiny": TINY}


def write_config(name, outdir, **overrides):
    """Write <outdir>/<name>.in and its particle file; returns the .in path."""
    os.makedirs(outdir, exist_ok=True)
    kw = dict(CONFIGS[name])
    kw.update(overrides)
    lam, _ = mdl.lam_grid(int(kw["nlam"]), float(kw["lam1"]), float(kw["lam2"]))
    part = os.path.join(outdir, f"analytic_{kw['nlam']}.particle")
    if not os.path.exists(part):
        mdl.write_analytic_particle(part, lam)
    kw["part1"] = os.path.basename(part)
    path = os.path.join(outdir, f"{name}.in")
    from . import keywords
    keywords.write_file(path, kw)
    return path


def load_config(name, workdir, **overrides):
    """Write the config into workdir and build its Model."""
    from . import keywords
    path = write_config(name, workdir, **overrides)
    kw = keywords.parse_file(path)
    return mdl.build_model(kw, basedir=workdir)
