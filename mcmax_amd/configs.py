"""Keyword files for the benchmark and test workloads.

small   -- plain method, CPU-runnable: 4000 K / 2 Rsun / 0.5 Msun star, dust
           mass 1e-4 Msun, Rin 0.1 AU, Rout 100 AU, Sigma ~ r^-1, MIN vertical
           profile exp(-(z/h)^2) with h/r = 0.05 at 1 AU and flaring 1.125
           (the reference's MIN convention, Init.f:3163: sh1AU is the h of
           exp(-(z/h)^2)), 100 r x 60 theta cells, 100 wavelengths 0.1-1000 um,
           one analytic grain (kappa_abs = 1e3 cm^2/g (lam/um)^-1 above 1 um,
           flat below; albedo 0.5; HG g = 0.5), FULL scattering.
production -- the same star; dust mass 1e-3 Msun; 200 r x 100 theta cells,
           200 wavelengths; hydrostatic iteration (iter, maxiter = 5).  Dust:
           one computepart particle from a constant refractive index
           m = 1.7 + 0.03i, sizes 0.05-1000 um with n(a) ~ a^-3.5, DHS with
           fmax = 0.8, no porous blend, split into ngrains = 10 species whose
           mass fractions follow the MRN distribution (abun01 = 1 keeps the
           split's MRN weights, Init.f:1452-1471); each species samples 10
           sizes and 10 vacuum fractions (the input file's ns and nf).
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
PRODUCTION = dict(SMALL, mdust=1e-3, nrad=200, ntheta=100, nlam=200, iter=True, maxiter=5, abun01=1.0, **{
    "computepart01:file": "dhs_const_n1.7_k0.03.dat", "computepart01:amin": 0.05, "computepart01:amax": 1000.0,
    "computepart01:apow": 3.5, "computepart01:fmax": 0.8, "computepart01:ngrains": 10,
    "computepart01:blend": False})
DHS_NK = (1.7, 0.03)  # the production dust's refractive index
DHS_NS, DHS_NF = 10, 10
TINY = dict(SMALL, nrad=20, ntheta=10, nlam=40, mdust=1e-4)

CONFIGS = {"small": SMALL, "production": PRODUCTION, "tiny": TINY}


def write_config(name, outdir, **overrides):
    """Write <outdir>/<name>.in and its particle file; returns the .in path."""
    os.makedirs(outdir, exist_ok=True)
    kw = dict(CONFIGS[name])
    kw.update(overrides)
    cp = kw.get("computepart01:file")
    if cp:
        # the computepart input (ComputePart.f:64-90) and a constant-n,k
        # lnk table that spans every wavelength of the grid
        lnk = "const_n{:g}_k{:g}.lnk".format(*DHS_NK)
        with open(os.path.join(outdir, lnk), "w") as f:
            f.write("# constant refractive index\n")
            for l in (1e-3, 1e5):
                f.write(f"{l:.6e} {DHS_NK[0]:.6e} {DHS_NK[1]:.6e}\n")
        with open(os.path.join(outdir, cp), "w") as f:
            f.write(f"{DHS_NS}\t# ns\n{DHS_NF}\t# nf\n'{lnk}'\n1.0 3.0\n")
    else:
        lam, _ = mdl.lam_grid(int(kw["nlam"]), float(kw["lam1"]), float(kw["lam2"]))
        part = os.path.join(outdir, f"analytic_{kw['nlam']}.particle")
        if not os.path.exists(part):
            mdl.write_analytic_particle(part, lam)
        kw["part1"] = os.path.basename(part)
    path = os.path.join(outdir, f"{name}.in")
    from . import keywords
    keywords.write_file(path, kw)
    return path


def load_config(name, workdir, tr=None, **overrides):
    """Write the config into workdir and build its Model (a computepart
    particle is computed on the device, through tr when given)."""
    from . import keywords
    path = write_config(name, workdir, **overrides)
    kw = keywords.parse_file(path)
    return mdl.build_model(kw, basedir=workdir, tr=tr)
