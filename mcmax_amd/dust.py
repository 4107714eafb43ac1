"""Dust species from optical constants: the host side of MCMax's
``computepartNN:`` particles (Init.f:755-790, 1318-1487, 3372; ComputePart.f
for the FILE standard).  The opacities themselves are computed on the device
by ``Transport.compute_parts`` (mcx_compute_parts).

Covered here:
- the keywords of one computepart particle and their defaults
  (Init.f:342-352, 755-790);
- its input file for standard = FILE (ComputePart.f:64-90): ns, nf, then per
  material an lnk file and "mass fraction, density";
- the lnk reader with its interpolation and extrapolation rules
  (readrefindCP, ComputePart.f:1102-1175);
- the effective-medium blend with vacuum (Blender, ComputePart.f:1177-1200,
  and its use at ComputePart.f:228-246);
- the split of a particle into ngrains size bins (Init.f:1318-1487) and the
  MRN mass fractions of the bins (gsd_MRN, MRN.f:21-230).

Single-precision places of the reference (its default REAL) are kept in
numpy float32.
"""
import math
import os

import numpy as np

from . import keywords as kwmod


class DustError(ValueError):
    pass


F32 = np.float32

# Init.f:342-352 (apow above 100 means "use mrn_index", Init.f:1322)
CP_DEFAULTS = dict(amin=0.1, amax=0.1, apow=1e5, fmax=0.0, blend=True, porosity=0.25, ngrains=1,
                   nsubgrains=1, norm_abun=-1.0, standard="FILE")
MRN_INDEX = 3.5  # Init.f:174


def _tokens(line):
    return line.replace(",", " ").split()


def _num(tok):
    v = kwmod.parse_value(tok)
    if not isinstance(v, (int, float)):
        raise DustError(f"not a number: {tok!r}")
    return float(v)


# ---------------------------------------------------------------------------
def read_lnk(path, grid, loglog=False):
    """readrefindCP: n and k (float32) on the wavelength grid.  Leading lines
    that do not read as three numbers are skipped; k below 1e-8 is raised to
    1e-8; grid points at or below the first wavelength take its values;
    inside the table log-log interpolation; beyond its end the dielectric
    rule (n constant, k ~ 1/lambda), or log-log extrapolation of both when
    loglog is set (keyword lnkloglog)."""
    grid = np.asarray(grid, dtype=np.float64)
    n = len(grid)
    rows = []
    with open(path) as f:
        started = False
        for line in f:
            t = _tokens(line)
            try:
                vals = [float(kwmod.parse_value(x)) for x in t[:3]]
                if len(vals) < 3:
                    raise ValueError
            except (ValueError, TypeError):
                if started:
                    raise DustError(f"{path}: unreadable line {line!r}")
                continue
            started = True
            rows.append(vals)
    if not rows:
        raise DustError(f"{path}: no wavelength, n, k rows")
    e1 = np.zeros(n, F32)
    e2 = np.zeros(n, F32)
    x0, y01, y02 = (F32(v) for v in rows[0])
    if y02 < F32(1e-8):
        y02 = F32(1e-8)
    i = 0
    while i < n and float(x0) >= grid[i]:
        e1[i], e2[i] = y01, y02
        i += 1
    for r in rows[1:]:
        if i >= n:
            break
        x1, y11, y12 = (F32(v) for v in r)
        if y12 < F32(1e-8):
            y12 = F32(1e-8)
        while i < n and grid[i] <= float(x1) and grid[i] > float(x0):
            b = math.log10(grid[i] / float(x1))
            d = float(np.log10(F32(x0 / x1)))
            e1[i] = F32(10.0 ** (float(np.log10(y11)) + b * float(np.log10(F32(y01 / y11))) / d))
            e2[i] = F32(10.0 ** (float(np.log10(y12)) + b * float(np.log10(F32(y02 / y12))) / d))
            i += 1
        x0, y01, y02 = x1, y11, y12
    if i < n:
        if i == 0:
            raise DustError(f"{path}: the table starts beyond the wavelength grid")
        if loglog:
            if i < 2:
                raise DustError(f"{path}: log-log extrapolation needs two grid points inside the table")
            for j in range(i, n):
                lg = math.log10(grid[i - 2] / grid[j]) / math.log10(grid[i - 2] / grid[i - 1])
                e1[j] = F32(10.0 ** (float(np.log10(e1[i - 2])) + float(np.log10(F32(e1[i - 1] / e1[i - 2]))) * lg))
                e2[j] = F32(10.0 ** (float(np.log10(e2[i - 2])) + float(np.log10(F32(e2[i - 1] / e2[i - 2]))) * lg))
        else:
            for j in range(i, n):
                e1[j] = e1[i - 1]
                e2[j] = F32(float(e2[i - 1]) * grid[i - 1] / grid[j])
    return e1, e2


def blend(e1s, e2s, abun):
    """Blender: the Bruggeman effective medium of materials (n, k) with
    volume fractions abun, by 100 fixed-point iterations from m = 1."""
    m = [complex(float(a), float(b)) for a, b in zip(e1s, e2s)]
    ab = [float(F32(x)) for x in abun]
    mm = complex(1.0, 0.0)
    me = mm
    for _ in range(100):
        s = complex(0.0, 0.0)
        for mj, aj in zip(m, ab):
            s = s + ((mj * mj - mm * mm) / (mj * mj + 2.0 * (mm * mm))) * aj
        me = (2.0 * s + 1.0) / (1.0 - s)
        me = mm * complex(np.sqrt(np.complex128(me)))
        mm = me
    return F32(me.real), F32(me.imag)


def read_computepart_file(path):
    """The FILE standard's input: (ns, nf, [(lnk file, mass fraction,
    density), ...]).  Lines starting with '*' are skipped (ignorestar)."""
    with open(path) as f:
        lines = [l for l in f if l.strip() and not l.lstrip().startswith("*")]
    if len(lines) < 2:
        raise DustError(f"{path}: needs ns and nf")
    ns = int(_num(_tokens(lines[0])[0]))
    nf = int(_num(_tokens(lines[1])[0]))
    mats = []
    for k in range(2, len(lines), 2):
        text = lines[k].strip()
        name = text[1:].split(text[0])[0] if text[0] in "'\"" else text.split()[0]
        if k + 1 >= len(lines):
            raise DustError(f"{path}: material {name!r} without fraction and density")
        t = _tokens(lines[k + 1])
        mats.append((name, _num(t[0]), _num(t[1])))
    if not mats:
        raise DustError(f"{path}: no materials")
    return ns, nf, mats


# ---------------------------------------------------------------------------
def split_sizes(amin, amax, ngrains):
    """Init.f:1379-1387, 1446-1449: the size bins of a particle split into
    ngrains species (ratios real(i)/real(ngrains) in single precision)."""
    bins = []
    for i in range(1, ngrains + 1):
        lo = amin if i == 1 else 10.0 ** (math.log10(amin) + math.log10(amax / amin) *
                                          float(F32(i - 1) / F32(ngrains)))
        hi = 10.0 ** (math.log10(amin) + math.log10(amax / amin) * float(F32(i) / F32(ngrains)))
        bins.append((lo, hi))
    return bins


def mrn_weights(rgrain_um, rmin, rmax, index):
    """gsd_MRN (MRN.f:21-230): mass fractions of grains of radii rgrain_um
    (log-centred bins) for n(a) ~ a^-index between rmin and rmax."""
    r = np.asarray(rgrain_um, dtype=np.float64)
    n = len(r)
    if n < 2:
        raise DustError("the MRN split needs at least two grains")
    lr = np.zeros(n + 1)
    for ii in range(1, n):
        lr[ii] = (math.log(r[ii - 1]) + math.log(r[ii])) / 2.0
    lr[0] = 2 * math.log(r[0]) - lr[1]
    lr[n] = 2 * math.log(r[n - 1]) - lr[n - 1]
    rint = np.exp(lr)
    for ii in range(n):
        if rint[ii] >= r[ii] or rint[ii + 1] <= r[ii]:
            raise DustError("grain sizes do not fit the MRN grid (MRN.f:99-107)")
    r_index = 4.0 - index
    r_factor = 3.0 / r_index if r_index != 0.0 else 0.0

    def mass(a, b):
        if abs(index - 4.0) <= 1e-6:
            return math.log(b) - math.log(a)
        return r_factor * (b ** r_index - a ** r_index)

    ab = np.zeros(n)
    for ii in range(n):
        lo, hi = rint[ii], rint[ii + 1]
        if lo >= rmax or hi <= rmin:
            continue
        ab[ii] = mass(max(lo, rmin), min(hi, rmax))
    lo, hi = rint[0], rint[n]
    if lo >= rmax or hi <= rmin:
        raise DustError("no grains between the MRN limits (MRN.f:184)")
    norm = mass(max(lo, rmin), min(hi, rmax))
    ab = ab / norm
    ab[ab <= 1e-100] = 0.0
    if abs(ab.sum() - 1.0) >= 1e-6:
        raise DustError("MRN abundances do not sum to one (MRN.f:205-212)")
    return ab


# ---------------------------------------------------------------------------
def particle_keywords(p):
    """The computepart particles of a keyword dict: {index: settings}."""
    out = {}
    for k, v in p.items():
        if not k.startswith("computepart") or ":" not in k:
            continue
        idx = int(k[len("computepart"):k.index(":")])
        sub = k[k.index(":") + 1:]
        d = out.setdefault(idx, dict(CP_DEFAULTS, abun={}))
        if sub == "file":
            d["file"] = str(v)
        elif sub in ("amin", "amax", "apow", "fmax", "porosity", "norm_abun"):
            d[sub] = float(v)
        elif sub in ("ngrains", "nsubgrains"):
            d[sub] = int(v)
        elif sub == "blend":
            d["blend"] = bool(v)
        elif sub == "standard":
            d["standard"] = str(v).upper()
        elif sub.startswith("abun") and sub[4:].isdigit():
            d["abun"][int(sub[4:])] = float(v)
        else:
            raise DustError(f"computepart keyword not supported: {sub!r} (ComputePart keywords, Init.f:755-790)")
    return out


def species_specs(d, lam, basedir=".", mrn_index=MRN_INDEX, loglog=False):
    """The device specs (Transport.compute_parts) of one computepart
    particle after its split into ngrains species, and the bins' relative
    mass weights (MRN over the bins; one weight for ngrains = 1)."""
    if d["standard"] != "FILE":
        raise DustError(f"computepart standard {d['standard']!r} is not supported (FILE only)")
    if "file" not in d:
        raise DustError("computepart needs a file")
    path = d["file"] if os.path.isabs(d["file"]) else os.path.join(basedir, d["file"])
    ns, nf, mats = read_computepart_file(path)
    apow = d["apow"] if d["apow"] <= 100.0 else mrn_index
    fmax = float(F32(d["fmax"]))
    nlam = len(lam)
    e1 = np.zeros((len(mats), nlam), F32)
    e2 = np.zeros((len(mats), nlam), F32)
    frac = np.zeros(len(mats), F32)
    rho = np.zeros(len(mats), F32)
    fdir = os.path.dirname(path)
    for j, (name, fr, rh) in enumerate(mats):
        lp = name if os.path.isabs(name) else os.path.join(fdir, name)
        if not os.path.exists(lp):
            lp = os.path.join(basedir, name)
        e1[j], e2[j] = read_lnk(lp, lam, loglog=loglog)
        if d["abun"].get(j + 1, -1.0) >= 0.0:
            fr = d["abun"][j + 1]
        rho[j] = F32(rh)
        frac[j] = F32(F32(fr) / rho[j])
    tot = F32(0.0)
    for x in frac:
        tot = F32(tot + x)
    mscale = rscale = 1.0
    if d["norm_abun"] > 0.0:
        mscale = 0.0
        for x, rh in zip(frac, rho):
            mscale = mscale + float(F32(x * rh)) / d["norm_abun"]
        rscale = mscale ** (1.0 / 3.0)
    frac = (frac / tot).astype(F32)
    if d["blend"]:
        por = F32(d["porosity"])
        fr = np.concatenate([frac * F32(1.0 - por), [por]]).astype(F32)
        rh = np.concatenate([rho, [F32(0.0)]]).astype(F32)
        b1 = np.zeros(nlam, F32)
        b2 = np.zeros(nlam, F32)
        for i in range(nlam):
            b1[i], b2[i] = blend(list(e1[:, i]) + [F32(1.0)], list(e2[:, i]) + [F32(0.0)], fr)
        rho_av = F32(0.0)
        for a, b in zip(fr, rh):
            rho_av = F32(rho_av + a * b)
        e1, e2 = b1[None, :], b2[None, :]
        frac = np.array([1.0], F32)
        rho = np.array([rho_av], F32)
    ng = int(d["ngrains"])
    bins = split_sizes(d["amin"], d["amax"], ng)
    specs = [dict(e1=e1, e2=e2, frac=frac, rho=rho, amin=lo, amax=hi, apow=apow, fmax=fmax, ns=ns, nf=nf,
                  rscale=rscale, mscale=mscale) for lo, hi in bins]
    if ng == 1:
        weights = np.ones(1)
    else:
        rg = [math.sqrt(lo * hi) * 1e-4 * 1e4 for lo, hi in bins]
        weights = mrn_weights(rg, d["amin"], d["amax"], apow)
    return specs, weights
