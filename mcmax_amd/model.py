"""Model setup for the transport stage: restates the part of MCMax's
``initialize`` (Init.f) that produces the tables MCRadiation reads.

Covered: keyword defaults (Init.f:66-438), the logarithmic wavelength grid
(Init.f:2581-2584, 2606-2614), the Planck table BB(nlam,0:TMAX) and BBint
(Init.f:2616-2626), the Planck star (Init.f:2628-2632), the theta grid
(Init.f:2788-2815, 2875-2890), the radial grid (Init.f:2787-2852, 2872-2874),
cell volumes and the MIN density with its mass renormalisation
(Init.f:2933-2939, 3158-3166, 3242-3267), particle files (readparticle,
Init.f:4664-4872), the Mueller normalisation and integrals
(Init.f:4800-4814, 3766-3781), Kp (Init.f:3784-3790), the Kabs floor
(Init.f:4069-4078) and the composition w = warg/sum(warg) (Init.f:3458-3463).

Arithmetic keeps the reference's association and its single-precision
places (REAL(i)/REAL(n) in the theta grid, default-REAL constants AU, Rsun,
Msun), so the tables agree with an independent restatement to rounding.
"""
import hashlib
import os

import numpy as np

from . import dust
from . import keywords as kwmod

TMAX = 10000
MAXNG = 16  # dust species the device kernels hold
PI = 3.14159265358979323846264338328
AU = float(np.float32(1.49598e13))       # Modules.f:8 (default REAL literal)
PARSEC = float(np.float32(3.08568025e18))
RSUN = float(np.float32(6.9599e10))
MSUN = float(np.float32(1.98892e33))
LSUN = float(np.float32(3.846e33))

DEFAULTS = {  # Init.f:66-438 (the keys this build reads)
    "startype": "PLANCK", "tstar": 10000.0, "rstar": 2.1, "mstar": 2.5, "distance": 100.0,
    "av": 0.0, "incangle": 45.0, "rin": 0.5, "rout": 1000.0, "nrad": 70, "ntheta": 60,
    "gridrefine": True, "denstype": "POW", "denspow": 1.0, "shpow": 1.3, "sh1au": 0.03,
    "mdust": None, "iter": False, "maxiter": 10, "epsiter": 3.0, "fweight": 0.5,
    "lam1": float(np.float32(0.1)), "lam2": 10000.0, "nlam": 100, "scattype": "ISOTROPIC",
    "tcontact": True, "maxinteract": 0, "nphotdiffuse": 25, "dtdiffuse": 1.0, "idum": -42,
    "ntspan": 0, "ntlev": 0, "nruns": 1, "forcefirst": False, "randomwalk": False,
    "factrw": 10.0,
}

SCAT_HOW = {"NONE": 0, "ISOTROPIC": 1, "FULL": 2, "RAYLEIGH": 2}


class SetupError(ValueError):
    pass


# ---------------------------------------------------------------------------
def planck(T, lam):
    """Subroutines.f:4-22, vectorised.  x is REAL*16 there: long double here."""
    k, c, h = 1.3807e-16, 2.9979e10, 6.6261e-27
    T = np.asarray(T, dtype=np.float64)
    lam = np.asarray(lam, dtype=np.float64)
    nu = c / (lam * 1e-4)
    x = (h * nu) / (k * T)
    num = ((2.0 * h) * ((nu * nu) * nu)) / (c * c)
    xl = np.asarray(x, dtype=np.longdouble)
    with np.errstate(over="ignore"):
        den = np.exp(np.minimum(xl, np.longdouble(1e4))) - np.longdouble(1.0)
    out = (np.asarray(num, dtype=np.longdouble) / den).astype(np.float64)
    return np.where(x > 1e3, 0.0, out)


def integrate(nu, spec):
    """Radiation.f:1-18 trapezoid, summed in wavelength order.  spec may carry
    leading axes; the sum runs over the last one."""
    spec = np.asarray(spec, dtype=np.float64)
    L = np.zeros(spec.shape[:-1])
    for i in range(len(nu) - 1):
        L = L + (abs(nu[i] - nu[i + 1]) * 0.5) * (spec[..., i] + spec[..., i + 1])
    return L


def lam_grid(nlam, lam1, lam2):
    i = np.arange(nlam, dtype=np.float64)
    e = np.log10(lam1) + ((np.log10(lam2) - np.log10(lam1)) * i) / float(nlam - 1)
    lam = np.array([pow(10.0, float(x)) for x in e])
    nu = 2.9979e14 / lam
    return lam, nu


def bb_table(lam, nu):
    T = np.arange(TMAX + 1, dtype=np.float64)
    BB = planck(T[1:, None], lam[None, :])
    BB = np.vstack([np.zeros((1, len(lam))), BB])
    BBint = integrate(nu, BB)
    BBint[0] = 0.0
    return BB, BBint


def theta_grid(ntheta_cells):
    """cos(theta) edges Theta[0..nt] (1 at the pole, 0 at the midplane) and
    cell centres theta_av[0..nt-1]."""
    NT = ntheta_cells + 1
    Th = [0.0] * (NT + 2)
    tav = [0.0] * (NT + 2)
    i = 0
    k = NT // 2
    for j in range(1, k + 1):
        i += 1
        Th[i] = float(np.float32(j - 1) / np.float32(k - 1))
        i += 1
        Th[i] = float(np.cos(((0.5 * PI) * float(j)) / float(k + 1)))
    Th[1:i + 1] = sorted(Th[1:i + 1])
    if i < NT:
        for j in range(i + 1, NT + 1):
            Th[j] = float(np.float32(j - i + 1) / np.float32(NT - i + 1))
    Th[1:NT + 1] = sorted(Th[1:NT + 1])
    for i in range(1, NT):
        tav[i] = float(np.arccos((Th[NT - i + 1] + Th[NT - i]) / 2.0))
    Th[1] = 1.0
    thet = [0.0] * (NT + 2)
    for i in range(2, NT):
        thet[i] = (tav[i - 1] + tav[i]) / 2.0
        Th[i] = float(np.cos((tav[i - 1] + tav[i]) / 2.0))
    Th[NT] = 0.0
    thet[NT] = PI / 2.0
    for i in range(1, NT):
        tav[i] = float(np.arccos((Th[i] + Th[i + 1]) / 2.0))
    return np.array(Th[1:NT + 1]), np.array(tav[1:NT]), np.array(thet[1:NT + 1])


def r_grid(nrad, Rin, Rout, Rstar_cm):
    nR = nrad + 1
    R = np.empty(nR + 1)
    R[0] = Rstar_cm / AU
    for i in range(1, nR + 1):
        R[i] = pow(10.0, float(np.log10(Rin) + ((np.log10(Rout) - np.log10(Rin)) * float(i - 1)) / float(nR - 1)))
    R_av = np.array([AU * pow(10.0, float((np.log10(R[i]) + np.log10(R[i + 1])) / 2.0)) for i in range(nR)])
    return R, R_av


def cell_volumes(R, Theta):
    AU3 = (AU * AU) * AU
    R3 = (R * R) * R
    dR3 = R3[1:] - R3[:-1]
    dTh = Theta[:-1] - Theta[1:]
    return (((4.0 * PI / 3.0) * dR3[:, None]) * dTh[None, :]) * AU3


def min_density(R, Theta, R_av, theta_av, Mdust_g, Rin, Rout, sh1AU, shpow, denspow):
    """denstype MIN (Init.f:3158-3166) renormalised to Mdust (Init.f:3261-3267);
    the void cell i=0 gets 1e-60 (Init.f:2933-2939)."""
    nr, nt = len(R) - 1, len(Theta) - 1
    AU3 = (AU * AU) * AU
    V = cell_volumes(R, Theta)
    dens = np.empty((nr, nt))
    dens[0, :] = 1e-60
    scale = Mdust_g / (AU3 * 2.0 * PI * sh1AU * np.sqrt(PI) * (Rout - Rin))
    r = (R_av[1:, None] * np.sin(theta_av)[None, :]) / AU
    z = (R_av[1:, None] * np.cos(theta_av)[None, :]) / AU
    hr = sh1AU * np.power(r, shpow)
    f1 = np.power(r, -denspow)
    t = z / hr
    f2 = np.exp(-(t * t))
    dens[1:, :] = ((scale * f1) * f2) * np.power(r, -shpow)
    mass_tot = 0.0
    for i in range(1, nr):
        for j in range(nt):
            mass_tot = mass_tot + V[i, j] * dens[i, j]
    dens[1:, :] = (dens[1:, :] * Mdust_g) / mass_tot
    return dens, V


# ---------------------------------------------------------------------------
def hg_bin(g, theta):
    """Init.f:5594-5613: HG averaged over +-0.5 degree (single-precision
    offsets as in the reference)."""
    n = 1000
    a = np.arange(n, dtype=np.float32) / np.float32(n - 1)
    b = (a - np.float32(0.5)) * np.float32(3.1415926536)
    t = theta + b.astype(np.float64) / 180.0
    f = (1.0 - g * g) / np.power((1.0 - (2.0 * g) * np.cos(t)) + g * g, 1.5) / 2.0
    hg = 0.0
    tot = 0.0
    st = np.sin(t)
    fs = f * st
    for k in range(n):
        hg = hg + fs[k]
        tot = tot + st[k]
    return hg / tot


def read_particle(path, lam):
    """readparticle (Init.f:4664-4872) without the Rayleigh/nspike/wasym2
    options: linear interpolation in wavelength onto the grid, power-law
    extrapolation beyond the last tabulated wavelength, then the F11
    normalisation.  Returns Kabs, Ksca, Kext [nlam] and F [nlam,6,180]."""
    with open(path) as f:
        toks = f.read().split()
    vals = np.array([t.replace("d", "e").replace("D", "e") for t in toks], dtype=np.float64)
    rec = 4 + 6 * 180
    if len(vals) % rec:
        raise SetupError(f"{path}: not a particle file ({len(vals)} numbers)")
    nrec = len(vals) // rec
    A = np.array(vals).reshape(nrec, rec)
    ls, Kext_f, Kabs_f, Ksca_f = A[:, 0], A[:, 1], A[:, 2], A[:, 3]
    F_f = A[:, 4:].reshape(nrec, 6, 180)
    nlam = len(lam)
    Kabs = np.zeros(nlam)
    Ksca = np.zeros(nlam)
    Kext = np.zeros(nlam)
    F = np.zeros((nlam, 6, 180))
    i = 0
    l0 = ls[0]
    p0 = (Kext_f[0], Kabs_f[0], Ksca_f[0], F_f[0])
    while i < nlam and l0 >= lam[i]:
        Kext[i], Kabs[i], Ksca[i], F[i] = p0[0], p0[1], p0[2], p0[3]
        i += 1
    rec_i = 1
    done = i >= nlam
    while not done and rec_i < nrec:
        l1 = ls[rec_i]
        p1 = (Kext_f[rec_i], Kabs_f[rec_i], Ksca_f[rec_i], F_f[rec_i])
        while lam[i] <= l1 and lam[i] >= l0:
            d = (lam[i] - l1) / (l0 - l1)
            Kext[i] = p1[0] + (lam[i] - l1) * (p0[0] - p1[0]) / (l0 - l1)
            Ksca[i] = p1[2] + (lam[i] - l1) * (p0[2] - p1[2]) / (l0 - l1)
            Kabs[i] = p1[1] + (lam[i] - l1) * (p0[1] - p1[1]) / (l0 - l1)
            F[i] = p1[3] + (lam[i] - l1) * (p0[3] - p1[3]) / (l0 - l1)
            del d
            i += 1
            if i >= nlam:
                done = True
                break
        l0, p0 = l1, p1
        rec_i += 1
    if i < nlam:  # Init.f:4780-4786
        for j in range(i, nlam):
            Ksca[j] = Ksca[i - 1] * (lam[i - 1] / lam[j]) ** 4
            Kabs[j] = Kabs[i - 1] * (lam[i - 1] / lam[j]) ** 2
            Kext[j] = Kabs[j] + Ksca[j]
            F[j] = F[i - 1]
    normalize_mueller(F)
    return Kabs, Ksca, Kext, F


def normalize_mueller(F):
    """Init.f:4800-4814, in place on F[nlam,6,180]."""
    s = np.sin((PI * (np.arange(1, 181, dtype=np.float64) - 0.5)) / 180.0)
    for l in range(F.shape[0]):
        tot = 0.0
        tot2 = 0.0
        for a in range(180):
            tot = tot + F[l, 0, a] * s[a]
            tot2 = tot2 + s[a]
        F[l] = (tot2 * F[l]) / tot


def mueller_integrals(F):
    """IF11, IF12 (Init.f:3766-3776) for F[..., 6, 180]."""
    thet = (PI * (np.arange(1, 181, dtype=np.float64) - 0.5)) / 180.0
    ps = PI * np.sin(thet)
    IF11 = np.zeros(F.shape[:-2])
    IF12 = np.zeros(F.shape[:-2])
    for a in range(180):
        IF11 = IF11 + (ps[a] * F[..., 0, a]) / 180.0
        IF12 = IF12 + (ps[a] * F[..., 1, a]) / 180.0
    return IF11, IF12


def asymmetry(F):
    """g = sum F11 cos sin / sum F11 sin per (species, lambda) (Init.f:4037-4050)."""
    th = (PI * (np.arange(1, 181, dtype=np.float64) - 0.5)) / 180.0
    c, s = np.cos(th), np.sin(th)
    g = np.zeros(F.shape[:-2])
    tot = np.zeros(F.shape[:-2])
    for a in range(180):
        g = g + (F[..., 0, a] * c[a]) * s[a]
        tot = tot + F[..., 0, a] * s[a]
    return g / tot


def phi_tables():
    i = np.arange(361, dtype=np.float64)
    phi = (PI * (i - 1.0)) / 179.5
    return np.cos(2.0 * phi), np.sin(2.0 * phi)


def write_analytic_particle(path, lam, kappa0=1e3, albedo=0.5, g=0.5):
    """The analytic grain of the small plain-method case: kappa_abs =
    kappa0 (lam/1um)^-1 above 1 um and flat below, kappa_sca =
    kappa_abs*albedo/(1-albedo), F11 = HG(g) bin averages, F12 = F34 = 0
    (unpolarised), F22 = F11, F33 = F44 = F11, tabulated on `lam`."""
    thet = (np.arange(1, 181) - 0.5) * PI / 180.0
    F11 = np.array([hg_bin(g, t) for t in thet])
    with open(path, "w") as f:
        for l in lam:
            l = float(l)
            ka = kappa0 * (l if l > 1.0 else 1.0) ** -1
            ks = ka * albedo / (1.0 - albedo)
            f.write(f"{l!r} {ka + ks!r} {ka!r} {ks!r}\n")
            for k in range(6):
                row = F11 if k in (0, 2, 3, 5) else np.zeros(180)
                f.write(" ".join(repr(float(x)) for x in row) + "\n")


# ---------------------------------------------------------------------------
# computepart results by (wavelength grid, species inputs): the models of a
# sweep share their dust, which is then computed once per process
_PART_CACHE = {}
_PART_CACHE_MAX = 4


def _parts_key(lam, specs):
    h = hashlib.sha256(np.ascontiguousarray(lam, dtype=np.float64).tobytes())
    for sp in specs:
        for k in sorted(sp):
            v = sp[k]
            h.update(k.encode())
            h.update(np.ascontiguousarray(v).tobytes() if isinstance(v, (np.ndarray, list, tuple)) else repr(v).encode())
    return h.hexdigest()


def compute_parts_cached(tr, lam, specs):
    """tr.compute_parts(lam, specs), remembered for the same inputs"""
    key = _parts_key(lam, specs)
    if key not in _PART_CACHE:
        if len(_PART_CACHE) >= _PART_CACHE_MAX:
            _PART_CACHE.pop(next(iter(_PART_CACHE)))
        _PART_CACHE[key] = tr.compute_parts(lam, specs)
    return _PART_CACHE[key]


class Model:
    """All setup arrays of one disk model (numpy, float64)."""

    def summary(self):
        return (f"nlam={self.nlam} ng={self.ng} nr={self.nr} nt={self.nt} "
                f"Lstar={self.Lstar:.6e} Mdust={self.Mdust_g:.4e} g")


def build_model(kw, basedir=".", tr=None, device=0):
    """Build a Model from a keyword dict (see keywords.parse_file).
    computepart particles are computed on the device: through the Transport
    tr when given, else through one opened on `device` for the call."""
    p = dict(DEFAULTS)
    for k, v in kw.items():
        p[k] = v
    if str(p["startype"]).upper() != "PLANCK":
        raise SetupError("this build supports startype=PLANCK")
    if str(p["denstype"]).upper() != "MIN":
        raise SetupError("this build supports denstype=MIN")
    if p.get("gridrefine", True):
        raise SetupError("gridrefine must be .false. (the build keeps the grid fixed)")
    if not p.get("tcontact", True):
        raise SetupError("tcontact=.false. is not supported by this build")
    if p.get("mdust") is None:
        raise SetupError("mdust is required")
    m = Model()
    m.keywords = p
    m.nlam = int(p["nlam"])
    m.lam, m.nu = lam_grid(m.nlam, float(p["lam1"]), float(p["lam2"]))
    m.BB, m.BBint = bb_table(m.lam, m.nu)
    m.Tstar = float(p["tstar"])
    m.Rstar_cm = float(p["rstar"]) * RSUN
    m.Mstar_g = float(p["mstar"]) * MSUN
    m.distance_cm = float(p["distance"]) * PARSEC
    m.Fstar = (PI * (m.Rstar_cm * m.Rstar_cm)) * planck(m.Tstar, m.lam)
    m.Lstar = float(integrate(m.nu, m.Fstar))
    m.incangle = float(p["incangle"])
    m.Mdust_g = float(p["mdust"]) * MSUN
    m.Rin, m.Rout = float(p["rin"]), float(p["rout"])
    if int(p.get("ntspan", 0)) * int(p.get("ntlev", 0)) != 0:
        raise SetupError("ntspan/ntlev refinement is not supported")
    if bool(p.get("radpress", False)):
        # the pass would also have to tally the radiation force per cell
        # (MCRadiation.f:842-880) for DiskStructure
        raise SetupError("radpress (radiation pressure from the Monte Carlo pass) is not supported")
    m.Theta, m.theta_av, m.thet = theta_grid(int(p["ntheta"]))
    m.R, m.R_av = r_grid(int(p["nrad"]), m.Rin, m.Rout, m.Rstar_cm)
    m.nr, m.nt = len(m.R) - 1, len(m.Theta) - 1
    m.dens, m.V = min_density(m.R, m.Theta, m.R_av, m.theta_av, m.Mdust_g, m.Rin, m.Rout,
                              float(p["sh1au"]), float(p["shpow"]), float(p["denspow"]))
    m.scat_how = SCAT_HOW[str(p["scattype"]).upper()]
    # particles: part<NN> files and computepart<NN> particles (each split
    # into its ngrains size bins, Init.f:1318-1487), in index order
    files = {int(k[4:]): v for k, v in p.items() if k.startswith("part") and k[4:].isdigit()}
    cps = dust.particle_keywords(p)
    idx = sorted(set(files) | set(cps))
    if not idx:
        raise SetupError("no part<NN> or computepart<NN> keyword")
    if idx != list(range(1, len(idx) + 1)) or set(files) & set(cps):
        raise SetupError("particles must be numbered 1..n, one keyword kind per number")
    arg_abun = any(k.startswith("abun") and k[4:].isdigit() for k in p)
    species, warg = [], []
    for k in idx:
        wk = float(p.get(f"abun{k}", 1.0))
        if k in files:
            species.append(("file", files[k]))
            warg.append(wk)
        else:
            specs, wj = dust.species_specs(cps[k], m.lam, basedir)
            species += [("compute", s_) for s_ in specs]
            warg += [wk * w_ for w_ in wj]
    ng = len(species)
    if ng > MAXNG:
        raise SetupError(f"{ng} dust species; this build holds at most {MAXNG}")
    m.ng = ng
    m.Kabs = np.zeros((ng, m.nlam))
    m.Ksca = np.zeros((ng, m.nlam))
    m.Kext = np.zeros((ng, m.nlam))
    m.F = np.zeros((ng, m.nlam, 6, 180))
    m.rho_grain = np.zeros(ng)
    comp = [i for i, (kind, _) in enumerate(species) if kind == "compute"]
    m.dust_stats = None
    if comp:
        own = tr is None and _parts_key(m.lam, [species[i][1] for i in comp]) not in _PART_CACHE
        if own:
            from .transport import Transport
            tr = Transport(device)
        try:
            res, m.dust_stats = compute_parts_cached(tr, m.lam, [species[i][1] for i in comp])
        finally:
            if own:
                tr.close()
        for i, r in zip(comp, res):
            m.Kabs[i], m.Ksca[i], m.F[i], m.rho_grain[i] = r["Kabs"], r["Ksca"], r["F"], r["rho"]
    for i, (kind, fname) in enumerate(species):
        if kind == "file":
            path = fname if os.path.isabs(fname) else os.path.join(basedir, fname)
            m.Kabs[i], m.Ksca[i], _, m.F[i] = read_particle(path, m.lam)
    m.Kabs = np.maximum(m.Kabs, 0.0)
    m.Ksca = np.maximum(m.Ksca, 0.0)
    m.Kext = m.Kabs + m.Ksca  # Init.f:3428-3436
    finish_opacities(m, p)
    # keyword maxinteract != 0 switches on the reference's interaction
    # roulette (Init.f:637-639, MCRadiation.f:568-571); 0 leaves it off
    m.maxinteract = int(p.get("maxinteract", 0))
    m.randomwalk = bool(p.get("randomwalk", False))
    m.factRW = float(p.get("factrw", 10.0))
    # Init.f:2266-2272: without any abundance keyword every species weighs the same
    warg = np.array(warg) if arg_abun else np.ones(ng)
    m.w = np.broadcast_to(warg / warg.sum(), (m.nr, m.nt, ng)).copy()
    return m


def finish_opacities(m, p=None):
    """Scattering options, Mueller integrals, Kp, Kabs floor (Init.f:3752-3799,
    4069-4078) for opacities already on the grid."""
    if m.scat_how == 0:
        m.Ksca[:] = 0.0
        m.Kext[:] = m.Kabs
    elif m.scat_how == 1:
        m.F[:, :, 0, :] = 1.0
    m.IF11, m.IF12 = mueller_integrals(m.F)
    m.cos2phi, m.sin2phi = phi_tables()
    m.gsym = asymmetry(m.F)
    m.Kp = np.zeros((m.ng, TMAX + 1))
    for s in range(m.ng):
        m.Kp[s] = integrate(m.nu, m.BB * m.Kabs[s][None, :])
    floor = m.Kabs == 0.0
    m.Kabs[floor] = 1e-8
    m.Kext[floor] = m.Kabs[floor] + m.Ksca[floor]
