"""ctypes binding of the oracle (liboracle.so, built from mcx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the mcmax_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")
TMAX = 10000
NANGLE = 30

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class OrcModel(C.Structure):
    _fields_ = [
        ("nlam", C.c_int), ("lam", _dp), ("nu", _dp), ("BB", _dp), ("ng", C.c_int),
        ("Kabs", _dp), ("Ksca", _dp), ("Kext", _dp), ("F", _dp), ("IF11", _dp), ("IF12", _dp),
        ("Kp", _dp), ("Fstar", _dp), ("Lstar", C.c_double), ("Rstar_cm", C.c_double),
        ("Tstar", C.c_double), ("nr", C.c_int), ("nt", C.c_int), ("R", _dp), ("Theta", _dp),
        ("dens", _dp), ("V", _dp), ("w", _dp), ("scat_how", C.c_int), ("cos2phi", _dp),
        ("sin2phi", _dp), ("rndw", C.c_int), ("factRW", C.c_double), ("gsym", _dp), ("thet", _dp),
    ]


class OrcRun(C.Structure):
    _fields_ = [("nphot", C.c_int64), ("rng", C.c_int), ("seed", C.c_uint64), ("maxinteract", C.c_int),
                ("inflight", C.c_int64), ("staged", C.c_int), ("stage0", C.c_int64), ("stage_ratio", C.c_int),
                ("epoch_deposits", C.c_int64)]


class OrcTally(C.Structure):
    _fields_ = [
        ("EJv", _dp), ("EJv2", _dp), ("Eabs", _dp), ("T", _dp), ("Ni", _ip),
        ("spec", _dp), ("scatspec", _dp), ("ispec", _ip), ("iscatspec", _ip),
        ("nemit", C.c_int64), ("ninteract", C.c_int64), ("nabs", C.c_int64), ("nscat", C.c_int64),
        ("nescape", C.c_int64), ("nstarhit", C.c_int64), ("nsteps", C.c_int64), ("nraar", C.c_int64),
        ("nmaxinteract", C.c_int64), ("tautot", C.c_double), ("Eescape", C.c_double),
        ("nev1", _ip), ("nev2", _ip), ("nrw", C.c_int64), ("pkt_steps", C.POINTER(C.c_int64)),
    ]


class Ran1State(C.Structure):
    _fields_ = [("idum", C.c_int32), ("iy", C.c_int32), ("iv", C.c_int32 * 32)]


class OrcStruct(C.Structure):
    _fields_ = [("nr", C.c_int), ("nt", C.c_int), ("ng", C.c_int), ("R_av", _dp), ("theta_av", _dp), ("V", _dp),
                ("Mstar", C.c_double), ("shscale", _dp), ("gshscale", _dp), ("settle", _ip), ("gascoupled", _ip),
                ("mpstr", C.c_int), ("FradZ", _dp)]


class OrcObs(C.Structure):
    _fields_ = [("angle", C.c_double), ("nphi", C.c_int), ("nj0", C.c_int), ("nphot", C.c_int),
                ("nphotstar", C.c_int), ("nexits", C.c_int), ("usepol", C.c_int), ("tracestar", C.c_int),
                ("traceemis", C.c_int), ("tracescat", C.c_int), ("pfstop", C.c_double)]


def temp_average(T, Temp, fw, first=False):
    """TempAverage (DiskStructure.f:25-79): returns the new averaged field."""
    T = np.ascontiguousarray(T, dtype=np.float64)
    out = np.ascontiguousarray(Temp if Temp is not None else np.zeros_like(T), dtype=np.float64).copy()
    load().orc_temp_average(T.size, p(T), p(out), float(fw), int(first))
    return out


def disk_structure(m, state, Temp, shscale=None, gshscale=None, settle=None, gascoupled=None, mpstr=False,
                   FradZ=None):
    """DiskStructure (DiskStructure.f:86-453) + CheckCells on a copy of
    state for the grid of m (nr, nt, ng, R_av, theta_av, V, Mstar_g),
    state = dict(dens, dens0, gasdens [nr,nt], w, w0 [nr,nt,ng]).
    Returns (new state, rho [ng+1,nr,nt], Runge-Kutta steps)."""
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    out = {k: f(state[k]).copy() for k in ("dens", "dens0", "gasdens", "w", "w0")}
    keep = dict(Rav=f(m.R_av), tav=f(m.theta_av), V=f(m.V),
                sh=f(np.ones(m.nr) if shscale is None else shscale),
                gsh=f(np.ones((m.ng, m.nr)) if gshscale is None else gshscale),
                st=i32(np.zeros(m.ng) if settle is None else settle),
                gc=i32(np.zeros(m.ng) if gascoupled is None else gascoupled),
                T=f(Temp), rho=np.zeros((m.ng + 1, m.nr, m.nt)))
    fz = None if FradZ is None else f(FradZ)
    st = OrcStruct(nr=m.nr, nt=m.nt, ng=m.ng, R_av=p(keep["Rav"]), theta_av=p(keep["tav"]), V=p(keep["V"]),
                   Mstar=m.Mstar_g, shscale=p(keep["sh"]), gshscale=p(keep["gsh"]), settle=p(keep["st"]),
                   gascoupled=p(keep["gc"]), mpstr=int(mpstr), FradZ=p(fz) if fz is not None else None)
    n = load().orc_disk_structure(C.byref(st), p(keep["T"]), p(out["dens"]), p(out["dens0"]),
                                    p(out["gasdens"]), p(out["w"]), p(out["w0"]), p(keep["rho"]))
    if n < 0:
        raise RuntimeError("oracle disk structure failed")
    return out, keep["rho"], n


def vertical_step(y0, z1, z2, T1, T2, scale, Mstar, R_av, F1=0.0, F2=0.0):
    """One odeint call of DiskStructure: returns (y(z2), Runge-Kutta steps)."""
    n = C.c_int(0)
    y = load().orc_vertical_step(y0, z1, z2, T1, T2, F1, F2, scale, Mstar, R_av, C.byref(n))
    return y, n.value


def compare_cells(nr, nt, a, b):
    """compareCells (DiskStructure.f:553-584) of two states, each a dict of
    T, dT, dens, ddens, EJv, dEJv [nr,nt]; returns (errT, errRho, errE)."""
    keys = ("T", "dT", "dens", "ddens", "EJv", "dEJv")
    arr = [np.ascontiguousarray(d[k], dtype=np.float64) for d in (a, b) for k in keys]
    e = [C.c_double() for _ in range(3)]
    load().orc_compare_cells(nr, nt, *[p(x) for x in arr], *[C.cast(C.byref(x), _dp) for x in e])
    return tuple(x.value for x in e)


def make_obs(angle_deg=45.0, nphi=45, nj0=2, nphot=1000, nphotstar=1000, nexits=5, usepol=True,
             tracestar=True, traceemis=True, tracescat=True, pfstop=0.25):
    """Observation defaults of determineoutput (Telescope.f:590-657)."""
    import math
    return OrcObs(angle=math.pi * angle_deg / 180.0, nphi=nphi, nj0=nj0, nphot=nphot, nphotstar=nphotstar,
                  nexits=nexits, usepol=int(usepol), tracestar=int(tracestar), traceemis=int(traceemis),
                  tracescat=int(tracescat), pfstop=pfstop)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def gauleg2(x1, x2, n):
    """gauleg2 nodes and weights (float32)."""
    lib = load()
    x, w = np.zeros(n, np.float32), np.zeros(n, np.float32)
    fp = C.POINTER(C.c_float)
    lib.orc_gauleg2(x1, x2, x.ctypes.data_as(fp), w.ctypes.data_as(fp), n)
    return x, w


def dhs_mu(nh=90):
    """The cosines ComputePart hands the layered sphere (ComputePart.f:311-314)."""
    i = np.arange(1, nh + 1, dtype=np.float32)
    th = ((i - np.float32(0.5)) / np.float32(nh) * np.float32(3.1415926536)).astype(np.float64) / 2.0
    return np.cos(th).astype(np.float32)


def dmilay(rcore, rshell, wvno, m_shell, m_core=1.0 + 0.0j, mu=None):
    """DMiLay: returns dict(err, qext, qsca, gqsc, m1, m2, s21, d21 [numang, 2]);
    refractive indices are given as n - i k (imaginary part <= 0)."""
    lib = load()
    mu = _f32(dhs_mu() if mu is None else mu)
    na = len(mu)
    fp = C.POINTER(C.c_float)
    out = {k: np.zeros((na, 2), np.float32) for k in ("m1", "m2", "s21", "d21")}
    q = [C.c_float() for _ in range(3)]
    err = lib.orc_dmilay(rcore, rshell, wvno, m_shell.real, m_shell.imag, m_core.real, m_core.imag,
                         mu.ctypes.data_as(fp), na, *[C.byref(x) for x in q],
                         *[out[k].ctypes.data_as(fp) for k in ("m1", "m2", "s21", "d21")])
    out.update(err=err, qext=q[0].value, qsca=q[1].value, gqsc=q[2].value)
    return out


def meerhoff_mie(rad, lam, e1, e2, na=180):
    """MeerhoffMie: (Csca, Cext, F11, F12, F33, F34) for one homogeneous
    sphere, m = e1 - i e2; F from the forward direction in 1-degree bins."""
    lib = load()
    F = [np.zeros(na) for _ in range(4)]
    cs, ce = C.c_double(), C.c_double()
    rc = lib.orc_meerhoff_mie(rad, lam, e1, e2, C.byref(cs), C.byref(ce), *[p(x) for x in F], na)
    if rc:
        raise ValueError(f"orc_meerhoff_mie returned {rc}")
    return (cs.value, ce.value, *F)


def compute_part(lam, e1, e2, frac, rho, amin, amax, apow, fmax, ns, nf, rscale=1.0, mscale=1.0):
    """ComputePart (FILE standard, DHS): e1, e2 [nm, nlam]; returns dict with
    Kabs, Ksca, Kext [nlam], F [nlam, 6, 180], rho and the evaluation counts."""
    lib = load()
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    nlam = len(lam)
    e1, e2 = _f32(np.atleast_2d(e1)), _f32(np.atleast_2d(e2))
    nm = e1.shape[0]
    frac, rho = _f32(frac), _f32(rho)
    fp = C.POINTER(C.c_float)
    K = [np.zeros(nlam) for _ in range(3)]
    F = np.zeros((nlam, 6, 180))
    ra = C.c_double()
    counts = (C.c_long * 3)()
    rc = lib.orc_compute_part(nlam, p(lam), nm, e1.ctypes.data_as(fp), e2.ctypes.data_as(fp),
                              frac.ctypes.data_as(fp), rho.ctypes.data_as(fp), amin, amax, apow, fmax, ns, nf,
                              rscale, mscale, p(K[0]), p(K[1]), p(K[2]), p(F), C.byref(ra), counts)
    if rc:
        raise ValueError(f"orc_compute_part returned {rc}")
    return dict(Kabs=K[0], Ksca=K[1], Kext=K[2], F=F, rho=ra.value, ndhs=counts[0], nhom=counts[1],
                nerr=counts[2])


def build(force=False):
    """Compile liboracle.so with gcc (no FMA contraction: plain IEEE FP64)."""
    srcs = [os.path.join(HERE, f) for f in ("mcx_oracle.c", "mcx_oracle_dust.c", "mcx_oracle.h")]
    if force or not os.path.exists(SO) or os.path.getmtime(SO) < max(os.path.getmtime(s) for s in srcs):
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", SO] + srcs[:2] + ["-lm"])
    return SO


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(SO):
        build()
    lib = C.CDLL(SO)
    lib.orc_ran1.argtypes = [C.POINTER(Ran1State)]
    lib.orc_ran1.restype = C.c_double
    lib.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.orc_philox_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    lib.orc_philox_uniform.restype = C.c_double
    for n, a in (("orc_planck", [C.c_double, C.c_double]), ("orc_luminosity", [C.c_double, C.c_double]),
                 ("orc_hg", [C.c_double, C.c_double]), ("orc_const", [C.c_int])):
        getattr(lib, n).argtypes = a
        getattr(lib, n).restype = C.c_double
    lib.orc_integrate.argtypes = [_dp, _dp, C.c_int]
    lib.orc_integrate.restype = C.c_double
    lib.orc_lam_grid.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp]
    lib.orc_bb_table.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    lib.orc_kp_table.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    lib.orc_star.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp]
    lib.orc_theta_grid.argtypes = [C.c_int, _dp, _dp]
    lib.orc_theta_grid.restype = C.c_int
    lib.orc_r_grid.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp]
    lib.orc_min_density.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp] + [C.c_double] * 6 + [_dp, _dp]
    lib.orc_normalize_mueller.argtypes = [C.c_int, _dp]
    lib.orc_mueller_integrals.argtypes = [C.c_int, _dp, _dp, _dp]
    lib.orc_phi_tables.argtypes = [_dp, _dp]
    lib.orc_mcradiation.argtypes = [C.POINTER(OrcModel), C.POINTER(OrcRun), C.POINTER(Ran1State),
                                    C.POINTER(OrcTally)]
    lib.orc_mcradiation.restype = C.c_int
    lib.orc_temperature.argtypes = [C.POINTER(OrcModel), _dp, _dp, _ip, _dp, _dp, _dp, _dp]
    lib.orc_determineT.argtypes = [C.POINTER(OrcModel), C.c_int, C.c_double, C.c_double]
    lib.orc_determineT.restype = C.c_double
    lib.orc_trace_ray.argtypes = [C.POINTER(OrcModel)] + [C.c_double] * 6 + [C.c_int] * 5 + [_ip, _dp,
                                                                                             C.POINTER(C.c_int)]
    lib.orc_trace_ray.restype = C.c_int
    lib.orc_theta_edges.argtypes = [C.c_int, _dp]
    lib.orc_tau01_midplane.argtypes = [C.POINTER(OrcModel), C.c_double]
    lib.orc_tau01_midplane.restype = C.c_double
    lib.orc_image_radii.argtypes = [C.POINTER(OrcModel), _dp, C.POINTER(OrcObs), C.c_double, _dp, C.c_int]
    lib.orc_image_radii.restype = C.c_int
    lib.orc_trace_flux.argtypes = [C.POINTER(OrcModel), _dp, _dp, C.POINTER(OrcObs), _dp, C.c_int, C.c_double,
                                   C.c_int, C.c_uint64, C.POINTER(Ran1State), _dp, _dp, _dp]
    lib.orc_trace_flux.restype = C.c_int
    lib.orc_trace_image.argtypes = [C.POINTER(OrcModel), _dp, _dp, C.POINTER(OrcObs), _dp, C.c_int, C.c_double,
                                    C.c_uint64, _dp]
    lib.orc_trace_image.restype = C.c_int
    lib.orc_make_image.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_uint64, _dp, _dp, _dp, _dp]
    lib.orc_make_image.restype = None
    lib.orc_arrinterpol_log.argtypes = [C.c_double, _dp, _dp, C.c_int]
    lib.orc_arrinterpol_log.restype = C.c_double
    lib.orc_diffcoeff.argtypes = [C.POINTER(OrcModel), C.c_int, C.c_int, C.c_int, _dp, _dp]
    lib.orc_init_random_walk.argtypes = [C.POINTER(OrcModel), _dp, _dp, _ip]
    lib.orc_rw_opacity_set.argtypes = [C.c_int]
    lib.orc_lu_solve.argtypes = [C.c_int, _dp, _dp]
    lib.orc_diffuse.argtypes = [C.POINTER(OrcModel), _dp, _dp, _dp, _dp, _ip, _ip, _dp, C.c_int, C.c_double,
                                C.POINTER(C.c_int)]
    lib.orc_min_dist.argtypes = [C.POINTER(OrcModel), C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
    lib.orc_min_dist.restype = C.c_double
    lib.orc_temp_average.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int]
    lib.orc_vertical_step.argtypes = [C.c_double] * 10 + [C.POINTER(C.c_int)]
    lib.orc_vertical_step.restype = C.c_double
    lib.orc_disk_structure.argtypes = [C.POINTER(OrcStruct), _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.orc_disk_structure.restype = C.c_long
    lib.orc_compare_cells.argtypes = [C.c_int, C.c_int] + [_dp] * 12 + [_dp] * 3
    _fp = C.POINTER(C.c_float)
    lib.orc_gauleg2.argtypes = [C.c_float, C.c_float, _fp, _fp, C.c_int]
    lib.orc_dmilay.argtypes = [C.c_float] * 7 + [_fp, C.c_int] + [_fp] * 7
    lib.orc_dmilay.restype = C.c_int
    lib.orc_meerhoff_mie.argtypes = [C.c_double] * 4 + [_dp] * 6 + [C.c_int]
    lib.orc_meerhoff_mie.restype = C.c_int
    lib.orc_compute_part.argtypes = ([C.c_int, _dp, C.c_int, _fp, _fp, _fp, _fp] + [C.c_double] * 4 +
                                     [C.c_int, C.c_int, C.c_double, C.c_double] + [_dp] * 5 +
                                     [C.POINTER(C.c_long)])
    lib.orc_compute_part.restype = C.c_int
    lib.orc_sample_reemission.argtypes = [C.POINTER(OrcModel), C.c_int, C.c_double, C.c_double, C.c_uint64,
                                          C.c_int, _ip, _dp]
    lib.orc_sample_scatangle.argtypes = [C.POINTER(OrcModel), C.c_int, C.c_int] + [C.c_double] * 4 + [
        C.c_int, C.c_uint64, C.c_int, _dp, _dp]
    _lib = lib
    return lib


def p(a):
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    return a.ctypes.data_as(_ip)


class Oracle:
    """The oracle bound to one model (arrays as built by mcmax_amd.model)."""

    def __init__(self, m):
        self.lib = load()
        self.m = m
        c = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        self.keep = dict(lam=c(m.lam), nu=c(m.nu), BB=c(m.BB), Kabs=c(m.Kabs), Ksca=c(m.Ksca),
                         Kext=c(m.Kext), F=c(m.F), IF11=c(m.IF11), IF12=c(m.IF12), Kp=c(m.Kp),
                         Fstar=c(m.Fstar), R=c(m.R), Th=c(m.Theta), dens=c(m.dens), V=c(m.V), w=c(m.w),
                         c2=c(m.cos2phi), s2=c(m.sin2phi), gsym=c(m.gsym), thet=c(m.thet))
        k = self.keep
        self.om = OrcModel(nlam=m.nlam, lam=p(k["lam"]), nu=p(k["nu"]), BB=p(k["BB"]), ng=m.ng,
                           Kabs=p(k["Kabs"]), Ksca=p(k["Ksca"]), Kext=p(k["Kext"]), F=p(k["F"]),
                           IF11=p(k["IF11"]), IF12=p(k["IF12"]), Kp=p(k["Kp"]), Fstar=p(k["Fstar"]),
                           Lstar=m.Lstar, Rstar_cm=m.Rstar_cm, Tstar=m.Tstar, nr=m.nr, nt=m.nt,
                           R=p(k["R"]), Theta=p(k["Th"]), dens=p(k["dens"]), V=p(k["V"]), w=p(k["w"]),
                           scat_how=m.scat_how, cos2phi=p(k["c2"]), sin2phi=p(k["s2"]),
                           rndw=int(m.randomwalk), factRW=float(m.factRW), gsym=p(k["gsym"]),
                           thet=p(k["thet"]))
        self.ran1 = Ran1State(idum=-42, iy=0)

    def sample_reemission(self, c, T0, T1, n, seed=1):
        """n reemission wavelengths (ilam1 1-based, wl1) of an absorption in
        cell c taking it from T0 to T1."""
        il = np.zeros(n, dtype=np.int32)
        wl = np.zeros(n)
        self.lib.orc_sample_reemission(C.byref(self.om), int(c), float(T0), float(T1), int(seed), int(n), p(il), p(wl))
        return il, wl

    def sample_scatangle(self, s, l, n, E=1.0, Q=0.0, U=0.0, V=0.0, pol=False, seed=1, with_phi=False):
        """cos theta of n scatterings off species s at wavelength index l
        (and the azimuths phi of the scattering with with_phi)."""
        out, ph = np.zeros(n), np.zeros(n)
        self.lib.orc_sample_scatangle(C.byref(self.om), int(s), int(l), float(E), float(Q), float(U), float(V),
                                      int(pol), int(seed), int(n), p(out), p(ph))
        return (out, ph) if with_phi else out

    def random_walk_tables(self):
        """(y, phi, thick) of InitRandomWalk for this model."""
        m = self.m
        y, ph = np.zeros(1000), np.zeros(1000)
        th = np.zeros(m.nr * m.nt, dtype=np.int32)
        self.lib.orc_init_random_walk(C.byref(self.om), p(y), p(ph), p(th))
        return y, ph, th.reshape(m.nr, m.nt).astype(bool)

    def diffcoeff(self, c, iT, opacity_set):
        a, b = C.c_double(), C.c_double()
        self.lib.orc_diffcoeff(C.byref(self.om), int(c), int(iT), int(opacity_set), C.byref(a), C.byref(b))
        return a.value, b.value

    def diffuse(self, T, dT, Ni, thick, EJv, nphotmin=25, dtmax=1.0):
        """Diffuse (Diffusion.f:9-329) on copies of T, thick, EJv; returns
        (T, thick, EJv, ncells)."""
        m = self.m
        T = np.ascontiguousarray(T, dtype=np.float64).ravel().copy()
        EJv = np.ascontiguousarray(EJv, dtype=np.float64).ravel().copy()
        dT = np.ascontiguousarray(dT, dtype=np.float64).ravel()
        Ni = np.ascontiguousarray(Ni, dtype=np.int32).ravel()
        th = np.ascontiguousarray(thick, dtype=np.int32).ravel().copy()
        Rav = np.ascontiguousarray(m.R_av, dtype=np.float64)
        tav = np.ascontiguousarray(m.theta_av, dtype=np.float64)
        nc = C.c_int()
        rc = self.lib.orc_diffuse(C.byref(self.om), p(Rav), p(tav), p(T), p(dT), p(Ni), p(th), p(EJv),
                                  int(nphotmin), float(dtmax), C.byref(nc))
        if rc != 0:
            raise RuntimeError(f"oracle diffuse failed: {rc}")
        sh = (m.nr, m.nt)
        return T.reshape(sh), th.reshape(sh).astype(bool), EJv.reshape(sh), nc.value

    def disk_structure(self, state, Temp, **kw):
        return disk_structure(self.m, state, Temp, **kw)

    def min_dist(self, x, y, z, i, jj):
        return self.lib.orc_min_dist(C.byref(self.om), x, y, z, int(i), int(jj))

    def mcradiation(self, nphot, rng="philox", seed=1, maxinteract=0, per_packet=False, inflight=0, staged=False,
                    stage0=64, stage_ratio=2, epoch_deposits=0):
        """One pass: the serial loop, or with staged=True the staged
        Bjorkman-Wood schedule of the device's concurrent pass (orc_run in
        mcx_oracle.h)."""
        m = self.m
        n = m.nr * m.nt
        nb = (NANGLE + 1) * m.nlam
        out = {k: np.zeros(n) for k in ("EJv", "EJv2", "Eabs", "T")}
        out["Ni"] = np.zeros(n, dtype=np.int32)
        out["spec"], out["scatspec"] = np.zeros(nb), np.zeros(nb)
        out["ispec"], out["iscatspec"] = np.zeros(nb, dtype=np.int32), np.zeros(nb, dtype=np.int32)
        out["nev1"], out["nev2"] = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        t = OrcTally(**{k: p(v) for k, v in out.items()})
        if per_packet:
            out["pkt_steps"] = np.zeros(int(nphot), dtype=np.int64)
            t.pkt_steps = out["pkt_steps"].ctypes.data_as(C.POINTER(C.c_int64))
        run = OrcRun(nphot=int(nphot), rng=0 if rng == "ran1" else 1, seed=int(seed), maxinteract=int(maxinteract),
                     inflight=int(inflight), staged=int(staged), stage0=int(stage0),
                     stage_ratio=int(stage_ratio), epoch_deposits=int(epoch_deposits))
        rc = self.lib.orc_mcradiation(C.byref(self.om), C.byref(run), C.byref(self.ran1), C.byref(t))
        if rc != 0:
            raise RuntimeError(f"oracle transport failed: {rc}")
        for k in ("nemit", "ninteract", "nabs", "nscat", "nescape", "nstarhit", "nsteps", "nraar",
                  "nmaxinteract", "tautot", "Eescape", "nrw"):
            out[k] = getattr(t, k)
        out["Tbw"] = out.pop("T")
        for k in ("EJv", "EJv2", "Eabs", "Tbw", "Ni", "nev1", "nev2"):
            out[k] = out[k].reshape(m.nr, m.nt)
        for k in ("spec", "scatspec", "ispec", "iscatspec"):
            out[k] = out[k].reshape(NANGLE + 1, m.nlam)
        return out

    def temperature(self, tal):
        m = self.m
        n = m.nr * m.nt
        T, dT, dE = np.zeros(n), np.zeros(n), np.zeros(n)
        EJv = np.ascontiguousarray(tal["EJv"].ravel(), dtype=np.float64)
        EJv2 = np.ascontiguousarray(tal["EJv2"].ravel(), dtype=np.float64)
        Ni = np.ascontiguousarray(tal["Ni"].ravel(), dtype=np.int32)
        Tbw = np.ascontiguousarray(tal["Tbw"].ravel(), dtype=np.float64)
        self.lib.orc_temperature(C.byref(self.om), p(EJv), p(EJv2), p(Ni), p(Tbw), p(T), p(dT), p(dE))
        return T.reshape(m.nr, m.nt), dT.reshape(m.nr, m.nt), dE.reshape(m.nr, m.nt)

    def image_radii(self, obs, lam_paths=0.55):
        m = self.m
        R01 = self.lib.orc_tau01_midplane(C.byref(self.om), lam_paths)
        nmax = 4 * (m.nr + 4 * m.nt + 400) + 10 * max(obs.nj0, 1) * m.nr
        R = np.zeros(nmax)
        tav = np.ascontiguousarray(m.theta_av, dtype=np.float64)
        n = self.lib.orc_image_radii(C.byref(self.om), p(tav), C.byref(obs), R01, p(R), nmax)
        if n < 0:
            raise RuntimeError("image radii buffer too small")
        return R[:n].copy(), R01

    def sed(self, Tcell, obs, lams, rng="philox", seed=1):
        """spectrum(lam) of Observe's SPECTRUM branch (Telescope.f:138-164):
        returns flux, scatflux, fluxQ arrays (erg/s/Hz/sr-type units before
        the 1e23/d^2 conversion)."""
        m = self.m
        Rimg, _ = self.image_radii(obs)
        thet = np.ascontiguousarray(m.thet, dtype=np.float64)
        T = np.ascontiguousarray(np.asarray(Tcell).ravel(), dtype=np.float64)
        out = np.zeros((len(lams), 3))
        for k, lam0 in enumerate(lams):
            f, s, q = C.c_double(), C.c_double(), C.c_double()
            rc = self.lib.orc_trace_flux(C.byref(self.om), p(thet), p(T), C.byref(obs), p(Rimg), len(Rimg),
                                         float(lam0), 0 if rng == "ran1" else 1, int(seed) + k,
                                         C.byref(self.ran1), C.byref(f), C.byref(s), C.byref(q))
            if rc != 0:
                raise RuntimeError(f"oracle SED failed at {lam0}: {rc}")
            out[k] = (f.value, s.value, q.value)
        return out

    def ray_image(self, Tcell, obs, lam0, seed=1):
        """TraceFlux's ray images at one wavelength (Telescope.f:165-169:
        paths traced at the image wavelength): returns (Rimg, rays) with
        rays [4, nimg, nphi] = I, Q, U, V."""
        m = self.m
        Rimg, _ = self.image_radii(obs, lam_paths=float(lam0))
        thet = np.ascontiguousarray(m.thet, dtype=np.float64)
        T = np.ascontiguousarray(np.asarray(Tcell).ravel(), dtype=np.float64)
        rays = np.zeros(4 * len(Rimg) * obs.nphi)
        rc = self.lib.orc_trace_image(C.byref(self.om), p(thet), p(T), C.byref(obs), p(Rimg), len(Rimg),
                                      float(lam0), int(seed), p(rays))
        if rc != 0:
            raise RuntimeError(f"oracle ray image failed at {lam0}: {rc}")
        return Rimg, rays.reshape(4, len(Rimg), obs.nphi)

    def make_image(self, Rimg, rays, scat_how, npix, nint, Rmax, pa_deg=0.0, seed=1):
        """MakeImage's spread of ray images onto npix x npix pixels; returns
        the raw images [4, npix, npix] (I, Q, U, V), pixel (ix, iy) at
        [:, ix-1, iy-1]."""
        lib = self.lib
        return make_image(lib, Rimg, rays, scat_how, npix, nint, Rmax, pa_deg, seed)

    def trace_ray(self, x, y, z, vx, vy, vz, i, jj, onEdge=0, edgeNr=0, maxseg=100000):
        cells = np.zeros(maxseg, dtype=np.int32)
        lens = np.zeros(maxseg)
        code = C.c_int()
        n = self.lib.orc_trace_ray(C.byref(self.om), x, y, z, vx, vy, vz, i, jj, onEdge, edgeNr, maxseg,
                                   p(cells), p(lens), C.byref(code))
        return cells[:max(n, 0)], lens[:max(n, 0)], code.value


def make_image(lib, Rimg, rays, scat_how, npix, nint, Rmax, pa_deg=0.0, seed=1):
    """orc_make_image on given ray images (Rimg [nimg], rays [4, nimg, nphi])."""
    Rimg = np.ascontiguousarray(Rimg, dtype=np.float64)
    rays = np.ascontiguousarray(rays, dtype=np.float64)
    nimg, nphi = rays.shape[1], rays.shape[2]
    assert rays.shape[0] == 4 and len(Rimg) == nimg
    out = np.zeros((4, npix, npix))
    lib.orc_make_image(nimg, nphi, p(Rimg), p(rays), int(scat_how), int(npix), int(nint), float(Rmax), float(pa_deg),
                       int(seed), p(out[0]), p(out[1]), p(out[2]), p(out[3]))
    return out
