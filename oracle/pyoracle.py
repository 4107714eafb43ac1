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
        ("sin2phi", _dp),
    ]


class OrcRun(C.Structure):
    _fields_ = [("nphot", C.c_int64), ("rng", C.c_int), ("seed", C.c_uint64), ("maxinteract", C.c_int)]


class OrcTally(C.Structure):
    _fields_ = [
        ("EJv", _dp), ("EJv2", _dp), ("Eabs", _dp), ("T", _dp), ("Ni", _ip),
        ("spec", _dp), ("scatspec", _dp), ("ispec", _ip), ("iscatspec", _ip),
        ("nemit", C.c_int64), ("ninteract", C.c_int64), ("nabs", C.c_int64), ("nscat", C.c_int64),
        ("nescape", C.c_int64), ("nstarhit", C.c_int64), ("nsteps", C.c_int64), ("nraar", C.c_int64),
        ("nmaxinteract", C.c_int64), ("tautot", C.c_double), ("Eescape", C.c_double),
        ("nev1", _ip), ("nev2", _ip),
    ]


class Ran1State(C.Structure):
    _fields_ = [("idum", C.c_int32), ("iy", C.c_int32), ("iv", C.c_int32 * 32)]


def build(force=False):
    """Compile liboracle.so with gcc (no FMA contraction: plain IEEE FP64)."""
    src = os.path.join(HERE, "mcx_oracle.c")
    if force or not os.path.exists(SO) or os.path.getmtime(SO) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", SO, src, "-lm"])
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
                         c2=c(m.cos2phi), s2=c(m.sin2phi))
        k = self.keep
        self.om = OrcModel(nlam=m.nlam, lam=p(k["lam"]), nu=p(k["nu"]), BB=p(k["BB"]), ng=m.ng,
                           Kabs=p(k["Kabs"]), Ksca=p(k["Ksca"]), Kext=p(k["Kext"]), F=p(k["F"]),
                           IF11=p(k["IF11"]), IF12=p(k["IF12"]), Kp=p(k["Kp"]), Fstar=p(k["Fstar"]),
                           Lstar=m.Lstar, Rstar_cm=m.Rstar_cm, Tstar=m.Tstar, nr=m.nr, nt=m.nt,
                           R=p(k["R"]), Theta=p(k["Th"]), dens=p(k["dens"]), V=p(k["V"]), w=p(k["w"]),
                           scat_how=m.scat_how, cos2phi=p(k["c2"]), sin2phi=p(k["s2"]))
        self.ran1 = Ran1State(idum=-42, iy=0)

    def mcradiation(self, nphot, rng="philox", seed=1, maxinteract=0):
        m = self.m
        n = m.nr * m.nt
        nb = (NANGLE + 1) * m.nlam
        out = {k: np.zeros(n) for k in ("EJv", "EJv2", "Eabs", "T")}
        out["Ni"] = np.zeros(n, dtype=np.int32)
        out["spec"], out["scatspec"] = np.zeros(nb), np.zeros(nb)
        out["ispec"], out["iscatspec"] = np.zeros(nb, dtype=np.int32), np.zeros(nb, dtype=np.int32)
        out["nev1"], out["nev2"] = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        t = OrcTally(**{k: p(v) for k, v in out.items()})
        run = OrcRun(nphot=int(nphot), rng=0 if rng == "ran1" else 1, seed=int(seed), maxinteract=int(maxinteract))
        rc = self.lib.orc_mcradiation(C.byref(self.om), C.byref(run), C.byref(self.ran1), C.byref(t))
        if rc != 0:
            raise RuntimeError(f"oracle transport failed: {rc}")
        for k in ("nemit", "ninteract", "nabs", "nscat", "nescape", "nstarhit", "nsteps", "nraar",
                  "nmaxinteract", "tautot", "Eescape"):
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

    def trace_ray(self, x, y, z, vx, vy, vz, i, jj, onEdge=0, edgeNr=0, maxseg=100000):
        cells = np.zeros(maxseg, dtype=np.int32)
        lens = np.zeros(maxseg)
        code = C.c_int()
        n = self.lib.orc_trace_ray(C.byref(self.om), x, y, z, vx, vy, vz, i, jj, onEdge, edgeNr, maxseg,
                                   p(cells), p(lens), C.byref(code))
        return cells[:max(n, 0)], lens[:max(n, 0)], code.value
