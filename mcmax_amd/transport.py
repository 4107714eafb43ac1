"""Host interface of the transport stage, mirroring the reference's call
sequence: ``MCRadiation(NphotTot, ...)`` (MCRadiation.f:6) followed by the
Lucy temperature inversion (Main.f:498-558).  Everything runs in libmcx.so on
the GPU; this module only moves arrays across the C-ABI.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import NANGLE, TMAX, Obs, ptr


class MCXError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mcx error {code}: {msg}")
        self.code = code


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Transport:
    """One libmcx context on one GPU."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        self.ctx = C.c_void_p()
        rc = self.lib.mcx_create(C.byref(self.ctx), int(device))
        if rc != 0:
            raise MCXError(rc, "mcx_create failed (no GPU visible?)")
        self._keep = []
        self.model = None
        # the largest per-species-mix table (over integer T) that set_model
        # builds, in bytes; None: the library's 4 GiB.  Many contexts on one
        # GPU set it low: the Planck means and diffusion coefficients are then
        # evaluated at each read (the same values) and the reemission sums
        # the species' cumulatives at each probe (the same up to rounding).
        self.class_table_bytes = None

    def close(self):
        if self.ctx:
            self.lib.mcx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise MCXError(rc, self.lib.mcx_last_error(self.ctx).decode())
        return rc

    def set_model(self, m, tables=True):
        """Hand a model's tables and grid to the device (mcx_set_tables,
        mcx_set_model).  tables=False re-sends only the grid, densities and
        abundances (after a new vertical structure)."""
        keep = dict(
            lam=_c(m.lam), nu=_c(m.nu), BB=_c(m.BB), Kabs=_c(m.Kabs), Ksca=_c(m.Ksca),
            F=_c(m.F), IF11=_c(m.IF11), IF12=_c(m.IF12), Kp=_c(m.Kp), Fstar=_c(m.Fstar),
            c2=_c(m.cos2phi), s2=_c(m.sin2phi), R=_c(m.R), Th=_c(m.Theta), dens=_c(m.dens),
            V=_c(m.V), w=_c(m.w), Kext=_c(m.Kext), gsym=_c(m.gsym), thet=_c(m.thet))
        t = _lib.Tables(nlam=m.nlam, ng=m.ng, tmax=TMAX, lam=ptr(keep["lam"]), nu=ptr(keep["nu"]),
                        BB=ptr(keep["BB"]), Kabs=ptr(keep["Kabs"]), Ksca=ptr(keep["Ksca"]),
                        F=ptr(keep["F"]), IF11=ptr(keep["IF11"]), IF12=ptr(keep["IF12"]),
                        Kp=ptr(keep["Kp"]), Fstar=ptr(keep["Fstar"]), Lstar=float(m.Lstar),
                        cos2phi=ptr(keep["c2"]), sin2phi=ptr(keep["s2"]), scat_how=int(m.scat_how),
                        Kext=ptr(keep["Kext"]), gsym=ptr(keep["gsym"]))
        if tables:
            self._check(self.lib.mcx_set_tables(self.ctx, C.byref(t)))
        g = _lib.Grid(nr=m.nr, nt=m.nt, R=ptr(keep["R"]), Theta=ptr(keep["Th"]), dens=ptr(keep["dens"]),
                      V=ptr(keep["V"]), w=ptr(keep["w"]), Rstar_cm=float(m.Rstar_cm), Tstar=float(m.Tstar),
                      thet=ptr(keep["thet"]))
        cap = self.class_table_bytes
        old = os.environ.get("MCX_CLASS_TABLE_BYTES")
        if cap is not None:
            os.environ["MCX_CLASS_TABLE_BYTES"] = str(int(cap))
        try:
            self._check(self.lib.mcx_set_model(self.ctx, C.byref(g)))
        finally:
            if cap is not None:
                if old is None:
                    del os.environ["MCX_CLASS_TABLE_BYTES"]
                else:
                    os.environ["MCX_CLASS_TABLE_BYTES"] = old
        self.model = m

    def transport(self, nphot, seed=1, nlanes=0, maxinteract=0, randomwalk=None, factRW=None, coop_lanes=-1,
                  stage0=0, stage_ratio=0, precision=0, max_ms=0.0, epoch_deposits=0):
        """One MC pass (mcx_transport); tallies stay on the device.  The
        random-walk switches default to the model's keywords.  nlanes=1 runs
        the serial loop on one lane; otherwise the staged schedule (stage0,
        stage_ratio, epoch_deposits; 0 picks 64, 2 and 16000).  precision 1: FP32 direction
        arithmetic (MCX_FP32).  max_ms > 0: the pass's time limit (MCX_ETIME
        raises)."""
        m = self.model
        rw = bool(getattr(m, "randomwalk", False)) if randomwalk is None else bool(randomwalk)
        f = float(getattr(m, "factRW", 10.0)) if factRW is None else float(factRW)
        r = _lib.Run(nphot=int(nphot), seed=int(seed), maxinteract=int(maxinteract), nlanes=int(nlanes),
                     stage0=int(stage0), randomwalk=int(rw), factRW=f, coop_lanes=int(coop_lanes),
                     stage_ratio=int(stage_ratio), precision=int(precision), max_ms=float(max_ms),
                     epoch_deposits=int(epoch_deposits))
        rc = self.lib.mcx_transport(self.ctx, C.byref(r))
        if rc not in (0, _lib.MCX_EGEOM):
            self._check(rc)
        return rc

    def tallies(self):
        m = self.model
        n = m.nr * m.nt
        nb = (NANGLE + 1) * m.nlam
        out = {k: np.zeros(n) for k in ("EJv", "EJv2", "Eabs", "Tbw")}
        out["Ni"] = np.zeros(n, dtype=np.int32)
        out["thick"] = np.zeros(n, dtype=np.int32)
        for k in ("spec", "scatspec"):
            out[k] = np.zeros(nb)
        for k in ("ispec", "iscatspec"):
            out[k] = np.zeros(nb, dtype=np.int32)
        t = _lib.Tallies(**{k: ptr(v) for k, v in out.items()})
        self._check(self.lib.mcx_get_tallies(self.ctx, C.byref(t)))
        for k in ("EJv", "EJv2", "Eabs", "Tbw", "Ni"):
            out[k] = out[k].reshape(m.nr, m.nt)
        out["thick"] = out["thick"].reshape(m.nr, m.nt).astype(bool)
        for k in ("spec", "scatspec", "ispec", "iscatspec"):
            out[k] = out[k].reshape(NANGLE + 1, m.nlam)
        for k in ("nemit", "ninteract", "nabs", "nscat", "nescape", "nstarhit", "nsteps", "nraar",
                  "nmaxinteract", "ngeomfail", "tautot", "Eescape", "nrw", "noverflow", "maxsteps",
                  "nvisit_overflow", "ntimeout", "ntallyatomic"):
            out[k] = getattr(t, k)
        out["t_done_ms"] = list(t.t_done)
        return out

    def temperature(self):
        m = self.model
        n = m.nr * m.nt
        T, dT, dE = np.zeros(n), np.zeros(n), np.zeros(n)
        o = _lib.Temps(T=ptr(T), dT=ptr(dT), dEJv=ptr(dE))
        self._check(self.lib.mcx_temperature(self.ctx, C.byref(o)))
        return T.reshape(m.nr, m.nt), dT.reshape(m.nr, m.nt), dE.reshape(m.nr, m.nt)

    def sed(self, obs, lams, T=None):
        """Ray-traced spectrum at wavelengths lams (mcx_sed): returns an
        array [len(lams), 3] of flux, scattered flux and Q flux.  T: cell
        temperatures [nr, nt], or None for those of the last temperature()."""
        m = self.model
        lams = np.ascontiguousarray(lams, dtype=np.float64)
        out = np.zeros((3, len(lams)))
        tav = _c(m.theta_av)
        Tp = None
        if T is not None:
            Tc = _c(np.asarray(T).ravel())
            Tp = ptr(Tc)
        self._check(self.lib.mcx_sed(self.ctx, C.byref(obs), Tp, ptr(tav), len(lams), ptr(lams), ptr(out[0]),
                                     ptr(out[1]), ptr(out[2])))
        return out.T.copy()

    def image(self, obs, lam, npix, nint, Rmax, pa_deg=0.0, seed=1, T=None, rays=False):
        """Raw pixel image of an IMAGE observation at wavelength lam
        (mcx_image): array [4, npix, npix] of I, Q, U, V with pixel (ix, iy)
        at [:, ix-1, iy-1], before the 1e23 zscale / distance^2 scaling.
        With rays=True also returns (Rimg, ray images [4, nimg, nphi])."""
        m = self.model
        tav = _c(m.theta_av)
        Tp = None
        if T is not None:
            Tc = _c(np.asarray(T).ravel())
            Tp = ptr(Tc)
        out = np.zeros((4, npix, npix))
        opts = _lib.ImageOpts(int(npix), int(nint), float(Rmax), float(pa_deg), int(seed))
        nimg = C.c_int32()
        cap = 4 * (m.nr + 4 * m.nt + 400) + 10 * max(int(obs.nj0), 1) * m.nr
        Rimg = np.zeros(cap) if rays else None
        rbuf = np.zeros(4 * cap * int(obs.nphi)) if rays else None
        self._check(self.lib.mcx_image(self.ctx, C.byref(obs), Tp, ptr(tav), float(lam), C.byref(opts),
                                       ptr(out[0]), ptr(out[1]), ptr(out[2]), ptr(out[3]), C.byref(nimg),
                                       cap if rays else 0, ptr(Rimg) if rays else None,
                                       ptr(rbuf) if rays else None))
        if not rays:
            return out
        n = nimg.value
        return out, Rimg[:n].copy(), rbuf[:4 * n * int(obs.nphi)].reshape(4, n, int(obs.nphi)).copy()

    def last_image_timing(self):
        """(rays_ms, spread_ms) of the last image() (mcx_last_image_timing)."""
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.mcx_last_image_timing(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_temperature(self, T, dT=None, Ni=None):
        """Load a temperature state (mcx_set_temperature)."""
        Tc = _c(np.asarray(T).ravel())
        dTc = _c(np.asarray(dT).ravel()) if dT is not None else None
        Nic = np.ascontiguousarray(np.asarray(Ni).ravel(), dtype=np.int32) if Ni is not None else None
        self._check(self.lib.mcx_set_temperature(self.ctx, ptr(Tc), ptr(dTc) if dTc is not None else None,
                                                 ptr(Nic) if Nic is not None else None))

    def diffuse(self, nphotmin=25, dtmax=1.0, factRW=None, EJv=None):
        """Diffuse (mcx_diffuse) on the device temperatures; returns
        (T, thick, EJv, ncells)."""
        m = self.model
        n = m.nr * m.nt
        f = float(getattr(m, "factRW", 10.0)) if factRW is None else float(factRW)
        T = np.zeros(n)
        th = np.zeros(n, dtype=np.int32)
        E = _c(np.asarray(EJv).ravel()).copy() if EJv is not None else np.zeros(n)
        nc = C.c_int32()
        Rav, tav = _c(m.R_av), _c(m.theta_av)
        self._check(self.lib.mcx_diffuse(self.ctx, int(nphotmin), float(dtmax), f, ptr(Rav), ptr(tav), ptr(T),
                                         ptr(th), ptr(E), C.byref(nc)))
        sh = (m.nr, m.nt)
        return T.reshape(sh), th.reshape(sh).astype(bool), E.reshape(sh), nc.value

    def temp_average(self, fw, first=False):
        """TempAverage (mcx_temp_average) of the device temperatures;
        returns the averaged field [nr, nt]."""
        m = self.model
        out = np.zeros(m.nr * m.nt)
        self._check(self.lib.mcx_temp_average(self.ctx, float(fw), int(bool(first)), ptr(out)))
        return out.reshape(m.nr, m.nt)

    def disk_structure(self, state, Temp=None, shscale=None, gshscale=None, settle=None, gascoupled=None,
                       mpstr=False, FradZ=None):
        """DiskStructure (mcx_disk_structure) on a copy of state = dict(dens,
        dens0, gasdens [nr, nt], w, w0 [nr, nt, ng]); Temp None uses the
        device's averaged temperatures.  Returns (new state, Runge-Kutta
        steps)."""
        m = self.model
        sh = (m.nr, m.nt)
        out = {k: _c(state[k]).copy() for k in ("dens", "dens0", "gasdens", "w", "w0")}
        i32 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.int32)
        keep = dict(Rav=_c(m.R_av), tav=_c(m.theta_av),
                    sh=None if shscale is None else _c(shscale), gsh=None if gshscale is None else _c(gshscale),
                    st=i32(settle), gc=i32(gascoupled), fz=None if FradZ is None else _c(FradZ),
                    T=None if Temp is None else _c(np.asarray(Temp).ravel()))
        st = _lib.Struct(R_av=ptr(keep["Rav"]), theta_av=ptr(keep["tav"]), Mstar=float(m.Mstar_g),
                         shscale=ptr(keep["sh"]), gshscale=ptr(keep["gsh"]), settle=ptr(keep["st"]),
                         gascoupled=ptr(keep["gc"]), mpstr=int(bool(mpstr)), FradZ=ptr(keep["fz"]))
        d = _lib.Density(**{k: ptr(v) for k, v in out.items()})
        n = C.c_int64()
        self._check(self.lib.mcx_disk_structure(self.ctx, C.byref(st), ptr(keep["T"]), C.byref(d), C.byref(n)))
        for k in ("dens", "dens0", "gasdens"):
            out[k] = out[k].reshape(sh)
        return out, n.value

    def compare_cells(self, prev, cur):
        """compareCells (mcx_compare_cells) of two dicts of T, dT, dens,
        ddens, EJv, dEJv [nr, nt]; returns (errT, errRho, errE)."""
        keys = ("T", "dT", "dens", "ddens", "EJv", "dEJv")
        kp = [_c(np.asarray(prev[k]).ravel()) for k in keys]
        kc = [_c(np.asarray(cur[k]).ravel()) for k in keys]
        a = _lib.CellState(*[ptr(x) for x in kp])
        b = _lib.CellState(*[ptr(x) for x in kc])
        e = [C.c_double() for _ in range(3)]
        self._check(self.lib.mcx_compare_cells(self.ctx, C.byref(a), C.byref(b), *[C.byref(x) for x in e]))
        return tuple(x.value for x in e)

    def compute_parts(self, lam, parts):
        """ComputePart on the device (mcx_compute_parts) for a list of
        species, each a dict with e1, e2 [nm, nlam] (n, k), frac, rho [nm],
        amin, amax, apow, fmax, ns, nf and optionally rscale, mscale.  Returns
        (list of dicts with Kabs, Ksca, Kext [nlam], F [nlam, 6, 180], rho;
        stats dict)."""
        lam = _c(lam)
        nlam = len(lam)
        keep, specs, outs, res = [], [], [], []
        for p in parts:
            e1 = np.ascontiguousarray(np.atleast_2d(p["e1"]), dtype=np.float32)
            e2 = np.ascontiguousarray(np.atleast_2d(p["e2"]), dtype=np.float32)
            if e1.shape != e2.shape or e1.shape[1] != nlam:
                raise ValueError("e1, e2 must be [nm, nlam]")
            frac = np.ascontiguousarray(p["frac"], dtype=np.float32)
            rho = np.ascontiguousarray(p["rho"], dtype=np.float32)
            keep += [e1, e2, frac, rho]
            fp = C.POINTER(C.c_float)
            specs.append(_lib.PartSpec(nm=e1.shape[0], ns=int(p["ns"]), nf=int(p["nf"]), amin=float(p["amin"]),
                                       amax=float(p["amax"]), apow=float(p["apow"]), fmax=float(p["fmax"]),
                                       rscale=float(p.get("rscale", 1.0)), mscale=float(p.get("mscale", 1.0)),
                                       e1=e1.ctypes.data_as(fp), e2=e2.ctypes.data_as(fp),
                                       frac=frac.ctypes.data_as(fp), rho=rho.ctypes.data_as(fp)))
            r = dict(Kabs=np.zeros(nlam), Ksca=np.zeros(nlam), Kext=np.zeros(nlam), F=np.zeros((nlam, 6, 180)),
                     rho=np.zeros(1))
            res.append(r)
            outs.append(_lib.PartOut(**{k: ptr(v) for k, v in r.items()}))
        sa = (_lib.PartSpec * len(specs))(*specs)
        oa = (_lib.PartOut * len(outs))(*outs)
        st = _lib.PartStats()
        self._check(self.lib.mcx_compute_parts(self.ctx, nlam, ptr(lam), len(specs), sa, oa, C.byref(st)))
        for r in res:
            r["rho"] = float(r["rho"][0])
        stats = {k: getattr(st, k) for k, _ in _lib.PartStats._fields_}
        return res, stats

    def timing(self):
        k, p = C.c_double(), C.c_double()
        self._check(self.lib.mcx_last_timing(self.ctx, C.byref(k), C.byref(p)))
        return k.value, p.value


def batch_transport(trs, nphot, seed=1, seed_stride=1, blocks_per_model=0, randomwalk=None, factRW=None,
                    maxinteract=0, precision=0, max_ms=0.0, stage0=0, stage_ratio=0, graph=False, epoch_deposits=0,
                    model_blocks=None):
    """One pass of nphot packets for every model of the Transport objects
    trs (one model each) in common launches (mcx_batch_transport); model m
    uses the key seed + m seed_stride.  graph=True: captured once, then
    replayed (mcx_batch_transport's use_graph).  model_blocks: each model's
    workgroups (mcx_batch_transport_blocks) instead of blocks_per_model for
    all.  Returns the per-model status codes; the tallies stay on each
    model's context."""
    lib = trs[0].lib
    m0 = trs[0].model
    rw = bool(getattr(m0, "randomwalk", False)) if randomwalk is None else bool(randomwalk)
    f = float(getattr(m0, "factRW", 10.0)) if factRW is None else float(factRW)
    r = _lib.Run(nphot=int(nphot), seed=int(seed), maxinteract=int(maxinteract), nlanes=0, stage0=int(stage0),
                 randomwalk=int(rw), factRW=f, coop_lanes=-1, stage_ratio=int(stage_ratio), precision=int(precision),
                 max_ms=float(max_ms), epoch_deposits=int(epoch_deposits))
    arr = (C.c_void_p * len(trs))(*[t.ctx for t in trs])
    st = np.zeros(len(trs), dtype=np.int32)
    if model_blocks is not None:
        mb = np.ascontiguousarray(model_blocks, dtype=np.int32)
        if mb.shape != (len(trs),):
            raise ValueError("model_blocks: one count per model")
        rc = lib.mcx_batch_transport_blocks(arr, len(trs), C.byref(r), mb.ctypes.data_as(C.POINTER(C.c_int32)),
                                            int(seed_stride), int(bool(graph)), st.ctypes.data_as(C.POINTER(C.c_int32)))
    else:
        rc = lib.mcx_batch_transport(arr, len(trs), C.byref(r), int(blocks_per_model), int(seed_stride),
                                     int(bool(graph)), st.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc not in (0, _lib.MCX_EGEOM):
        raise MCXError(rc, lib.mcx_last_error(trs[0].ctx).decode())
    return st


def make_obs(angle_deg=45.0, nphi=45, nj0=2, nphot=1000, nphotstar=1000, nexits=5, usepol=True, tracestar=True,
             traceemis=True, tracescat=True, pfstop=0.25, seed=1):
    """An observation with the reference's defaults (Telescope.f:590-657,
    pfstop Init.f:378)."""
    import math
    return Obs(angle=math.pi * angle_deg / 180.0, nphi=nphi, nj0=nj0, nphot=nphot, nphotstar=nphotstar,
               nexits=nexits, usepol=int(usepol), tracestar=int(tracestar), traceemis=int(traceemis),
               tracescat=int(tracescat), pfstop=pfstop, seed=seed)


def mc_radiation(tr, nphot, seed=1, nlanes=0, maxinteract=0, coop_lanes=-1, stage0=0, stage_ratio=0, precision=0,
                 max_ms=0.0, epoch_deposits=0):
    """MCRadiation + temperature inversion; returns a dict with the tallies
    and T, dT, dEJv."""
    rc = tr.transport(nphot, seed=seed, nlanes=nlanes, maxinteract=maxinteract, coop_lanes=coop_lanes,
                      stage0=stage0, stage_ratio=stage_ratio, precision=precision, max_ms=max_ms,
                      epoch_deposits=epoch_deposits)
    out = tr.tallies()
    out["T"], out["dT"], out["dEJv"] = tr.temperature()
    out["rc"] = rc
    out["kernel_ms"], out["pass_ms"] = tr.timing()
    return out
