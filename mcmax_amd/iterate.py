"""The driver's structure iteration over libmcx (Main.f:202-791, the
struct_iter path): per iteration an MC pass and the temperature inversion,
the diffusion approximation for thick cells with poor statistics, then
TempAverage, DiskStructure and compareCells with the reference's
convergence rule; after the last pass the ray-traced SED.

Everything numeric runs on the device through ``transport.Transport``; this
module keeps the loop's bookkeeping (iteration weights, convergence flags,
the previous state compareCells reads).
"""
import time

import numpy as np


def density_state(m):
    """C%dens0 = C%gasdens = C%dens and C%w0 = C%w at set-up (Init.f:3899-3904)."""
    return dict(dens=m.dens.copy(), dens0=m.dens.copy(), gasdens=m.dens.copy(), w=m.w.copy(), w0=m.w.copy())


def average_weight(niter, f_weight, niter0):
    """TempAverage's weight in the driver (Main.f:684-690)."""
    if niter > niter0:
        return 1.0 / (niter - niter0)
    if niter > 1:
        return f_weight
    return 1.0


def initial_temperature(m):
    """The temperature the set-up assigns before the first structure solve
    (Init.f:3935): Tstar sqrt(R_av(1)/R_av(i)) / 10."""
    T = np.zeros((m.nr, m.nt))
    T[1:] = (m.Tstar * np.sqrt(m.R_av[1] / m.R_av[1:]) / 10.0)[:, None]
    return T


class Iteration:
    """One run of the loop; `history` holds one record per MC pass."""

    def __init__(self, tr, m, nphot, seed=1, maxiter=None, epsiter=None, f_weight=None, niter0=1000,
                 nphotdiffuse=None, dtdiffuse=None, struct_iter=True, structure_kw=None, log=None):
        p = getattr(m, "keywords", {})
        self.tr, self.m = tr, m
        self.nphot, self.seed = int(nphot), int(seed)
        self.maxiter = int(p.get("maxiter", 10) if maxiter is None else maxiter)
        self.epsiter = float(p.get("epsiter", 3.0) if epsiter is None else epsiter)
        self.f_weight = float(p.get("fweight", 0.5) if f_weight is None else f_weight)
        self.niter0 = int(niter0)
        self.nphotdiffuse = int(p.get("nphotdiffuse", 25) if nphotdiffuse is None else nphotdiffuse)
        self.dtdiffuse = float(p.get("dtdiffuse", 1.0) if dtdiffuse is None else dtdiffuse)
        self.struct_iter = bool(struct_iter)
        self.structure_kw = dict(structure_kw or {})
        self.log = log
        self.history = []

    def _say(self, msg):
        if self.log:
            self.log(msg)

    def run(self):
        tr, m = self.tr, self.m
        nr, nt = m.nr, m.nt
        state = density_state(m)
        tr.set_model(m)
        zeros = np.zeros((nr, nt))
        # Cprev before the first pass: the set-up's T; dT, EJv, dEJv unset (0),
        # so the first compareCells counts no cell (Main.f:458, DiskStructure.f:565)
        cur = dict(T=initial_temperature(m), dT=zeros, dens=state["dens"].copy(), ddens=zeros, EJv=zeros,
                   dEJv=zeros)
        niter, lastiter, finaliter = 0, False, False
        while True:
            prev = cur
            t0 = time.perf_counter()
            rc = tr.transport(self.nphot, seed=self.seed + len(self.history),
                              maxinteract=int(getattr(m, "maxinteract", 0)))
            tal = tr.tallies()
            self.last_tallies = tal
            T, dT, dEJv = tr.temperature()
            EJv = tal["EJv"] / m.V  # C%EJv per volume (Main.f:505)
            ncells = 0
            if self.nphotdiffuse > 0 or self.dtdiffuse < 1.0:
                T, _, EJv, ncells = tr.diffuse(self.nphotdiffuse, self.dtdiffuse, EJv=EJv)
            kernel_ms, pass_ms = tr.timing()
            rec = dict(niter=niter, rc=rc, T=T, dT=dT, EJv=EJv, dEJv=dEJv, ndiffuse=ncells, kernel_ms=kernel_ms,
                       pass_ms=pass_ms, Eescape=tal["Eescape"], nsteps=tal["nsteps"])
            self.history.append(rec)
            cur = dict(T=T, dT=dT, dens=state["dens"], ddens=zeros, EJv=EJv, dEJv=dEJv)
            if not (self.struct_iter and not lastiter and self.maxiter > 0):
                rec["wall_s"] = time.perf_counter() - t0
                break
            niter += 1
            fw = average_weight(niter, self.f_weight, self.niter0)
            if niter == 1:
                tr.temp_average(fw, first=True)  # the set-up call allocates the average
            Tav = tr.temp_average(fw)
            state, nrk = tr.disk_structure(state, **self.structure_kw)
            m.dens, m.w = state["dens"], state["w"]
            tr.set_model(m, tables=False)
            cur = dict(cur, dens=state["dens"])
            errT, errRho, errE = tr.compare_cells(prev, cur)
            rec.update(Tav=Tav, rk_steps=nrk, errT=errT, errE=errE, fw=fw, wall_s=time.perf_counter() - t0)
            self._say(f"iteration {niter}: errT {errT:.3f} errE {errE:.3f} sigma, {ncells} cells diffused, "
                      f"pass {pass_ms / 1e3:.2f} s")
            if (errT < self.epsiter and finaliter) or niter > self.maxiter:
                lastiter = True
                continue
            finaliter = errT < self.epsiter
        self.state = state
        return self.history[-1]
