"""Sweeps of many disk models run together (mcx_batch_transport).

The reference holds one model in its module globals and runs the Monte
Carlo pass for it (Modules.f:141-150, MCRadiation.f:6); a parameter study is
many runs one after another.  Here every model has its own context on the
device and each structure iteration's transport pass runs for all of them
in common launches: the packet histories that end each stage run side by
side instead of model after model.  Each model's pass is bit-identical to
its own single-model pass with the same key, so the results do not depend
on which models share a launch.
"""
import os
import time

import numpy as np

from . import autotune, configs, iterate, transport

# the parameter ranges of the sweep (uniform unless noted)
RANGES = dict(nrad=(60, 240), ntheta=(40, 120), mdust=(1e-6, 1e-2, "log"), rin=(0.05, 1.0, "log"),
              rout=(50.0, 300.0), denspow=(0.5, 1.5), shpow=(1.0, 1.3), tstar=(3000.0, 10000.0),
              rstar=(1.0, 3.0))


def draw(n, seed=20261203):
    """n models' keyword overrides drawn from RANGES with a seeded generator
    (grids nr x ntheta uneven; dust mass, inner radius log-uniform)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        p = {}
        for k, r in RANGES.items():
            if k in ("nrad", "ntheta"):
                p[k] = int(rng.integers(r[0], r[1] + 1))
            elif len(r) == 3:
                p[k] = float(np.exp(rng.uniform(np.log(r[0]), np.log(r[1]))))
            else:
                p[k] = float(rng.uniform(r[0], r[1]))
        out.append(p)
    return out


class Sweep:
    """Models of the production dust (ten DHS species) with drawn star, disk
    and grid, each on its own context."""

    def __init__(self, params, workdir, nphot, randomwalk=True, device=0, base="production", log=None):
        self.nphot = int(nphot)
        self.log = log
        self.models, self.trs = [], []
        dust = transport.Transport(device)
        t0 = time.perf_counter()
        try:
            for k, p in enumerate(params):
                m = configs.load_config(base, os.path.join(workdir, f"m{k:03d}"), tr=dust, randomwalk=randomwalk, **p)
                self.models.append(m)
                if log and (k + 1) % 16 == 0:
                    log(f"{k + 1} of {len(params)} models built ({time.perf_counter() - t0:.1f} s)")
        finally:
            dust.close()
        for k, m in enumerate(self.models):
            t = transport.Transport(device)
            # DiskStructure gives most cells a species mix of their own: the
            # per-mix tables over T (80-240 kB each, 16 MB with the
            # reemission cumulative) would take GBs per model.  At most 1 MiB
            # per table: every model sums the species' cumulatives per probe
            # (one kernel variant for the whole batch in every iteration).
            t.class_table_bytes = 1 << 20
            t.set_model(m)
            self.trs.append(t)
            if log and (k + 1) % 16 == 0:
                log(f"{k + 1} of {len(self.models)} models on the device ({time.perf_counter() - t0:.1f} s)")

    def close(self):
        for t in self.trs:
            t.close()
        self.trs = []

    def _after_pass(self, k, states, niter):
        """the rest of one structure iteration for model k (Main.f:458-791):
        inversion, Diffuse, TempAverage, DiskStructure, the new structure"""
        t, m = self.trs[k], self.models[k]
        T, dT, dEJv = t.temperature()
        tal = t.tallies()
        EJv = tal["EJv"] / m.V
        T, _, EJv, _ = t.diffuse(25, 1.0, EJv=EJv)
        fw = iterate.average_weight(niter, 0.5, 1000)
        if niter == 1:
            t.temp_average(fw, first=True)
        t.temp_average(fw)
        states[k], _ = t.disk_structure(states[k])
        m.dens, m.w = states[k]["dens"], states[k]["w"]
        t.set_model(m, tables=False)
        return T

    maxinteract = 0  # > 0: the reference's interaction roulette (keyword maxinteract)

    @staticmethod
    def shares(work, total, minimum=2):
        """Workgroups per model in proportion to its work (at least minimum
        each), summing to total."""
        work = np.maximum(np.asarray(work, dtype=float), 1.0)
        n = len(work)
        if total < minimum * n:
            raise ValueError("fewer workgroups than the minimum shares")
        extra = total - minimum * n
        raw = extra * work / work.sum()
        out = minimum + np.floor(raw).astype(np.int64)
        left = total - int(out.sum())
        out[np.argsort(-(raw - np.floor(raw)))[:left]] += 1
        return out.astype(np.int32)

    @staticmethod
    def cohorts(nm, size):
        """The models' launches: consecutive groups of at most size models
        (0: one launch for all)."""
        size = size if size and size < nm else nm
        return [list(range(a, min(a + size, nm))) for a in range(0, nm, size)]

    @staticmethod
    def cohort_key(seed, it, seed_stride, first):
        """The base key of a launch whose first model is `first`: its model q
        then draws with seed + 1000 it + seed_stride (first + q), the key
        the model has whatever its launch (and alone, run_one)."""
        return seed + 1000 * it + seed_stride * first

    def run(self, niter=3, seed=1, graph=True, blocks_per_model=0, seed_stride=1, balance=True, cohort=0):
        """niter structure iterations of every model, each pass batched:
        the models in cohorts of at most `cohort` per launch (0: all in
        one), launch after launch.  balance: from the second pass on, each
        model's share of its launch follows the lane time it took in the
        pass before (its workgroups times its finish; the launch's total
        stays that of the first pass, blocks_per_model each, 0: the
        device's resident workgroups shared out).  Model k's key is
        seed + 1000 iteration + seed_stride k whatever its cohort.  Returns
        dict(T: last temperatures, transport_s, total_s, shares, finish_ms:
        per pass the models' finish times)."""
        nm = len(self.models)
        groups = self.cohorts(nm, cohort)
        states = [iterate.density_state(m) for m in self.models]
        t_tr, t0 = 0.0, time.perf_counter()
        T = [None] * nm
        shares = [None] * len(groups)
        hist, finish = [], []
        for it in range(1, niter + 1):
            fin = [0.0] * nm
            hist.append([None if sh is None else sh.tolist() for sh in shares])
            for g, ks in enumerate(groups):
                t1 = time.perf_counter()
                trs = [self.trs[k] for k in ks]
                st = transport.batch_transport(trs, self.nphot, seed=self.cohort_key(seed, it, seed_stride, ks[0]),
                                               seed_stride=seed_stride, blocks_per_model=blocks_per_model,
                                               graph=graph, maxinteract=self.maxinteract, model_blocks=shares[g])
                if np.any(st != 0):
                    raise RuntimeError(f"batched pass: status {st}")
                t_tr += time.perf_counter() - t1
                f = [trs[q].tallies()["t_done_ms"][4] for q in range(len(ks))]
                for q, k in enumerate(ks):
                    fin[k] = f[q]
                if self.log:
                    what = f"cohort {g + 1} of {len(groups)}: " if len(groups) > 1 else ""
                    self.log(f"iteration {it}: {what}batched pass of {len(ks)} models "
                             f"{time.perf_counter() - t1:.1f} s (models done after {min(f) / 1e3:.1f} s median "
                             f"{np.median(f) / 1e3:.1f} s last {max(f) / 1e3:.1f} s)")
                if balance and it < niter:
                    # the lane time each model took: its workgroups x its finish
                    bpm = blocks_per_model or max(1, autotune.RESIDENT // len(ks))
                    prev = np.full(len(ks), bpm) if shares[g] is None else shares[g]
                    shares[g] = self.shares(prev * np.asarray(f), int(prev.sum()), minimum=min(2, bpm))
            finish.append(fin)
            for k in range(nm):
                T[k] = self._after_pass(k, states, it)
                if self.log and (k + 1) % 64 == 0:
                    self.log(f"iteration {it}: {k + 1} models' structure solved")
            if self.log:
                self.log(f"iteration {it}: {time.perf_counter() - t0:.1f} s")
        return dict(T=T, transport_s=t_tr, total_s=time.perf_counter() - t0, shares=hist, finish_ms=finish)

    def run_one(self, k, niter=3, seed=1, seed_stride=1):
        """model k alone, its passes model by model (mcx_transport) with the
        keys the batched sweep gives it.  Returns dict(T, transport_s,
        total_s)."""
        t, m = self.trs[k], self.models[k]
        states = {k: iterate.density_state(m)}
        t_tr, t0 = 0.0, time.perf_counter()
        T = None
        for it in range(1, niter + 1):
            t1 = time.perf_counter()
            t.transport(self.nphot, seed=seed + 1000 * it + seed_stride * k, maxinteract=self.maxinteract)
            t_tr += time.perf_counter() - t1
            T = self._after_pass(k, states, it)
        return dict(T=T, transport_s=t_tr, total_s=time.perf_counter() - t0)
