"""Learned warm start of the structure iteration.

The reference starts every model from the fixed guess of Init.f:3935 and
needs several Monte Carlo passes, each followed by a vertical-structure solve
(DiskStructure.f:19, 36-44, 71), before the temperatures settle.  Here a
network (mcmax_amd.mlp, BF16 matrix-core layers) trained on temperature
fields that the plain method produced maps each cell's features to log T.
The warm start solves the vertical structure for the network's temperatures
first, then runs the plain iteration; the first Monte Carlo pass judges the
network: when the median |T_pass/T_net - 1| over the disk exceeds `judge`
(5%) the model falls back to the cold start.

Features per cell (radial index i >= 1; the void row is left out), from the
model's set-up structure:
  log10 r/AU, theta, log10 tau_radial(1 um), log10 tau_vertical(1 um),
  Teff/1000 K, log10 L*/Lsun, flaring index, log10 Mdust/Msun
with tau_radial from the inner edge outwards along the cell's theta and
tau_vertical from the pole down along its radius, each to the cell centre.
"""
import math
import time

import numpy as np

from . import iterate, mlp, sweep, transport

AU = 1.49598e13
LSUN = 3.846e33
MSUN = 1.989e33
FEATURES = ("log10 r/AU", "theta", "log10 tau_r(1um)", "log10 tau_v(1um)", "Teff/1000K", "log10 L*/Lsun",
            "flaring index", "log10 Mdust/Msun")


def kext_at(m, lam0=1.0):
    """each species' extinction opacity at lam0 micron (log-log interpolation)"""
    ll = np.log(m.lam)
    return np.array([math.exp(np.interp(math.log(lam0), ll, np.log(np.maximum(k, 1e-300)))) for k in m.Kext])


def features(m, dens=None):
    """[(nr-1)*nt, 8] features of the disk cells (row-major over i >= 1, j);
    dens: the density the optical depths are taken from (default m.dens)"""
    dens = m.dens if dens is None else dens
    kx = kext_at(m)
    krho = (m.w @ kx) * dens  # [nr, nt] 1/cm
    dR = np.diff(m.R) * AU  # [nr]
    tr = np.cumsum(krho * dR[:, None], axis=0) - 0.5 * krho * dR[:, None]
    th_edge = np.arccos(np.clip(m.Theta, -1.0, 1.0))  # [nt+1], 0 at the pole
    dth = np.diff(th_edge)  # [nt]
    ds = m.R_av[:, None] * dth[None, :]
    tv = np.cumsum(krho * ds, axis=1) - 0.5 * krho * ds
    nr, nt = m.nr, m.nt
    r = np.broadcast_to((m.R_av / AU)[:, None], (nr, nt))
    th = np.broadcast_to(m.theta_av[None, :], (nr, nt))
    shpow = float(m.keywords["shpow"])
    X = np.empty((nr - 1, nt, 8))
    X[..., 0] = np.log10(r[1:])
    X[..., 1] = th[1:]
    X[..., 2] = np.log10(tr[1:] + 1e-10)
    X[..., 3] = np.log10(tv[1:] + 1e-10)
    X[..., 4] = m.Tstar / 1000.0
    X[..., 5] = math.log10(m.Lstar / LSUN)
    X[..., 6] = shpow
    X[..., 7] = math.log10(m.Mdust_g / MSUN)
    return X.reshape(-1, 8)


def median_rel(Ta, Tb):
    """median |Ta/Tb - 1| over the disk cells (radial index >= 1)"""
    a, b = np.asarray(Ta)[1:].ravel(), np.asarray(Tb)[1:].ravel()
    ok = b > 0
    return float(np.median(np.abs(a[ok] / b[ok] - 1.0)))


def field_from_logT(m, y):
    """[nr, nt] temperature field from per-cell log10 T of the disk cells; the
    void row gets the set-up's value there"""
    T = iterate.initial_temperature(m)
    T[1:] = (10.0 ** np.asarray(y, dtype=np.float64)).reshape(m.nr - 1, m.nt)
    return T


class Batch(sweep.Sweep):
    """Models on their own contexts whose passes run batched, with per-model
    histories and an optional warm start."""

    maxinteract = 0  # > 0: the reference's interaction roulette (keyword maxinteract)

    def iterate(self, niter, seed=1, T_init=None, judge=0.05, graph=False, blocks_per_model=0):
        """niter Monte Carlo passes per model, each followed by the rest of
        the structure iteration (inversion, Diffuse, TempAverage,
        DiskStructure).  T_init[k] (or None): the warm start of model k --
        its vertical structure is solved for T_init[k] before the first
        pass, and that pass judges it.  Returns dict(T=[model][pass] fields,
        accepted=[model] bool or None, judged=[model] median rel. change,
        transport_s, total_s)."""
        nm = len(self.models)
        init = [(m.dens.copy(), m.w.copy()) for m in self.models]
        states = [iterate.density_state(m) for m in self.models]
        hist = [[] for _ in range(nm)]
        it_of = [0] * nm  # structure iterations done by each model
        accepted, judged = [None] * nm, [None] * nm
        if T_init is not None:
            for k in range(nm):
                if T_init[k] is None:
                    continue
                t = self.trs[k]
                t.set_temperature(T_init[k])
                t.temp_average(1.0, first=True)
                t.temp_average(1.0)
                states[k], _ = t.disk_structure(states[k])
                m = self.models[k]
                m.dens, m.w = states[k]["dens"], states[k]["w"]
                t.set_model(m, tables=False)
                it_of[k] = 1
        t_tr, t0 = 0.0, time.perf_counter()
        for p in range(1, niter + 1):
            t1 = time.perf_counter()
            st = transport.batch_transport(self.trs, self.nphot, seed=seed + 1000 * p, blocks_per_model=blocks_per_model,
                                           graph=graph, maxinteract=self.maxinteract)
            if np.any(st != 0):
                raise RuntimeError(f"batched pass {p}: status {st}")
            t_tr += time.perf_counter() - t1
            if self.log:
                self.log(f"pass {p}: transport of {nm} models {time.perf_counter() - t1:.1f} s")
            for k in range(nm):
                t, m = self.trs[k], self.models[k]
                T, _, _ = t.temperature()
                tal = t.tallies()
                T, _, _, _ = t.diffuse(25, 1.0, EJv=tal["EJv"] / m.V)
                hist[k].append(T)
                if p == 1 and T_init is not None and T_init[k] is not None:
                    judged[k] = median_rel(T, T_init[k])
                    accepted[k] = judged[k] <= judge
                    if not accepted[k]:
                        # fall back: the set-up structure, a cold first iteration
                        m.dens, m.w = init[k][0].copy(), init[k][1].copy()
                        states[k] = iterate.density_state(m)
                        t.set_model(m, tables=False)
                        it_of[k] = 0
                        hist[k][-1] = None  # this pass does not count as a cold pass
                        continue
                it_of[k] += 1
                fw = iterate.average_weight(it_of[k], 0.5, 1000)
                if it_of[k] == 1:
                    t.temp_average(fw, first=True)
                t.temp_average(fw)
                states[k], _ = t.disk_structure(states[k])
                m.dens, m.w = states[k]["dens"], states[k]["w"]
                t.set_model(m, tables=False)
            if self.log:
                self.log(f"pass {p}: {time.perf_counter() - t0:.1f} s")
        return dict(T=hist, accepted=accepted, judged=judged, transport_s=t_tr, total_s=time.perf_counter() - t0)


def dataset(params, workdir, nphot, niter=3, seed=1, cells_per_model=None, rng_seed=0, log=None, chunk=64,
            randomwalk=True, maxinteract=0, save=None):
    """Features and log10 T targets of the plain method's temperature after
    niter structure iterations, for the models of params (run in batches of
    `chunk`); cells_per_model > 0 keeps a random subset of each model's
    cells; maxinteract > 0 plays the reference's interaction roulette
    (MCRadiation.f:568-571), which bounds the longest histories of the
    thickest models.  save(X, y, k), when given, is called after every
    chunk with the rows so far.  Returns (X [rows, 8], y [rows], model index
    [rows])."""
    rng = np.random.default_rng(rng_seed)
    Xs, ys, ks = [], [], []
    for a in range(0, len(params), chunk):
        sub = params[a:a + chunk]
        B = Batch(sub, f"{workdir}/d{a:05d}", nphot, randomwalk=randomwalk, log=log)
        B.maxinteract = int(maxinteract)
        try:
            setup = [features(m) for m in B.models]
            res = B.iterate(niter, seed=seed + 7919 * a)
        finally:
            B.close()
        for k, (X, m) in enumerate(zip(setup, B.models)):
            y = np.log10(np.maximum(res["T"][k][-1][1:].ravel(), 1.0))
            idx = np.arange(len(y))
            if cells_per_model and cells_per_model < len(y):
                idx = np.sort(rng.choice(len(y), size=cells_per_model, replace=False))
            Xs.append(X[idx].astype(np.float32))
            ys.append(y[idx].astype(np.float32))
            ks.append(np.full(len(idx), a + k, dtype=np.int32))
        if log:
            log(f"dataset: {a + len(sub)} of {len(params)} models, pass time {res['transport_s']:.1f} s")
        if save:
            save(np.concatenate(Xs), np.concatenate(ys), np.concatenate(ks))
    return np.concatenate(Xs), np.concatenate(ys), np.concatenate(ks)


def train(X, y, steps=20000, batch=8192, lr=1e-3, lr_final=1e-5, seed=1, device=0, log=None, chunk=2000):
    """A network trained on (X, y) with Adam, the learning rate decaying
    geometrically from lr to lr_final; returns (network, loss history)."""
    net = mlp.Network(device, seed=seed)
    net.fit_normalisation(X, y)
    net.set_data(X, y)
    losses = []
    nchunk = max(1, steps // chunk)
    for c in range(nchunk):
        f = c / max(1, nchunk - 1)
        rate = lr * (lr_final / lr) ** f
        losses.append(net.train(chunk, batch=batch, lr=rate, seed=seed * 1000003 + c))
        if log:
            log(f"train: step {(c + 1) * chunk}, lr {rate:.2e}, loss {losses[-1][-200:].mean():.3e} (normalised)")
    return net, np.concatenate(losses)


def predict_fields(net, models):
    """the network's temperature field [nr, nt] for each model (one forward
    call for all of their cells)"""
    Xs = [features(m) for m in models]
    y = net.predict(np.concatenate(Xs))
    out, a = [], 0
    for m, X in zip(models, Xs):
        out.append(field_from_logT(m, y[a:a + len(X)]))
        a += len(X)
    return out


def draw_models(n, seed, base_seed=20261203):
    """n model parameter sets drawn as the sweep draws them, from the stream
    `seed` (the training set and the evaluation sweep use different ones)"""
    return sweep.draw(n, seed=base_seed + int(seed))


def _main(argv=None):
    import argparse
    import os
    import sys
    import tempfile
    ap = argparse.ArgumentParser(prog="python -m mcmax_amd.warmstart",
                                 description="training data and training of the warm-start network")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a1 = sub.add_parser("data", help="plain-method temperature fields of drawn models -> features and targets")
    a1.add_argument("--models", type=int, default=256)
    a1.add_argument("--first", type=int, default=0, help="index of the first model of the stream")
    a1.add_argument("--stream", type=int, default=1, help="model stream (the evaluation sweep uses 0)")
    a1.add_argument("--nphot", type=int, default=2_000_000)
    a1.add_argument("--niter", type=int, default=3)
    a1.add_argument("--cells", type=int, default=0, help="cells kept per model (0: all)")
    a1.add_argument("--chunk", type=int, default=64)
    a1.add_argument("--maxinteract", type=int, default=0, help="the reference's interaction roulette (0: off)")
    a1.add_argument("--out", required=True)
    a3 = sub.add_parser("forward", help="time the network's forward pass on random rows (matrix-core throughput)")
    a3.add_argument("--rows", type=int, default=1 << 22)
    a3.add_argument("--repeat", type=int, default=3)
    a3.add_argument("--network", default="")
    a2 = sub.add_parser("train", help="train the network on data files")
    a2.add_argument("data", nargs="+")
    a2.add_argument("--steps", type=int, default=20000)
    a2.add_argument("--batch", type=int, default=8192)
    a2.add_argument("--lr", type=float, default=1e-3)
    a2.add_argument("--seed", type=int, default=1)
    a2.add_argument("--holdout", type=float, default=0.1,
                    help="fraction of the models kept out of training to judge the network")
    a2.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    log = lambda s: print(f"[warmstart] {s}", file=sys.stderr, flush=True)
    if a.cmd == "forward":
        import json
        net = mlp.Network.load(a.network) if a.network else mlp.Network(0, seed=1)
        X = np.random.default_rng(1).normal(size=(a.rows, 8)).astype(np.float32)
        best = None
        for _ in range(a.repeat):
            net.forward_raw(X)
            t = net.timing()
            best = t["forward_ms"] if best is None else min(best, t["forward_ms"])
        flop = 2.0 * (8 * 256 + 3 * 256 * 256 + 256) * a.rows
        print(json.dumps({"kernel": "k_mlp_fwd", "rows": a.rows, "ms": best, "tflops": flop / (best * 1e-3) / 1e12,
                          "frac_of_bf16_dense_peak": flop / (best * 1e-3) / 2.5e15}))
        return
    if a.cmd == "data":
        params = draw_models(a.first + a.models, a.stream)[a.first:]
        t0 = time.perf_counter()
        def save(X, y, k):
            np.savez(a.out + ".tmp.npz", X=X, y=y, model=k + a.first, nphot=a.nphot, niter=a.niter, stream=a.stream,
                     maxinteract=a.maxinteract)
            os.replace(a.out + ".tmp.npz", a.out)
        X, y, k = dataset(params, tempfile.mkdtemp(prefix="mcx_ws_"), a.nphot, niter=a.niter,
                          cells_per_model=a.cells, rng_seed=a.first, log=log, chunk=a.chunk, maxinteract=a.maxinteract,
                          save=save)
        log(f"{len(y)} rows from {a.models} models in {time.perf_counter() - t0:.1f} s -> {a.out}")
    else:
        Xs, ys, ks = [], [], []
        for p in a.data:
            d = np.load(p, allow_pickle=False)
            Xs.append(d["X"])
            ys.append(d["y"])
            ks.append(d["model"])
        X, y, k = np.concatenate(Xs), np.concatenate(ys), np.concatenate(ks)
        models = np.unique(k)
        held = models[len(models) - int(round(a.holdout * len(models))):] if a.holdout > 0 else models[:0]
        test = np.isin(k, held)
        t0 = time.perf_counter()
        net, loss = train(X[~test], y[~test], steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        net.save(a.out)
        log(f"trained on {int((~test).sum())} rows of {len(models) - len(held)} models, {a.steps} steps in "
            f"{time.perf_counter() - t0:.1f} s ({net.timing()['step_ms']:.3f} ms per step); final loss "
            f"{loss[-500:].mean():.3e} (normalised) -> {a.out}")
        if test.any():
            rel = np.abs(10.0 ** (net.predict(X[test]) - y[test]) - 1.0)
            per_model = [float(np.median(rel[k[test] == m])) for m in held]
            log(f"held-out models ({len(held)}): |dT|/T median {np.median(rel):.4f}, 90th percentile "
                f"{np.percentile(rel, 90):.4f}; per-model medians: median {np.median(per_model):.4f}, "
                f"worst {max(per_model):.4f}")


if __name__ == "__main__":
    _main()
