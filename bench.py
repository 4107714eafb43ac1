"""Benchmark: packets/s of the MCMax transport pass on one MI355X.

A step is one Monte Carlo pass (MCRadiation.f:6) of `--nphot` packets over the
production disk model (200 r x 100 theta cells, 200 wavelengths, dust mass
1e-3 Msun, ten DHS dust species of 0.05-1000 um grains from m = 1.7 + 0.03i,
computed on the device by ComputePart's algorithm before the timed region)
followed by the Lucy temperature inversion (Main.f:498-549), with all tables
resident in HBM before the timed region.  The pass uses the modified
random walk in optically thick cells (keyword randomwalk, factRW = 10), as the
reference's example disks do; `--plain` switches it off.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "packets/s on 1 MI355X + time to converged T-field at ≤1% median |ΔT|/T"
HBM_PEAK = 8.0e12


def algorithmic_bytes(t, ng, nlam):
    """Bytes the transport must move per launch, from on-device event counts:
    per cell step rho, w(ng), EJv add, EJv2 add, Ni bit; per absorption the
    Eabs add, Kp reads and two CDF searches; per scattering the Mueller pick
    and the angle search; per random-walk step the Eabs and EJv adds, the cell
    T load and update, the (KDext, KDabs) pair, rho, two Kp entries and the
    survival-table search; 8 B per escape."""
    Bx = 8 + 8 * ng + 8 + 8 + 4
    Ba = 8 + 24 * ng + 16 * ng * math.ceil(math.log2(nlam)) + 16
    Bs = 8 * ng + 8 * math.ceil(math.log2(180))
    Brw = 8 + 8 + 16 + 16 + 8 + 16 + 8 * math.ceil(math.log2(1000)) + 8
    total = t["nsteps"] * Bx + t["nabs"] * Ba + t["nscat"] * Bs + t.get("nrw", 0) * Brw + 8 * t["nescape"]
    return total, (Bx, Ba, Bs, Brw)


def pass_atomics(tal, nphot, kern_s):
    """Device-scope atomics of the pass: the EJv/EJv2 limbs flushed at cell
    exits (counted by the kernel), one Eabs add per absorption and
    random-walk deposit, one Ni add per (packet, cell) entry, two spectrum
    adds per escape, one stage-readiness credit per packet; and their rate
    over the kernel's duration, against the measured peak of random int64
    atomic adds (profiles/atomics_microbench.json, bench.py --atomics)."""
    import numpy as np
    n = (int(tal["ntallyatomic"]) + int(tal["nabs"]) + int(tal.get("nrw", 0)) + int(np.sum(tal["Ni"])) +
         2 * int(tal["nescape"]) + int(nphot))
    out = {"issued": n, "per_s": n / kern_s}
    path = os.path.join(ROOT, "profiles", "atomics_microbench.json")
    if os.path.exists(path):
        with open(path) as f:
            peak = json.load(f).get("value")
        if peak:
            out["peak_per_s"] = peak
            out["frac"] = out["per_s"] / peak
    return out


def measured_traffic(config, nphot, ng, schedule):
    """HBM bytes of the pass's k_transport launch for this workload, from
    the rocprofv3 FETCH_SIZE/WRITE_SIZE runs summarised in profiles/ (scaled
    by what the counters read for a plain copy of known size); None when no
    measurement of this workload and stage schedule exists."""
    path = os.path.join(ROOT, "profiles", "k_transport_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if (t.get("config") != config or int(t.get("nphot", -1)) != int(nphot) or int(t.get("ng", -1)) != int(ng)
            or t.get("schedule") != schedule):
        return None
    return t.get("bytes_per_launch")


def _cpu_worker(args):
    """one host process: the oracle's serial loop on batches of the model
    until `seconds` of its own CPU time (fresh seeds per batch)"""
    m, seconds, seed0, n = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(m)
    steps, npk, b = 0, 0, 0
    t0 = time.process_time()
    while time.process_time() - t0 < seconds and b < 64:
        r = o.mcradiation(n, rng="philox", seed=seed0 + b)
        steps += r["nsteps"]
        npk += n
        b += 1
    return steps, npk, time.process_time() - t0


def host_cores():
    """the host cores this process may use (at most 16: the box's CPU share
    per GPU), and the CPU's model name"""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    name = "?"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(16, n)), n, name


def cpu_baseline(m, pass_steps, nphot, seconds=15.0):
    """The oracle (FP64 C restatement of the reference's serial loop) on the
    GPU box's host cores, on a bounded sample of the same workload: one
    process per core, each an independent serial pass over batches of
    packets (the reference's nruns sub-runs, MCRadiation.f:192, 269-668)
    for about `seconds` of CPU time.  Packet histories are heavy-tailed,
    so a sample's packets/s overstates a full pass: the rate is taken per
    cell step and applied to the cell steps of the measured pass
    (pass_steps for nphot packets)."""
    import multiprocessing as mp
    cores, avail, name = host_cores()
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(m)
    t0 = time.perf_counter()
    o.mcradiation(2000, rng="philox", seed=99)
    dt = time.perf_counter() - t0
    n = max(2000, int(2000 * 0.5 * seconds / max(dt, 1e-3)) // 8)
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, [(m, seconds, 1000 * (k + 1), n) for k in range(cores)])
    wall = time.perf_counter() - t0
    steps = sum(r[0] for r in res)
    npk = sum(r[1] for r in res)
    cpu1 = sum(r[0] / r[2] for r in res) / cores
    steps_per_s = steps / wall
    pass_s = pass_steps / steps_per_s
    return {"value": nphot / pass_s, "unit": "packets/s", "cores": cores, "kind": "port",
            "sample": f"{cores} host processes ({name}; {avail} cores visible) each running the oracle's serial "
                      f"loop on batches of {n} packets of the same model for ~{seconds:.0f} s CPU: {npk} packets, "
                      f"{steps} cell steps in {wall:.1f} s wall ({steps_per_s:.3e} steps/s; {cpu1:.3e} steps/s per "
                      f"core); at that rate the measured pass ({pass_steps} cell steps) takes {pass_s:.1f} s",
            "one_core_packets_per_s": nphot / (pass_steps / cpu1)}


def dust_bench(a, work, lam_stride=10):
    """The dust-opacity generator (ComputePart on the device) for the config's
    computepart particle: a warmup call, then `--steps` timed calls; the CPU
    oracle on one core for every species at every lam_stride-th wavelength,
    scaled to the full grid."""
    from mcmax_amd import configs, dust, keywords, model, transport
    path = configs.write_config(a.config, work)
    kw = keywords.parse_file(path)
    lam, _ = model.lam_grid(int(kw["nlam"]), float(kw["lam1"]), float(kw["lam2"]))
    specs = []
    for d in dust.particle_keywords(kw).values():
        specs += dust.species_specs(d, lam, work)[0]
    if not specs:
        raise SystemExit(f"config {a.config} has no computepart particle")
    tr = transport.Transport(0)
    for _ in range(a.warmup):
        tr.compute_parts(lam, specs)
    ms = []
    for _ in range(a.steps):
        t0 = time.perf_counter()
        _, st = tr.compute_parts(lam, specs)
        ms.append(1e3 * (time.perf_counter() - t0))
    tr.close()
    rows = len(specs) * len(lam)
    out = {"metric": "dust opacities: (species, wavelength) rows per second", "value": rows / (min(ms) / 1e3),
           "unit": "rows/s", "steps": a.steps, "warmup": a.warmup, "ms_per_call": min(ms),
           "device_ms": st["ms_total"], "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{a.config} computepart particle", "species": len(specs), "nlam": len(lam),
                      "ns": specs[0]["ns"], "nf": specs[0]["nf"]},
           "work": st}
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        sel = list(range(0, len(lam), lam_stride))
        t0 = time.perf_counter()
        for p in specs:
            pyoracle.compute_part(lam[sel], p["e1"][:, sel], p["e2"][:, sel], p["frac"], p["rho"], p["amin"],
                                  p["amax"], p["apow"], p["fmax"], p["ns"], p["nf"])
        dt = time.perf_counter() - t0
        n = len(specs) * len(sel)
        out["cpu_baseline"] = {"value": n / dt, "unit": "rows/s", "cores": 1, "kind": "port",
                               "sample": f"every species at {len(sel)} of {len(lam)} wavelengths ({n} rows) "
                                         f"in {dt:.1f} s on one host core"}
    return out


def stage_schedule(n, ratio=0, s0=64):
    """the stage ends of a concurrent pass (the library's schedule,
    mcx_transport): s0, then ratio x (ratio 0: 2; -1: 1 + sqrt(P/s0), at
    most 16)"""
    r0 = 2 if ratio == 0 else ratio
    P, out = min(s0, n), [min(s0, n)]
    while P < n:
        r = r0 if r0 >= 2 else min(16.0, 1.0 + math.sqrt(P / s0))
        nx = math.ceil(P * r)
        P = n if nx >= n else nx
        out.append(P)
    return out


def _median_rel(Ta, Tb):
    """median |Ta/Tb - 1| over the disk cells (radial index >= 1)."""
    import numpy as np
    a, b = np.asarray(Ta)[1:].ravel(), np.asarray(Tb)[1:].ravel()
    ok = b > 0
    return float(np.median(np.abs(a[ok] / b[ok] - 1.0)))


def converge_bench(a, work):
    """Wall time of the structure iteration (Main.f:202-791: pass, inversion,
    Diffuse, TempAverage, DiskStructure, compareCells) on the config, with
    --nphot packets per pass, until the reference's rule ends it (errT <
    epsiter on two iterations in a row, or maxiter, then one last pass).
    Reports per pass the wall time, errT and the median |dT|/T against the
    previous pass, and the time until that median first drops to 1%."""
    from mcmax_amd import configs, iterate, transport
    tr = transport.Transport(0)
    m = configs.load_config(a.config, work, tr=tr, randomwalk=not a.plain)
    log = lambda s_: print(f"[converge] {s_}", file=sys.stderr, flush=True)
    it = iterate.Iteration(tr, m, a.nphot, seed=a.seed, struct_iter=True, log=log)
    t0 = time.perf_counter()
    it.run()
    wall = time.perf_counter() - t0
    tr.close()
    passes, cum, t_conv, prev = [], 0.0, None, None
    for k, r in enumerate(it.history):
        cum += r["wall_s"]
        rel = _median_rel(r["T"], prev) if prev is not None else None
        passes.append({"pass": k, "wall_s": r["wall_s"], "cumulative_s": cum, "kernel_ms": r["kernel_ms"],
                       "errT": r.get("errT"), "median_dT_over_T_vs_previous": rel, "ndiffuse": r["ndiffuse"]})
        if t_conv is None and rel is not None and rel <= 0.01:
            t_conv = cum
        prev = r["T"]
    final = it.history[-1]["T"]
    for q, r in zip(passes, it.history):
        q["median_dT_over_T_vs_final"] = _median_rel(r["T"], final)
    return {"metric": "time to converged T-field at <=1% median |dT|/T (structure iteration)",
            "value": t_conv, "unit": "s", "higher_is_better": False, "n_gpus": 1, "wall_s": wall,
            "passes": passes, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{a.config} disk model, structure iteration", "nphot": a.nphot,
                       "maxiter": it.maxiter, "epsiter": it.epsiter, "ng": m.ng,
                       "randomwalk": not a.plain},
            "converged_by": "median |dT|/T between consecutive passes <= 1% (the reference's own rule, "
                            "errT < epsiter twice, decides when the loop stops)"}


def converged_small(work, seed=11):
    """The metric's second half on the CPU-runnable plain-method workload
    (small model: 100 r x 60 theta cells, 100 wavelengths, one analytic
    grain, no vertical-structure iteration, 1e5 packets, the modified random
    walk as in the production pass): wall time of the
    device's passes (each pass + Lucy inversion, 1e5 packets, doubled until
    the field is within 1%) until the median |dT|/T against the FP64
    oracle's converged field (tests/golden/converged_small.npz, 2e6 packets
    in 8 sub-runs) is at most 1%."""
    import numpy as np
    from mcmax_amd import configs, transport
    path = os.path.join(ROOT, "tests", "golden", "converged_small.npz")
    if not os.path.exists(path):
        return None
    ref = np.load(path)
    tr = transport.Transport(0)
    try:
        t0 = time.perf_counter()
        m = configs.load_config("small", os.path.join(work, "small"), tr=tr, randomwalk=True)
        tr.set_model(m)
        setup_s = time.perf_counter() - t0
        passes, total, n = [], 0.0, 100000
        while True:
            t1 = time.perf_counter()
            tr.transport(n, seed=seed + len(passes))
            T, _, _ = tr.temperature()
            dt = time.perf_counter() - t1
            total += dt
            med = float(np.median(np.abs(T / ref["T"] - 1.0)))
            passes.append({"nphot": n, "wall_s": dt, "median_dT_over_T_vs_oracle": med})
            if med <= 0.01 or n >= 3200000:
                break
            n *= 2
    finally:
        tr.close()
    return {"workload": "small plain-method model (100 r x 60 theta, 100 wavelengths, one analytic grain), "
                        "modified random walk (factRW=10), passes of 1e5 packets doubled until converged",
            "reference": f"FP64 oracle field, {int(ref['nphot'])} packets in {int(ref['subruns'])} sub-runs",
            "criterion": "median |dT|/T <= 1% against the reference", "converged": passes[-1][
                "median_dT_over_T_vs_oracle"] <= 0.01,
            "time_s": total, "setup_s": setup_s, "passes": passes}


def image_bench(a, work, lam0=1.6):
    """An IMAGE observation with the reference's defaults (determineoutput,
    Telescope.f:590-657: 500 x 500 pixels, 1000 samples per polar cell,
    nphi 45, nrad 2, 1000 + 1000 scattering packets, the field the disk's
    diameter) at lam0 micron, inclination the model's, on the config's
    disk with power-law temperatures: a warmup call, then `--steps` timed
    calls of mcx_image (rays and spread).  The CPU leg runs the oracle's
    MakeImage on the rings of a bounded sample and scales to all rings."""
    import numpy as np
    from mcmax_amd import configs, transport
    tr = transport.Transport(0)
    m = configs.load_config(a.config, work, tr=tr)
    tr.set_model(m)
    T = np.empty((m.nr, m.nt))
    T[:] = np.minimum(1500.0, 400.0 * (m.R_av / 1.496e13) ** -0.5)[:, None]
    tr.set_temperature(T)
    npix, nint = 500, 1000
    Rmax = float(m.R[-1])
    ob = transport.make_obs(angle_deg=float(m.incangle), nphi=45, nj0=2, nphot=1000, nphotstar=1000, seed=a.seed)
    for _ in range(a.warmup):
        tr.image(ob, lam0, npix, nint, Rmax, seed=a.seed)
    walls, rays_ms, spread_ms = [], [], []
    for s in range(a.steps):
        t0 = time.perf_counter()
        raw, Rimg, rays = tr.image(ob, lam0, npix, nint, Rmax, seed=a.seed + s, rays=True)
        walls.append(time.perf_counter() - t0)
        r_ms, s_ms = tr.last_image_timing()
        rays_ms.append(r_ms)
        spread_ms.append(s_ms)
        print(f"[bench] image {s}: {1e3 * walls[-1]:.1f} ms (rays {r_ms:.1f} ms, spread {s_ms:.2f} ms)",
              file=sys.stderr, flush=True)
    tr.close()
    nimg, nphi = rays.shape[1], rays.shape[2]
    samples = (nimg - 1) * nphi * nint
    # per sample: the four neighbouring rays' Stokes entries read, two
    # deposits of (1 or 4) atomic adds; all of it L2-resident
    nst = 4 if m.scat_how == 2 else 1
    b_sample = 4 * nst * 8 + 2 * nst * 8
    spread_s = sum(spread_ms) / len(spread_ms) * 1e-3
    wall = sum(walls) / len(walls)
    out = {"metric": "IMAGE observations per second (rays + MakeImage spread)", "value": 1.0 / wall,
           "unit": "images/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * wall,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"{a.config} disk model, IMAGE at {lam0} um, 500x500 px, nint 1000, nphi 45",
                      "nimg": nimg, "nphi": nphi, "samples": samples, "scat_how": m.scat_how, "ng": m.ng},
           "rays_ms": sum(rays_ms) / len(rays_ms), "spread_ms": 1e3 * spread_s,
           "spread": {"kernel": "k_image", "samples_per_s": samples / spread_s,
                      "bytes_per_sample": b_sample, "l2_bytes_per_s": samples * b_sample / spread_s},
           "flux_check": float(raw[0].sum())}
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        lib = pyoracle.load()
        # the whole spread of the last image's rays (a few seconds on one core)
        t0 = time.perf_counter()
        ref = pyoracle.make_image(lib, Rimg, rays, m.scat_how, npix, nint, Rmax, 0.0, a.seed + a.steps - 1)
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": samples / dt, "unit": "samples/s (spread)", "cores": 1, "kind": "port",
                               "sample": f"MakeImage restated in C on all {nimg - 1} rings ({samples} samples) in "
                                         f"{dt:.1f} s on one host core, against the device's "
                                         f"{1e3 * spread_s:.2f} ms; the two images' I totals differ by "
                                         f"{abs(ref[0].sum() / raw[0].sum() - 1):.1e}"}
    return out


def precision_study(a, work):
    """The precision study: the config with --mdust (default 1e-2 Msun:
    midplane radial tau ~1e5), `--steps` pairs of passes with the same seeds
    in FP64 and in FP32 (MCX_FP32); per-cell T error and MC SED error of the
    FP32 passes against the FP64 passes, and the time of each."""
    import numpy as np
    from mcmax_amd import configs, transport
    tr = transport.Transport(0)
    m = configs.load_config(a.config, work, tr=tr, randomwalk=not a.plain, mdust=a.mdust)
    tr.set_model(m)
    res = {0: [], 1: []}
    times = {0: [], 1: []}
    for s in range(a.steps):
        for prec in (0, 1):
            g = transport.mc_radiation(tr, a.nphot, seed=a.seed + s, precision=prec, max_ms=1e3 * a.budget)
            res[prec].append(g)
            times[prec].append(g["kernel_ms"])
            print(f"[bench] precision {prec} seed {a.seed + s}: {g['kernel_ms']:.0f} ms, maxsteps {g['maxsteps']}",
                  file=sys.stderr, flush=True)
    tr.close()
    T64 = np.array([r["T"][1:] for r in res[0]])
    T32 = np.array([r["T"][1:] for r in res[1]])
    ok = np.array([r["Ni"][1:] for r in res[0] + res[1]]).min(axis=0) > 20
    rel = np.abs(T32.mean(0) / T64.mean(0) - 1)[ok]
    S64 = np.array([r["spec"].sum(axis=0) for r in res[0]]).mean(0)
    S32 = np.array([r["spec"].sum(axis=0) for r in res[1]]).mean(0)
    n64 = np.array([r["ispec"].sum(axis=0) for r in res[0]]).sum(0)
    well = n64 >= 10000
    sed = np.abs(S32[well] / S64[well] - 1)
    return {"metric": "FP32 transport: per-cell T and MC SED error against FP64 passes with the same seeds",
            "value": float(np.median(rel)), "unit": "median |dT|/T", "higher_is_better": False,
            "steps": a.steps, "nphot": a.nphot, "data": "synthetic",
            "config": {"workload": f"{a.config} disk model at mdust {a.mdust} Msun", "nr": m.nr, "nt": m.nt,
                       "ng": m.ng, "randomwalk": not a.plain},
            "T_rel_p90": float(np.percentile(rel, 90)), "T_rel_max": float(rel.max()), "cells": int(ok.sum()),
            "sed_max_rel_well_sampled": float(sed.max()) if well.any() else None, "sed_bins": int(well.sum()),
            "ms_fp64": times[0], "ms_fp32": times[1]}


def sweep_bench(a, work):
    """The many-model sweep: --sweep models drawn from the sweep ranges
    (uneven grids 60-240 x 40-120, dust masses 1e-6..1e-2 Msun, stars
    3000-10000 K), the production dust, --nphot packets per model per
    iteration, 3 structure iterations with every pass batched (graph
    replay, blocks per model from the autotuner; with --cohort N the models
    go N to a launch, launch after launch).  Then --sweep-check models
    run alone, model by model (mcx_transport) with the same keys: their
    final temperatures must equal the batched sweep's bit for bit, and their
    mean time per model gives the model-by-model time of the whole sweep."""
    import numpy as np
    from mcmax_amd import autotune, sweep
    log = lambda s_: print(f"[sweep] {s_}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    params = sweep.draw(a.sweep, seed=20261203)
    S = sweep.Sweep(params, work, a.nphot, randomwalk=not a.plain, log=log)
    S.maxinteract = a.maxinteract
    setup_s = time.perf_counter() - t0
    log(f"{a.sweep} models set up in {setup_s:.1f} s")
    init = [(m.dens.copy(), m.w.copy()) for m in S.models]
    size = a.cohort if a.cohort and a.cohort < a.sweep else a.sweep
    if a.alone_only:
        # only the model-by-model timing: --sweep-check models alone
        rng = np.random.default_rng(11)
        check = sorted(rng.choice(a.sweep, size=min(a.sweep_check, a.sweep), replace=False).tolist())
        alone = []
        for k in check:
            r = S.run_one(k, niter=3, seed=a.seed)
            alone.append(r["total_s"])
            log(f"model {k} alone: {r['total_s']:.1f} s")
        S.close()
        return {"metric": f"{a.sweep}-model sweep, model by model: wall time per model (3 structure iterations)",
                "value": float(np.mean(alone)), "unit": "s", "higher_is_better": False, "n_gpus": 1,
                "models": a.sweep, "checked_models": check, "model_by_model_s_per_model": alone,
                "model_by_model_estimate_s": float(np.mean(alone)) * a.sweep, "nphot": a.nphot,
                "maxinteract": a.maxinteract}
    bpm, tune_rec = autotune.tune(S.trs[:size], a.nphot, log=log, maxinteract=a.maxinteract)
    res = S.run(niter=3, seed=a.seed, graph=True, blocks_per_model=bpm, cohort=size)
    log(f"batched sweep: {res['total_s']:.1f} s ({res['transport_s']:.1f} s in the passes)")
    rng = np.random.default_rng(7)
    check = sorted(rng.choice(a.sweep, size=min(a.sweep_check, a.sweep), replace=False).tolist())
    alone, identical = [], []
    for k in check:
        m = S.models[k]
        m.dens, m.w = init[k][0].copy(), init[k][1].copy()
        S.trs[k].set_model(m, tables=False)
        r = S.run_one(k, niter=3, seed=a.seed)
        alone.append(r["total_s"])
        identical.append(bool(np.array_equal(r["T"], res["T"][k])))
        log(f"model {k} alone: {r['total_s']:.1f} s, identical {identical[-1]}")
    S.close()
    seq_est = float(np.mean(alone)) * a.sweep
    ncell = [int(p["nrad"]) * int(p["ntheta"]) for p in params]
    return {"metric": f"{a.sweep}-model sweep: wall time to 3 structure iterations", "value": res["total_s"],
            "unit": "s", "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{a.sweep} disk models drawn from the sweep ranges (seed 20261203), production "
                                   f"dust, {a.nphot} packets per model per iteration, 3 iterations, modified random "
                                   f"walk {not a.plain}", "models": a.sweep, "cells_min": min(ncell),
                       "cells_max": max(ncell), "maxinteract": a.maxinteract, "models_per_launch": size},
            "packets_per_s": 3 * a.nphot * a.sweep / res["transport_s"],
            "transport_s": res["transport_s"], "setup_s": setup_s, "blocks_per_model": bpm, "autotune": tune_rec,
            "model_by_model_s_per_model": alone, "checked_models": check, "identical_to_alone": identical,
            "model_by_model_estimate_s": seq_est, "speedup_vs_model_by_model": seq_est / res["total_s"]}


def warmstart_bench(a, work):
    """The learned warm start on a fresh sweep: --sweep models (default 64)
    drawn from the sweep ranges with the evaluation stream (the network was
    trained on another stream), --nphot packets per model per pass, K =
    --passes passes each, batched: two independent cold starts (A: the
    reference field, its last pass; B: the cold accuracy per pass) and the
    warm start W (the network's field first, judged by the first pass with
    fallback above 5%).  Reports per pass the median over models of the
    median |dT|/T against A's final field, the passes cold and warm need to
    reach 1% (the metric's convergence threshold), the packets that saves,
    and the network's forward throughput on the matrix cores."""
    import numpy as np
    from mcmax_amd import mlp, warmstart
    log = lambda s_: print(f"[warmstart] {s_}", file=sys.stderr, flush=True)
    path = a.network or os.path.join(ROOT, "mcmax_amd", "data", "warmstart_net.npz")
    net = mlp.Network.load(path)
    nm, K = (a.sweep or 64), a.passes
    params = warmstart.draw_models(nm, seed=0)
    runs = {}
    for name, seed, warm in (("A", a.seed, False), ("B", a.seed + 7, False), ("W", a.seed + 13, True)):
        B = warmstart.Batch(params, os.path.join(work, name), a.nphot, randomwalk=not a.plain, log=log)
        B.maxinteract = a.maxinteract
        try:
            T0 = None
            t_net = 0.0
            if warm:
                t1 = time.perf_counter()
                T0 = warmstart.predict_fields(net, B.models)
                t_net = time.perf_counter() - t1
            t1 = time.perf_counter()
            r = B.iterate(K, seed=seed, T_init=T0)
            r["wall_s"] = time.perf_counter() - t1
            r["net_s"] = t_net
            r["T0"] = T0
            runs[name] = r
            log(f"run {name}: {r['wall_s']:.1f} s ({r['transport_s']:.1f} s in the passes)")
        finally:
            B.close()
    ref = [runs["A"]["T"][k][-1] for k in range(nm)]

    def curve(r):
        out = []
        for p in range(K):
            v = [warmstart.median_rel(r["T"][k][p], ref[k]) for k in range(nm) if r["T"][k][p] is not None]
            out.append(float(np.median(v)) if v else None)
        return out
    cold, warm = curve(runs["B"]), curve(runs["W"])
    net0 = float(np.median([warmstart.median_rel(runs["W"]["T0"][k], ref[k]) for k in range(nm)]))
    floor = cold[-1]
    # converged: the metric's threshold, median |dT|/T <= 1% against the
    # reference field (a second cold start, its last pass)
    target = 0.01
    reach = lambda c: next((p + 1 for p, v in enumerate(c) if v is not None and v <= target), None)
    pc, pw = reach(cold), reach(warm)
    acc = [x for x in runs["W"]["accepted"] if x is not None]
    # network throughput on a large batch of cells
    X = np.random.default_rng(1).normal(size=(1 << 22, 8)).astype(np.float32)
    net.forward_raw(X)
    t = net.timing()
    flop = 2.0 * (8 * 256 + 3 * 256 * 256 + 256) * t["forward_rows"]
    tfs = flop / (t["forward_ms"] * 1e-3) / 1e12
    return {"metric": "learned warm start: passes to the cold start's final accuracy", "value": pw, "unit": "passes",
            "higher_is_better": False, "n_gpus": 1, "dtype": "bf16 network, f64 transport", "data": "synthetic",
            "config": {"workload": f"{nm} disk models from the sweep ranges (evaluation stream), production dust, "
                                   f"{a.nphot} packets per model per pass, {K} passes", "models": nm,
                       "maxinteract": a.maxinteract, "network": os.path.relpath(path, ROOT)},
            "cold_passes": pc, "warm_passes": pw,
            "passes_saved": (pc - pw) if (pc and pw) else None,
            "packets_saved": ((pc - pw) * a.nphot * nm) if (pc and pw) else None,
            "target_median_dT_over_T": target, "cold_final_accuracy": floor,
            "median_dT_over_T_per_pass": {"cold": cold, "warm": warm, "network_alone": net0},
            "accepted": int(sum(acc)), "judged": len(acc),
            "judge_median": float(np.median([x for x in runs["W"]["judged"] if x is not None])),
            "wall_s": {k: runs[k]["wall_s"] for k in runs}, "network_s": runs["W"]["net_s"],
            "roofline": {"bound": "mfma", "kernel": "k_mlp_fwd", "achieved": tfs, "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": tfs / 2500.0, "rows": t["forward_rows"], "ms": t["forward_ms"],
                         "flop_per_row": 2.0 * (8 * 256 + 3 * 256 * 256 + 256)}}


def atomics_bench(a):
    """Peak rate of random 64-bit integer atomic adds on this GPU, the
    reference for the pass's tally atomics: torch's scatter_add_ (one
    device atomicAdd per element) of 2^26 random indices into int64 arrays
    of the production grid's size (20100 cells: the tallies' working set)
    and of 2^27 elements (HBM-resident)."""
    import torch
    out = {"metric": "random int64 atomic adds per second (torch scatter_add_)", "unit": "atomics/s", "cases": []}
    g = torch.Generator(device="cuda").manual_seed(1)
    n = 1 << 26
    for size in (20100, 1 << 27):
        dst = torch.zeros(size, dtype=torch.int64, device="cuda")
        idx = torch.randint(0, size, (n,), generator=g, device="cuda")
        val = torch.ones(n, dtype=torch.int64, device="cuda")
        for _ in range(2):
            dst.scatter_add_(0, idx, val)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record()
        for _ in range(reps):
            dst.scatter_add_(0, idx, val)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        assert int(dst.sum().item()) == (reps + 2) * n
        out["cases"].append({"array_elements": size, "atomics": n, "ms": ms, "atomics_per_s": n / (ms * 1e-3)})
    out["value"] = max(c["atomics_per_s"] for c in out["cases"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nphot", type=int, default=10_000_000)
    ap.add_argument("--mdust", type=float, default=1e-2, help="dust mass of --precision-study (Msun)")
    ap.add_argument("--config", default="production")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain", action="store_true", help="no modified random walk")
    ap.add_argument("--lanes", type=int, default=0,
                    help="packets in flight (0: the library's default; 1: one lane, packets in order)")
    ap.add_argument("--converge", action="store_true",
                    help="time the structure iteration (passes until the T-field is converged) instead")
    ap.add_argument("--dust", action="store_true",
                    help="measure the dust-opacity generator on the model's computepart particle instead")
    ap.add_argument("--image", action="store_true",
                    help="measure an IMAGE observation (ray images and MakeImage's spread) instead")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="transport precision mode (fp32: MCX_FP32 direction arithmetic)")
    ap.add_argument("--precision-study", action="store_true",
                    help="FP32 against FP64 passes with the same seeds (per-cell T and MC SED errors) instead")
    ap.add_argument("--sweep", type=int, default=0,
                    help="time the batched many-model sweep of this many models instead (--nphot per model)")
    ap.add_argument("--alone-only", action="store_true",
                    help="with --sweep: only time --sweep-check models alone (model by model), no batched sweep")
    ap.add_argument("--cohort", type=int, default=0,
                    help="models per batched launch in --sweep (0: all in one)")
    ap.add_argument("--sweep-check", type=int, default=4,
                    help="models of --sweep also run alone (bit-identity and the model-by-model time)")
    ap.add_argument("--atomics", action="store_true",
                    help="measure the GPU's random 64-bit atomic-add rate instead (the tallies' reference)")
    ap.add_argument("--warmstart", action="store_true",
                    help="evaluate the learned warm start on a fresh sweep instead (--sweep models, --passes)")
    ap.add_argument("--passes", type=int, default=4, help="passes per model of --warmstart")
    ap.add_argument("--maxinteract", type=int, default=0,
                    help="--warmstart: the reference's interaction roulette (keyword maxinteract; 0: off)")
    ap.add_argument("--network", default="", help="network file of --warmstart (default: the committed one)")
    ap.add_argument("--epoch-deposits", type=int, default=0,
                    help="staged schedule: deposits per epoch step (0: the library's default; -1: none)")
    ap.add_argument("--stage-ratio", type=int, default=0,
                    help="staged schedule: stage ratio (0: the library's default 2; -1: growing ratio)")
    ap.add_argument("--budget", type=float, default=400.0,
                    help="wall-clock seconds for warmup and timed passes together: warmup stops after a quarter "
                         "of it, timed passes when the next would overrun it (the steps run are reported)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a plain device copy of this many bytes (rocprof counter calibration)")
    a = ap.parse_args()
    if a.calibrate_copy:
        import torch
        n = a.calibrate_copy // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n * 4}))
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: every run here uses a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not run)", file=sys.stderr)
        sys.exit(2)

    from mcmax_amd import configs, transport
    work = tempfile.mkdtemp(prefix="mcx_bench_")
    if a.dust:
        print(json.dumps(dust_bench(a, work)))
        return
    if a.converge:
        print(json.dumps(converge_bench(a, work)))
        return
    if a.image:
        print(json.dumps(image_bench(a, work)))
        return
    if a.atomics:
        print(json.dumps(atomics_bench(a)))
        return
    if a.warmstart:
        print(json.dumps(warmstart_bench(a, work)))
        return
    if a.sweep:
        print(json.dumps(sweep_bench(a, work)))
        return
    if a.precision_study:
        print(json.dumps(precision_study(a, work)))
        return
    prec = 1 if a.precision == "fp32" else 0
    tr = transport.Transport(0)
    m = configs.load_config(a.config, work, tr=tr, randomwalk=not a.plain)
    print(f"[bench] model {a.config}: {m.summary()}", file=sys.stderr, flush=True)
    if m.dust_stats:
        print(f"[bench] dust opacities on the device: {m.dust_stats['ms_total']:.1f} ms", file=sys.stderr, flush=True)
    tr.set_model(m)
    # wall-clock budget: the passes of a heavy-tailed workload vary, so the
    # number of passes run is capped by time and reported
    t_start = time.perf_counter()
    warm, longest = 0, 0.0
    for w in range(a.warmup):
        if w > 0 and time.perf_counter() - t_start > 0.25 * a.budget:
            break
        t1 = time.perf_counter()
        tr.transport(a.nphot, seed=a.seed + 1000 + w, nlanes=a.lanes, stage_ratio=a.stage_ratio, precision=prec,
                     max_ms=1e3 * a.budget, epoch_deposits=a.epoch_deposits)
        tr.temperature()
        longest = max(longest, time.perf_counter() - t1)
        warm += 1
        print(f"[bench] warmup {w} done ({time.perf_counter() - t1:.1f} s)", file=sys.stderr, flush=True)
    kms = []
    t0 = time.perf_counter()
    steps = 0
    for s in range(a.steps):
        if s > 0 and time.perf_counter() - t_start + longest > a.budget:
            break
        t1 = time.perf_counter()
        rc = tr.transport(a.nphot, seed=a.seed + s, nlanes=a.lanes, stage_ratio=a.stage_ratio, precision=prec,
                          max_ms=1e3 * a.budget, epoch_deposits=a.epoch_deposits)
        k, _ = tr.timing()
        kms.append(k)
        tr.temperature()
        longest = max(longest, time.perf_counter() - t1)
        steps += 1
        print(f"[bench] step {s}: kernel {k:.1f} ms rc={rc}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    tal = tr.tallies()
    ms_step = 1e3 * wall / steps
    value = a.nphot * steps / wall
    kern_s = sum(kms) / len(kms) * 1e-3
    nbytes, (Bx, Ba, Bs, Brw) = algorithmic_bytes(tal, m.ng, m.nlam)
    # the whole pass is one k_transport launch (its stages open inside it)
    launches = 1
    achieved = nbytes / kern_s
    nstage = len(stage_schedule(a.nphot, a.stage_ratio))
    B = a.epoch_deposits if a.epoch_deposits else 16000
    schedule = (f"one launch, {nstage} stages (stage0 64, ratio {a.stage_ratio if a.stage_ratio else 2}), "
                f"epoch deposits {B if B > 0 else 'none'}")
    out = {
        "metric": METRIC, "value": value, "unit": "packets/s", "n_gpus": 1, "steps": steps,
        "warmup": warm, "steps_requested": a.steps, "warmup_requested": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == 0 else "f32 directions / f64 geometry", "data": "synthetic",
        "config": {"workload": f"{a.config} disk model, one MC pass + T inversion per step"
                               + ("" if a.plain else ", modified random walk (factRW=10)"),
                   "nphot": a.nphot, "nr": m.nr, "nt": m.nt, "nlam": m.nlam, "ng": m.ng,
                   "randomwalk": not a.plain, "lanes": a.lanes or "default"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(a.config, a.nphot, m.ng, schedule),
                     "kernel": "k_transport", "launches": launches, "schedule": schedule,
                     "kernel_ms_per_launch": kern_s * 1e3 / launches,
                     "bytes_per_launch": nbytes / launches, "kernel_ms": kern_s * 1e3,
                     "bytes_per_event": {"step": Bx, "absorption": Ba, "scattering": Bs, "rw_step": Brw,
                                         "escape": 8},
                     "events": {k: int(tal[k]) for k in ("nsteps", "nabs", "nscat", "nrw", "nescape")}},
        "atomics": pass_atomics(tal, a.nphot, kern_s),
        "energy_error": abs(tal["Eescape"] / m.Lstar - 1.0),
        "dust": m.dust_stats,
        "timeline": {"maxsteps": int(tal["maxsteps"]), "t_done_ms": tal["t_done_ms"],
                     "done_fractions": [0.5, 0.9, 0.99, 0.999, 1.0]},
    }
    tr.close()
    out["time_to_converged"] = converged_small(work)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, int(tal["nsteps"]), a.nphot)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
