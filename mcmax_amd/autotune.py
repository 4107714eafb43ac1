"""The batched pass's launch shape, picked by measurement.

For a set of models the blocks each model's packets are served by (the tile
schedule of mcx_batch_transport) are chosen by timing a short probe pass for
each candidate; the choice is cached per device, model count, grid sizes and
packet count in a JSON file (~/.cache/mcmax_amd/autotune.json, or
$MCMAX_AMD_CACHE).  Results do not depend on the choice (a model's pass is
bit-identical whatever the launch shape); only the time does.
"""
import hashlib
import json
import os
import time

from . import transport


def _cache_path():
    d = os.environ.get("MCMAX_AMD_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "mcmax_amd")
    os.makedirs(d, exist_ok=True)
    return os.path.join(d, "autotune.json")


def _key(trs, nphot, maxinteract=0):
    h = hashlib.sha256()
    for t in trs:
        m = t.model
        h.update(f"{m.nr}x{m.nt}x{m.nlam}x{m.ng};".encode())
    return f"{len(trs)}:{nphot}:{maxinteract}:{h.hexdigest()[:16]}"


RESIDENT = 2 * 256  # resident workgroups of a batched launch: two per CU on the 256 CUs


def tune(trs, nphot, candidates=None, probe_nphot=None, randomwalk=None, log=None, use_cache=True, maxinteract=0):
    """blocks per model for a batched pass of nphot packets over the models
    of trs: a probe pass of probe_nphot packets (default nphot/16, at least
    4096) per candidate; the fewest workgroups within 3% of the fastest
    kept."""
    nm = len(trs)
    resident = RESIDENT
    if candidates is None:
        # launches that fill the device: the resident workgroups shared out,
        # or twice as many (the later ones start as the first retire); fewer
        # would leave CUs idle once the probe's short tails are gone
        base = max(1, resident // nm)
        candidates = [base, 2 * base]
    key = _key(trs, nphot, maxinteract)
    path = _cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if use_cache and key in cache:
        return cache[key]["blocks_per_model"], cache[key]
    probe = int(probe_nphot or max(4096, nphot // 16))
    times = {}
    for bpm in candidates:
        t0 = time.perf_counter()
        transport.batch_transport(trs, probe, seed=987654, blocks_per_model=bpm, randomwalk=randomwalk,
                                  maxinteract=maxinteract)
        times[bpm] = time.perf_counter() - t0
        if log:
            log(f"autotune: {bpm} blocks per model: {times[bpm]:.2f} s")
    # a probe is short and its time noisy: the fewest workgroups within 3%
    # of the fastest
    fastest = min(times.values())
    best = min(k for k, v in times.items() if v <= 1.03 * fastest)
    rec = {"blocks_per_model": best, "probe_nphot": probe, "seconds": {str(k): v for k, v in times.items()}}
    cache[key] = rec
    if use_cache:
        with open(path, "w") as f:
            json.dump(cache, f, indent=1)
    return best, rec
