"""Write the converged temperature field of the small plain-method model (the
CPU-runnable workload: 4000 K star, 1e-4 Msun, 100 r x 60 theta cells, 100
wavelengths, one analytic grain, no vertical-structure iteration, the
modified random walk with factRW = 10 as the reference's example disks run
it) from the oracle's FP64 restatement of the reference's serial packet loop
(MCRadiation.f:320-648) and Lucy inversion (Main.f:498-549).  (Without the
random walk a few packets of this model take ~1e7 cell steps in the
optically thick inner midplane.)

The field comes from K independent sub-runs of N packets with their own
Philox seeds, the reference's nruns scheme (MCRadiation.f:192, 269-668): the
path-length tallies EJv of the sub-runs are averaged (each is normalised to
the star's luminosity) and inverted once.  At K N = 2e6 packets the field's
own noise is a few tenths of a percent in the median cell, below the 1%
criterion that the device's time to a converged field is measured against.

    python tests/golden/make_converged_small.py [K] [N]    (default 8 x 250000)
"""
import multiprocessing as mp
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(HERE, "converged_small.npz")
SEED0 = 70001


def _model(workdir):
    from mcmax_amd import configs
    return configs.load_config("small", workdir, randomwalk=True)


def _subrun(args):
    workdir, seed, n = args
    import pyoracle
    m = _model(os.path.join(workdir, f"s{seed}"))
    r = pyoracle.Oracle(m).mcradiation(n, rng="philox", seed=seed)
    return r


def main():
    k = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    n = int(sys.argv[2]) if len(sys.argv) > 2 else 250000
    with tempfile.TemporaryDirectory() as wd:
        m = _model(wd)
        with mp.Pool(min(k, os.cpu_count() or 1)) as pool:
            runs = pool.map(_subrun, [(wd, SEED0 + j, n) for j in range(k)])
        import pyoracle
        o = pyoracle.Oracle(m)
        merged = dict(runs[0])
        for key in ("EJv", "EJv2", "Eabs"):
            merged[key] = np.mean([r[key] for r in runs], axis=0)
        # the start of each cell's T search (the result does not depend on it)
        merged["Tbw"] = np.max([r["Tbw"] for r in runs], axis=0)
        merged["Ni"] = np.sum([r["Ni"] for r in runs], axis=0)
        T, dT, _ = o.temperature(merged)
        # the sub-runs' own fields, for the spread at N packets
        Tsub = np.array([o.temperature(r)[0] for r in runs])
    np.savez_compressed(OUT, T=T, dT=dT, Tsub=Tsub, nphot=k * n, subruns=k, seed0=SEED0, randomwalk=True,
                        nsteps=np.array([r["nsteps"] for r in runs]))
    rel = np.abs(Tsub / T[None] - 1.0)
    print(f"wrote {OUT}: {k} x {n} packets; a sub-run's median |dT|/T against the field: "
          f"{np.median(rel, axis=(1, 2))}")


if __name__ == "__main__":
    main()
