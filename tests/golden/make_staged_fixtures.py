"""Write the serial-loop fixtures that the concurrent (staged) device pass is
checked against: the oracle's restatement of the reference's serial packet
loop (MCRadiation.f:320-648) run with K Philox seeds on a small optically
thick model, with per seed the Lucy temperature field (Main.f:498-549), the
distinct-packet counts, the escaping MC spectrum (MCRadiation.f:576-646) and
the event counts.  The model is the 20 x 10 cell, 40-wavelength grid with the
production dust (ten DHS species, m = 1.7 + 0.03i, 0.05-1000 um, MRN
weights) at the production dust mass 1e-3 Msun; the dust opacities come from
the oracle's ComputePart restatement (bit-identical to the device's).

    python tests/golden/make_staged_fixtures.py [small] [large]

(small: 48 seeds of 20000 packets; large: 16 seeds of 200000 packets)
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from mcmax_amd import configs  # noqa: E402

NPHOT = 20000
SEEDS = list(range(1001, 1049))
# the same model at 2e5 packets, 16 seeds
NPHOT_L, SEEDS_L = 200000, list(range(2001, 2017))
DHS10 = {"abun01": 1.0, "computepart01:file": "dhs.dat", "computepart01:amin": 0.05,
         "computepart01:amax": 1000.0, "computepart01:apow": 3.5, "computepart01:fmax": 0.8,
         "computepart01:ngrains": 10, "computepart01:blend": False}
CASES = {"dhs10_thick": dict(mdust=1e-3, **DHS10)}


class OracleDust:
    """compute_parts through the oracle's ComputePart restatement."""

    def compute_parts(self, lam, specs):
        out = []
        for p in specs:
            r = pyoracle.compute_part(lam, p["e1"], p["e2"], p["frac"], p["rho"], p["amin"], p["amax"], p["apow"],
                                      p["fmax"], p["ns"], p["nf"], p.get("rscale", 1.0), p.get("mscale", 1.0))
            out.append(dict(Kabs=r["Kabs"], Ksca=r["Ksca"], Kext=r["Kext"], F=r["F"], rho=r["rho"]))
        return out, {}

    def close(self):
        pass


def model_hash(m):
    h = hashlib.sha256()
    for a in (m.Kabs, m.Ksca, m.F, m.dens, m.w, m.R, m.Theta):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def run_case(name, overrides, workdir, nphot=NPHOT, seeds=SEEDS, tag=""):
    m = configs.load_config("tiny", os.path.join(workdir, name), tr=OracleDust(), **overrides)
    o = pyoracle.Oracle(m)
    keys = ("T", "Ni", "spec", "ispec", "ninteract", "nabs", "nsteps", "Eescape")
    acc = {k: [] for k in keys}
    for s in seeds:
        r = o.mcradiation(nphot, rng="philox", seed=s)
        r["T"], r["dT"], _ = o.temperature(r)
        for k in keys:
            acc[k].append(r[k])
        print(f"{name} seed {s}: {r['ninteract']} interactions", flush=True)
    out = {k: np.asarray(v) for k, v in acc.items()}
    # statistics only need single precision (the spectrum's values exceed
    # float32's range)
    out["T"] = out["T"].astype(np.float32)
    out.update(nphot=nphot, seeds=np.asarray(seeds), model_hash=model_hash(m))
    np.savez_compressed(os.path.join(HERE, f"serial_{name}{tag}.npz"), **out)


if __name__ == "__main__":
    pyoracle.build()
    work = tempfile.mkdtemp(prefix="mcx_golden_")
    which = sys.argv[1:] or ["small", "large"]
    for name, ov in CASES.items():
        if "small" in which:
            run_case(name, ov, work)
        if "large" in which:
            run_case(name, ov, work, NPHOT_L, SEEDS_L, "_2e5")
