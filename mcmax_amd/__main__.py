"""Command line of the stage, as the reference's (Main.f:60-160):

    python -m mcmax_amd <input file> <Nphot> [observation files ...] [-o outdir] [-s key=value ...]

Reads the keyword file (with -s overrides), builds the model, runs the
Monte Carlo passes with the structure iteration when `iter=.true.`
(Main.f:202-791), writes denstemp.dat (denstemp.fits.gz with the keyword
outputfits) and the MCSpec spectra, then the ray-traced spectrum of every
SPECTRUM observation.  Nphot = 0 reads the temperatures of that structure
file in outdir instead of running passes (Main.f:380).
Everything numeric runs on the GPU through libmcx.so.
"""
import argparse
import os
import sys

import numpy as np

from . import iterate, keywords, model, observe, output
from .transport import Transport


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m mcmax_amd")
    ap.add_argument("input")
    ap.add_argument("nphot", type=int)
    ap.add_argument("obs", nargs="*")
    ap.add_argument("-o", dest="outdir", default=".")
    ap.add_argument("-s", dest="overrides", action="append", default=[])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    os.makedirs(a.outdir, exist_ok=True)
    kw = keywords.parse_file(a.input, overrides=a.overrides)
    m = model.build_model(kw, basedir=os.path.dirname(os.path.abspath(a.input)))
    print(f"model: {m.summary()}", flush=True)
    # Init.f:1556-1560: the keyword outputfits picks the FITS structure file
    fits_out = bool(kw.get("outputfits", False))
    structure_name = "denstemp.fits.gz" if fits_out else "denstemp.dat"
    tr = Transport(a.device)
    try:
        if a.nphot > 0:
            it = iterate.Iteration(tr, m, a.nphot, seed=a.seed, struct_iter=bool(kw.get("iter", False)),
                                   log=lambda s: print(s, flush=True))
            last = it.run()
            T, state = last["T"], it.state
            output.write_mcspec(a.outdir, m, it.last_tallies)
        else:
            d = output.read_denstemp(os.path.join(a.outdir, structure_name))
            if d["T"].shape != (m.nr - 1, m.nt):
                raise SystemExit(f"{structure_name} does not match the model's grid")
            T = np.zeros((m.nr, m.nt))
            T[1:] = d["T"]
            T[0] = 1.0
            state = iterate.density_state(m)
            tr.set_model(m)
            tr.set_temperature(T)
        if fits_out:
            output.write_denstemp_fits(os.path.join(a.outdir, structure_name), m, T, state)
        else:
            output.write_denstemp(os.path.join(a.outdir, structure_name), m, T, state)
        for path in a.obs:
            for tel in observe.parse_observation_file(path, m):
                for p in observe.observe(tr, m, tel, a.outdir, T=T, seed=a.seed):
                    print(f"wrote {p}", flush=True)
    finally:
        tr.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
