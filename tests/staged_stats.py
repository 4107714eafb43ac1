"""Statistical comparison of a set of passes against the serial loop's
fixtures (tests/golden/serial_*.npz): the criteria of a matched-packet-count
comparison with the reference's serial loop (median |dT|/T <= 1% of the mean
fields, per-cell differences within 3 sigma of the seed-to-seed MC noise, the
MC spectrum within 2% where the counts allow, energy conserved)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_fixture(name):
    return dict(np.load(os.path.join(GOLDEN, f"serial_{name}.npz")))


def compare(runs, fx, min_ni=20):
    """runs: list of pass dicts (T [nr, nt], Ni, spec [31, nlam]); fx: the
    serial fixture.  Returns the figures the criteria use."""
    Tg = np.array([r["T"][1:] for r in runs])
    To = fx["T"][:, 1:]
    Nig = np.array([r["Ni"][1:] for r in runs]).min(axis=0)
    sel = (Nig > min_ni) & (fx["Ni"][:, 1:].min(axis=0) > min_ni)
    mg, mo = Tg.mean(axis=0)[sel], To.mean(axis=0)[sel]
    vg, vo = Tg.var(axis=0, ddof=1)[sel], To.var(axis=0, ddof=1)[sel]
    # every (gpu seed, oracle seed) pair of single runs, cell by cell: the
    # difference against the seed-to-seed noise of one run of each
    k = min(len(Tg), len(To))
    z = (Tg[:k][:, sel] - To[:k][:, sel]) / np.sqrt(vg + vo + 1e-300)
    # the MC spectrum summed over the escape directions, per wavelength bin
    Sg = np.array([r["spec"].sum(axis=0) for r in runs])
    So = fx["spec"].sum(axis=1)
    sg, so = Sg.mean(axis=0), So.mean(axis=0)
    dg = Sg.std(axis=0, ddof=1) / np.sqrt(len(Sg))
    do = So.std(axis=0, ddof=1) / np.sqrt(len(So))
    dd = np.sqrt(dg ** 2 + do ** 2)
    ng = np.array([r["ispec"].sum(axis=0) for r in runs]).mean(axis=0)
    no = fx["ispec"].sum(axis=1).mean(axis=0)
    populated = (ng >= 200) & (no >= 200)
    zs = np.abs(sg - so) / np.maximum(dd, 1e-300)
    well = populated & (dd <= 0.0067 * np.abs(so))
    return dict(ncells=int(sel.sum()), median_rel=float(np.median(np.abs(mg / mo - 1))),
                bias=float(np.median(mg / mo - 1)), within3=float(np.mean(np.abs(z) < 3.0)),
                spec_zmax=float(zs[populated].max()) if populated.any() else 0.0,
                spec_populated=int(populated.sum()), spec_bins=int(well.sum()),
                spec_maxdev=float(np.max(np.abs(sg[well] / so[well] - 1))) if well.any() else 0.0,
                spec_ratio=sg / np.where(so != 0, so, 1.0))
