"""The boundary from Fortran: mcmax_amd/fortran/mcx_driver (amdflang, through
the ISO_C_BINDING module mcx_binding.f90) makes MCMax's call sequence
(mcx_create, mcx_set_tables, mcx_set_model, mcx_transport, mcx_get_tallies,
mcx_temperature) on the tiny model; its tallies and temperatures equal those
of the Python (ctypes) path bit for bit, for a one-lane pass and a
concurrent staged pass."""
import os
import subprocess

import numpy as np
import pytest

from mcmax_amd import _lib, build, transport

pytestmark = pytest.mark.gpu

COUNTERS = ("nemit", "ninteract", "nabs", "nscat", "nescape", "nstarhit", "nsteps", "nraar", "nmaxinteract",
            "ngeomfail")


def _write_input(path, m, nphot, seed, nlanes, randomwalk):
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
    with open(path, "wb") as f:
        np.array([m.nlam, m.ng, m.nr, m.nt, m.scat_how, int(randomwalk), nlanes, 0], dtype=np.int32).tofile(f)
        np.array([nphot, seed, 0, 0], dtype=np.int64).tofile(f)
        np.array([m.Lstar, m.Rstar_cm, m.Tstar, float(getattr(m, "factRW", 10.0))], dtype=np.float64).tofile(f)
        for a in (m.lam, m.nu, m.BB, m.Kabs, m.Ksca, m.F, m.IF11, m.IF12, m.Kp, m.Fstar, m.cos2phi, m.sin2phi,
                  m.Kext, m.gsym, m.R, m.Theta, m.dens, m.V, m.w, m.thet):
            f64(a).tofile(f)


def _read_output(path, ncell):
    with open(path, "rb") as f:
        rc = int(np.fromfile(f, np.int32, 1)[0])
        cnt = np.fromfile(f, np.int64, 10)
        out = dict(rc=rc, **{k: int(v) for k, v in zip(COUNTERS, cnt)})
        for k in ("EJv", "EJv2", "Eabs"):
            out[k] = np.fromfile(f, np.float64, ncell)
        out["Ni"] = np.fromfile(f, np.int32, ncell)
        out["T"] = np.fromfile(f, np.float64, ncell)
        out["dT"] = np.fromfile(f, np.float64, ncell)
    return out


@pytest.mark.parametrize("nlanes,nphot", [(1, 300), (0, 20000)])
def test_fortran_caller_matches_python(gpu, tiny_model, tmp_path, nlanes, nphot):
    if not os.path.exists(build.DRIVER):
        pytest.skip("the Fortran caller is not built (no amdflang)")
    m = tiny_model
    inp, outp = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(inp, m, nphot, 29, nlanes, False)
    r = subprocess.run([build.DRIVER, inp, outp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    fo = _read_output(outp, m.nr * m.nt)
    gpu.set_model(m)
    py = transport.mc_radiation(gpu, nphot, seed=29, nlanes=nlanes)
    assert fo["rc"] == py["rc"] == 0
    for k in COUNTERS:
        assert fo[k] == py[k], k
    for k in ("EJv", "EJv2", "Eabs", "Ni", "T", "dT"):
        assert np.array_equal(fo[k], np.asarray(py[k]).ravel()), k
    assert fo["nescape"] == nphot
