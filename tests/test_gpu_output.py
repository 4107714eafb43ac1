"""The command line end to end on the device: a pass, denstemp.dat, the
MCSpec spectra and an observation's spectrum; then the Nphot = 0 restart
from denstemp.dat, which must reproduce the observation's spectrum."""
import os
import shutil

import numpy as np
import pytest

from mcmax_amd import __main__ as cli
from mcmax_amd import configs, output

pytestmark = pytest.mark.gpu


def test_cli_pass_outputs_and_restart(tmp_path):
    inp = configs.write_config("tiny", str(tmp_path / "in"))
    obs = tmp_path / "obs.in"
    obs.write_text("type=SPECTRUM\ninc=60\nnphot=200\nnphotstar=200\nnlam=12\n")
    out1 = tmp_path / "run1"
    print("cli: tiny, 20000 packets", flush=True)
    assert cli.main([inp, "20000", str(obs), "-o", str(out1)]) == 0
    names = set(os.listdir(out1))
    assert {"denstemp.dat", "MCSpec.dat", "spectrum60.0.dat"} <= names
    assert sum(n.startswith("MCSpec") for n in names) == 31
    d = output.read_denstemp(str(out1 / "denstemp.dat"))
    assert np.all(np.isfinite(d["T"])) and np.all(d["T"] > 0)
    s1 = np.loadtxt(out1 / "spectrum60.0.dat")
    assert s1.shape[0] == 12 and np.all(np.isfinite(s1)) and np.all(s1[:, 1] > 0)
    # flux >= scattered flux; the stellar column is the star's Planck curve
    assert np.all(s1[:, 1] >= s1[:, 2])
    # restart: Nphot = 0 reads the temperatures back and traces the same rays
    out2 = tmp_path / "run2"
    os.makedirs(out2)
    shutil.copy(out1 / "denstemp.dat", out2 / "denstemp.dat")
    print("cli: restart from denstemp.dat", flush=True)
    assert cli.main([inp, "0", str(obs), "-o", str(out2)]) == 0
    s2 = np.loadtxt(out2 / "spectrum60.0.dat")
    assert np.allclose(s1, s2, rtol=1e-12, atol=0)


def test_cli_fits_structure_and_restart(tmp_path):
    """outputfits: the structure file is denstemp.fits.gz, and Nphot = 0
    restarts from it."""
    inp = configs.write_config("tiny", str(tmp_path / "in"))
    out = tmp_path / "run"
    print("cli: tiny, outputfits", flush=True)
    assert cli.main([inp, "20000", "-o", str(out), "-s", "outputfits=.true."]) == 0
    assert os.path.exists(out / "denstemp.fits.gz") and not os.path.exists(out / "denstemp.dat")
    d = output.read_denstemp(str(out / "denstemp.fits.gz"))
    assert np.all(d["T"] > 0)
    assert cli.main([inp, "0", "-o", str(out), "-s", "outputfits=.true."]) == 0
    e = output.read_denstemp(str(out / "denstemp.fits.gz"))
    assert np.array_equal(d["T"], e["T"])
