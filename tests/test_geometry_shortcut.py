"""The device's ray-sphere / ray-cone crossings skip the square root and the
divisions when the signs of the quadratic's coefficients rule out a positive
root.  This compiles the device header for the host and checks, bit for bit,
that every crossing equals the plain quadratic of MCRadiation.f:1274-1355 on
random inputs mixed with signed zeros, denormals and extreme values."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_quadratic_shortcut_is_exact(tmp_path):
    exe = str(tmp_path / "quadratics_check")
    src = os.path.join(ROOT, "tests", "native", "quadratics_check.hip")
    subprocess.check_call([HIPCC, "-O2", "--cuda-host-only", "-ffp-contract=off",
                           "-I", os.path.join(ROOT, "mcmax_amd", "csrc"), "-o", exe, src],
                          stderr=subprocess.DEVNULL)
    r = subprocess.run([exe, "4000000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
