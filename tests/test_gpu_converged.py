"""The device's pass of the small plain-method model (modified random walk)
at its configured 1e5 packets against the oracle's converged field
(tests/golden/converged_small.npz): within the metric's 1% median |dT|/T."""
import os

import numpy as np
import pytest

from mcmax_amd import transport

pytestmark = pytest.mark.gpu

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "converged_small.npz")


def test_small_pass_converged_against_oracle(gpu, workdir):
    from mcmax_amd import configs
    ref = np.load(PATH)
    m = configs.load_config("small", workdir + "/conv_small", tr=gpu, randomwalk=True)
    gpu.set_model(m)
    meds = []
    for seed in (11, 12):
        gpu.transport(100000, seed=seed)
        T, _, _ = gpu.temperature()
        meds.append(float(np.median(np.abs(T / ref["T"] - 1.0))))
    print("median |dT|/T against the oracle's field:", meds)
    assert max(meds) <= 0.01, meds
    # and not suspiciously close: 1e5 packets carry their own noise
    assert min(meds) > 5e-4, meds
