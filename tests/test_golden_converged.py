"""The converged field of the small plain-method model with the modified
random walk (tests/golden/converged_small.npz, written by
make_converged_small.py from the oracle's serial loop in 8 sub-runs of
250000 packets): its shape, its sub-runs' spread, and a fresh oracle run of
1e4 packets against it."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "converged_small.npz")


@pytest.fixture(scope="module")
def ref():
    return np.load(PATH)


def test_fixture_shape_and_spread(ref):
    T, Tsub = ref["T"], ref["Tsub"]
    assert T.shape == (101, 60) and Tsub.shape == (8, 101, 60)
    assert int(ref["nphot"]) == 8 * 250000 and bool(ref["randomwalk"])
    assert np.all(np.isfinite(T)) and np.all(T >= 1.0)
    # each sub-run (1/8 of the packets) within ~0.25% of the field in the
    # median cell: the field's own noise is ~sqrt(8) smaller
    rel = np.median(np.abs(Tsub / T[None] - 1.0), axis=(1, 2))
    assert np.all(rel < 0.004), rel
    # the sub-runs are independent: their mean is the field to within their
    # spread / sqrt(8) (the inversion is not linear, so only close)
    assert np.median(np.abs(Tsub.mean(axis=0) / T - 1.0)) < 0.002


def test_fresh_oracle_run_against_fixture(ref, tmp_path):
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
    import pyoracle
    from mcmax_amd import configs
    m = configs.load_config("small", str(tmp_path), randomwalk=True)
    o = pyoracle.Oracle(m)
    r = o.mcradiation(10000, rng="philox", seed=5)
    T, _, _ = o.temperature(r)
    med = np.median(np.abs(T / ref["T"] - 1.0))
    # 1e4 packets: ~5x the 2.5e5-packet sub-runs' 0.22%
    assert 0.003 < med < 0.03, med
