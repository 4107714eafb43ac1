"""The sweep's shares of a batched launch: workgroups per model in
proportion to the lane time of its previous pass, a minimum each, summing to the
launch's total (host logic, no GPU)."""
import numpy as np
import pytest

from mcmax_amd.sweep import Sweep


def test_shares_sum_minimum_and_order():
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = int(rng.integers(1, 300))
        total = int(rng.integers(2 * n, 2 * n + 2000))
        work = np.exp(rng.uniform(0, 20, n))
        s = Sweep.shares(work, total)
        assert s.sum() == total and s.min() >= 2 and s.dtype == np.int32
        # more work never gets fewer workgroups
        o = np.argsort(work)
        assert np.all(np.diff(s[o]) >= 0)


def test_shares_proportional_and_even():
    assert Sweep.shares([1, 1, 1, 1], 512).tolist() == [128] * 4
    s = Sweep.shares([3e8, 1e8], 402)
    assert s.tolist() == [2 + 299, 2 + 99]
    assert Sweep.shares([5, 5, 5], 7).sum() == 7


def test_shares_too_few():
    with pytest.raises(ValueError):
        Sweep.shares([1, 2, 3], 5)



def test_cohorts_cover_every_model_once():
    for nm in (1, 5, 128, 256, 257):
        for size in (0, 1, 7, 128, 300):
            groups = Sweep.cohorts(nm, size)
            flat = [k for g in groups for k in g]
            assert flat == list(range(nm))
            cap = size if size and size < nm else nm
            assert all(1 <= len(g) <= cap for g in groups)
    assert Sweep.cohorts(256, 128) == [list(range(128)), list(range(128, 256))]


def test_cohort_keys_match_model_keys():
    # model k of launch g draws with the launch's base key + stride * q,
    # q its place in the launch: the key run_one gives it alone
    seed, stride, it = 17, 3, 2
    for g in Sweep.cohorts(10, 4):
        base = Sweep.cohort_key(seed, it, stride, g[0])
        for q, k in enumerate(g):
            assert base + stride * q == seed + 1000 * it + stride * k
