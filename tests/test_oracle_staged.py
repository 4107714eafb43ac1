"""The oracle's staged Bjorkman-Wood schedule (the device's concurrent pass,
orc_mcradiation in oracle/mcx_oracle.h) against the reference's serial loop:
the oracle's restatement of MCRadiation.f:320-648 run with 48 seeds
(tests/golden/serial_dhs10_thick.npz, make_staged_fixtures.py), on the
production dust at the production dust mass."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))


def test_stage_schedule():
    import ctypes as C
    import pyoracle
    lib = pyoracle.load()
    lib.orc_stage_next.restype = C.c_int64
    lib.orc_stage_next.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_int]

    def sched(N, s0, r):
        P, out = min(s0, N), [min(s0, N)]
        while P < N:
            P = lib.orc_stage_next(P, N, s0, r)
            out.append(P)
        return out
    assert sched(20000, 64, 2) == [64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 20000]
    assert sched(1000, 100, 3) == [100, 300, 900, 1000]
    assert sched(20000, 64, -1) == [64, 128, 310, 993, 4905, 20000]
    g = sched(10 ** 7, 64, -1)
    assert all(b <= 16 * a for a, b in zip(g, g[1:])) and g[-1] == 10 ** 7


@pytest.fixture(scope="module")
def dhs10(workdir):
    import pyoracle
    import staged_stats
    from make_staged_fixtures import CASES, OracleDust, model_hash
    from mcmax_amd import configs
    fx = staged_stats.load_fixture("dhs10_thick")
    m = configs.load_config("tiny", workdir + "/orc_dhs10", tr=OracleDust(), **CASES["dhs10_thick"])
    assert model_hash(m) == str(fx["model_hash"])
    return m, fx, pyoracle.Oracle(m)


def test_staged_matches_serial_statistics(dhs10):
    """Eight staged passes against the 48 serial runs: the temperature
    criteria, and the MC spectrum within 4 sigma of the noise of the mean in
    every bin with 200 escapes per run or more."""
    import staged_stats
    m, fx, o = dhs10
    runs = []
    for s in range(1, 9):
        r = o.mcradiation(int(fx["nphot"]), seed=9000 + s, staged=True)
        r["T"], _, _ = o.temperature(r)
        assert abs(r["Eescape"] / m.Lstar - 1) < 1e-9
        runs.append(r)
    res = staged_stats.compare(runs, fx)
    print({k: v for k, v in res.items() if k != "spec_ratio"})
    assert res["median_rel"] <= 0.01 and abs(res["bias"]) <= 0.003
    assert res["within3"] >= 0.975
    assert res["spec_zmax"] <= 4.0


def test_staged_order_free(dhs10):
    """A staged history depends only on its own draws and its stage's
    snapshot, so the order in which a stage's packets run does not matter:
    the stages run in reverse order give the identical pass (the device runs
    them in any order)."""
    m, fx, o = dhs10
    a = o.mcradiation(3000, seed=5, staged=1)
    b = o.mcradiation(3000, seed=5, staged=2)
    for k in ("nabs", "nscat", "nsteps", "nescape", "nemit"):
        assert a[k] == b[k], k
    for k in ("Ni", "Eabs", "Tbw", "ispec"):
        assert np.array_equal(a[k], b[k]), k
    # EJv is a double sum in the oracle: the order changes only its rounding
    assert np.allclose(a["EJv"], b["EJv"], rtol=1e-12, atol=0)
    # while the serial loop's order is its result
    c = o.mcradiation(3000, seed=5)
    assert not np.array_equal(a["Eabs"], c["Eabs"])
