"""Transport in FP32 (precision mode MCX_FP32: directions, rotations,
Stokes vector and optical-depth draws in single precision; positions,
crossing quadratics, optical-depth sums, tallies and temperatures FP64)
against the all-FP64 device pass with the same Philox seeds.  Rounding makes
the histories part, so the comparison is statistical: the temperature field
and the MC spectrum of 10 passes of each precision, by the criteria of the
serial-loop comparison (tests/staged_stats.py)."""
import numpy as np
import pytest

from mcmax_amd import transport

pytestmark = pytest.mark.gpu


def _runs(gpu, m, n, seeds, precision):
    gpu.set_model(m)
    out = []
    for s in seeds:
        # a pass that overruns two minutes stops and raises (MCX_ETIME)
        g = transport.mc_radiation(gpu, n, seed=s, precision=precision, max_ms=120000.0)
        assert g["noverflow"] == 0 and g["ngeomfail"] == 0 and g["nvisit_overflow"] == 0
        assert g["nescape"] == n and abs(g["Eescape"] / m.Lstar - 1) < 1e-9
        out.append(g)
        print(f"  precision {precision} seed {s}: {g['kernel_ms']:.1f} ms", flush=True)
    return out


def _as_fixture(runs):
    return {k: np.asarray([r[k] for r in runs]) for k in ("T", "Ni", "spec", "ispec")}


@pytest.mark.parametrize("case", ["dhs10_thick", "analytic_rw"])
def test_fp32_against_fp64(gpu, workdir, case):
    import staged_stats
    from mcmax_amd import configs
    if case == "dhs10_thick":
        import os, sys
        sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
        from make_staged_fixtures import CASES
        m = configs.load_config("tiny", workdir + "/prec_dhs10", tr=gpu, **CASES["dhs10_thick"])
        n = 20000
    else:
        m = configs.load_config("small", workdir + "/prec_rw", randomwalk=True, mdust=1e-3)
        n = 50000
    seeds = range(300, 310)
    f64 = _runs(gpu, m, n, seeds, 0)
    f32 = _runs(gpu, m, n, seeds, 1)
    # the same seeds start the same histories; rounding parts them
    assert not np.array_equal(f64[0]["EJv"], f32[0]["EJv"])
    res = staged_stats.compare(f32, _as_fixture(f64))
    print({k: v for k, v in res.items() if k != "spec_ratio"})
    assert res["median_rel"] <= 0.01 and abs(res["bias"]) <= 0.003, res
    assert res["within3"] >= 0.975, res
    assert res["spec_zmax"] <= 4.0, res
    assert res["spec_bins"] == 0 or res["spec_maxdev"] <= 0.02, res
