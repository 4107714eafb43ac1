"""The production workload (200 r x 100 theta cells, 200 wavelengths,
1e-3 Msun, ten DHS species from the device's ComputePart, modified random
walk with factRW = 10) through the driver at reduced packet counts: the
first pass's temperature field against the oracle's serial loop on the same
model and packet count, within the spread of two oracle passes with other
seeds; the keyword file's structure iteration (iter, maxiter = 5); and the
ray-traced SED at 45 degrees on the final field against the oracle's
formal solution."""
import numpy as np
import pytest

from mcmax_amd import configs, iterate, transport

pytestmark = pytest.mark.gpu

NPHOT = 100000


@pytest.fixture(scope="module")
def prod(gpu, workdir):
    return configs.load_config("production", workdir + "/prod_cfg", tr=gpu, randomwalk=True)


def test_production_first_pass_against_oracle(gpu, oracle_lib, prod):
    m = prod
    assert (m.nr, m.nt, m.nlam, m.ng) == (201, 100, 200, 10)
    gpu.set_model(m)
    gpu.transport(NPHOT, seed=5)
    tal = gpu.tallies()
    assert tal["nescape"] == NPHOT and abs(tal["Eescape"] / m.Lstar - 1.0) < 1e-9
    Tg, _, _ = gpu.temperature()
    o = oracle_lib.Oracle(m)
    T1 = o.temperature(o.mcradiation(NPHOT, rng="philox", seed=6))[0]
    T2 = o.temperature(o.mcradiation(NPHOT, rng="philox", seed=7))[0]
    disk = np.s_[1:, :]
    d_go = np.median(np.abs(Tg[disk] / T1[disk] - 1.0))
    d_oo = np.median(np.abs(T2[disk] / T1[disk] - 1.0))
    print(f"median |dT|/T: device vs oracle {d_go:.4f}, oracle vs oracle {d_oo:.4f}")
    assert d_go <= 1.5 * d_oo + 0.002, (d_go, d_oo)


def test_production_structure_iteration_and_sed(gpu, oracle_lib, prod, workdir):
    m = configs.load_config("production", workdir + "/prod_cfg_it", tr=gpu, randomwalk=True)
    col0 = (iterate.density_state(m)["dens0"] * m.V)[1:].sum(axis=1)
    it = iterate.Iteration(gpu, m, NPHOT, seed=9)
    assert it.maxiter == 5
    it.run()
    h = it.history
    assert 3 <= len(h) <= 7
    assert all(abs(r["Eescape"] / m.Lstar - 1.0) < 1e-9 for r in h)
    assert all(np.isfinite(r["errT"]) for r in h[:-1])
    # DiskStructure keeps every column's dust mass
    st = it.state
    assert np.allclose((st["dens0"] * m.V)[1:].sum(axis=1), col0, rtol=1e-8)
    # the SED at 45 degrees of the final field: thermal emission and
    # extinction along every image ray, deterministic on both sides
    T = h[-1]["T"]
    lams = np.array([m.lam[60], m.lam[100], m.lam[140], m.lam[180]])
    gpu.set_model(m, tables=False)
    g = gpu.sed(transport.make_obs(angle_deg=45.0, nphot=0, nphotstar=0), lams, T=T)
    ob = oracle_lib.Oracle(m).sed(T, oracle_lib.make_obs(angle_deg=45.0, nphot=0, nphotstar=0), lams)
    assert np.all(g[:, 0] > 0)
    assert np.allclose(g[:, 0], ob[:, 0], rtol=1e-9), np.max(np.abs(g[:, 0] / ob[:, 0] - 1))
