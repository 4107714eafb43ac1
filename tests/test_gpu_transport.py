"""GPU parity of the transport pass against the oracle.

Serial tests run the device kernel with a single lane, so packets are traced
one after another in packet order; the oracle in Philox mode draws from the
same per-packet streams, so both passes follow the same histories and agree to
floating-point rounding (the device uses fixed-point tallies and binary-search
CDFs, the oracle double sums and the reference's linear scans).
Parallel tests compare statistics.
"""
import numpy as np
import pytest

from mcmax_amd import transport

pytestmark = pytest.mark.gpu

COUNTERS = ("nemit", "ninteract", "nabs", "nscat", "nescape", "nstarhit", "nsteps", "nraar", "nrw")


def fixed_point_units(m, nphot):
    """Per-cell resolution of the device's EJv and EJv2 integer sums: one
    event of the cell's bound B = E0 AU (2 sqrt(R2^2-R1^2)) max(KabsTot)
    maps to 2^31 (EJv) and B^2 to 2^40 (EJv2)."""
    AU = float(np.float32(1.49598e13))
    E0 = m.Lstar / float(np.float32(nphot))
    KabsTot = np.einsum("ijs,sl->ijl", m.w, m.Kabs)
    chord = 2.0 * np.sqrt(np.maximum(m.R[1:] ** 2 - m.R[:-1] ** 2, 0.0))
    B = E0 * AU * chord[:, None] * KabsTot.max(axis=2) * (1 + 1 / 64)
    return B * 2.0 ** -31, B * B * 2.0 ** -40


def _serial_compare(tr, m, oracle_lib, n, seed, rtol=1e-6, coop_lanes=-1):
    tr.set_model(m)
    g = transport.mc_radiation(tr, n, seed=seed, nlanes=1, coop_lanes=coop_lanes)
    orc = oracle_lib.Oracle(m)
    o = orc.mcradiation(n, rng="philox", seed=seed)
    assert g["noverflow"] == 0 and g["ngeomfail"] == 0
    if m.randomwalk:
        assert np.array_equal(g["thick"], orc.random_walk_tables()[2])
    for k in COUNTERS:
        if k == "nsteps":
            # the z=0 plane is the cone cos^2(theta)=0, whose discriminant
            # (MCRadiation.f:1318) is zero in exact arithmetic: whether a step
            # registers the midplane crossing is decided by rounding, in the
            # reference as well.  A crossing there keeps the cell and draws no
            # random number, so it only splits a segment.
            assert abs(g[k] - o[k]) <= 0.005 * o[k], (k, g[k], o[k])
        else:
            assert g[k] == o[k], (k, g[k], o[k])
    assert np.array_equal(g["Ni"], o["Ni"])
    u1, u2 = fixed_point_units(m, n)
    # EJv: each device addition rounds to half a unit (split segments add at
    # most one more addition each).  The reemission weight wl1 of the chosen
    # bin comes from cumulative rows on the device and from the running sum
    # of the difference spectrum in the oracle: the same bin, but where the
    # bin's increment is tiny against the cumulative the weight agrees only
    # to ~1e-9 relative, and so do the segment contributions after it.
    bound = 0.5 * u1 * (2 * o["nev1"] + 2) + 1e-8 * np.abs(o["EJv"]) + 1e-300
    assert np.all(np.abs(g["EJv"] - o["EJv"]) <= bound), np.max(np.abs(g["EJv"] - o["EJv"]) / bound)
    # EJv2 sums squares of segments: exact away from the midplane, only of the
    # same order in the midplane cells where a split changes the squares
    off = np.ones_like(o["EJv2"], dtype=bool)
    off[:, -1] = False
    bound2 = 0.5 * u2 * o["nev2"] + 2e-8 * np.abs(o["EJv2"]) + 1e-300
    d2 = np.abs(g["EJv2"] - o["EJv2"])
    bad = np.argwhere(off & (d2 > bound2))
    for i, j in bad[:8]:
        print("EJv2 cell", i, j, "gpu", g["EJv2"][i, j], "oracle", o["EJv2"][i, j], "nev2", o["nev2"][i, j],
              "EJv", g["EJv"][i, j], o["EJv"][i, j], "u2", u2[i, j])
    assert np.all(d2[off] <= bound2[off]), np.max(d2[off] / bound2[off])
    # midplane cells: a visit's running sum is squared at every registered
    # crossing of z = 0, and which crossings register is decided by
    # rounding; a visit split into k parts changes its square by up to k
    mid = o["EJv2"][:, -1] > 0
    ratio = g["EJv2"][:, -1][mid] / o["EJv2"][:, -1][mid]
    assert np.all((ratio > 0.25) & (ratio < 4.0)), ratio
    row = g["EJv2"][:, -1].sum() / o["EJv2"][:, -1].sum()
    assert 0.5 < row < 2.0, row
    # Eabs: absorptions add the packet energy exactly; random-walk deposits
    # round to half a unit of E0 2^-24 each (at most nev1 of them)
    E0 = m.Lstar / float(np.float32(n))
    ebound = 0.5 * E0 * 2.0 ** -24 * o["nev1"] + 1e-9 * np.abs(o["Eabs"])
    erel_abs = ebound / np.maximum(np.abs(o["Eabs"]), 1e-300)
    for k in ("Eabs", "Tbw"):
        a, b = g[k], o[k]
        mask = np.abs(b) > 0
        assert np.array_equal(a == 0, b == 0), k
        rel = np.abs(a[mask] - b[mask]) / np.abs(b[mask])
        assert np.all(rel <= erel_abs[mask] + 1e-12), (k, np.max(rel - erel_abs[mask]))
    assert np.array_equal(g["ispec"], o["ispec"])
    assert np.array_equal(g["iscatspec"], o["iscatspec"])
    mask = o["spec"] != 0
    assert np.max(np.abs(g["spec"][mask] / o["spec"][mask] - 1)) < 1e-9
    To, dTo, dEo = oracle_lib.Oracle(m).temperature(o)
    # T ~ EJv^(1/4) at worst: a quarter of EJv's relative bound (the
    # piecewise-linear T^4 interpolation keeps dT/T <= dE/E)
    erel = (0.5 * u1 * o["nev1"] + 1e-9 * o["EJv"]) / np.maximum(o["EJv"], 1e-300)
    trel = np.abs(g["T"] / To - 1)
    assert np.all(trel <= np.minimum(erel, 1.0) + 1e-9), np.max(trel - erel)
    return g, o


def test_serial_parity_tiny(gpu, tiny_model, oracle_lib):
    _serial_compare(gpu, tiny_model, oracle_lib, 400, seed=5)


def test_serial_parity_small(gpu, small_model, oracle_lib):
    _serial_compare(gpu, small_model, oracle_lib, 150, seed=17)


def test_serial_parity_lane_lookups(gpu, small_model, oracle_lib):
    """The per-lane binary-search lookups (coop_lanes=0) follow the same
    histories as the wavefront-cooperative ones used by a lone lane."""
    g, _ = _serial_compare(gpu, small_model, oracle_lib, 150, seed=17, coop_lanes=0)
    h = transport.mc_radiation(gpu, 150, seed=17, nlanes=1)
    for k in ("EJv", "EJv2", "Eabs", "Tbw", "Ni", "ispec"):
        assert np.array_equal(g[k], h[k]), k


@pytest.fixture(scope="module")
def tiny_rw_model(workdir):
    from mcmax_amd import configs
    return configs.load_config("tiny", workdir + "/gpu_tiny_rw", mdust=1e-3, randomwalk=True, factrw=2.0)


@pytest.fixture(scope="module")
def small_rw_model(workdir):
    from mcmax_amd import configs
    return configs.load_config("small", workdir + "/gpu_small_rw", randomwalk=True)


def test_serial_parity_tiny_random_walk(gpu, tiny_rw_model, oracle_lib):
    g, o = _serial_compare(gpu, tiny_rw_model, oracle_lib, 400, seed=9)
    assert o["nrw"] > 100


def test_serial_parity_small_random_walk(gpu, small_rw_model, oracle_lib):
    g, o = _serial_compare(gpu, small_rw_model, oracle_lib, 300, seed=23)
    assert o["nrw"] > 50


def test_energy_conservation(gpu, small_rw_model):
    """Every packet leaves the grid with its energy: sum over escapes = L*/4pi."""
    n = 200000
    gpu.set_model(small_rw_model)
    g = transport.mc_radiation(gpu, n, seed=2)
    assert g["ngeomfail"] == 0 and g["noverflow"] == 0
    assert g["nescape"] == n
    assert abs(g["Eescape"] / small_rw_model.Lstar - 1) < 1e-9


def _statistics(gpu, m, oracle_lib, n, min_ni=50):
    """A full-device pass against the oracle at matched packet count: the
    temperature criteria (median |dT|/T <= 1%, no systematic shift, per-cell
    T within 3 sigma of the MC noise by Lucy's dT of both runs, Main.f:527-533)
    and energy conservation.  Packets in flight together take their
    Bjorkman-Wood slices in arrival order: that leaves the temperatures
    within the serial run's statistics at the default concurrency (at most
    nphot/64 in flight), while interaction counts and the escaping MC
    spectrum of optically thick, unresolved cells do depend on the order,
    so they are only compared in the one-lane tests."""
    gpu.set_model(m)
    g = transport.mc_radiation(gpu, n, seed=101)
    orc = oracle_lib.Oracle(m)
    o = orc.mcradiation(n, rng="philox", seed=202)
    To, dTo, _ = orc.temperature(o)
    assert g["noverflow"] == 0 and g["ngeomfail"] == 0
    assert g["nescape"] == n and abs(g["Eescape"] / m.Lstar - 1) < 1e-9
    sel = (o["Ni"][1:] > min_ni) & (g["Ni"][1:] > min_ni)
    Tg, dTg = g["T"][1:][sel], g["dT"][1:][sel]
    rel = Tg / To[1:][sel] - 1
    assert np.median(np.abs(rel)) < 0.01, np.median(np.abs(rel))
    assert abs(np.median(rel)) < 0.003, np.median(rel)
    z = np.abs(Tg - To[1:][sel]) / np.sqrt(dTg ** 2 + dTo[1:][sel] ** 2)
    assert np.mean(z < 3.0) > 0.975, np.sort(z)[-10:]
    print("nemit gpu/oracle", g["nemit"] / o["nemit"])
    return g, o


def test_parallel_statistics_tiny(gpu, tiny_model, oracle_lib):
    _statistics(gpu, tiny_model, oracle_lib, 100000)


def test_parallel_statistics_small_random_walk(gpu, small_rw_model, oracle_lib):
    g, o = _statistics(gpu, small_rw_model, oracle_lib, 100000)
    assert g["nrw"] > 0


def _dhs_overrides(ngrains):
    return {"abun01": 1.0, "computepart01:file": "dhs.dat", "computepart01:amin": 0.05,
            "computepart01:amax": 1000.0, "computepart01:apow": 3.5, "computepart01:fmax": 0.8,
            "computepart01:ngrains": ngrains, "computepart01:blend": False}


@pytest.fixture(scope="module")
def tiny_dhs_model(workdir, gpu):
    from mcmax_amd import configs
    return configs.load_config("tiny", workdir + "/gpu_tiny_dhs", tr=gpu, mdust=1e-3, **_dhs_overrides(3))


@pytest.fixture(scope="module")
def tiny_dhs_rw_model(workdir, gpu):
    from mcmax_amd import configs
    return configs.load_config("tiny", workdir + "/gpu_tiny_dhs_rw", tr=gpu, mdust=1e-2, randomwalk=True,
                               factrw=2.0, **_dhs_overrides(3))


def test_serial_parity_dhs_species(gpu, tiny_dhs_model, oracle_lib):
    """Three DHS species with different opacities, albedos and Mueller
    matrices: the per-species choices of reemission and scattering."""
    m = tiny_dhs_model
    assert m.ng == 3 and not np.allclose(m.Kabs[0], m.Kabs[2])
    _serial_compare(gpu, m, oracle_lib, 300, seed=11)


def test_serial_parity_dhs_species_random_walk(gpu, tiny_dhs_rw_model, oracle_lib):
    g, o = _serial_compare(gpu, tiny_dhs_rw_model, oracle_lib, 200, seed=13)
    assert o["nrw"] > 20


def test_parallel_statistics_dhs_species(gpu, workdir, oracle_lib):
    """Three DHS species (class reemission tables, species pick) at full
    device concurrency against the oracle's serial loop: the temperature
    criteria and energy conservation, at the tiny model's own dust mass.
    (With 10x more dust, one species or three alike, the row above the
    midplane runs ~1% cool at full concurrency, the arrival-order lag of
    the Bjorkman-Wood slices, and Lucy's dT understates the noise of the
    random-walking midplane row, so the 3-sigma rule does not hold there;
    the one-lane tests above cover that model.)"""
    from mcmax_amd import configs
    m = configs.load_config("tiny", workdir + "/gpu_tiny_dhs_thin", tr=gpu, **_dhs_overrides(3))
    assert m.ng == 3
    _statistics(gpu, m, oracle_lib, 100000)


def test_serial_parity_interaction_roulette(gpu, workdir, oracle_lib):
    """Keyword maxinteract != 0: after every interaction the packet dies with
    probability 1/maxinteract, the survivors carry E maxinteract/(maxinteract-1)
    (MCRadiation.f:567-571).  One lane against the oracle: the same packets
    die, and every counter but the midplane-split step count agrees."""
    from mcmax_amd import configs
    m = configs.load_config("tiny", workdir + "/gpu_tiny_roulette", mdust=1e-3, maxinteract=40)
    assert m.maxinteract == 40
    gpu.set_model(m)
    n = 300
    g = transport.mc_radiation(gpu, n, seed=17, nlanes=1, maxinteract=m.maxinteract)
    o = oracle_lib.Oracle(m).mcradiation(n, rng="philox", seed=17, maxinteract=m.maxinteract)
    assert o["nmaxinteract"] > 10
    assert g["nmaxinteract"] == o["nmaxinteract"]
    for k in COUNTERS:
        if k == "nsteps":
            assert abs(g[k] - o[k]) <= 0.005 * o[k], (k, g[k], o[k])
        else:
            assert g[k] == o[k], (k, g[k], o[k])
    assert np.array_equal(g["Ni"], o["Ni"])
    # killed packets take their energy with them: less escapes than L*
    assert g["nescape"] + g["nmaxinteract"] == n
    assert g["Eescape"] < m.Lstar
