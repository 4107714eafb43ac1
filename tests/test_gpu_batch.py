"""Many models in common launches (mcx_batch_transport): each model's tallies
are bit-identical to its own single-model pass with the same Philox key,
whatever the other models of the launch (their grids differ in size, dust
mass and star)."""
import numpy as np
import pytest

from mcmax_amd import transport

pytestmark = pytest.mark.gpu

KEYS = ("EJv", "EJv2", "Eabs", "Tbw", "Ni", "spec", "ispec", "scatspec")
COUNTS = ("nemit", "ninteract", "nabs", "nscat", "nescape", "nstarhit", "nsteps", "nrw", "maxsteps")


@pytest.fixture(scope="module")
def models(workdir):
    from mcmax_amd import configs
    specs = [dict(mdust=1e-4), dict(nrad=30, ntheta=14, mdust=1e-3, tstar=6000.0, rstar=1.5),
             dict(nrad=24, ntheta=8, mdust=3e-4, rin=0.3, randomwalk=True, factrw=2.0),
             dict(nrad=16, ntheta=12, mdust=1e-5, tstar=9000.0)]
    return [configs.load_config("tiny", f"{workdir}/batch{k}", **sp) for k, sp in enumerate(specs)]


@pytest.mark.parametrize("bpm", [1, 3, (5, 1, 2, 3)])
def test_batch_equals_single(models, bpm):
    """bpm: workgroups per model, or each model's own share (block map)."""
    trs = []
    try:
        for m in models:
            t = transport.Transport(0)
            t.set_model(m)
            trs.append(t)
        n, seed, stride = 20000, 77, 11
        shares = dict(model_blocks=list(bpm)) if isinstance(bpm, tuple) else dict(blocks_per_model=bpm)
        st = transport.batch_transport(trs, n, seed=seed, seed_stride=stride, randomwalk=True, factRW=2.0,
                                       max_ms=120000.0, **shares)
        assert np.all(st == 0), st
        batch = [t.tallies() for t in trs]
        # every model's timeline starts with its own first workgroup
        assert all(0.0 < b["t_done_ms"][4] < 120000.0 for b in batch), [b["t_done_ms"] for b in batch]
        Tb = [t.temperature()[0] for t in trs]
        for k, (t, m) in enumerate(zip(trs, models)):
            t.transport(n, seed=seed + stride * k, randomwalk=True, factRW=2.0, max_ms=120000.0)
            g = t.tallies()
            g["T"] = t.temperature()[0]
            assert g["nescape"] == n and abs(g["Eescape"] / m.Lstar - 1) < 1e-9
            for key in KEYS:
                assert np.array_equal(batch[k][key], g[key]), (k, key)
            for key in COUNTS:
                assert batch[k][key] == g[key], (k, key)
            assert np.array_equal(Tb[k], g["T"]), k
        # the models differ: different energies and grids
        assert len({b["EJv"].shape for b in batch}) > 1
    finally:
        for t in trs:
            t.close()


def test_batch_graph_replays(models):
    """The pass captured in a HIP graph and replayed: five replays with
    different keys, eager passes and synchronisations between them, a new
    density for one model (same grid) before the third, the first model
    (whose context holds the stage table) set up again without class tables
    (its buffers move), then two replays with explicit shares of the
    workgroups per model that change between them; every replay equals the
    eager batched pass with the same keys, bit for bit."""
    trs = []
    try:
        for m in models[:3]:
            t = transport.Transport(0)
            t.set_model(m)
            trs.append(t)
        n = 12000
        kw = dict(seed_stride=5, blocks_per_model=2, randomwalk=True, factRW=2.0)
        for rep, seed in enumerate((101, 202, 303, 404, 505)):
            if rep >= 3:
                # each model's own share; the total stays 6, the shares move
                kw = dict(seed_stride=5, model_blocks=[(3, 2, 1), (1, 1, 4)][rep - 3], randomwalk=True,
                          factRW=2.0)
            if rep == 2:
                m2 = models[1]
                dens = m2.dens.copy()
                m2.dens = dens * 1.5
                trs[1].set_model(m2, tables=False)
                m2.dens = dens
                trs[0].class_table_bytes = 0
                trs[0].set_model(models[0], tables=False)
            st = transport.batch_transport(trs, n, seed=seed, graph=True, **kw)
            assert np.all(st == 0), st
            got = [t.tallies() for t in trs]
            # eager work between replays
            transport.batch_transport(trs, n, seed=seed, graph=False, **kw)
            want = [t.tallies() for t in trs]
            for k in range(len(trs)):
                for key in KEYS:
                    assert np.array_equal(got[k][key], want[k][key]), (rep, k, key)
                for key in COUNTS:
                    assert got[k][key] == want[k][key], (rep, k, key)
    finally:
        for t in trs:
            t.close()
