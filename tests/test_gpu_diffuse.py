"""GPU parity of the diffusion approximation (mcx_diffuse) with the oracle's
Diffuse (Diffusion.f:9-329): the same cells are selected and numbered, the
matrices agree to rounding, and the dense LU solves (rocSOLVER on the
device, textbook partial pivoting in the oracle) agree to the conditioning
of the system."""
import numpy as np
import pytest

import pyoracle
from mcmax_amd import configs, transport
from test_oracle_rw import _grey

pytestmark = pytest.mark.gpu


def test_diffuse_closed_form(gpu, oracle_lib, workdir):
    m = configs.load_config("tiny", workdir + "/gdiff", mdust=1e-3)
    q = _grey(m, 3.0, 2.0, 0.4)
    q.dens = np.zeros_like(m.dens)
    q.dens[1:] = 1e-12 * (m.R_av[1:, None] / m.R_av[1])
    x = np.log(q.R_av[1:])
    T4 = 1500.0 ** 4 + (60.0 ** 4 - 1500.0 ** 4) * (x - x[0]) / (x[-1] - x[0])
    Tprof = T4 ** 0.25
    T = np.full((q.nr, q.nt), 100.0)
    T[1] = Tprof[0]
    T[-1] = Tprof[-1]
    T[:, 0] = np.concatenate([[100.0], Tprof])
    Ni = np.zeros((q.nr, q.nt), dtype=np.int32)
    dT = np.ones((q.nr, q.nt))
    gpu.set_model(q)
    gpu.set_temperature(T, dT, Ni)
    Tg, thg, Eg, ng = gpu.diffuse(factRW=2.0)
    o = pyoracle.Oracle(q)
    _, _, thick = o.random_walk_tables() if q.randomwalk else (None, None, None)
    thick0 = thg  # the device flags after the solve are compared below
    To, tho, Eo, no = o.diffuse(T, dT, Ni, _thick_flags(o, q, 2.0), np.zeros((q.nr, q.nt)))
    assert ng == no
    assert np.array_equal(thg, tho)
    assert np.allclose(Tg, To, rtol=1e-9), np.max(np.abs(Tg / To - 1))
    expect = np.repeat(Tprof[1:-1, None], q.nt - 1, axis=1)
    assert np.allclose(Tg[2:-1, 1:], expect, rtol=1e-8)
    solved = Tg != T
    assert np.allclose(Eg[solved], Eo[solved], rtol=1e-14)


def _thick_flags(o, m, factRW):
    """C%thick of InitRandomWalk with the given factRW"""
    saved = o.om.factRW
    o.om.factRW = factRW
    _, _, th = o.random_walk_tables()
    o.om.factRW = saved
    return th


def test_diffuse_after_pass(gpu, oracle_lib, workdir):
    """A thick model after a real pass: the device diffusion of the device
    temperatures against the oracle's diffusion of the same inputs."""
    m = configs.load_config("small", workdir + "/gdiff2", mdust=1e-3, randomwalk=True)
    gpu.set_model(m)
    g = transport.mc_radiation(gpu, 20000, seed=3)
    T0, dT0, Ni0 = g["T"].copy(), g["dT"].copy(), g["Ni"].copy()
    Tg, thg, Eg, ng = gpu.diffuse(nphotmin=25, dtmax=1.0)
    o = pyoracle.Oracle(m)
    To, tho, Eo, no = o.diffuse(T0, dT0, Ni0, _thick_flags(o, m, m.factRW), np.zeros((m.nr, m.nt)))
    assert no > 50 and ng == no
    assert np.array_equal(thg, tho)
    rel = np.abs(Tg / To - 1)
    assert rel.max() < 1e-7, rel.max()
