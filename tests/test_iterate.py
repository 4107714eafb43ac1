"""Control flow of the structure iteration (Main.f:202-791) with a stand-in
for the device: which calls happen in which order, the TempAverage weights,
the convergence rule (errT < epsiter on two iterations in a row, or
niter > maxiter, then one last pass), and that the first compareCells
counts no cell."""
import types

import numpy as np

from mcmax_amd import iterate


class FakeTransport:
    def __init__(self, m, errT):
        self.m, self.errT, self.calls = m, list(errT), []
        self.prev_seen = []

    def set_model(self, m, tables=True):
        self.calls.append(("set_model", tables))

    def transport(self, nphot, seed=1, maxinteract=0):
        self.calls.append(("transport", seed))
        return 0

    def tallies(self):
        n = (self.m.nr, self.m.nt)
        return dict(EJv=np.ones(n), Eescape=1.0, nsteps=10)

    def temperature(self):
        n = (self.m.nr, self.m.nt)
        return np.full(n, 100.0), np.full(n, 1.0), np.full(n, 0.1)

    def diffuse(self, nphotmin, dtmax, EJv=None):
        self.calls.append(("diffuse", nphotmin, dtmax))
        return np.full((self.m.nr, self.m.nt), 100.0), None, EJv, 0

    def timing(self):
        return 1.0, 2.0

    def temp_average(self, fw, first=False):
        self.calls.append(("temp_average", fw, first))
        return np.full((self.m.nr, self.m.nt), 100.0)

    def disk_structure(self, state, **kw):
        self.calls.append(("disk_structure",))
        return dict(state), 5

    def compare_cells(self, prev, cur):
        self.prev_seen.append(prev)
        self.calls.append(("compare_cells",))
        return self.errT.pop(0), 0.0, 0.0


def _model():
    nr, nt = 4, 3
    return types.SimpleNamespace(nr=nr, nt=nt, ng=1, dens=np.ones((nr, nt)), w=np.ones((nr, nt, 1)),
                                 V=np.ones((nr, nt)), Tstar=4000.0, R_av=np.arange(1.0, nr + 1), keywords={})


def _names(calls):
    return [c[0] for c in calls]


def test_converges_after_two_quiet_iterations():
    m = _model()
    tr = FakeTransport(m, [0.0, 5.0, 2.0, 1.0])
    it = iterate.Iteration(tr, m, 100, epsiter=3.0, maxiter=10)
    it.run()
    n = _names(tr.calls)
    assert n.count("transport") == 5 and n.count("disk_structure") == 4
    assert n.count("diffuse") == 5  # nphotdiffuse 25 (Init.f default) on every pass
    # TempAverage: the set-up copy, then 1, f_weight, f_weight, ...
    ta = [c for c in tr.calls if c[0] == "temp_average"]
    assert ta[0] == ("temp_average", 1.0, True) and ta[1] == ("temp_average", 1.0, False)
    assert [c[1] for c in ta[2:]] == [0.5, 0.5, 0.5]
    # the first comparison sees the set-up state: no dT, so no cell counts
    assert np.all(tr.prev_seen[0]["dT"] == 0.0)
    # seeds differ per pass; the last pass is not followed by a structure solve
    seeds = [c[1] for c in tr.calls if c[0] == "transport"]
    assert len(set(seeds)) == 5
    assert n[-1] != "disk_structure" and "transport" in n[-4:]


def test_maxiter_caps_the_loop():
    m = _model()
    tr = FakeTransport(m, [10.0] * 10)
    iterate.Iteration(tr, m, 100, epsiter=3.0, maxiter=2).run()
    n = _names(tr.calls)
    # niter = 1, 2, 3 solve the structure; niter = 3 > maxiter sets lastiter
    assert n.count("disk_structure") == 3 and n.count("transport") == 4


def test_no_structure_iteration():
    m = _model()
    tr = FakeTransport(m, [])
    iterate.Iteration(tr, m, 100, struct_iter=False, nphotdiffuse=0, dtdiffuse=1.0).run()
    assert _names(tr.calls) == ["set_model", "transport"]


def test_average_weight():
    assert iterate.average_weight(1, 0.5, 1000) == 1.0
    assert iterate.average_weight(2, 0.5, 1000) == 0.5
    assert iterate.average_weight(7, 0.5, 4) == 1.0 / 3.0
