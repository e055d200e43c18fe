"""GPU tests of the ensembles (conatus_amd/ensembles.py) against single runs.

The reference's own claim (python/conatus/ensembles/ensemble.py:8-20, tests/ensembles): a member of
a batch is bit-identical to the same run alone.  Here every member's arithmetic is its single run's
(the batched kernels never mix members), so members are checked bit for bit against IVPSolver runs
with the same kernel mode, eager and graph-replayed, for parameter ensembles (different Ra per
member) and for sets of uneven sizes; the single runs themselves are pinned to the CPU oracle by
tests/test_gpu_rb2d.py."""
import numpy as np
import pytest

from conatus_amd.problems import RayleighBenard2D

pytestmark = pytest.mark.gpu


def _pb(Nx, Nz, Ra):
    return RayleighBenard2D(Nx=Nx, Nz=Nz, Lx=4.0, Ra=Ra, Pr=1.0)


def _single(pb, X0, steps, dt, mode, graph):
    from conatus_amd.timestep import IVPSolver
    s = IVPSolver(pb, "SBDF2", dt, mode=mode)
    s.set_state(X0)
    (s.run_graph if graph else s.run)(steps)
    return s.state()


@pytest.mark.parametrize("graph", [False, True])
def test_parameter_ensemble_members_equal_single_runs(graph):
    from conatus_amd.ensembles import Ensemble
    from conatus_amd.timestep import IVPSolver
    ras = [2e4, 1e5, 4e5]
    pbs = [_pb(32, 16, ra) for ra in ras]
    solvers = [IVPSolver(pb, "SBDF2", 1e-3, mode="wave") for pb in pbs]
    ens = Ensemble.from_solvers(solvers, member_of=[0, 1, 2, 1])
    X0 = np.stack([pbs[i].initial_state(seed=10 + b) for b, i in enumerate([0, 1, 2, 1])])
    ens.set_states(X0)
    print(ens.explain(), flush=True)
    (ens.run_graph if graph else ens.run)(25)
    got = ens.states()
    for b, i in enumerate([0, 1, 2, 1]):
        ref = _single(pbs[i], X0[b], 25, 1e-3, "wave", graph)
        assert np.array_equal(got[b], ref), (b, np.max(abs(got[b] - ref)))
    assert abs(ens.t - 25e-3) < 1e-12 and ens.iteration == 25


def test_loop_rhs_equals_batched_rhs():
    from conatus_amd.ensembles import Ensemble
    pbs = [_pb(16, 12, ra) for ra in (3e4, 6e4)]
    X0 = np.stack([pb.initial_state(seed=b) for b, pb in enumerate(pbs)])
    out = []
    for rhs in ("batched", "loop"):
        e = Ensemble.from_problems(pbs, "SBDF2", 1e-3, rhs=rhs, mode="wave")
        e.set_states(X0)
        e.run(12)
        out.append(e.states())
    assert np.array_equal(out[0], out[1])


def test_callback_every():
    from conatus_amd.ensembles import Ensemble
    from conatus_amd.timestep import IVPSolver
    e = Ensemble(IVPSolver(_pb(16, 12, 1e4), "SBDF2", 1e-3), 2)
    pb = e.problems[0]
    e.set_states(np.stack([pb.initial_state(seed=1)] * 2))
    seen = []
    e.run(6, callback=lambda en: seen.append(en.iteration), every=2)
    assert seen == [2, 4, 6]
    s = e.states()
    assert np.array_equal(s[0], s[1])


def test_uneven_set_in_one_graph_equals_single_runs():
    from conatus_amd.ensembles import EnsembleSet
    spec = [(32, 16, 2e4), (16, 12, 5e4), (32, 16, 1e5), (24, 12, 3e4), (16, 12, 8e4)]
    pbs = [_pb(nx, nz, ra) for nx, nz, ra in spec]
    X0 = [pb.initial_state(seed=20 + i) for i, pb in enumerate(pbs)]
    es = EnsembleSet(pbs, "SBDF2", 1e-3)
    print(es.explain(), flush=True)
    assert len(es.groups) == 3
    for tu in es.tuning:
        assert tu["mode"] in ("lane", "wave8", "wave32", "part") and tu["ms_per_solve"]
    es.set_states(X0)
    es.run_graph(10)
    es.run_graph(8)                 # continuation of the same trajectories
    got = es.states()
    for i, pb in enumerate(pbs):
        gi, _ = es.where[i]
        ref = _single(pb, X0[i], 18, 1e-3, es.groups[gi].mode, graph=False)
        assert np.array_equal(got[i], ref), (i, np.max(abs(got[i] - ref)))


def test_large_ensemble_members_equal_single_runs():
    """130 members x 32 modes: the explicit term takes its separate z-backward launch (one pass per
    mode instead of one per grid line); members stay bit-identical to their single runs."""
    from conatus_amd.ensembles import Ensemble
    pb = _pb(64, 16, 5e4)
    P = 130
    e = Ensemble.from_problems([pb] * P, "SBDF2", 1e-3, mode="wave")
    X0 = np.stack([pb.initial_state(seed=200 + b) for b in range(P)])
    e.set_states(X0)
    e.run(6)
    got = e.states()
    for b in (0, 77, P - 1):
        ref = _single(pb, X0[b], 6, 1e-3, "wave", graph=False)
        assert np.array_equal(got[b], ref), (b, np.max(abs(got[b] - ref)))
