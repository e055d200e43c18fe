"""GPU tests of the continuous-batching server (conatus_amd/serving.py).

Each result must be the same alone and in the mix: every request's final state is compared bit
for bit with the same run stepped alone by IVPSolver (same scheme, dt and kernel mode), for
requests of several sizes, Rayleigh numbers and step counts arriving at different ticks, so that
runs start while others run, join the resident slot ensemble of their size class (state, history
and factors copied in), leave it when they finish, and the slot ensembles grow when full."""
import numpy as np
import pytest

from conatus_amd.problems import RayleighBenard2D

pytestmark = pytest.mark.gpu


def _alone(q, dt, mode):
    from conatus_amd.timestep import IVPSolver
    s = IVPSolver(q.problem, "SBDF2", dt, mode=mode)
    s.set_state(q.X0)
    s.run(q.steps)
    return s.state()


def _requests(spec):
    from conatus_amd.serving import Request
    out = []
    for i, (nx, nz, ra, steps, arrival) in enumerate(spec):
        pb = RayleighBenard2D(Nx=nx, Nz=nz, Lx=4.0, Ra=ra, Pr=1.0)
        out.append(Request(pb, steps, pb.initial_state(seed=100 + i), arrival=arrival))
    return out


def test_each_result_is_the_same_alone_and_in_the_mix():
    from conatus_amd.ensembles import discretization_key
    from conatus_amd.serving import Server
    spec = [(16, 12, 2e4, 9, 0), (32, 16, 5e4, 14, 0), (16, 12, 8e4, 5, 0), (16, 12, 3e4, 7, 2),
            (24, 12, 1e5, 6, 3), (32, 16, 2e5, 4, 6), (16, 12, 6e4, 3, 20)]
    reqs = _requests(spec)
    srv = Server("SBDF2", 1e-3)
    st = srv.serve(reqs, clock="ticks")
    print(st, flush=True)
    assert st["requests"] == len(spec) and st["steps"] == sum(s[3] for s in spec)
    assert st["joins"] == len(spec)                           # every run joined its class after startup
    for q in reqs:
        assert len(q.step_latencies) == q.steps
        ref = _alone(q, 1e-3, srv._modes[discretization_key(q.problem)])
        assert np.array_equal(q.result, ref), (q.rid, np.max(abs(q.result - ref)))


def test_pages_built_on_demand_keep_results():
    from conatus_amd.serving import Server
    reqs = _requests([(16, 12, 2e4, 9, 0), (16, 12, 5e4, 7, 1), (16, 12, 9e4, 6, 2), (16, 12, 7e4, 5, 2)])
    srv = Server("SBDF2", 1e-3, autotune=False, mode="wave", page_size=2)
    st = srv.serve(reqs, clock="ticks")
    assert st["pages"]["16x12"] == 2                          # two stepping pages of two slots
    for q in reqs:
        assert np.array_equal(q.result, _alone(q, 1e-3, "wave"))


def test_sbdf3_startup_pages():
    from conatus_amd.serving import Server
    from conatus_amd.timestep import IVPSolver
    reqs = _requests([(16, 12, 2e4, 8, 0), (16, 12, 5e4, 2, 0), (16, 12, 9e4, 6, 3)])
    srv = Server("SBDF3", 1e-3, autotune=False, mode="wave", page_size=4)
    srv.serve(reqs, clock="ticks")
    for q in reqs:
        s = IVPSolver(q.problem, "SBDF3", 1e-3, mode="wave")
        s.set_state(q.X0)
        s.run(q.steps)
        assert np.array_equal(q.result, s.state())


def test_admission_by_footprint_serializes_when_the_budget_is_small():
    from conatus_amd.serving import Server, footprint_bytes
    reqs = _requests([(16, 12, 2e4, 4, 0), (16, 12, 4e4, 3, 0), (16, 12, 6e4, 5, 0)])
    budget = footprint_bytes(reqs[0].problem) * 1.5          # one run at a time
    srv = Server("SBDF2", 1e-3, hbm_budget=budget, autotune=False, mode="wave")
    srv.serve(reqs, clock="ticks")
    adm = [q.admitted for q in reqs]
    assert adm == sorted(adm) and len(set(adm)) == 3          # admitted one after another, in order
    for q in reqs:
        assert np.array_equal(q.result, _alone(q, 1e-3, "wave"))
    assert srv.resident == 0


def test_channel_runs_in_the_mix_equal_their_runs_alone():
    """3D channel requests served beside RB 2D ones (the served configuration mixes both): each
    result bit-identical to its run alone."""
    from conatus_amd.problems import Channel3D
    from conatus_amd.serving import Request, Server
    reqs = _requests([(16, 12, 2e4, 9, 0), (16, 12, 5e4, 6, 2)])
    for i, (re_, steps, arrival) in enumerate([(800.0, 8, 0), (1500.0, 5, 1), (800.0, 6, 3)]):
        pb = Channel3D(Nx=8, Ny=13, Nz=8, Re=re_, G=2.0 / 1000.0)
        reqs.append(Request(pb, steps, pb.initial_state(seed=i, energy=1e-3), arrival=arrival))
    srv = Server("SBDF2", 1e-3, autotune=False, mode="wave", channel_page_size=2)
    st = srv.serve(reqs, clock="ticks")
    print(st, flush=True)
    assert st["requests"] == len(reqs) and "channel 8x13x8" in st["pages"]
    for q in reqs:
        assert np.array_equal(q.result, _alone(q, 1e-3, "wave")), q.rid
