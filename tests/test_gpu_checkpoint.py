"""GPU tests of checkpoint/restart (conatus_amd/checkpoint.py): a run continued from a checkpoint
is bit-identical to the uninterrupted run (the reference's rule, tests/output/test_io_hdf5.py),
eager and graph-replayed, at every SBDF order and for the Crank-Nicolson family (L X history
under the LX<i> keys); mismatched numerics or problems are refused, and an incomplete payload is
refused before the solver changes."""
import numpy as np
import pytest

from conatus_amd.problems import RayleighBenard2D

pytestmark = pytest.mark.gpu


def _pb(Ra=5e4):
    return RayleighBenard2D(Nx=32, Nz=16, Lx=4.0, Ra=Ra, Pr=1.0)


@pytest.mark.parametrize("scheme,graph", [("SBDF2", False), ("SBDF2", True), ("SBDF3", True), ("SBDF4", False),
                                          ("CNAB2", False), ("CNLF2", False), ("RK443", True)])
def test_restart_is_bit_for_bit(scheme, graph, tmp_path):
    from conatus_amd.timestep import IVPSolver
    pb = _pb()
    X0 = pb.initial_state(seed=5)
    a = IVPSolver(pb, scheme, 1e-3)
    a.set_state(X0)
    (a.run_graph if graph else a.run)(24)
    b = IVPSolver(pb, scheme, 1e-3)
    b.set_state(X0)
    (b.run_graph if graph else b.run)(11)
    path = b.save_checkpoint(tmp_path / "state")
    assert path.suffix == ".npz"
    c = IVPSolver(pb, scheme, 1e-3)
    c.load_checkpoint(tmp_path / "state")
    assert c.iteration == 11 and c.t == b.t
    (c.run_graph if graph else c.run)(13)
    assert np.array_equal(c.state(), a.state())
    assert c.t == a.t


def test_restart_refuses_other_numerics_and_problems(tmp_path):
    from conatus_amd.errors import OptionError
    from conatus_amd.timestep import IVPSolver
    pb = _pb()
    s = IVPSolver(pb, "SBDF2", 1e-3)
    s.set_state(pb.initial_state(seed=1))
    s.run(3)
    s.save_checkpoint(tmp_path / "c.npz")
    with pytest.raises(OptionError, match="dt"):
        IVPSolver(pb, "SBDF2", 2e-3).load_checkpoint(tmp_path / "c.npz")
    with pytest.raises(OptionError, match="Ra"):
        IVPSolver(_pb(Ra=6e4), "SBDF2", 1e-3).load_checkpoint(tmp_path / "c.npz")
    with pytest.raises(OptionError, match="h5py"):
        s.save_checkpoint(tmp_path / "c.h5")


def test_record_has_the_reference_keys_and_cn_history(tmp_path):
    """The meta record carries the keys the reference's loader reads (integrator.py:977-1000) and
    a CN run stores nhist_LX levels of L X as LX<i> (integrator.py:1005-1008)."""
    import json
    from conatus_amd.timestep import IVPSolver
    pb = _pb()
    s = IVPSolver(pb, "CNAB2", 1e-3)
    s.set_state(pb.initial_state(seed=2))
    s.run(4)
    path = s.save_checkpoint(tmp_path / "cn.npz")
    with np.load(path, allow_pickle=False) as z:
        files = set(z.files)
        meta = json.loads(str(z["meta"]))
        X = z["X"]
    for key in ("scheme", "dt", "startup", "startup_substeps", "n", "iteration", "t", "t_origin", "it_origin",
                "reduced", "ordering", "problem", "nhist", "nhist_LX", "identity"):
        assert key in meta, key
    assert meta["nhist"] == meta["nhist_LX"] == 2
    assert {"X", "MX0", "MX1", "F0", "F1", "LX0", "LX1"} <= files
    assert X.shape == (pb.K * pb.n,) and meta["n"] == pb.K * pb.n


def test_meta_without_format_and_incomplete_payload(tmp_path):
    """A record without 'format' (the reference writes none) restarts; a payload that lacks a
    declared level is refused and leaves the solver as it was (the reference's check_payload)."""
    from conatus_amd.checkpoint import payload, restore
    from conatus_amd.errors import OptionError
    from conatus_amd.timestep import IVPSolver
    pb = _pb()
    a = IVPSolver(pb, "SBDF3", 1e-3)
    a.set_state(pb.initial_state(seed=3))
    a.run(5)
    meta, arrays = payload(a)
    del meta["format"]
    b = IVPSolver(pb, "SBDF3", 1e-3)
    restore(b, meta, arrays)
    a.run(3)
    b.run(3)
    assert np.array_equal(a.state(), b.state()) and a.t == b.t
    c = IVPSolver(pb, "SBDF3", 1e-3)
    X1 = pb.initial_state(seed=4)
    c.set_state(X1)
    broken = dict(arrays)
    del broken["F2"]
    with pytest.raises(OptionError, match="incomplete"):
        restore(c, meta, broken)
    assert np.array_equal(c.state(), X1) and c.iteration == 0 and c.imex.history_length == 0
