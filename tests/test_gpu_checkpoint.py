"""GPU tests of checkpoint/restart (conatus_amd/checkpoint.py): a run continued from a checkpoint
is bit-identical to the uninterrupted run (the reference's rule, tests/output/test_io_hdf5.py),
eager and graph-replayed, at every SBDF order; mismatched numerics or problems are refused."""
import numpy as np
import pytest

from conatus_amd.problems import RayleighBenard2D

pytestmark = pytest.mark.gpu


def _pb(Ra=5e4):
    return RayleighBenard2D(Nx=32, Nz=16, Lx=4.0, Ra=Ra, Pr=1.0)


@pytest.mark.parametrize("scheme,graph", [("SBDF2", False), ("SBDF2", True), ("SBDF3", True), ("SBDF4", False)])
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
