"""GPU parity of the poloidal-toroidal channel (conatus_amd/channel_pt.py, csrc/chpt.hip) against its
CPU oracle (oracle/channel_pt.py, pinned by tests/test_oracle_channel_pt.py).

Bars: the explicit term (six launches) within 1e-13 of the oracle's numpy evaluation relative to
its largest entry, every entry of F written (boundary rows and unused columns exactly zero,
the kx = 0 plane exactly conjugate-symmetric); states after k SBDF steps within 1e-11 relative
of the oracle (real LAPACK solves of the same systems); the laminar Poiseuille flow steady to
rounding; graph replays and split runs equal to eager single runs bit for bit; at the
128 x 97 x 128 configuration the explicit term at 1e-13 and three SBDF3 steps at 1e-11."""
import numpy as np
import pytest

from conatus_amd.channel_pt import ChannelPT3D
from oracle.channel_pt import ChannelPTOracle

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(abs(a - b)) / max(np.max(abs(b)), 1e-300))


def _oracle(pb, scheme="SBDF3", dt=5e-4):
    return ChannelPTOracle(pb.Nx, pb.Ny, pb.Nz, pb.Re, pb.Lx, pb.Lz, pb.G, pb.dealias, scheme=scheme, dt=dt)


def _rough(pb, seed):
    """Perturbed Poiseuille flow plus smooth random values in every phi, psi and mean coefficient
    (taus and boundary-row slots included)."""
    rng = np.random.default_rng(seed)
    X = pb.initial_state(seed=seed, energy=1e-2, kmax=2)
    A, B = pb.split(X)
    A += 0.02 * rng.standard_normal(A.shape) * np.exp(-0.25 * np.arange(pb.nP))
    B += 0.02 * rng.standard_normal(B.shape) * np.exp(-0.25 * np.arange(pb.nS))
    return X


def _device_F(pb, X):
    import torch
    from conatus_amd.channel_pt import ChannelPTExplicit
    ex = ChannelPTExplicit(pb)
    out = torch.full((pb.size,), float("nan"), dtype=torch.float64, device="cuda")
    ex.write(torch.from_numpy(X).cuda(), out)
    return out.cpu().numpy()


@pytest.mark.parametrize("shape", [(8, 9, 8), (16, 17, 12), (32, 33, 32), (64, 65, 64), (256, 9, 4)])
def test_explicit_term_matches_oracle(shape):
    """Grids 12 .. 96 points take the register-split transforms; (256, 9, 4) has Mx = 384 and
    Mz = 6, which take the Stockham kernels.  Its bar is 1e-12: with kx up to 127 the combination
    a = kx^2 P_uu + .. - k^2 P_vv cancels terms ~1.6e4 times larger than the result."""
    nx, ny, nz = shape
    print(f"[chpt explicit] {shape}", flush=True)
    pb = ChannelPT3D(Nx=nx, Ny=ny, Nz=nz, Re=400.0, Lx=2 * np.pi, Lz=np.pi)
    X = _rough(pb, seed=nx + ny)
    F = _device_F(pb, X)
    assert np.all(np.isfinite(F))                       # every entry written
    ref = _oracle(pb).explicit(X)
    r = rel(F, ref)
    print(f"relative difference {r:.2e}", flush=True)
    assert r < (1e-12 if nx > 128 else 1e-13)
    A, B = pb.split(F)
    assert not A[:, :, :4].any() and not B[:, :, :2].any()        # boundary rows
    for s in range(0, pb.P, pb.Kzh):                             # kz = 0 slots: no (kx, -0) mode
        assert not A[2:, s].any() and not B[2:, s].any()
    for kz in range(1, pb.Kzh):                                  # kx = 0: exact conjugate pairs
        assert np.array_equal(A[2, kz], A[0, kz]) and np.array_equal(A[3, kz], -A[1, kz])
        assert np.array_equal(B[2, kz], B[0, kz]) and np.array_equal(B[3, kz], -B[1, kz])
    assert not A[:, 0].any()                                      # the (0, 0) phi pin


@pytest.mark.parametrize("scheme,steps", [("SBDF1", 6), ("SBDF2", 10), ("SBDF3", 20), ("SBDF4", 12)])
def test_trajectory_matches_oracle(scheme, steps):
    from conatus_amd.channel_pt import ChannelPTSolver
    print(f"[chpt trajectory] {scheme}", flush=True)
    pb = ChannelPT3D(Nx=16, Ny=17, Nz=12, Re=800.0)
    X0 = pb.initial_state(seed=3, energy=1e-3, kmax=3)
    s = ChannelPTSolver(pb, scheme, 2e-3)
    s.set_state(X0)
    s.run(steps)
    ref = _oracle(pb, scheme, 2e-3).run(X0, steps)
    r = rel(s.state(), ref)
    print(f"{scheme}: relative difference after {steps} steps {r:.2e}", flush=True)
    assert r < 1e-11
    assert s.t == steps * 2e-3 and s.iteration == steps


def test_graph_split_runs_and_poiseuille():
    from conatus_amd.channel_pt import ChannelPTSolver
    pb = ChannelPT3D(Nx=16, Ny=17, Nz=16, Re=1000.0)
    X0 = pb.initial_state(seed=7, energy=1e-3)
    a = ChannelPTSolver(pb, "SBDF3", 1e-3)
    a.set_state(X0)
    a.run(20)
    b = ChannelPTSolver(pb, "SBDF3", 1e-3)
    b.set_state(X0)
    b.run_graph(9)                  # 2 startup steps, 2 ring turns, 1 single phase
    b.run_graph(11)                 # 3 ring turns, 2 single phases
    c = ChannelPTSolver(pb, "SBDF3", 1e-3)
    c.set_state(X0)
    c.run(7)
    c.run(13)
    assert np.array_equal(a.state(), b.state()) and np.array_equal(a.state(), c.state())
    assert a.t == b.t == c.t == 20 * 1e-3
    p = ChannelPTSolver(pb, "SBDF3", 1e-3)
    P = pb.poiseuille()
    p.set_state(P)
    p.run_graph(16)
    assert rel(p.state(), P) < 1e-13


def test_config_channel_128x97x128():
    """The configuration: Lx = 4 pi, Lz = 2 pi, Re = 2000, Poiseuille + a seeded divergence-free
    perturbation of energy 1e-4, SBDF3, dt = 5e-4."""
    from conatus_amd.channel_pt import ChannelPTSolver
    print("[chpt config] building 128x97x128", flush=True)
    pb = ChannelPT3D(Nx=128, Ny=97, Nz=128, Re=2000.0)
    X0 = pb.initial_state(seed=0, energy=1e-4)
    o = _oracle(pb, "SBDF3", 5e-4)
    F = _device_F(pb, X0)
    r = rel(F, o.explicit(X0))
    print(f"[chpt config] explicit term relative difference {r:.2e}", flush=True)
    assert r < 1e-13
    s = ChannelPTSolver(pb, "SBDF3", 5e-4)
    s.set_state(X0)
    s.run_graph(3)
    ref = o.run(X0, 3)
    r = rel(s.state(), ref)
    print(f"[chpt config] state after 3 SBDF3 steps: relative difference {r:.2e}", flush=True)
    assert r < 1e-11
