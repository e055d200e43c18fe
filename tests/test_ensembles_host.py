"""Host-side checks of the ensemble and serving layers (no GPU): grouping keys, footprint model,
option errors raised before any device work."""
import numpy as np
import pytest

from conatus_amd.problems import RayleighBenard2D


def test_discretization_key_ignores_parameters():
    from conatus_amd.ensembles import discretization_key
    a = RayleighBenard2D(Nx=16, Nz=12, Lx=4.0, Ra=1e4, Pr=1.0)
    b = RayleighBenard2D(Nx=16, Nz=12, Lx=4.0, Ra=1e6, Pr=7.0)
    c = RayleighBenard2D(Nx=32, Nz=12, Lx=4.0, Ra=1e4, Pr=1.0)
    assert discretization_key(a) == discretization_key(b) != discretization_key(c)


def test_footprint_grows_with_size_and_order():
    from conatus_amd.serving import footprint_bytes
    small = RayleighBenard2D(Nx=64, Nz=32, Lx=4.0, Ra=1e5, Pr=1.0)
    big = RayleighBenard2D(Nx=256, Nz=128, Lx=4.0, Ra=1e5, Pr=1.0)
    assert footprint_bytes(big) > 8 * footprint_bytes(small)
    assert footprint_bytes(small, 3) > footprint_bytes(small, 2)
    # a 128 x 64 run needs well under 1 GB: thousands fit a 200 GB budget
    assert footprint_bytes(RayleighBenard2D(Nx=128, Nz=64, Lx=4.0, Ra=1e5, Pr=1.0)) < 1e9


def test_server_refuses_bad_requests():
    from conatus_amd.errors import OptionError
    from conatus_amd.serving import Request, Server
    pb = RayleighBenard2D(Nx=16, Nz=12, Lx=4.0, Ra=1e4, Pr=1.0)
    srv = Server("SBDF2", 1e-3, autotune=False, mode="wave")
    with pytest.raises(OptionError):
        srv.submit(Request(pb, 0, pb.initial_state()))
    with pytest.raises(OptionError):
        srv.submit(Request(pb, 3, np.zeros((2, 2))))
    with pytest.raises(OptionError):
        Server("SBDF2", 1e-3, hbm_budget=10.0, autotune=False).submit(Request(pb, 3, pb.initial_state()))
    with pytest.raises(OptionError):
        Server(autotune=True, mode="wave")


def test_channel_keys_parameters_and_footprint():
    from conatus_amd.ensembles import discretization_key
    from conatus_amd.problems import Channel3D
    from conatus_amd.serving import footprint_bytes
    a = Channel3D(Nx=8, Ny=9, Nz=8, Re=500.0, G=1e-3)
    b = Channel3D(Nx=8, Ny=9, Nz=8, Re=900.0, G=1e-3)
    c = Channel3D(Nx=8, Ny=9, Nz=8, Re=900.0)
    assert discretization_key(a) == discretization_key(b) != discretization_key(c)   # G is part of F
    assert a.params() != b.params()
    big = Channel3D(Nx=16, Ny=17, Nz=16, Re=500.0)
    assert footprint_bytes(big) > 2 * footprint_bytes(a) > 0       # systems padded to 64 per handle
    assert footprint_bytes(a) > a.explicit_scratch_bytes() > 0


def test_memory_plan_handle_layout():
    """A factor handle's bytes by hand for B = 100 systems, n = 50, kl = 3 (kernel width 4),
    ku = 5 (kl + ku = 9 padded to the kernel width 16, U rows of 17), r = 2, complex: systems
    padded to 128; 100 systems of 50 rows are a partitioned batch (four chunks of at most 13
    rows, coupling widths 4 forward and 9 backward), whose spikes count too; 300 systems are
    not."""
    from conatus_amd.memory import handle_bytes, plan
    from conatus_amd.problems import RayleighBenard2D
    n, es = 50, 16

    def fixed(B):
        Bp = (B + 63) // 64 * 64
        band = Bp * n * (4 + 5 + 1) * es + Bp * n * 17 * es + Bp * n * 4 * es + Bp * n * 4
        border = 3 * Bp * 2 * n * es + 2 * (Bp * 4 * es + Bp * 2 * 4 + Bp * 8 + Bp * 4) + n * 4
        return band + border + B * n * es

    spikes = 100 * 4 * (4 * (13 + 4) + 9 * (13 + 9)) * es
    assert handle_bytes(100, n, 3, 5, 2) == fixed(100) + spikes
    assert handle_bytes(300, n, 3, 5, 2) == fixed(300)
    pb = RayleighBenard2D(Nx=64, Nz=32, Lx=4.0, Ra=1e5, Pr=1.0)
    p2, p3 = plan(pb, "SBDF2"), plan(pb, "SBDF3")
    assert p3.lines["factors"] == 3 * p2.lines["factors"] // 2   # one handle per startup order
    assert plan(pb, "RK443").lines["factors"] == p2.lines["factors"] // 2
    assert plan(pb, "MCNAB2").lines["factors"] == p2.lines["factors"]   # CNAB1 startup + MCNAB2
