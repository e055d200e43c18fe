"""The device-memory plan (conatus_amd/memory.py) against the device: the free-memory delta of
building and stepping a run (factorizations and graphs included) is the plan's total within the
allocators' rounding; the process floor (context, code objects) is measured in a fresh process."""
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLOOR_PROBE = r"""
import torch, numpy as np
from conatus_amd.runtime.hip import BandedBatch
import scipy.sparse as sp
torch.cuda.init()
A = sp.diags([np.full(64, 4.0), np.ones(63), np.ones(63)], [0, -1, 1], format="csr")
bb = BandedBatch([A] * 8, [])
d = torch.ones((8, 64), dtype=torch.float64, device="cuda")
bb.solve(d)
torch.cuda.synchronize()
free, total = torch.cuda.mem_get_info()
print(free)
"""


def _measure(build):
    import gc
    import torch
    gc.collect()                          # buffers of earlier tests' runs released first
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free0, _ = torch.cuda.mem_get_info()
    keep = build()
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    return free0 - free1, keep


@pytest.mark.parametrize("case", ["channel 64x65x64 SBDF3", "RB 2D 256x128 SBDF2"])
def test_plan_matches_the_device(case):
    from conatus_amd.memory import plan
    from conatus_amd.problems import Channel3D, RayleighBenard2D
    from conatus_amd.timestep import IVPSolver
    if case.startswith("channel"):
        pb, scheme = Channel3D(Nx=64, Ny=65, Nz=64, Re=2000.0), "SBDF3"
    else:
        pb, scheme = RayleighBenard2D(Nx=256, Nz=128, Lx=4.0, Ra=1e5, Pr=1.0), "SBDF2"
    X0 = pb.initial_state(seed=0)

    def build():
        s = IVPSolver(pb, scheme, 1e-4)
        s.set_state(X0)
        s.run_graph(12)
        return s
    import torch
    r0 = torch.cuda.memory_reserved()
    used, s = _measure(build)
    p = plan(pb, scheme, 1e-4)
    torch_part = torch.cuda.memory_reserved() - r0
    print(f"[memory] {case}: device {used / 2**20:.1f} MiB, plan {p.total / 2**20:.1f} MiB; torch allocator "
          f"{torch_part / 2**20:.1f} MiB, library {(used - torch_part) / 2**20:.1f} MiB, factorizations "
          f"{s.imex.n_factorizations} of {[f.props().factor_bytes / 2**20 for f in s.imex._factors.values()]} MiB "
          f"(mode {[f.props().solve_mode for f in s.imex._factors.values()]})",
          flush=True)
    print(p.explain(), flush=True)
    assert np.all(np.isfinite(s.state()))
    assert abs(used - p.total) <= p.tolerance_bytes()


def test_process_floor():
    """Free memory before a fresh process minus free memory once it has loaded torch and the
    library and run one solve (this process's own buffers stay where they are meanwhile)."""
    import gc
    import torch
    from conatus_amd.memory import PROCESS_FLOOR_BYTES
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free0, _ = torch.cuda.mem_get_info()
    out = subprocess.run([sys.executable, "-c", FLOOR_PROBE], capture_output=True, text=True, timeout=240,
                         cwd=__import__("os").path.dirname(__import__("os").path.dirname(__file__)))
    assert out.returncode == 0, out.stderr[-2000:]
    floor = free0 - int(out.stdout.strip().splitlines()[-1])
    print(f"[memory] process floor {floor / 2**20:.1f} MiB (declared {PROCESS_FLOOR_BYTES / 2**20:.0f} MiB)", flush=True)
    assert 0 < floor <= PROCESS_FLOOR_BYTES
