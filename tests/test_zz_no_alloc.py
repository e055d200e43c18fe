"""The allocation-free solve (cn_bb_solve_noalloc) against the allocating one (cn_bb_solve), both
called through the C ABI on the same factored handle and the same right-hand sides: the same
kernels run, so the solutions must be equal bit for bit.  Two calls in a row (the second on a
smaller batch, so the shared work-space array still holds the first call's data), in every
kernel mapping that uses a work space (lane streamed, lane direct) and in wave mode (none);
sizes beyond the fixed 64 MiB are refused before any device work."""
import ctypes
import os

import numpy as np
import pytest

from conatus_amd import _lib
from tests.banded_cases import microbench_csr

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _handle(B, n, mode, seed):
    from conatus_amd.runtime.banded import BatchHandle
    rng = np.random.default_rng(seed)
    ip, ix, datas = microbench_csr(rng, B, n)
    return BatchHandle([(n, ip, ix, d) for d in datas], [0, 1], np.complex128, mode=mode), rng


def _both(h, B, n, nrhs, rng):
    """(allocating, allocation-free) solutions of the same right-hand sides on one handle."""
    import torch
    b = rng.standard_normal((nrhs, B, n)) + 1j * rng.standard_normal((nrhs, B, n))
    ref = torch.from_numpy(b).cuda()
    got = ref.clone()
    L = _lib.lib()
    s = torch.cuda.current_stream().cuda_stream
    v = ctypes.c_void_p
    print(f"case B={B} n={n} nrhs={nrhs}", flush=True)
    _lib.check(L.cn_bb_solve(h._h, v(ref.data_ptr()), nrhs, B * n, 0, v(s)), "cn_bb_solve")
    _lib.check(L.cn_bb_solve_noalloc(h._h, v(got.data_ptr()), nrhs, B * n, v(s)), "cn_bb_solve_noalloc")
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("mode,direct", [("lane", False), ("lane", True), ("wave", False)])
def test_noalloc_equals_the_allocating_solve(mode, direct, monkeypatch):
    if direct:
        monkeypatch.setenv("CNHIP_SOLVE", "direct")     # read when a handle is created
    big, rng = _handle(2048, 256, mode, 11)
    a, b = _both(big, 2048, 256, 4, rng)                  # 32 MiB of work space
    assert np.array_equal(a, b)
    small, rng = _handle(300, 96, mode, 12)
    a, b = _both(small, 300, 96, 1, rng)
    assert np.array_equal(a, b)
    # the first handle again, after the smaller call left its data in the shared array
    a, b = _both(big, 2048, 256, 2, rng)
    assert np.array_equal(a, b)


def test_noalloc_refuses_what_does_not_fit():
    import torch
    from conatus_amd.errors import LayoutError
    B, n = 4096, 256
    h, _ = _handle(B, n, "lane", 13)
    nrhs = SCRATCH // (B * n * 16) + 1                     # one right-hand side too many
    d = torch.zeros((nrhs, B, n), dtype=torch.complex128, device="cuda")
    v = ctypes.c_void_p
    with pytest.raises(LayoutError, match="64 MiB|67108864"):
        _lib.check(_lib.lib().cn_bb_solve_noalloc(h._h, v(d.data_ptr()), nrhs, B * n,
                                                  v(torch.cuda.current_stream().cuda_stream)),
                   "cn_bb_solve_noalloc")
    assert os.environ.get("CNHIP_SOLVE") is None
