"""GPU check of shared factors (conatus_amd/runtime/shared.py): systems B = D_r A D_c solved with
A's factors are bit for bit the solves of B with its own factors (the reference's claim,
python/conatus/timestep/shared_factors.py:8-13), in both kernel modes, with half the factor
memory; unpaired primaries keep their own results."""
import numpy as np
import pytest
import scipy.sparse as sp

from tests.banded_cases import conditioned

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["lane", "wave"])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_shared_solves_are_bitwise_the_own_factor_solves(mode, dtype, rng):
    import torch
    from conatus_amd.runtime.banded import BatchHandle, csr_args
    from conatus_amd.runtime.shared import SharedBandedBatch, SharedFactor, sign_similarity
    n, r = 60, 2
    P = 1030 if mode == "lane" else 128                 # lane mode needs a batch of >= 1024 systems
    cplx = dtype == np.complex128
    print(f"case shared mode={mode} dtype={np.dtype(dtype).name} P={P}", flush=True)
    As = [conditioned(rng, n, 2, 3, r, dtype=complex if cplx else float) for _ in range(P)]
    mats, shares = list(As), []
    for p in range(0, P, 2):                             # every other system gets a sign-similar partner
        dr = rng.choice([-1.0, 1.0], n)
        dc = rng.choice([-1.0, 1.0], n)
        dc[:r] = dr[:r]
        B = sp.csr_matrix(sp.diags(dr) @ As[p] @ sp.diags(dc))
        got = sign_similarity(As[p], B, list(range(r)))
        assert got is not None
        shares.append(SharedFactor(p, len(mats), got[0], got[1]))
        mats.append(B)
    tdt = torch.complex128 if cplx else torch.float64
    s = torch.cuda.current_stream().cuda_stream
    own = BatchHandle([csr_args(A, dtype) for A in mats], list(range(r)), dtype, True, mode=mode)
    own.reserve(1)
    sh = SharedBandedBatch(mats, list(range(r)), shares, dtype, True, mode=mode)
    assert sh.factor_bytes < 0.75 * own.props().factor_bytes
    b = rng.standard_normal((len(mats), n)) + (1j * rng.standard_normal((len(mats), n)) if cplx else 0)
    d1 = torch.from_numpy(b).to("cuda", tdt)
    d2 = d1.clone()
    own.solve_ptr(d1.data_ptr(), 1, d1.numel(), False, s)
    sh.solve(d2)
    assert torch.equal(d1, d2)
