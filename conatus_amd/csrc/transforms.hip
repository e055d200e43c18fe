// Batched grid <-> coefficient transforms of libcnhip along the last axis of (rows, n) arrays:
// the standalone Chebyshev, RealFourier and ComplexFourier transforms of the reference's device
// runtime (python/conatus/runtime/cuda.py:93-125 over the cheb_*/realfourier_*/complexfourier_*
// bindings of core/bindings/module.cpp:300-495).
//
// Every one of them is a fixed linear map from the n_in values of a row to the n_out values of the
// result, so a batch is one product (rows x n_in) x (n_in x n_out): the caller passes the map as
// a dense matrix (built once per size on the host, with the reference's conventions: Gauss-
// Chebyshev nodes, the (-1)^n DCT-II/M scaling with c_0 halved, RealFourier [cos_k, msin_k] pairs,
// ComplexFourier fftfreq order with the Nyquist entry zero), and the product runs on FP64 MFMA.
// For the sizes of this path (n up to a few hundred) the product is cheap and a batch is bounded
// by its HBM traffic; no FFT plan is needed, and any grid size works.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "cn_common.h"
#include "cnhip.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int TR_KC = 8;   // 16-output chunks accumulated per pass over a row tile

// out[r][k][p] = sum_m Ar[k][m] in[r][m][p] (+ i sum_m Ai[k][m] in[r][m] for a complex map, P = 2),
// r < rows, k < nout, m < nin, p < P (P = 1: real rows; P = 2: complex rows, (re, im) interleaved).
// One wavefront per tile of 16 columns c = (row, part): rows 16 / P per tile.  The products are
// v_mfma_f64_16x16x4f64 with A = 16 outputs x 4 inputs of the map (lane l: A[l % 16][l / 16]),
// B = 4 inputs x 16 columns (lane l: B[l / 16][l % 16]) and D = 16 outputs x 16 columns (lane l:
// D[4 i + l / 16][l % 16] in register i).  For a complex map the second product takes B' = i B:
// column (row, re) gets -Im, column (row, im) gets +Re of the same row, read from the partner lane.
template <bool CA>
__global__ void __launch_bounds__(64) k_tr_rows(const double* __restrict__ in, double* __restrict__ out, int64_t rows,
                                                int nin, int nout, int P, const double* __restrict__ Ar,
                                                const double* __restrict__ Ai) {
    const int l = threadIdx.x;
    const int col = l & 15, kq = l >> 4;
    const int CR = 16 / P;
    const int64_t r = (int64_t)blockIdx.x * CR + col / P;
    const int p = col % P;
    const bool rv = r < rows;
    const double* xr = in + (rv ? r : 0) * (int64_t)nin * P + p;
    const double sg = p ? 1.0 : -1.0;
    for (int kb = 0; kb < nout; kb += 16 * TR_KC) {
        d4 acc[TR_KC];
#pragma unroll
        for (int c = 0; c < TR_KC; ++c) acc[c] = d4{0.0, 0.0, 0.0, 0.0};
        for (int m0 = 0; m0 < nin; m0 += 4) {
            const int m = m0 + kq;
            const double b = (rv && m < nin) ? xr[(int64_t)m * P] : 0.0;
            double bp = 0.0;
            if constexpr (CA) bp = sg * __shfl_xor(b, 1, 64);
#pragma unroll
            for (int c = 0; c < TR_KC; ++c) {
                if (kb + 16 * c >= nout) break;       // uniform over the wavefront
                const int k = kb + 16 * c + col;
                const bool ok = k < nout && m < nin;
                const double a = ok ? Ar[(int64_t)k * nin + m] : 0.0;
                acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
                if constexpr (CA) {
                    const double ai = ok ? Ai[(int64_t)k * nin + m] : 0.0;
                    acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bp, acc[c], 0, 0, 0);
                }
            }
        }
        if (!rv) continue;
        double* orow = out + r * (int64_t)nout * P + p;
#pragma unroll
        for (int c = 0; c < TR_KC; ++c) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = kb + 16 * c + 4 * i + kq;
                if (k < nout) orow[(int64_t)k * P] = acc[c][i];
            }
        }
    }
}

}  // namespace

extern "C" {

int cn_tr_rows(const double* in, double* out, int64_t rows, int nin, int nout, int parts, const double* Ar,
               const double* Ai, void* stream) {
    if (!in || !out || !Ar) {
        cn::set_error("cn_tr_rows: null pointer");
        return CN_EINVAL;
    }
    if (rows < 0 || nin < 1 || nout < 1 || (parts != 1 && parts != 2)) {
        cn::set_error("cn_tr_rows: rows must be >= 0, nin and nout >= 1, parts 1 (real) or 2 (complex)");
        return CN_EINVAL;
    }
    if (Ai && parts != 2) {
        cn::set_error("cn_tr_rows: a complex map (Ai) needs complex rows (parts = 2)");
        return CN_EINVAL;
    }
    if (static_cast<const void*>(in) == static_cast<const void*>(out)) {
        cn::set_error("cn_tr_rows: the transform is out of place (in and out must differ)");
        return CN_EINVAL;
    }
    if (rows == 0) return CN_OK;
    const int64_t per = 16 / parts;
    const int64_t blocks = (rows + per - 1) / per;
    if (blocks > 0x7fffffffLL) {
        cn::set_error("cn_tr_rows: too many rows for one launch");
        return CN_EOVERFLOW;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (Ai)
        hipLaunchKernelGGL(k_tr_rows<true>, dim3((unsigned)blocks), dim3(64), 0, s, in, out, rows, nin, nout, parts, Ar, Ai);
    else
        hipLaunchKernelGGL(k_tr_rows<false>, dim3((unsigned)blocks), dim3(64), 0, s, in, out, rows, nin, nout, parts, Ar,
                           nullptr);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        cn::set_error(std::string("cn_tr_rows: launch failed: ") + hipGetErrorString(e));
        return CN_EHIP;
    }
    return CN_OK;
}

}  // extern "C"
