// libcnhip IMEX stage kernels: batched block-diagonal CSR products and the fused SBDF / RK
// right-hand-side assembly of  M dX/dt + L X = F(X, t)  (core/include/conatus/imex.hpp:1-8).
//
// The arithmetic follows the reference's host integrator operation by operation, with every
// product and sum rounded separately (this file is built with -ffp-contract=off and writes no
// fma), so that the assembled right-hand side is bitwise the one of the CPU oracle:
//   CSR row:  s = 0; s += a_p * x_{col_p} in storage order; y = alpha * s   (banded.hpp:70-76)
//   SBDF:     rhs = 0; for j = 1..q: rhs += c_j F^{n-j} - (a_j / dt) M X^{n-j}  (imex.hpp:167-172)
//   RK:       rhs = M X_0; for k < i: rhs += (dt A_ik) F_k; rhs -= (dt H_ik) L X_k  (imex.hpp:193-199)
//   F:        g * X + h  (pointwise model of the reference's device integrator)
// Complex products are the plain (ac - bd, ad + bc) form of std::complex / numpy.
// One thread per row; a fused stage reads X and the history once and writes the new history
// entries and the right-hand side in the same pass (HBM-bound streaming).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/cnhip.h"
#include "cn_common.h"

using cn::zc;

namespace {


__device__ __forceinline__ double pmul(double a, double b) { return a * b; }
__device__ __forceinline__ zc pmul(zc a, zc b) {
    const double rr = a.re * b.re, ii = a.im * b.im, ri = a.re * b.im, ir = a.im * b.re;
    return zc{rr - ii, ri + ir};
}
__device__ __forceinline__ double padd(double a, double b) { return a + b; }
__device__ __forceinline__ zc padd(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ double psub(double a, double b) { return a - b; }
__device__ __forceinline__ zc psub(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }
// real scalar times value: T(c) * v with T(c) = c + 0i
__device__ __forceinline__ double sc(double c, double v) { return c * v; }
__device__ __forceinline__ zc sc(double c, zc v) { return pmul(zc{c, 0.0}, v); }

template <class T>
__device__ __forceinline__ T csr_row(const int* __restrict__ rp, const int* __restrict__ idx,
                                     const T* __restrict__ val, const T* __restrict__ xm, size_t i) {
    T s = cn::zero<T>();
    for (int p = rp[i]; p < rp[i + 1]; ++p) s = padd(s, pmul(val[p], xm[idx[p]]));
    return s;
}

struct Csr {
    const int* rowptr;   // batch * (n + 1) row pointers (global into idx/val)
    const int* idx;      // local column indices
    const void* val;
};

template <class T>
__global__ void __launch_bounds__(256)
k_csr_mv(Csr A, const T* __restrict__ x, T* __restrict__ y, int n, size_t N, double alpha, int acc) {
    const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (g >= N) return;
    const size_t m = g / n, i = g - m * n;
    const T s = csr_row<T>(A.rowptr + m * (n + 1), A.idx, static_cast<const T*>(A.val), x + m * n, i);
    const T v = sc(alpha, s);
    y[g] = acc ? padd(y[g], v) : v;
}

constexpr int MAXH = 4;

struct SbdfArgs {
    const void* mx_hist[MAXH];   // M X^{n-j}, j = 1..q-1 (index j-1 = previous steps), may be unused
    const void* f_hist[MAXH];
    double c[MAXH + 1];          // c_1 .. c_q at index 1..q
    double adt[MAXH + 1];        // a_j / dt at index 1..q
    int q;
};

// One SBDF stage: mx0 = M x, f0 = F (external, or g x + h), rhs = sum_j (c_j F_j - a_j/dt MX_j)
// with F_1 = f0, MX_1 = mx0 and the older levels from the history pointers.
template <class T>
__global__ void __launch_bounds__(256)
k_sbdf_stage(Csr M, const T* __restrict__ x, const T* __restrict__ g, const T* __restrict__ h,
             const T* __restrict__ fext, T* __restrict__ mx0, T* __restrict__ f0, T* __restrict__ rhs,
             SbdfArgs a, int n, size_t N) {
    const size_t gi = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (gi >= N) return;
    const size_t m = gi / n, i = gi - m * n;
    const T mx = csr_row<T>(M.rowptr + m * (n + 1), M.idx, static_cast<const T*>(M.val), x + m * n, i);
    T fx;
    if (fext) {
        fx = fext[gi];
    } else {
        fx = cn::zero<T>();
        if (g) fx = pmul(g[gi], x[gi]);
        if (h) fx = padd(fx, h[gi]);
    }
    mx0[gi] = mx;
    if (f0) f0[gi] = fx;
    T r = psub(sc(a.c[1], fx), sc(a.adt[1], mx));
    r = padd(cn::zero<T>(), r);
#pragma unroll
    for (int j = 2; j <= MAXH; ++j) {
        if (j > a.q) break;
        const T fj = static_cast<const T*>(a.f_hist[j - 2])[gi];
        const T mj = static_cast<const T*>(a.mx_hist[j - 2])[gi];
        r = padd(r, psub(sc(a.c[j], fj), sc(a.adt[j], mj)));
    }
    rhs[gi] = r;
}

constexpr int MAXS = 4;

struct RkArgs {
    const void* F[MAXS];
    const void* LX[MAXS];
    double dtA[MAXS];   // dt * A_ik (0: skipped)
    double dtH[MAXS];   // dt * H_ik (0: skipped)
    int stages;         // k < stages
};

// rhs = MX0 + sum_k (dt A_ik) F_k - (dt H_ik) L X_k, in the reference's order.
template <class T>
__global__ void __launch_bounds__(256)
k_rk_rhs(const T* __restrict__ mx0, T* __restrict__ rhs, RkArgs a, size_t N) {
    const size_t gi = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (gi >= N) return;
    T r = mx0[gi];
#pragma unroll
    for (int k = 0; k < MAXS; ++k) {
        if (k >= a.stages) break;
        if (a.dtA[k] != 0.0) r = padd(r, sc(a.dtA[k], static_cast<const T*>(a.F[k])[gi]));
        if (a.dtH[k] != 0.0) r = psub(r, sc(a.dtH[k], static_cast<const T*>(a.LX[k])[gi]));
    }
    rhs[gi] = r;
}

template <class T>
__global__ void __launch_bounds__(256)
k_pointwise(T* __restrict__ out, const T* __restrict__ x, const T* __restrict__ g, const T* __restrict__ h,
            size_t N) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= N) return;
    T v = cn::zero<T>();
    if (g) v = pmul(g[i], x[i]);
    if (h) v = padd(v, h[i]);
    out[i] = v;
}

int fail_i(int code, const std::string& m) {
    cn::set_error(m);
    return code;
}

int launched(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_i(CN_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return CN_OK;
}

dim3 grid_for(size_t N) { return dim3((unsigned)((N + 255) / 256)); }

}  // namespace

struct cn_csr {
    int dtype = CN_F64, batch = 0, n = 0;
    int* rowptr = nullptr;
    int* idx = nullptr;
    void* val = nullptr;
    int64_t nnz = 0;
};

extern "C" {

int cn_csr_create(cn_csr_t* out, int dtype, int batch, int n, const int32_t* const* indptr,
                  const int32_t* const* indices, const void* const* data, const int64_t* nnz) {
    if (!out) return fail_i(CN_EINVAL, "cn_csr_create: null output");
    *out = nullptr;
    if ((dtype != CN_F64 && dtype != CN_C128) || batch < 1 || n < 1)
        return fail_i(CN_EINVAL, "cn_csr_create: bad dtype, batch or n");
    const size_t es = dtype == CN_C128 ? 16 : 8;
    int64_t total = 0;
    for (int m = 0; m < batch; ++m) {
        if (!indptr[m] || (nnz[m] && (!indices[m] || !data[m]))) return fail_i(CN_EINVAL, "cn_csr_create: null array");
        if (indptr[m][0] != 0 || indptr[m][n] != nnz[m]) return fail_i(CN_EINVAL, "cn_csr_create: inconsistent CSR");
        for (int i = 0; i < n; ++i)
            if (indptr[m][i + 1] < indptr[m][i]) return fail_i(CN_EINVAL, "cn_csr_create: indptr decreases");
        for (int64_t p = 0; p < nnz[m]; ++p)
            if (indices[m][p] < 0 || indices[m][p] >= n) return fail_i(CN_EINVAL, "cn_csr_create: column out of range");
        total += nnz[m];
    }
    if (total > INT32_MAX) return fail_i(CN_EOVERFLOW, "cn_csr_create: more than 2^31 entries");
    auto* c = new cn_csr;
    c->dtype = dtype;
    c->batch = batch;
    c->n = n;
    c->nnz = total;
    std::string err;
    int* hrp = new int[(size_t)batch * (n + 1)];
    int* hix = new int[total > 0 ? total : 1];
    char* hv = new char[(total > 0 ? total : 1) * es];
    int64_t off = 0;
    for (int m = 0; m < batch; ++m) {
        for (int i = 0; i <= n; ++i) hrp[(size_t)m * (n + 1) + i] = (int)(off + indptr[m][i]);
        for (int64_t p = 0; p < nnz[m]; ++p) hix[off + p] = indices[m][p];
        std::memcpy(hv + off * es, data[m], nnz[m] * es);
        off += nnz[m];
    }
    hipError_t e = hipMalloc(&c->rowptr, sizeof(int) * (size_t)batch * (n + 1));
    if (e == hipSuccess) e = hipMalloc(&c->idx, sizeof(int) * (total > 0 ? total : 1));
    if (e == hipSuccess) e = hipMalloc(&c->val, es * (total > 0 ? total : 1));
    if (e == hipSuccess) e = hipMemcpy(c->rowptr, hrp, sizeof(int) * (size_t)batch * (n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && total) e = hipMemcpy(c->idx, hix, sizeof(int) * total, hipMemcpyHostToDevice);
    if (e == hipSuccess && total) e = hipMemcpy(c->val, hv, es * total, hipMemcpyHostToDevice);
    delete[] hrp;
    delete[] hix;
    delete[] hv;
    if (e != hipSuccess) {
        cn_csr_destroy(c);
        return fail_i(CN_EHIP, std::string("cn_csr_create: ") + hipGetErrorString(e));
    }
    *out = c;
    cn::set_error("");
    return CN_OK;
}

void cn_csr_destroy(cn_csr_t c) {
    if (!c) return;
    if (c->rowptr) (void)hipFree(c->rowptr);
    if (c->idx) (void)hipFree(c->idx);
    if (c->val) (void)hipFree(c->val);
    delete c;
}

int cn_csr_matvec(cn_csr_t A, const void* x, void* y, double alpha, int accumulate, void* stream) {
    if (!A || !x || !y) return fail_i(CN_EINVAL, "cn_csr_matvec: null argument");
    const size_t N = (size_t)A->batch * A->n;
    Csr c{A->rowptr, A->idx, A->val};
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (A->dtype == CN_C128)
        hipLaunchKernelGGL((k_csr_mv<zc>), grid_for(N), dim3(256), 0, s, c, static_cast<const zc*>(x),
                           static_cast<zc*>(y), A->n, N, alpha, accumulate);
    else
        hipLaunchKernelGGL((k_csr_mv<double>), grid_for(N), dim3(256), 0, s, c, static_cast<const double*>(x),
                           static_cast<double*>(y), A->n, N, alpha, accumulate);
    return launched("k_csr_mv");
}

int cn_sbdf_stage(cn_csr_t M, const void* x, const void* g, const void* h, const void* fext, void* mx0, void* f0,
                  const void* const* mx_hist, const void* const* f_hist, int q, const double* c,
                  const double* a_over_dt, void* rhs, void* stream) {
    if (!M || !x || !mx0 || !rhs || !c || !a_over_dt) return fail_i(CN_EINVAL, "cn_sbdf_stage: null argument");
    if (q < 1 || q > MAXH) return fail_i(CN_EINVAL, "cn_sbdf_stage: order q must be 1..4");
    SbdfArgs a{};
    for (int j = 2; j <= q; ++j) {
        if (!mx_hist || !f_hist || !mx_hist[j - 2] || !f_hist[j - 2])
            return fail_i(CN_EINVAL, "cn_sbdf_stage: missing history level");
        a.mx_hist[j - 2] = mx_hist[j - 2];
        a.f_hist[j - 2] = f_hist[j - 2];
    }
    for (int j = 1; j <= q; ++j) {
        a.c[j] = c[j];
        a.adt[j] = a_over_dt[j];
    }
    a.q = q;
    const size_t N = (size_t)M->batch * M->n;
    Csr cm{M->rowptr, M->idx, M->val};
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (M->dtype == CN_C128)
        hipLaunchKernelGGL((k_sbdf_stage<zc>), grid_for(N), dim3(256), 0, s, cm, static_cast<const zc*>(x),
                           static_cast<const zc*>(g), static_cast<const zc*>(h), static_cast<const zc*>(fext),
                           static_cast<zc*>(mx0), static_cast<zc*>(f0), static_cast<zc*>(rhs), a, M->n, N);
    else
        hipLaunchKernelGGL((k_sbdf_stage<double>), grid_for(N), dim3(256), 0, s, cm, static_cast<const double*>(x),
                           static_cast<const double*>(g), static_cast<const double*>(h),
                           static_cast<const double*>(fext), static_cast<double*>(mx0), static_cast<double*>(f0),
                           static_cast<double*>(rhs), a, M->n, N);
    return launched("k_sbdf_stage");
}

int cn_rk_rhs(int dtype, int64_t N, const void* mx0, const void* const* F, const void* const* LX,
              const double* dtA, const double* dtH, int stages, void* rhs, void* stream) {
    if (!mx0 || !rhs || !dtA || !dtH || stages < 0 || stages > MAXS) return fail_i(CN_EINVAL, "cn_rk_rhs: bad argument");
    RkArgs a{};
    for (int k = 0; k < stages; ++k) {
        a.dtA[k] = dtA[k];
        a.dtH[k] = dtH[k];
        a.F[k] = F ? F[k] : nullptr;
        a.LX[k] = LX ? LX[k] : nullptr;
        if ((a.dtA[k] != 0.0 && !a.F[k]) || (a.dtH[k] != 0.0 && !a.LX[k]))
            return fail_i(CN_EINVAL, "cn_rk_rhs: missing stage vector");
    }
    a.stages = stages;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == CN_C128)
        hipLaunchKernelGGL((k_rk_rhs<zc>), grid_for(N), dim3(256), 0, s, static_cast<const zc*>(mx0),
                           static_cast<zc*>(rhs), a, (size_t)N);
    else
        hipLaunchKernelGGL((k_rk_rhs<double>), grid_for(N), dim3(256), 0, s, static_cast<const double*>(mx0),
                           static_cast<double*>(rhs), a, (size_t)N);
    return launched("k_rk_rhs");
}

int cn_pointwise(int dtype, int64_t N, void* out, const void* x, const void* g, const void* h, void* stream) {
    if (!out || !x) return fail_i(CN_EINVAL, "cn_pointwise: null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == CN_C128)
        hipLaunchKernelGGL((k_pointwise<zc>), grid_for(N), dim3(256), 0, s, static_cast<zc*>(out),
                           static_cast<const zc*>(x), static_cast<const zc*>(g), static_cast<const zc*>(h), (size_t)N);
    else
        hipLaunchKernelGGL((k_pointwise<double>), grid_for(N), dim3(256), 0, s, static_cast<double*>(out),
                           static_cast<const double*>(x), static_cast<const double*>(g),
                           static_cast<const double*>(h), (size_t)N);
    return launched("k_pointwise");
}

}  // extern "C"
