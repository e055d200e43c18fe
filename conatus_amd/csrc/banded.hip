// libcnhip: batched bordered-banded factor and solve for MI355X (gfx950), FP64 real and complex.
//
// Algorithm (the reference's, core/include/conatus/banded.hpp:1-8,253-341): after moving the r
// boundary rows to the top, P A = B + U V^T with B banded (kl, ku) whose border rows are s_i e_i,
// U = [e_0 .. e_{r-1}], V^T = border rows - diag(s).  B = P_B L U (partial pivoting inside the
// band, LAPACK gbtf2 order), W = B^-1 U, S = I + V^T W, and A x = b is solved as
// y = B^-1 P b, z = S^-1 V^T y, x = y - W z.
//
// MI355X design (not a translation of the reference's warp kernels):
//  * Storage is interleaved by 16 systems: entry e of system m lives at
//    ((m/16)*len + e)*16 + m%16, so the solve's lanes (consecutive systems) read 256 contiguous
//    bytes per factor entry and a 16-system chunk is produced by one workgroup.
//  * Factor: one group of G (16/32/64) lanes per system, lane c owns band column q = c (mod M),
//    M = kl+ku+1, and holds the kl+1 active rows of that column in registers (a window that
//    shifts by one row per step, so that every register index is static).  The
//    owner of the pivot column finds the pivot, forms the reciprocal and the multipliers, and
//    broadcasts them over the wavefront (ds_bpermute); every other lane swaps and updates its own
//    column.  r extra lanes carry the columns of U (unit vectors) through the same elimination,
//    so the forward half of W = B^-1 U costs nothing extra.  Row j of U leaves the window as one
//    coalesced row store; the incoming band row is one coalesced row load.  U's diagonal is
//    stored as its reciprocal so that solves multiply instead of divide.
//  * Border: one lane per (system, border column) finishes W by back substitution and
//    accumulates the column of S = I + V^T W; one lane per system LU-factors S (partial pivoting)
//    and forms cond_1(S) from the explicit inverse, as banded.hpp:403-432.
//  * Solve: one lane per (system, right-hand side); forward sweep (pivots + L) and backward sweep
//    (U, window of the last wu solutions in registers), V^T y accumulated during the backward
//    sweep, z = S^-1 V^T y, then x -= W z.  Gather/scatter maps and the split complex mode of
//    the reference's device solver are folded into the loads and stores.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cnhip.h"
#include "cn_common.h"

using cn::zc;

namespace {

constexpr int IL = 16;     // systems per interleave chunk
constexpr int MAXR = 8;    // largest border rank handled by the kernels

thread_local std::string g_err;

struct CnError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw CnError{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(CN_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

size_t checked_mul(std::initializer_list<size_t> dims) {
    size_t p = 1;
    for (size_t d : dims) {
        if (d != 0 && p > std::numeric_limits<size_t>::max() / d) fail(CN_EOVERFLOW, "array dimensions overflow size_t");
        p *= d;
    }
    return p;
}

// ------------------------------------------------------------------------------------------
// Factor kernel: G lanes per system, 256/G systems per workgroup of 256 lanes.
// Ab: unfactored band rows, row i holds A(i, i-KL+d), d = 0..M-1 (row stride M).
// Uf: U rows, row j holds U(j, j+d), d = 0..US-1, d = 0 holding 1/U(j, j).
// Lf: multipliers of step j, Lf(j, i) = l(j+1+i, j), i < KL.  piv: pivot offset 0..KL of step j.
// Wf: forward half of W (L^-1 P_B e_k), rows k < r of length n.
// ------------------------------------------------------------------------------------------
template <class T, int KL>
__global__ void __launch_bounds__(256)
k_band_lu(const T* __restrict__ Ab, T* __restrict__ Uf, T* __restrict__ Lf, int* __restrict__ piv,
          T* __restrict__ Wf, int* __restrict__ info, int B, int n, int M, int r, int G, int US,
          int dopivot) {
    const int t = threadIdx.x;
    const int g = t / G;
    const int c = t - g * G;
    const int m = blockIdx.x * (256 / G) + g;
    const bool live = m < B;
    const size_t ch = m / IL;
    const int ml = m % IL;
    const T* A0 = Ab + ch * ((size_t)n * M * IL) + ml;
    T* U0 = Uf + ch * ((size_t)n * US * IL) + ml;
    T* L0 = Lf + ch * ((size_t)n * KL * IL) + ml;
    int* P0 = piv + ch * ((size_t)n * IL) + ml;
    T* W0 = Wf + ch * ((size_t)r * n * IL) + ml;
    const bool band = live && c < M;
    const bool wl = live && c >= M && c < M + r;
    const int k = c - M;
    const int gbase = (t & 63) - c;   // wavefront lane of the group's lane 0 (G divides 64)

    T win[KL + 1];
#pragma unroll
    for (int i = 0; i <= KL; ++i) {
        T v = cn::zero<T>();
        if (band && i < n) {
            const int d = c - i + KL;
            if (d >= 0 && d < M) v = A0[((size_t)i * M + d) * IL];
        }
        if (wl) v = (i == k) ? cn::one<T>() : cn::zero<T>();
        win[i] = v;
    }
    int first_singular = 0;
    for (int j = 0; j < n; ++j) {
        // win[i] holds row j+i of this lane's column
        const int owner = j % M;
        const int src = gbase + owner;
        const bool is_owner = (c == owner);
        // pivot search on the owner's column (first maximum of |re|+|im|, as i?amax)
        double best = cn::cabs1(win[0]);
        int ip = 0;
        if (dopivot) {
#pragma unroll
            for (int i = 1; i <= KL; ++i) {
                const double v = cn::cabs1(win[i]);
                if (v > best) { best = v; ip = i; }
            }
        }
        ip = __shfl(ip, src, 64);
        best = __shfl(best, src, 64);
        if (ip != 0) cn::swap_first<T, KL>(win, ip);   // row interchange in every column of the window and in W
        const bool singular = !(best != 0.0);
        T rc = cn::recip(win[0]);
        if (singular) rc = cn::zero<T>();
        T l[KL];
#pragma unroll
        for (int i = 1; i <= KL; ++i) l[i - 1] = cn::shfl(cn::mul(win[i], rc), src);
        const int d = (c - owner + M) % M;   // column offset of this lane's column from j
        if (band) {
            T ov = win[0];
            if (is_owner) ov = rc;
            U0[((size_t)j * US + d) * IL] = ov;
        }
        if (wl) W0[((size_t)k * n + j) * IL] = win[0];
        if (band && is_owner) {
#pragma unroll
            for (int i = 0; i < KL; ++i) L0[((size_t)j * KL + i) * IL] = l[i];
            P0[(size_t)j * IL] = ip;
        }
        if ((band && !is_owner) || wl) {
#pragma unroll
            for (int i = 1; i <= KL; ++i) win[i] = cn::fnms(l[i - 1], win[0], win[i]);
        }
        // row j leaves the window; row j+KL+1 enters at the bottom
        const int jn = j + KL + 1;
        T v = cn::zero<T>();
        if (band) {
            if (is_owner) {   // the owner takes column j+M: rows j+1..j+KL are zero there
#pragma unroll
                for (int i = 1; i <= KL; ++i) win[i] = cn::zero<T>();
            }
            const int dA = is_owner ? (M - 1) : (d - 1);
            if (jn < n) v = A0[((size_t)jn * M + dA) * IL];
        }
        if (wl && jn == k) v = cn::one<T>();
#pragma unroll
        for (int i = 0; i < KL; ++i) win[i] = win[i + 1];
        win[KL] = v;
        if (singular && first_singular == 0) first_singular = j + 1;
    }
    if (live && c == 0) info[m] = first_singular;
}

__device__ __forceinline__ double absv(double a) { return fabs(a); }
__device__ __forceinline__ double absv(zc a) { return hypot(a.re, a.im); }
__device__ __forceinline__ zc neg(zc a) { return zc{-a.re, -a.im}; }
__device__ __forceinline__ double neg(double a) { return -a; }

// One lane per system: LU of S in place (column-major, partial pivoting by |re|+|im|), and
// cond_1(S) = |S|_1 |S^-1|_1 with the inverse from the factors (banded.hpp:403-432).
template <class T>
__global__ void __launch_bounds__(256)
k_cap_lu(T* __restrict__ S, int* __restrict__ spiv, double* __restrict__ cond, int* __restrict__ info,
         int B, int r) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= B) return;
    T a[MAXR * MAXR];
    T* Sm = S + (size_t)m * r * r;
    for (int q = 0; q < r * r; ++q) a[q] = Sm[q];
    double nS = 0.0;
    for (int b = 0; b < r; ++b) {
        double s = 0.0;
        for (int i = 0; i < r; ++i) s += absv(a[b * r + i]);
        if (!(s <= nS)) nS = s;
    }
    int bad = 0;
    int p[MAXR];
    for (int jc = 0; jc < r; ++jc) {
        int ip = jc;
        double best = cn::cabs1(a[jc * r + jc]);
        for (int i = jc + 1; i < r; ++i) {
            const double v = cn::cabs1(a[jc * r + i]);
            if (v > best) { best = v; ip = i; }
        }
        p[jc] = ip;
        if (best == 0.0) { bad = 1; continue; }
        if (ip != jc)
            for (int q = 0; q < r; ++q) { T tmp = a[q * r + jc]; a[q * r + jc] = a[q * r + ip]; a[q * r + ip] = tmp; }
        const T rc = cn::recip(a[jc * r + jc]);
        for (int i = jc + 1; i < r; ++i) a[jc * r + i] = cn::mul(a[jc * r + i], rc);
        for (int q = jc + 1; q < r; ++q)
            for (int i = jc + 1; i < r; ++i) a[q * r + i] = cn::fnms(a[jc * r + i], a[q * r + jc], a[q * r + i]);
    }
    for (int q = 0; q < r * r; ++q) Sm[q] = a[q];
    for (int q = 0; q < r; ++q) spiv[(size_t)m * r + q] = p[q];
    double c = 1.0;
    if (bad) {
        if (info[m] == 0) info[m] = -1;
        c = INFINITY;
    } else {
        // inverse column by column: solve S x = e_b with the factors
        double ninv = 0.0;
        for (int b = 0; b < r; ++b) {
            T x[MAXR];
            for (int i = 0; i < r; ++i) x[i] = (i == b) ? cn::one<T>() : cn::zero<T>();
            for (int i = 0; i < r; ++i)
                if (p[i] != i) { T tmp = x[i]; x[i] = x[p[i]]; x[p[i]] = tmp; }
            for (int i = 0; i < r; ++i)
                for (int q = 0; q < i; ++q) x[i] = cn::fnms(a[q * r + i], x[q], x[i]);
            for (int i = r - 1; i >= 0; --i) {
                for (int q = i + 1; q < r; ++q) x[i] = cn::fnms(a[q * r + i], x[q], x[i]);
                x[i] = cn::mul(x[i], cn::recip(a[i * r + i]));
            }
            double s = 0.0;
            for (int i = 0; i < r; ++i) s += absv(x[i]);
            if (!(s <= ninv)) ninv = s;
        }
        c = nS * ninv;
    }
    cond[m] = c;
}

// ------------------------------------------------------------------------------------------
// Solve kernel.  IO modes: 0 plain in place (x holds b), 1 mapped (rows/cols), 2 split complex.
// ------------------------------------------------------------------------------------------
struct SolveIO {
    int mode;
    void* x;               // mode 0: in place blocks; mode 1/2: output
    const void* rhs;       // mode 1: rhs; mode 2: re
    const int64_t* rows;   // mode 1: rows (or null); mode 2: rows_c
    const int64_t* rows2;  // mode 2: rows_s
    const int64_t* cols;   // mode 1: cols (or null); mode 2: cols_c
    const int64_t* cols2;  // mode 2: cols_s
    int64_t rhs_stride;    // mode 0
};

template <class T>
__device__ __forceinline__ T io_load_b(const SolveIO& io, size_t mn, int k, int i);
template <class T>
__device__ __forceinline__ void io_store(const SolveIO& io, size_t mn, int k, int i, T v);
template <class T>
__device__ __forceinline__ T io_load_x(const SolveIO& io, size_t mn, int k, int i);

template <>
__device__ __forceinline__ double io_load_b<double>(const SolveIO& io, size_t mn, int k, int i) {
    if (io.mode == 0) return static_cast<const double*>(io.x)[k * io.rhs_stride + mn + i];
    const size_t q = io.rows ? (size_t)io.rows[mn + i] : mn + i;
    return static_cast<const double*>(io.rhs)[q];
}
template <>
__device__ __forceinline__ zc io_load_b<zc>(const SolveIO& io, size_t mn, int k, int i) {
    if (io.mode == 0) return static_cast<const zc*>(io.x)[k * io.rhs_stride + mn + i];
    if (io.mode == 1) {
        const size_t q = io.rows ? (size_t)io.rows[mn + i] : mn + i;
        return static_cast<const zc*>(io.rhs)[q];
    }
    const double* re = static_cast<const double*>(io.rhs);
    return zc{re[io.rows[mn + i]], re[io.rows2[mn + i]]};
}
template <>
__device__ __forceinline__ double io_load_x<double>(const SolveIO& io, size_t mn, int k, int i) {
    if (io.mode == 0) return static_cast<const double*>(io.x)[k * io.rhs_stride + mn + i];
    const size_t q = io.cols ? (size_t)io.cols[mn + i] : mn + i;
    return static_cast<const double*>(io.x)[q];
}
template <>
__device__ __forceinline__ zc io_load_x<zc>(const SolveIO& io, size_t mn, int k, int i) {
    if (io.mode == 0) return static_cast<const zc*>(io.x)[k * io.rhs_stride + mn + i];
    if (io.mode == 1) {
        const size_t q = io.cols ? (size_t)io.cols[mn + i] : mn + i;
        return static_cast<const zc*>(io.x)[q];
    }
    const double* x = static_cast<const double*>(io.x);
    return zc{x[io.cols[mn + i]], x[io.cols2[mn + i]]};
}
template <>
__device__ __forceinline__ void io_store<double>(const SolveIO& io, size_t mn, int k, int i, double v) {
    if (io.mode == 0) { static_cast<double*>(io.x)[k * io.rhs_stride + mn + i] = v; return; }
    const size_t q = io.cols ? (size_t)io.cols[mn + i] : mn + i;
    static_cast<double*>(io.x)[q] = v;
}
template <>
__device__ __forceinline__ void io_store<zc>(const SolveIO& io, size_t mn, int k, int i, zc v) {
    if (io.mode == 0) { static_cast<zc*>(io.x)[k * io.rhs_stride + mn + i] = v; return; }
    if (io.mode == 1) {
        const size_t q = io.cols ? (size_t)io.cols[mn + i] : mn + i;
        static_cast<zc*>(io.x)[q] = v;
        return;
    }
    double* x = static_cast<double*>(io.x);
    x[io.cols[mn + i]] = v.re;
    x[io.cols2[mn + i]] = v.im;
}

// z = S^-1 t in registers (S LU-factored column-major with pivots sp; r <= MAXR)
template <class T>
__device__ __forceinline__ void cap_solve(T (&t)[MAXR], const T* __restrict__ Sm, const int* __restrict__ sp, int r) {
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
        if (i >= r) break;
        const int pi = sp[i];
#pragma unroll
        for (int q = i + 1; q < MAXR; ++q) {
            const bool c = (q == pi);
            const T ti = t[i], tq = t[q];
            t[i] = cn::sel(c, tq, ti);
            t[q] = cn::sel(c, ti, tq);
        }
    }
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
        if (i >= r) break;
#pragma unroll
        for (int q = 0; q < i; ++q) t[i] = cn::fnms(Sm[q * r + i], t[q], t[i]);
    }
#pragma unroll
    for (int i = MAXR - 1; i >= 0; --i) {
        if (i < r) {
#pragma unroll
            for (int q = i + 1; q < MAXR; ++q)
                if (q < r) t[i] = cn::fnms(Sm[q * r + i], t[q], t[i]);
            t[i] = cn::mul(t[i], cn::recip(Sm[i * r + i]));
        }
    }
}

// Direct solve (any shape and IO mode).  Woodbury without a correction pass:
//   c = P b; t = Z^T c (Z^T = V^T B^-1, accumulated as c streams in); y_c = L^-1 P_B c;
//   z = S^-1 t; x = U^-1 (y_c - Wf z) with Wf = L^-1 P_B U (the factor's forward half of W).
// Mathematically x = B^-1 (c - U z) = (P A)^-1 c.
template <class T, int KL, int WU>
__global__ void __launch_bounds__(256)
k_solve(const T* __restrict__ Uf, const T* __restrict__ Lf, const int* __restrict__ piv,
        const T* __restrict__ Zt, const T* __restrict__ Wf, const T* __restrict__ S,
        const int* __restrict__ spiv, const int* __restrict__ perm, T* __restrict__ ws,
        SolveIO io, int B, int n, int r, int US, int nrhs) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = gid / nrhs;
    const int k = gid - m * nrhs;
    if (m >= B) return;
    const size_t NL = (size_t)B * nrhs;
    const size_t ch = m / IL;
    const int ml = m % IL;
    const T* U0 = Uf + ch * ((size_t)n * US * IL) + ml;
    const T* L0 = Lf + ch * ((size_t)n * KL * IL) + ml;
    const int* P0 = piv + ch * ((size_t)n * IL) + ml;
    const T* Z0 = Zt + ch * ((size_t)r * n * IL) + ml;
    const T* W0 = Wf + ch * ((size_t)r * n * IL) + ml;
    const size_t mn = (size_t)m * n;
    T* w0 = ws + gid;

    T t[MAXR];
#pragma unroll
    for (int a = 0; a < MAXR; ++a) t[a] = cn::zero<T>();
    // forward: y = L^-1 P_B c, window of KL+1 rows; t += Z^T c as c enters
    T y[KL + 1];
#pragma unroll
    for (int i = 0; i <= KL; ++i) {
        y[i] = cn::zero<T>();
        if (i < n) {
            y[i] = io_load_b<T>(io, mn, k, perm[i]);
#pragma unroll
            for (int a = 0; a < MAXR; ++a)
                if (a < r) t[a] = cn::fmac(Z0[((size_t)a * n + i) * IL], y[i], t[a]);
        }
    }
    for (int j = 0; j < n; ++j) {
        const int p = P0[(size_t)j * IL];
        if (p != 0) cn::swap_first<T, KL>(y, p);
        const T* Lr = L0 + (size_t)j * KL * IL;
#pragma unroll
        for (int i = 1; i <= KL; ++i) y[i] = cn::fnms(Lr[(i - 1) * IL], y[0], y[i]);
        w0[(size_t)j * NL] = y[0];
        const int jn = j + KL + 1;
#pragma unroll
        for (int i = 0; i < KL; ++i) y[i] = y[i + 1];
        T c = cn::zero<T>();
        if (jn < n) {
            c = io_load_b<T>(io, mn, k, perm[jn]);
#pragma unroll
            for (int a = 0; a < MAXR; ++a)
                if (a < r) t[a] = cn::fmac(Z0[((size_t)a * n + jn) * IL], c, t[a]);
        }
        y[KL] = c;
    }
    if (r > 0) cap_solve<T>(t, S + (size_t)m * r * r, spiv + (size_t)m * r, r);
    // backward: x = U^-1 (y_c - Wf z), window of the last WU solutions
    T xw[WU];
#pragma unroll
    for (int d = 0; d < WU; ++d) xw[d] = cn::zero<T>();
    for (int j = n - 1; j >= 0; --j) {
        const T* Ur = U0 + (size_t)j * US * IL;
        T acc = w0[(size_t)j * NL];
#pragma unroll
        for (int a = 0; a < MAXR; ++a)
            if (a < r) acc = cn::fnms(W0[((size_t)a * n + j) * IL], t[a], acc);
#pragma unroll
        for (int d = 1; d <= WU; ++d) acc = cn::fnms(Ur[d * IL], xw[d - 1], acc);
        const T x = cn::mul(acc, Ur[0]);
        io_store<T>(io, mn, k, j, x);
#pragma unroll
        for (int d = WU - 1; d > 0; --d) xw[d] = xw[d - 1];
        xw[0] = x;
    }
}

// ------------------------------------------------------------------------------------------
// Streamed solve (the hot kernel for many narrow-band systems): one wavefront per workgroup,
// 64 / NR systems x NR right-hand sides (NR in {1, 2, 4}); the per-system factor streams
// (L rows, pivots, Z^T for the forward sweep; U rows, Wf for the backward sweep) are copied in
// blocks of SP steps from HBM straight into LDS with global_load_lds_dwordx4, double-buffered:
// the copy of block b+1 is in flight while block b is computed from LDS.  The per-lane
// right-hand side and forward results are prefetched one block ahead in registers.  Arithmetic
// identical to k_solve.  IO: in place (mode 0) or contiguous out-of-place (mode 1, no maps).
// ------------------------------------------------------------------------------------------
constexpr int SP = 8;   // steps per streamed block

__device__ __forceinline__ void dma16(const char* __restrict__ src, char* lds_dst, int nbytes, int lane) {
    for (int off = 0; off < nbytes; off += 1024) {
        const int my = off + lane * 16;
        if (my < nbytes)
            __builtin_amdgcn_global_load_lds(static_cast<const void*>(src + my),
                                             (__attribute__((address_space(3))) void*)(lds_dst + off), 16, 0, 0);
    }
}

__device__ __forceinline__ void dma_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

template <class T>
__device__ __forceinline__ T lds_at(const char* p) { return *reinterpret_cast<const T*>(p); }

// forward block blk of the streamed solve into LDS buffer buf: per chunk [L: SP x KL][Z: r x SP][piv]
template <class T, int KL, int C, int CE>
__device__ __forceinline__ void solve_fwd_dma(const T* __restrict__ Lf, const T* __restrict__ Zt,
                                              const int* __restrict__ piv, char* lds, size_t ch0, int n, int r,
                                              int blk, int buf, int bufbytes, int fchunk, int fL, int fZ, int lane) {
    const int j0 = blk * SP;
    const int cnt = min(SP, n - j0);
    const int jz0 = j0 + KL + 1;
    const int cz = max(0, min(SP, n - jz0));
    char* base = lds + buf * bufbytes;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const size_t ch = ch0 + c;
        char* cb = base + c * fchunk;
        dma16(reinterpret_cast<const char*>(Lf + (ch * n * KL + (size_t)j0 * KL) * IL), cb, cnt * KL * CE, lane);
        for (int a = 0; a < r && cz > 0; ++a)
            dma16(reinterpret_cast<const char*>(Zt + (ch * r * n + (size_t)a * n + jz0) * IL), cb + fL + a * SP * CE,
                  cz * CE, lane);
        dma16(reinterpret_cast<const char*>(piv + (ch * n + j0) * IL), cb + fL + fZ, cnt * IL * 4, lane);
    }
}

// backward block blk (rows n-1-blk*SP downwards) into LDS buffer buf: per chunk [U: SP x US][W: r x SP]
template <class T, int C, int CE>
__device__ __forceinline__ void solve_bwd_dma(const T* __restrict__ Uf, const T* __restrict__ Wf, char* lds,
                                              size_t ch0, int n, int r, int US, int blk, int buf, int bufbytes,
                                              int bchunk, int bU, int lane) {
    const int jhi = n - 1 - blk * SP;
    const int jlo = max(0, jhi - SP + 1);
    const int cnt = jhi - jlo + 1;
    char* base = lds + buf * bufbytes;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const size_t ch = ch0 + c;
        char* cb = base + c * bchunk;
        dma16(reinterpret_cast<const char*>(Uf + (ch * n * US + (size_t)jlo * US) * IL), cb, cnt * US * CE, lane);
        for (int a = 0; a < r; ++a)
            dma16(reinterpret_cast<const char*>(Wf + (ch * r * n + (size_t)a * n + jlo) * IL), cb + bU + a * SP * CE,
                  cnt * CE, lane);
    }
}

template <class T, int KL, int WU, int NR>
__global__ void __launch_bounds__(64)
k_solve_s(const T* __restrict__ Uf, const T* __restrict__ Lf, const int* __restrict__ piv,
          const T* __restrict__ Zt, const T* __restrict__ Wf, const T* __restrict__ S,
          const int* __restrict__ spiv, const int* __restrict__ perm, T* __restrict__ ws,
          const T* __restrict__ bin, T* __restrict__ xout, int64_t rhs_stride, int B, int n, int r, int US,
          int bufbytes) {
    static_assert(KL + 1 <= SP, "the window lookahead must stay within the next block");
    constexpr int NS = 64 / NR;                 // systems per workgroup
    constexpr int C = NS / IL;                  // interleave chunks per workgroup
    constexpr int CE = IL * (int)sizeof(T);     // bytes of one entry across a chunk
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x;
    const int mloc = lane / NR, k = lane - mloc * NR;
    const int m = blockIdx.x * NS + mloc;
    const bool live = m < B;
    const int cl = mloc / IL, ml = mloc % IL;
    const size_t ch0 = (size_t)blockIdx.x * C;
    const size_t NL = (size_t)B * NR;
    const size_t gid = (size_t)m * NR + k;
    const T* bl = bin + k * rhs_stride + (size_t)m * n;
    T* xl = xout + k * rhs_stride + (size_t)m * n;
    T* wl = ws + gid;
    // forward layout of one buffer: per chunk [L: SP x KL][Z: r x SP][piv: SP x 16 ints]
    const int fL = SP * KL * CE, fZ = r * SP * CE, fP = SP * IL * 4;
    const int fchunk = fL + fZ + fP;
    // backward layout: per chunk [U: SP x US][W: r x SP]
    const int bU = SP * US * CE;
    const int bchunk = bU + r * SP * CE;


    T t[MAXR];
#pragma unroll
    for (int a = 0; a < MAXR; ++a) t[a] = cn::zero<T>();
    const T* Z0 = Zt + ((size_t)(m / IL) * r * n) * IL + ml;
    T y[KL + 1];
#pragma unroll
    for (int i = 0; i <= KL; ++i) {
        y[i] = cn::zero<T>();
        if (live && i < n) {
            y[i] = bl[perm[i]];
#pragma unroll
            for (int a = 0; a < MAXR; ++a)
                if (a < r) t[a] = cn::fmac(Z0[((size_t)a * n + i) * IL], y[i], t[a]);
        }
    }
    const int nblk = (n + SP - 1) / SP;
    // per-lane incoming values of block b: c_{b SP + KL + 1 + s}
    T cur[SP], nxt[SP];
#pragma unroll
    for (int s = 0; s < SP; ++s) {
        const int jn = KL + 1 + s;
        cur[s] = (live && jn < n) ? bl[perm[jn]] : cn::zero<T>();
    }
    solve_fwd_dma<T, KL, C, CE>(Lf, Zt, piv, lds, ch0, n, r, 0, 0, bufbytes, fchunk, fL, fZ, lane);
    dma_wait();
    for (int b = 0; b < nblk; ++b) {
        const int buf = b & 1;
        if (b + 1 < nblk) {
            solve_fwd_dma<T, KL, C, CE>(Lf, Zt, piv, lds, ch0, n, r, b + 1, buf ^ 1, bufbytes, fchunk, fL, fZ, lane);
#pragma unroll
            for (int s = 0; s < SP; ++s) {
                const int jn = (b + 1) * SP + KL + 1 + s;
                nxt[s] = (live && jn < n) ? bl[perm[jn]] : cn::zero<T>();
            }
        }
        const char* fb = lds + buf * bufbytes + cl * fchunk;
#pragma unroll
        for (int s = 0; s < SP; ++s) {
            const int j = b * SP + s;
            if (j >= n) break;
            const int p = *reinterpret_cast<const int*>(fb + fL + fZ + (s * IL + ml) * 4);
            if (p != 0) cn::swap_first<T, KL>(y, p);
#pragma unroll
            for (int i = 1; i <= KL; ++i)
                y[i] = cn::fnms(lds_at<T>(fb + ((s * KL + i - 1) * IL + ml) * sizeof(T)), y[0], y[i]);
            if (live) wl[(size_t)j * NL] = y[0];
            const int jn = j + KL + 1;
#pragma unroll
            for (int i = 0; i < KL; ++i) y[i] = y[i + 1];
            const T c = cur[s];
            y[KL] = c;
            if (jn < n) {
#pragma unroll
                for (int a = 0; a < MAXR; ++a)
                    if (a < r) t[a] = cn::fmac(lds_at<T>(fb + fL + ((a * SP + s) * IL + ml) * sizeof(T)), c, t[a]);
            }
        }
#pragma unroll
        for (int s = 0; s < SP; ++s) cur[s] = nxt[s];
        dma_wait();
    }
    if (r > 0 && live) cap_solve<T>(t, S + (size_t)m * r * r, spiv + (size_t)m * r, r);
    // backward sweep, blocks of SP rows from the bottom
    T xw[WU];
#pragma unroll
    for (int d = 0; d < WU; ++d) xw[d] = cn::zero<T>();
    T wc[SP], wn[SP];
#pragma unroll
    for (int s = 0; s < SP; ++s) {
        const int j = n - 1 - s;
        wc[s] = (live && j >= 0) ? wl[(size_t)j * NL] : cn::zero<T>();
    }
    solve_bwd_dma<T, C, CE>(Uf, Wf, lds, ch0, n, r, US, 0, 0, bufbytes, bchunk, bU, lane);
    dma_wait();
    for (int b = 0; b < nblk; ++b) {
        const int buf = b & 1;
        const int jhi = n - 1 - b * SP;
        const int jlo = max(0, jhi - SP + 1);
        if (b + 1 < nblk) {
            solve_bwd_dma<T, C, CE>(Uf, Wf, lds, ch0, n, r, US, b + 1, buf ^ 1, bufbytes, bchunk, bU, lane);
#pragma unroll
            for (int s = 0; s < SP; ++s) {
                const int j = jhi - SP - s;
                wn[s] = (live && j >= 0) ? wl[(size_t)j * NL] : cn::zero<T>();
            }
        }
        const char* bb = lds + buf * bufbytes + cl * bchunk;
#pragma unroll
        for (int s = 0; s < SP; ++s) {
            const int j = jhi - s;
            if (j < jlo) break;
            const int row = j - jlo;
            T acc = wc[s];
#pragma unroll
            for (int a = 0; a < MAXR; ++a)
                if (a < r) acc = cn::fnms(lds_at<T>(bb + bU + ((a * SP + row) * IL + ml) * sizeof(T)), t[a], acc);
            const char* ur = bb + (size_t)row * US * CE + ml * sizeof(T);
#pragma unroll
            for (int d = 1; d <= WU; ++d) acc = cn::fnms(lds_at<T>(ur + d * CE), xw[d - 1], acc);
            const T x = cn::mul(acc, lds_at<T>(ur));
            if (live) xl[j] = x;
#pragma unroll
            for (int d = WU - 1; d > 0; --d) xw[d] = xw[d - 1];
            xw[0] = x;
        }
#pragma unroll
        for (int s = 0; s < SP; ++s) wc[s] = wn[s];
        dma_wait();
    }
}

// ------------------------------------------------------------------------------------------
// Transposed solve A^T x = b (plain transpose), the discrete adjoint path (banded.hpp:346-372).
// A^T = (B^T + V U^T) P, so y = P x solves (B^T + V U^T) y = b:
//   y0 = B^-T b = P_B L^-T U^-T b,  z = St^-1 y0[:r],  y = y0 - Wt z,  x[perm[i]] = y[i],
// with Wt = B^-T V and St = I + U^T Wt built once (prepare).  Lane = (system, rhs).
//   phase 1: U^T w = b (forward; column j of U is read as U(j-d, j) = Uf[j-d][d]);
//   phase 2: L^T and the row interchanges, backward (gbtrs 'T' order: dot, then swap);
//   phase 3: Woodbury correction and the scatter through perm.
// prep = 1 computes Wt instead: lane = (system, border index a), input row a of V^T, output and
// work space row a of Wt, no phase 3.
// ------------------------------------------------------------------------------------------
template <class T, int KL, int WU>
__global__ void __launch_bounds__(256)
k_solve_t(const T* __restrict__ Uf, const T* __restrict__ Lf, const int* __restrict__ piv,
          const T* __restrict__ Vt, T* __restrict__ Wt, const T* __restrict__ St,
          const int* __restrict__ stpiv, const int* __restrict__ perm, T* __restrict__ ws,
          T* __restrict__ x, int64_t rhs_stride, int B, int n, int r, int US, int nrhs, int prep) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = gid / nrhs;
    const int k = gid - m * nrhs;
    if (m >= B) return;
    const size_t ch = m / IL;
    const int ml = m % IL;
    const T* U0 = Uf + ch * ((size_t)n * US * IL) + ml;
    const T* L0 = Lf + ch * ((size_t)n * KL * IL) + ml;
    const int* P0 = piv + ch * ((size_t)n * IL) + ml;
    // input row, work row (stride between consecutive entries)
    const T* in;
    size_t in_st;
    T* w;
    size_t w_st;
    if (prep) {
        in = Vt + ch * ((size_t)r * n * IL) + ml + (size_t)k * n * IL;
        in_st = IL;
        w = Wt + ch * ((size_t)r * n * IL) + ml + (size_t)k * n * IL;
        w_st = IL;
    } else {
        in = x + k * rhs_stride + (size_t)m * n;
        in_st = 1;
        w = ws + gid;
        w_st = (size_t)B * nrhs;
    }
    // phase 1: U^T w = b
    T uw[WU];
#pragma unroll
    for (int d = 0; d < WU; ++d) uw[d] = cn::zero<T>();   // uw[d-1] = w_{j-d}
    for (int j = 0; j < n; ++j) {
        T acc = in[(size_t)j * in_st];
#pragma unroll
        for (int d = 1; d <= WU; ++d)
            if (j - d >= 0) acc = cn::fnms(U0[((size_t)(j - d) * US + d) * IL], uw[d - 1], acc);
        const T wj = cn::mul(acc, U0[(size_t)j * US * IL]);
        w[(size_t)j * w_st] = wj;
#pragma unroll
        for (int d = WU - 1; d > 0; --d) uw[d] = uw[d - 1];
        uw[0] = wj;
    }
    // phase 2: L^T with interchanges, j = n-1 .. 0; window y[i] = entry j+i
    T y[KL + 1];
#pragma unroll
    for (int i = 0; i <= KL; ++i) y[i] = cn::zero<T>();
    for (int j = n - 1; j >= 0; --j) {
        // entry j+1+KL leaves the window: no later step touches it, it is final
        if (j + 1 + KL <= n - 1) w[(size_t)(j + 1 + KL) * w_st] = y[KL];
#pragma unroll
        for (int i = KL; i > 0; --i) y[i] = y[i - 1];
        T v = w[(size_t)j * w_st];
        const T* Lr = L0 + (size_t)j * KL * IL;
#pragma unroll
        for (int i = 1; i <= KL; ++i) v = cn::fnms(Lr[(i - 1) * IL], y[i], v);
        y[0] = v;
        const int p = P0[(size_t)j * IL];
        if (p != 0) cn::swap_first<T, KL>(y, p);
    }
#pragma unroll
    for (int i = 0; i <= KL; ++i)
        if (i < n) w[(size_t)i * w_st] = y[i];
    if (prep) return;
    // phase 3: z = St^-1 y0[:r]; y = y0 - Wt z; x[perm[i]] = y[i]
    T z[MAXR];
#pragma unroll
    for (int a = 0; a < MAXR; ++a) z[a] = (a < r) ? w[(size_t)a * w_st] : cn::zero<T>();
    if (r > 0) {
        const T* Sm = St + (size_t)m * r * r;
        const int* sp = stpiv + (size_t)m * r;
#pragma unroll
        for (int i = 0; i < MAXR; ++i) {
            if (i >= r) break;
            const int pi = sp[i];
#pragma unroll
            for (int q = i + 1; q < MAXR; ++q) {
                const bool c = (q == pi);
                const T zi = z[i], zq = z[q];
                z[i] = cn::sel(c, zq, zi);
                z[q] = cn::sel(c, zi, zq);
            }
        }
#pragma unroll
        for (int i = 0; i < MAXR; ++i) {
            if (i >= r) break;
#pragma unroll
            for (int q = 0; q < i; ++q) z[i] = cn::fnms(Sm[q * r + i], z[q], z[i]);
        }
#pragma unroll
        for (int i = MAXR - 1; i >= 0; --i) {
            if (i < r) {
#pragma unroll
                for (int q = i + 1; q < MAXR; ++q)
                    if (q < r) z[i] = cn::fnms(Sm[q * r + i], z[q], z[i]);
                z[i] = cn::mul(z[i], cn::recip(Sm[i * r + i]));
            }
        }
    }
    const T* W0 = Wt + ch * ((size_t)r * n * IL) + ml;
    T* xo = x + k * rhs_stride + (size_t)m * n;
    for (int i = 0; i < n; ++i) {
        T v = w[(size_t)i * w_st];
#pragma unroll
        for (int a = 0; a < MAXR; ++a)
            if (a < r) v = cn::fnms(W0[((size_t)a * n + i) * IL], z[a], v);
        xo[perm[i]] = v;
    }
}

// S = I + V^T B^-1 U (column-major r x r per system) from Z^T = V^T B^-1 (rows of Wt):
// S(a, b) = delta_ab + Z^T[a][b].
template <class T>
__global__ void __launch_bounds__(256)
k_build_s(const T* __restrict__ Zt, T* __restrict__ S, int B, int n, int r) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = gid / (r * r);
    const int q = gid - m * r * r;
    if (m >= B) return;
    const int b = q / r, a = q - b * r;
    const size_t ch = m / IL;
    const int ml = m % IL;
    T v = Zt[(ch * ((size_t)r * n) + (size_t)a * n + b) * IL + ml];
    if (a == b) v = cn::fmac(cn::one<T>(), cn::one<T>(), v);
    S[(size_t)m * r * r + (size_t)b * r + a] = v;
}

// St = I + U^T Wt (column-major r x r per system): St(a, b) = delta_ab + Wt[b][a].
template <class T>
__global__ void __launch_bounds__(256)
k_build_st(const T* __restrict__ Wt, T* __restrict__ St, int* __restrict__ info_t, int B, int n, int r) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = gid / (r * r);
    const int q = gid - m * r * r;
    if (m >= B) return;
    if (q == 0) info_t[m] = 0;
    const int b = q / r, a = q - b * r;
    const size_t ch = m / IL;
    const int ml = m % IL;
    T v = Wt[(ch * ((size_t)r * n) + (size_t)b * n + a) * IL + ml];
    if (a == b) v = cn::fmac(cn::one<T>(), cn::one<T>(), v);
    St[(size_t)m * r * r + (size_t)b * r + a] = v;
}

// ------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------
constexpr int KL_SET[] = {1, 2, 4, 6, 8, 10, 12, 14, 16};
constexpr int WU_SET[] = {4, 8, 16, 24, 32};

int pick_kl(int kl) {
    for (int v : KL_SET) if (v >= kl) return v;
    return -1;
}
int pick_wu(int w) {
    for (int v : WU_SET) if (v >= w) return v;
    return -1;
}

using FactorFn = void (*)(dim3, dim3, hipStream_t, const void*, void*, void*, int*, void*, int*, int, int,
                          int, int, int, int, int);
using SolveFn = void (*)(dim3, dim3, hipStream_t, const void*, const void*, const int*, const void*, const void*,
                         const void*, const int*, const int*, void*, SolveIO, int, int, int, int, int);

template <class T, int KL>
void launch_factor(dim3 g, dim3 b, hipStream_t s, const void* Ab, void* Uf, void* Lf, int* piv, void* Wf,
                   int* info, int B, int n, int M, int r, int G, int US, int pivot) {
    hipLaunchKernelGGL((k_band_lu<T, KL>), g, b, 0, s, static_cast<const T*>(Ab), static_cast<T*>(Uf),
                       static_cast<T*>(Lf), piv, static_cast<T*>(Wf), info, B, n, M, r, G, US, pivot);
}
template <class T, int KL, int WU>
void launch_solve(dim3 g, dim3 b, hipStream_t s, const void* Uf, const void* Lf, const int* piv, const void* Vt,
                  const void* W, const void* S, const int* spiv, const int* perm, void* ws, SolveIO io, int B,
                  int n, int r, int US, int nrhs) {
    hipLaunchKernelGGL((k_solve<T, KL, WU>), g, b, 0, s, static_cast<const T*>(Uf), static_cast<const T*>(Lf),
                       piv, static_cast<const T*>(Vt), static_cast<const T*>(W), static_cast<const T*>(S), spiv,
                       perm, static_cast<T*>(ws), io, B, n, r, US, nrhs);
}

using SolveTFn = void (*)(dim3, dim3, hipStream_t, const void*, const void*, const int*, const void*, void*,
                          const void*, const int*, const int*, void*, void*, int64_t, int, int, int, int, int, int);

template <class T, int KL, int WU>
void launch_solve_t(dim3 g, dim3 b, hipStream_t s, const void* Uf, const void* Lf, const int* piv, const void* Vt,
                    void* Wt, const void* St, const int* stpiv, const int* perm, void* ws, void* x, int64_t rs, int B,
                    int n, int r, int US, int nrhs, int prep) {
    hipLaunchKernelGGL((k_solve_t<T, KL, WU>), g, b, 0, s, static_cast<const T*>(Uf), static_cast<const T*>(Lf),
                       piv, static_cast<const T*>(Vt), static_cast<T*>(Wt), static_cast<const T*>(St), stpiv, perm,
                       static_cast<T*>(ws), static_cast<T*>(x), rs, B, n, r, US, nrhs, prep);
}

template <class T, int KL>
SolveTFn solve_t_fn_kl(int WU) {
    if (WU < KL) return nullptr;
    switch (WU) {
        case 4: if constexpr (KL <= 4) return &launch_solve_t<T, KL, 4>; else return nullptr;
        case 8: if constexpr (KL <= 8) return &launch_solve_t<T, KL, 8>; else return nullptr;
        case 16: return &launch_solve_t<T, KL, 16>;
        case 24: return &launch_solve_t<T, KL, 24>;
        case 32: return &launch_solve_t<T, KL, 32>;
    }
    return nullptr;
}
template <class T>
SolveTFn solve_t_fn(int KL, int WU) {
    switch (KL) {
        case 1: return solve_t_fn_kl<T, 1>(WU);
        case 2: return solve_t_fn_kl<T, 2>(WU);
        case 4: return solve_t_fn_kl<T, 4>(WU);
        case 6: return solve_t_fn_kl<T, 6>(WU);
        case 8: return solve_t_fn_kl<T, 8>(WU);
        case 10: return solve_t_fn_kl<T, 10>(WU);
        case 12: return solve_t_fn_kl<T, 12>(WU);
        case 14: return solve_t_fn_kl<T, 14>(WU);
        case 16: return solve_t_fn_kl<T, 16>(WU);
    }
    return nullptr;
}

template <class T>
FactorFn factor_fn(int KL) {
    switch (KL) {
        case 1: return &launch_factor<T, 1>;
        case 2: return &launch_factor<T, 2>;
        case 4: return &launch_factor<T, 4>;
        case 6: return &launch_factor<T, 6>;
        case 8: return &launch_factor<T, 8>;
        case 10: return &launch_factor<T, 10>;
        case 12: return &launch_factor<T, 12>;
        case 14: return &launch_factor<T, 14>;
        case 16: return &launch_factor<T, 16>;
    }
    return nullptr;
}
template <class T, int KL>
SolveFn solve_fn_kl(int WU) {
    if (WU < KL) return nullptr;
    switch (WU) {
        case 4: if constexpr (KL <= 4) return &launch_solve<T, KL, 4>; else return nullptr;
        case 8: if constexpr (KL <= 8) return &launch_solve<T, KL, 8>; else return nullptr;
        case 16: return &launch_solve<T, KL, 16>;
        case 24: return &launch_solve<T, KL, 24>;
        case 32: return &launch_solve<T, KL, 32>;
    }
    return nullptr;
}
using SolveSFn = void (*)(dim3, dim3, size_t, hipStream_t, const void*, const void*, const int*, const void*,
                          const void*, const void*, const int*, const int*, void*, const void*, void*, int64_t, int,
                          int, int, int, int);

template <class T, int KL, int WU, int NR>
void launch_solve_s(dim3 g, dim3 b, size_t lds, hipStream_t s, const void* Uf, const void* Lf, const int* piv,
                    const void* Zt, const void* Wf, const void* S, const int* spiv, const int* perm, void* ws,
                    const void* bin, void* xout, int64_t rs, int B, int n, int r, int US, int bufbytes) {
    hipLaunchKernelGGL((k_solve_s<T, KL, WU, NR>), g, b, lds, s, static_cast<const T*>(Uf), static_cast<const T*>(Lf),
                       piv, static_cast<const T*>(Zt), static_cast<const T*>(Wf), static_cast<const T*>(S), spiv, perm,
                       static_cast<T*>(ws), static_cast<const T*>(bin), static_cast<T*>(xout), rs, B, n, r, US,
                       bufbytes);
}

template <class T, int KL, int WU>
SolveSFn solve_s_nr(int nr) {
    switch (nr) {
        case 1: return &launch_solve_s<T, KL, WU, 1>;
        case 2: return &launch_solve_s<T, KL, WU, 2>;
        case 4: return &launch_solve_s<T, KL, WU, 4>;
    }
    return nullptr;
}
template <class T, int KL>
SolveSFn solve_s_wu(int WU, int nr) {
    switch (WU) {
        case 4: if constexpr (KL <= 4) return solve_s_nr<T, KL, 4>(nr); else return nullptr;
        case 8: return solve_s_nr<T, KL, 8>(nr);
        case 16: return solve_s_nr<T, KL, 16>(nr);
    }
    return nullptr;
}
template <class T>
SolveSFn solve_s_fn(int KL, int WU, int nr) {
    switch (KL) {
        case 1: return solve_s_wu<T, 1>(WU, nr);
        case 2: return solve_s_wu<T, 2>(WU, nr);
        case 4: return solve_s_wu<T, 4>(WU, nr);
        case 6: return solve_s_wu<T, 6>(WU, nr);
    }
    return nullptr;
}

template <class T>
SolveFn solve_fn(int KL, int WU) {
    switch (KL) {
        case 1: return solve_fn_kl<T, 1>(WU);
        case 2: return solve_fn_kl<T, 2>(WU);
        case 4: return solve_fn_kl<T, 4>(WU);
        case 6: return solve_fn_kl<T, 6>(WU);
        case 8: return solve_fn_kl<T, 8>(WU);
        case 10: return solve_fn_kl<T, 10>(WU);
        case 12: return solve_fn_kl<T, 12>(WU);
        case 14: return solve_fn_kl<T, 14>(WU);
        case 16: return solve_fn_kl<T, 16>(WU);
    }
    return nullptr;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void alloc(size_t b) {
        release();
        bytes = b;
        if (b) hip_check(hipMalloc(&p, b), "hipMalloc");
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
};

}  // namespace

void cn::set_error(const std::string& msg) { g_err = msg; }

struct cn_bb {
    int dtype = CN_C128, B = 0, Bp = 0, n = 0, r = 0, kl = 0, ku = 0, KLP = 0, M = 0, WU = 0, US = 0, G = 0;
    int pivot = 1;
    int streamed = 1;   // 0: always use the direct solve kernel (CNHIP_SOLVE=direct)
    int nrhs_cap = 0;
    size_t esz = 16;
    double cond_max = 1.0, factor_ms = 0.0;
    int device = 0;
    DevBuf Ab, Uf, Lf, piv, Vt, W, S, spiv, cond, info, perm, ws;
    DevBuf Wt, St, stpiv, cond_t, info_t;   // transposed Woodbury data (built on demand)
    bool tr_ready = false;
    std::vector<double> hcond;
};

namespace {

// Host layout of one system into the interleaved device arrays (banded.hpp:170-251 semantics).
template <class T>
void layout_one(const cn_bb& h, int m, const int32_t* ip, const int32_t* ix, const T* dv,
                const std::vector<int>& perm, const std::vector<int>& pos, std::vector<T>& Ab, std::vector<T>& Vt) {
    const int n = h.n, r = h.r, M = h.M, KLP = h.KLP;
    const size_t ch = m / IL, ml = m % IL;
    auto AB = [&](int i, int d) -> T& { return Ab[(ch * ((size_t)n * M) + (size_t)i * M + d) * IL + ml]; };
    auto VT = [&](int a, int j) -> T& { return Vt[(ch * ((size_t)r * n) + (size_t)a * n + j) * IL + ml]; };
    for (int i = 0; i < n; ++i) {
        const int o = perm[i];
        if (i < r) {
            double s = 0.0;
            for (int p = ip[o]; p < ip[o + 1]; ++p) {
                s = std::max(s, static_cast<double>(std::abs(dv[p])));
                VT(i, ix[p]) += dv[p];
            }
            if (s == 0.0) s = 1.0;
            AB(i, KLP) = T(s);
            VT(i, i) -= T(s);
        } else {
            for (int p = ip[o]; p < ip[o + 1]; ++p) {
                if (dv[p] == T(0)) continue;
                const int j = ix[p];
                AB(i, j - i + KLP) += dv[p];
            }
        }
    }
    (void)pos;
}

template <class T>
void run_factor(cn_bb& h, hipStream_t s) {
    const int spb = 256 / h.G;
    const dim3 fb(256), fg((h.Bp + spb - 1) / spb);
    FactorFn f = factor_fn<T>(h.KLP);
    if (!f) fail(CN_ELOGIC, "no factor kernel for kl " + std::to_string(h.KLP));
    // band LU, with the forward half Wf = L^-1 P_B U of W in the r extra lanes
    f(fg, fb, s, h.Ab.p, h.Uf.p, h.Lf.p, static_cast<int*>(h.piv.p), h.W.p, static_cast<int*>(h.info.p), h.B, h.n,
      h.M, h.r, h.G, h.US, h.pivot);
    hip_check(hipGetLastError(), "k_band_lu launch");
    if (h.r > 0) {
        // Z^T = V^T B^-1 (rows B^-T v_a), used by the forward solve (t = Z^T c) and, as Wt, by the
        // transposed solve; then S = I + Z^T U and St = I + U^T Wt, both LU-factored
        SolveTFn ft = solve_t_fn<T>(h.KLP, h.WU);
        if (!ft) fail(CN_ELOGIC, "no transposed solve kernel");
        const int lanes = h.B * h.r;
        ft(dim3((lanes + 255) / 256), dim3(256), s, h.Uf.p, h.Lf.p, static_cast<const int*>(h.piv.p), h.Vt.p, h.Wt.p,
           h.St.p, static_cast<const int*>(h.stpiv.p), static_cast<const int*>(h.perm.p), nullptr, nullptr, 0, h.B, h.n,
           h.r, h.US, h.r, 1);
        hip_check(hipGetLastError(), "k_solve_t (Z^T) launch");
        const int l2 = h.B * h.r * h.r;
        hipLaunchKernelGGL((k_build_s<T>), dim3((l2 + 255) / 256), dim3(256), 0, s, static_cast<const T*>(h.Wt.p),
                           static_cast<T*>(h.S.p), h.B, h.n, h.r);
        hip_check(hipGetLastError(), "k_build_s launch");
        hipLaunchKernelGGL((k_cap_lu<T>), dim3((h.B + 255) / 256), dim3(256), 0, s, static_cast<T*>(h.S.p),
                           static_cast<int*>(h.spiv.p), static_cast<double*>(h.cond.p), static_cast<int*>(h.info.p),
                           h.B, h.r);
        hip_check(hipGetLastError(), "k_cap_lu launch");
        hipLaunchKernelGGL((k_build_st<T>), dim3((l2 + 255) / 256), dim3(256), 0, s, static_cast<const T*>(h.Wt.p),
                           static_cast<T*>(h.St.p), static_cast<int*>(h.info_t.p), h.B, h.n, h.r);
        hip_check(hipGetLastError(), "k_build_st launch");
        hipLaunchKernelGGL((k_cap_lu<T>), dim3((h.B + 255) / 256), dim3(256), 0, s, static_cast<T*>(h.St.p),
                           static_cast<int*>(h.stpiv.p), static_cast<double*>(h.cond_t.p),
                           static_cast<int*>(h.info_t.p), h.B, h.r);
        hip_check(hipGetLastError(), "k_cap_lu (transposed) launch");
    }
    h.tr_ready = true;
}

void factor_and_check(cn_bb& h, hipStream_t s, bool check) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (check) {
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, s), "hipEventRecord");
    }
    if (h.dtype == CN_C128) run_factor<zc>(h, s); else run_factor<double>(h, s);
    if (!check) return;
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h.factor_ms = ms;
    std::vector<int> info(h.B);
    hip_check(hipMemcpy(info.data(), h.info.p, sizeof(int) * h.B, hipMemcpyDeviceToHost), "hipMemcpy info");
    h.hcond.assign(h.B, 1.0);
    if (h.r > 0)
        hip_check(hipMemcpy(h.hcond.data(), h.cond.p, sizeof(double) * h.B, hipMemcpyDeviceToHost), "hipMemcpy cond");
    h.cond_max = 1.0;
    for (int m = 0; m < h.B; ++m) {
        if (info[m] > 0)
            fail(CN_ESINGULAR, "BandedBatch: banded part of mode " + std::to_string(m) + " is singular at pivot " +
                                   std::to_string(info[m]) + (h.pivot ? "" : " (no pivoting was requested)") +
                                   "; the non-border rows must form an invertible banded operator");
        if (info[m] < 0)
            fail(CN_ESINGULAR, "BandedBatch: capacitance matrix of mode " + std::to_string(m) +
                                   " is singular: the boundary rows are not independent of the banded operator");
        if (std::isnan(h.hcond[m]))
            fail(CN_ESINGULAR, "BandedBatch: capacitance condition number of mode " + std::to_string(m) + " is NaN");
        h.cond_max = std::max(h.cond_max, h.hcond[m]);
    }
    if (h.r > 0) {
        std::vector<int> info_t(h.B);
        hip_check(hipMemcpy(info_t.data(), h.info_t.p, sizeof(int) * h.B, hipMemcpyDeviceToHost), "hipMemcpy info_t");
        for (int m = 0; m < h.B; ++m)
            if (info_t[m] != 0)
                fail(CN_ESINGULAR, "BandedBatch: transposed capacitance matrix of mode " + std::to_string(m) +
                                       " is singular");
    }
}

template <class T>
void create_impl(cn_bb& h, const int32_t* const* indptr, const int32_t* const* indices, const void* const* data,
                 const int64_t* nnz, const int32_t* border) {
    const int n = h.n, B = h.B, r = h.r;
    // CSR structure and finiteness (core/bindings/module.cpp:75-95, banded.hpp:173-180)
    for (int m = 0; m < B; ++m) {
        const int32_t* ip = indptr[m];
        const int32_t* ix = indices[m];
        const T* dv = static_cast<const T*>(data[m]);
        if (!ip || (nnz[m] > 0 && (!ix || !dv))) fail(CN_EINVAL, "CSR: null array for matrix " + std::to_string(m));
        if (ip[0] != 0 || (int64_t)ip[n] != nnz[m])
            fail(CN_EINVAL, "inconsistent CSR arrays (indptr[0] = 0, indptr[n] = nnz = len(indices) = len(data) "
                            "required) for matrix " + std::to_string(m));
        for (int i = 0; i < n; ++i)
            if (ip[i + 1] < ip[i]) fail(CN_EINVAL, "CSR: indptr decreases at row " + std::to_string(i));
        for (int64_t p = 0; p < nnz[m]; ++p) {
            if (ix[p] < 0 || ix[p] >= n)
                fail(CN_EINVAL, "CSR: column index " + std::to_string(ix[p]) + " outside [0, " + std::to_string(n) + ")");
            const std::complex<double> v(dv[p]);
            if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) {
                int row = 0;
                while (ip[row + 1] <= p) ++row;
                fail(CN_EINVAL, "BorderedBanded: non-finite matrix entry at (" + std::to_string(row) + ", " +
                                    std::to_string(ix[p]) + ") of matrix " + std::to_string(m) +
                                    ": the operator must be finite");
            }
        }
    }
    std::vector<char> isb(n, 0);
    for (int a = 0; a < r; ++a) {
        const int i = border[a];
        if (i < 0 || i >= n) fail(CN_EINVAL, "border row out of range");
        if (isb[i]) fail(CN_EINVAL, "duplicate border row");
        isb[i] = 1;
    }
    std::vector<int> perm(border, border + r), pos(n);
    for (int i = 0; i < n; ++i) if (!isb[i]) perm.push_back(i);
    for (int i = 0; i < n; ++i) pos[perm[i]] = i;
    // common band of the bulk rows (banded.hpp:144-168), max over the batch
    int kl = 0, ku = 0;
    for (int m = 0; m < B; ++m) {
        const int32_t* ip = indptr[m];
        const int32_t* ix = indices[m];
        const T* dv = static_cast<const T*>(data[m]);
        for (int i = r; i < n; ++i) {
            const int o = perm[i];
            for (int p = ip[o]; p < ip[o + 1]; ++p) {
                if (dv[p] == T(0)) continue;
                kl = std::max(kl, i - ix[p]);
                ku = std::max(ku, ix[p] - i);
            }
        }
    }
    h.kl = kl;
    h.ku = ku;
    h.KLP = pick_kl(std::max(kl, 1));
    if (h.KLP < 0) fail(CN_EUNSUPPORTED, "BandedBatch: kl = " + std::to_string(kl) + " exceeds the compiled kernels (<= 16)");
    h.M = h.KLP + ku + 1;
    int G = 16;
    while (G < h.M + r) G *= 2;
    if (G > 64)
        fail(CN_EUNSUPPORTED, "BandedBatch: band kl+ku+1 = " + std::to_string(h.M) + " plus border rank " +
                                  std::to_string(r) + " exceeds one wavefront (64 lanes)");
    h.G = G;
    h.WU = pick_wu(h.KLP + ku);
    if (h.WU < 0) fail(CN_EUNSUPPORTED, "BandedBatch: kl+ku = " + std::to_string(h.KLP + ku) + " exceeds 32");
    h.US = h.WU + 1;
    h.Bp = (B + 63) / 64 * 64;   // whole 64-system workgroups of the streamed solve stay in bounds
    const size_t es = sizeof(T);
    const size_t nAb = checked_mul({(size_t)h.Bp, (size_t)n, (size_t)h.M});
    const size_t nU = checked_mul({(size_t)h.Bp, (size_t)n, (size_t)h.US});
    const size_t nL = checked_mul({(size_t)h.Bp, (size_t)n, (size_t)h.KLP});
    const size_t nV = checked_mul({(size_t)h.Bp, (size_t)r, (size_t)n});
    std::vector<T> Ab(nAb, T(0)), Vt(nV, T(0));
    for (int m = 0; m < B; ++m)
        layout_one<T>(h, m, indptr[m], indices[m], static_cast<const T*>(data[m]), perm, pos, Ab, Vt);
    hip_check(hipGetDevice(&h.device), "hipGetDevice");
    h.Ab.alloc(nAb * es);
    h.Uf.alloc(nU * es);
    h.Lf.alloc(nL * es);
    h.piv.alloc((size_t)h.Bp * n * sizeof(int));
    h.Vt.alloc(std::max<size_t>(nV, 1) * es);
    h.W.alloc(std::max<size_t>(nV, 1) * es);
    h.S.alloc(std::max<size_t>((size_t)h.Bp * r * r, 1) * es);
    h.spiv.alloc(std::max<size_t>((size_t)h.Bp * r, 1) * sizeof(int));
    h.cond.alloc((size_t)h.Bp * sizeof(double));
    h.info.alloc((size_t)h.Bp * sizeof(int));
    h.perm.alloc((size_t)n * sizeof(int));
    h.Wt.alloc(std::max<size_t>(nV, 1) * es);
    h.St.alloc(std::max<size_t>((size_t)h.Bp * r * r, 1) * es);
    h.stpiv.alloc(std::max<size_t>((size_t)h.Bp * r, 1) * sizeof(int));
    h.cond_t.alloc((size_t)h.Bp * sizeof(double));
    h.info_t.alloc((size_t)h.Bp * sizeof(int));
    hip_check(hipMemset(h.info_t.p, 0, (size_t)h.Bp * sizeof(int)), "hipMemset info_t");
    hip_check(hipMemset(h.W.p, 0, std::max<size_t>(nV, 1) * es), "hipMemset W");
    hip_check(hipMemset(h.Wt.p, 0, std::max<size_t>(nV, 1) * es), "hipMemset Wt");
    hip_check(hipMemcpy(h.Ab.p, Ab.data(), nAb * es, hipMemcpyHostToDevice), "hipMemcpy Ab");
    if (nV) hip_check(hipMemcpy(h.Vt.p, Vt.data(), nV * es, hipMemcpyHostToDevice), "hipMemcpy Vt");
    hip_check(hipMemcpy(h.perm.p, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy perm");
    hip_check(hipMemset(h.Uf.p, 0, nU * es), "hipMemset U");
    hip_check(hipMemset(h.Lf.p, 0, nL * es), "hipMemset L");
    hip_check(hipMemset(h.info.p, 0, (size_t)h.Bp * sizeof(int)), "hipMemset info");
    hip_check(hipMemset(h.cond.p, 0, (size_t)h.Bp * sizeof(double)), "hipMemset cond");
}

void reserve_impl(cn_bb& h, int nrhs) {
    if (nrhs <= h.nrhs_cap) return;
    h.ws.alloc(checked_mul({(size_t)h.B, (size_t)nrhs, (size_t)h.n, h.esz}));
    h.nrhs_cap = nrhs;
}

void prepare_t(cn_bb& h) {
    // the transposed Woodbury data (Wt = Z^T, LU of St) are built with every factorization
    if (!h.tr_ready) fail(CN_ELOGIC, "BandedBatch: factorization missing");
}

void solve_t_impl(cn_bb& h, void* x, int nrhs, int64_t rs, hipStream_t s) {
    if (!h.tr_ready || nrhs > h.nrhs_cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
        if (cs != hipStreamCaptureStatusNone)
            fail(CN_EINVAL, "BandedBatch transposed solve: prepare_transpose / reserve must run before graph capture");
        prepare_t(h);
        reserve_impl(h, nrhs);
    }
    SolveTFn f = (h.dtype == CN_C128) ? solve_t_fn<zc>(h.KLP, h.WU) : solve_t_fn<double>(h.KLP, h.WU);
    if (!f) fail(CN_ELOGIC, "no transposed solve kernel");
    const int lanes = h.B * nrhs;
    f(dim3((lanes + 255) / 256), dim3(256), s, h.Uf.p, h.Lf.p, static_cast<const int*>(h.piv.p), h.Vt.p, h.Wt.p,
      h.St.p, static_cast<const int*>(h.stpiv.p), static_cast<const int*>(h.perm.p), h.ws.p, x, rs, h.B, h.n, h.r,
      h.US, nrhs, 0);
    hip_check(hipGetLastError(), "k_solve_t launch");
}

// LDS bytes of one buffer of the streamed solve (see k_solve_s), or -1 if it does not apply
int streamed_bufbytes(const cn_bb& h, const SolveIO& io, int nrhs) {
    if (!(io.mode == 0 || (io.mode == 1 && !io.rows && !io.cols))) return -1;
    if (!(nrhs == 1 || nrhs == 2 || nrhs == 4)) return -1;
    if (h.KLP + 1 > SP || h.WU > 16) return -1;
    const int C = (64 / nrhs) / IL;
    const int CE = IL * (int)h.esz;
    const int fchunk = SP * h.KLP * CE + h.r * SP * CE + SP * IL * 4;
    const int bchunk = SP * h.US * CE + h.r * SP * CE;
    const int buf = (std::max(C * fchunk, C * bchunk) + 15) / 16 * 16;
    if (2 * buf > 64 * 1024) return -1;
    return buf;
}

void solve_impl(cn_bb& h, SolveIO io, int nrhs, hipStream_t s) {
    if (nrhs > h.nrhs_cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
        if (cs != hipStreamCaptureStatusNone)
            fail(CN_EINVAL, "BandedBatch solve: work space for " + std::to_string(nrhs) +
                                " right-hand sides not reserved before graph capture (call cn_bb_reserve)");
        reserve_impl(h, nrhs);
    }
    const int bufbytes = h.streamed ? streamed_bufbytes(h, io, nrhs) : -1;
    if (bufbytes > 0) {
        SolveSFn fs = (h.dtype == CN_C128) ? solve_s_fn<zc>(h.KLP, h.WU, nrhs) : solve_s_fn<double>(h.KLP, h.WU, nrhs);
        if (fs) {
            const int NS = 64 / nrhs;
            const void* bin = io.mode == 0 ? io.x : io.rhs;
            fs(dim3((h.B + NS - 1) / NS), dim3(64), 2 * bufbytes, s, h.Uf.p, h.Lf.p, static_cast<const int*>(h.piv.p),
               h.Wt.p, h.W.p, h.S.p, static_cast<const int*>(h.spiv.p), static_cast<const int*>(h.perm.p), h.ws.p, bin,
               io.x, io.mode == 0 ? io.rhs_stride : 0, h.B, h.n, h.r, h.US, bufbytes);
            hip_check(hipGetLastError(), "k_solve_s launch");
            return;
        }
    }
    SolveFn f = (h.dtype == CN_C128) ? solve_fn<zc>(h.KLP, h.WU) : solve_fn<double>(h.KLP, h.WU);
    if (!f) fail(CN_ELOGIC, "no solve kernel for (kl, wu) = (" + std::to_string(h.KLP) + ", " + std::to_string(h.WU) + ")");
    const int lanes = h.B * nrhs;
    f(dim3((lanes + 255) / 256), dim3(256), s, h.Uf.p, h.Lf.p, static_cast<const int*>(h.piv.p), h.Wt.p, h.W.p,
      h.S.p, static_cast<const int*>(h.spiv.p), static_cast<const int*>(h.perm.p), h.ws.p, io, h.B, h.n, h.r, h.US,
      nrhs);
    hip_check(hipGetLastError(), "k_solve launch");
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        g_err.clear();
        return CN_OK;
    } catch (const CnError& e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return CN_EOVERFLOW;
    } catch (const std::exception& e) {
        g_err = e.what();
        return CN_ELOGIC;
    }
}

}  // namespace

extern "C" {

const char* cn_last_error(void) { return g_err.c_str(); }

const char* cn_version(void) { return "cnhip 0.1.0 gfx950"; }

int cn_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int cn_bb_create(cn_bb_t* out, int dtype, int batch, int n, const int32_t* const* indptr,
                 const int32_t* const* indices, const void* const* data, const int64_t* nnz,
                 const int32_t* border, int r, int pivot, int col_border) {
    return guarded([&] {
        if (!out) fail(CN_EINVAL, "cn_bb_create: null output handle");
        *out = nullptr;
        if (dtype != CN_F64 && dtype != CN_C128) fail(CN_EINVAL, "cn_bb_create: dtype must be CN_F64 or CN_C128");
        if (batch < 1) fail(CN_EINVAL, "BandedBatch: empty batch");
        if (n < 1) fail(CN_EINVAL, "BandedBatch: n must be >= 1, got " + std::to_string(n));
        if (r < 0 || r > n) fail(CN_EINVAL, "BandedBatch: border rank must be in [0, n]");
        if (r > MAXR)
            fail(CN_EUNSUPPORTED, "BandedBatch: border rank " + std::to_string(r) + " > " + std::to_string(MAXR) +
                                      " not supported on the device");
        if (col_border) fail(CN_EUNSUPPORTED, "BandedBatch: the column border is not supported by this library");
        if (!indptr || !indices || !data || !nnz || (r > 0 && !border)) fail(CN_EINVAL, "cn_bb_create: null array");
        auto h = std::make_unique<cn_bb>();
        h->dtype = dtype;
        h->B = batch;
        h->n = n;
        h->r = r;
        h->pivot = pivot ? 1 : 0;
        h->esz = dtype == CN_C128 ? 16 : 8;
        if (const char* e = std::getenv("CNHIP_SOLVE")) h->streamed = std::string(e) != "direct";
        if (dtype == CN_C128) create_impl<std::complex<double>>(*h, indptr, indices, data, nnz, border);
        else create_impl<double>(*h, indptr, indices, data, nnz, border);
        factor_and_check(*h, nullptr, true);
        reserve_impl(*h, 1);
        *out = h.release();
    });
}

int cn_bb_factor(cn_bb_t h, void* stream, int check) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_factor: null handle");
        h->tr_ready = false;   // the transposed data follow the factors
        factor_and_check(*h, static_cast<hipStream_t>(stream), check != 0);
    });
}

int cn_bb_reserve(cn_bb_t h, int nrhs) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_reserve: null handle");
        if (nrhs < 1) fail(CN_EINVAL, "cn_bb_reserve: nrhs must be >= 1");
        reserve_impl(*h, nrhs);
    });
}

int cn_bb_solve(cn_bb_t h, void* x, int nrhs, int64_t rhs_stride, int transpose, void* stream) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_solve: null handle");
        if (!x) fail(CN_EINVAL, "BandedBatch.solve_ptr: null device pointer");
        if (nrhs < 1) fail(CN_EINVAL, "cn_bb_solve: nrhs must be >= 1");
        if (nrhs > 1 && rhs_stride < (int64_t)h->B * h->n)
            fail(CN_EINVAL, "cn_bb_solve: rhs_stride must be >= batch * n");
        if (transpose) {
            solve_t_impl(*h, x, nrhs, rhs_stride, static_cast<hipStream_t>(stream));
            return;
        }
        SolveIO io{};
        io.mode = 0;
        io.x = x;
        io.rhs_stride = rhs_stride;
        solve_impl(*h, io, nrhs, static_cast<hipStream_t>(stream));
    });
}

int cn_bb_solve_mapped(cn_bb_t h, const void* rhs, const int64_t* rows, void* x, const int64_t* cols, void* stream) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_solve_mapped: null handle");
        if (!rhs || !x) fail(CN_EINVAL, "BandedBatch.solve_mapped_ptr: null device pointer");
        SolveIO io{};
        io.mode = 1;
        io.x = x;
        io.rhs = rhs;
        io.rows = rows;
        io.cols = cols;
        solve_impl(*h, io, 1, static_cast<hipStream_t>(stream));
    });
}

int cn_bb_solve_split(cn_bb_t h, const double* re, const int64_t* rows_c, const int64_t* rows_s, double* x,
                      const int64_t* cols_c, const int64_t* cols_s, void* stream) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_solve_split: null handle");
        if (h->dtype != CN_C128) fail(CN_EINVAL, "BandedBatch.solve_split_ptr: complex128 factors required");
        if (!re || !x || !rows_c || !rows_s || !cols_c || !cols_s)
            fail(CN_EINVAL, "BandedBatch.solve_split_ptr: null device pointer");
        SolveIO io{};
        io.mode = 2;
        io.x = x;
        io.rhs = re;
        io.rows = rows_c;
        io.rows2 = rows_s;
        io.cols = cols_c;
        io.cols2 = cols_s;
        solve_impl(*h, io, 1, static_cast<hipStream_t>(stream));
    });
}

int cn_bb_prepare_transpose(cn_bb_t h) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_prepare_transpose: null handle");
        prepare_t(*h);
    });
}

int cn_bb_props_get(cn_bb_t h, cn_bb_props* p) {
    return guarded([&] {
        if (!h || !p) fail(CN_EINVAL, "cn_bb_props_get: null argument");
        p->dtype = h->dtype;
        p->batch = h->B;
        p->n = h->n;
        p->kl = h->kl;
        p->ku = h->ku;
        p->rank = h->r;
        p->pivot = h->pivot;
        p->kl_kernel = h->KLP;
        p->wu_kernel = h->WU;
        p->lanes = h->G;
        p->nrhs_reserved = h->nrhs_cap;
        p->transpose_ready = h->tr_ready ? 1 : 0;
        p->cond_max = h->cond_max;
        p->factor_ms = h->factor_ms;
        p->factor_bytes = (int64_t)(h->Uf.bytes + h->Lf.bytes + h->piv.bytes + h->Vt.bytes + h->W.bytes + h->S.bytes +
                                    h->spiv.bytes);
    });
}

int cn_bb_cond(cn_bb_t h, double* out) {
    return guarded([&] {
        if (!h || !out) fail(CN_EINVAL, "cn_bb_cond: null argument");
        for (int m = 0; m < h->B; ++m) out[m] = m < (int)h->hcond.size() ? h->hcond[m] : 1.0;
    });
}

void cn_bb_destroy(cn_bb_t h) { delete h; }

}  // extern "C"
