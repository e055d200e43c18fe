// libcnhip device code of the bordered-banded solver (kernels and device helpers), shared by
// banded.hip and the lane-mode translation units banded_lane_*.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cn_common.h"

namespace cnk {

using cn::zc;

constexpr int IL = 16;     // systems per interleave chunk
constexpr int MAXR = 8;    // largest border rank handled by the kernels
constexpr int LDS_CAP = 160 * 1024;   // LDS of one workgroup on gfx950 (163840 bytes)

using SolveSFn = void (*)(dim3, dim3, size_t, hipStream_t, const void*, const void*, const int*, const void*,
                          const void*, const void*, const int*, const int*, void*, const void*, void*, int64_t, int,
                          int, int, int, int);
using BorderSFn = void (*)(dim3, dim3, size_t, hipStream_t, const void*, const void*, void*, void*, int, int, int, int,
                           int);

// launch tables of the streamed (lane mode) kernels, one object per (type, block length, border
// rows compiled for, kl) built from banded_lane.hip; the border tables live in the kl = 1 objects
#define CN_LANE_TABLES(tag)                                  \
    SolveSFn lane_solve_##tag##_k1(int WU, int nr);          \
    SolveSFn lane_solve_##tag##_k2(int WU, int nr);          \
    SolveSFn lane_solve_##tag##_k4(int WU, int nr);          \
    SolveSFn lane_solve_##tag##_k6(int WU, int nr);          \
    BorderSFn lane_border_##tag(int WU, int nr);
CN_LANE_TABLES(zc16_r2)
CN_LANE_TABLES(zc16_r8)
CN_LANE_TABLES(zc8_r2)
CN_LANE_TABLES(zc8_r8)
CN_LANE_TABLES(d16_r2)
CN_LANE_TABLES(d16_r8)
CN_LANE_TABLES(d8_r2)
CN_LANE_TABLES(d8_r8)
#undef CN_LANE_TABLES

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
          int dopivot, int il, int ucol) {
    const int t = threadIdx.x;
    const int g = t / G;
    const int c = t - g * G;
    const int m = blockIdx.x * (256 / G) + g;
    const bool live = m < B;
    const size_t ch = m / il;
    const int ml = m % il;
    const T* A0 = Ab + ch * ((size_t)n * M * il) + ml;
    T* U0 = Uf + ch * ((size_t)n * US * il) + ml;
    T* L0 = Lf + ch * ((size_t)n * KL * il) + ml;
    int* P0 = piv + ch * ((size_t)n * il) + ml;
    T* W0 = Wf + ch * ((size_t)r * n * il) + ml;
    const bool band = live && c < M;
    const bool wl = live && c >= M && c < M + r;
    const int k = c - M;
    const int gbase = (t & 63) - c;   // wavefront lane of the group's lane 0 (G divides 64)

    // Offsets within a chunk are int (the launch checks n * max(M, US, r) * il < 2^31), and the
    // step's owner and column offsets advance by one step at a time: no 64-bit multiply and no
    // division by M in the loop.
    T win[KL + 1];
#pragma unroll
    for (int i = 0; i <= KL; ++i) {
        T v = cn::zero<T>();
        if (band && i < n) {
            const int d = c - i + KL;
            if (d >= 0 && d < M) v = A0[(i * M + d) * il];
        }
        if (wl) v = (i == k) ? cn::one<T>() : cn::zero<T>();
        win[i] = v;
    }
    // The band element that enters the window at the end of step s (row s+KL+1 of this lane's
    // column at step s, or of column s+M for the lane own_s = s mod M that owned column s), read
    // PF steps before it is needed: a read issued in the step before its use would put one memory
    // latency on the chain of every step.
    auto entering = [&](int s, int own_s) -> T {
        T v = cn::zero<T>();
        const int jn = s + KL + 1;
        if (band && jn < n) {
            int dA = c - own_s - 1;   // d - 1 for the other lanes, M - 1 for the owner
            if (dA < 0) dA += M;
            v = A0[(jn * M + dA) * il];
        }
        return v;
    };
    constexpr int PF = 8;
    T ahead[PF];
    int own_pf = 0;               // (j + PF) mod M, advanced with j
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        ahead[u] = entering(u, own_pf);
        own_pf = own_pf + 1 == M ? 0 : own_pf + 1;
    }
    int owner = 0;                // j mod M
    int first_singular = 0;
    for (int j0 = 0; j0 < n; j0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            const int j = j0 + u;
            if (j >= n) break;
            // win[i] holds row j+i of this lane's column
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
            int d = c - owner;   // column offset of this lane's column from j
            if (d < 0) d += M;
            if (band) {
                T ov = win[0];
                if (is_owner) ov = rc;
                if (!ucol) U0[(j * US + d) * il] = ov;                   // row j of U: U(j, j+d)
                else if (j + d < n) U0[((j + d) * US + d) * il] = ov;   // column j+d: U(j, j+d)
            }
            if (wl) W0[(k * n + j) * il] = win[0];
            if (band && is_owner) {
#pragma unroll
                for (int i = 0; i < KL; ++i) L0[(j * KL + i) * il] = l[i];
                P0[j * il] = ip;
            }
            if ((band && !is_owner) || wl) {
#pragma unroll
                for (int i = 1; i <= KL; ++i) win[i] = cn::fnms(l[i - 1], win[0], win[i]);
            }
            // row j leaves the window; row j+KL+1 enters at the bottom
            const int jn = j + KL + 1;
            T v = ahead[u];
            ahead[u] = entering(j + PF, own_pf);
            own_pf = own_pf + 1 == M ? 0 : own_pf + 1;
            if (band && is_owner) {   // the owner takes column j+M: rows j+1..j+KL are zero there
#pragma unroll
                for (int i = 1; i <= KL; ++i) win[i] = cn::zero<T>();
            }
            if (wl && jn == k) v = cn::one<T>();
#pragma unroll
            for (int i = 0; i < KL; ++i) win[i] = win[i + 1];
            win[KL] = v;
            if (singular && first_singular == 0) first_singular = j + 1;
            owner = owner + 1 == M ? 0 : owner + 1;
        }
    }
    if (live && c == 0) info[m] = first_singular;
}

__device__ __forceinline__ double absv(double a) { return fabs(a); }
__device__ __forceinline__ double absv(zc a) { return hypot(a.re, a.im); }
__device__ __forceinline__ zc neg(zc a) { return zc{-a.re, -a.im}; }
__device__ __forceinline__ double neg(double a) { return -a; }

// One lane per system: LU of S in place (column-major, partial pivoting by |re|+|im|), and
// cond_1(S) = |S|_1 |S^-1|_1 with the inverse from the factors (banded.hpp:403-432).
// one lane per system; the r x r matrix, its pivots and the inverse column of each lane live in
// LDS, entry-major ([entry][lane], CAP_LANES lanes per workgroup), not in private arrays: a
// dynamically indexed private array is scratch memory, which the runtime backs for every wave
// the device can hold (about 0.8 GB at 1.2 KB per lane).
constexpr int CAP_LANES = 64;
template <class T>
__host__ __device__ constexpr size_t cap_lds_bytes() {
    return (size_t)(MAXR * MAXR + MAXR) * CAP_LANES * sizeof(T) + (size_t)MAXR * CAP_LANES * sizeof(int);
}

template <class T>
__global__ void __launch_bounds__(CAP_LANES)
k_cap_lu(T* __restrict__ S, int* __restrict__ spiv, double* __restrict__ cond, int* __restrict__ info,
         int B, int r) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x;
    const int m = blockIdx.x * CAP_LANES + lane;
    if (m >= B) return;
    T* La = reinterpret_cast<T*>(lds);                         // [MAXR * MAXR][CAP_LANES]
    T* Lx = La + MAXR * MAXR * CAP_LANES;                      // [MAXR][CAP_LANES]
    int* Lp = reinterpret_cast<int*>(Lx + MAXR * CAP_LANES);   // [MAXR][CAP_LANES]
#define CAP_A(q) La[(q) * CAP_LANES + lane]
#define CAP_X(i) Lx[(i) * CAP_LANES + lane]
#define CAP_P(i) Lp[(i) * CAP_LANES + lane]
    T* Sm = S + (size_t)m * r * r;
    for (int q = 0; q < r * r; ++q) CAP_A(q) = Sm[q];
    double nS = 0.0;
    for (int b = 0; b < r; ++b) {
        double s = 0.0;
        for (int i = 0; i < r; ++i) s += absv(CAP_A(b * r + i));
        if (!(s <= nS)) nS = s;
    }
    int bad = 0;
    for (int jc = 0; jc < r; ++jc) {
        int ip = jc;
        double best = cn::cabs1(CAP_A(jc * r + jc));
        for (int i = jc + 1; i < r; ++i) {
            const double v = cn::cabs1(CAP_A(jc * r + i));
            if (v > best) { best = v; ip = i; }
        }
        CAP_P(jc) = ip;
        if (best == 0.0) { bad = 1; continue; }
        if (ip != jc)
            for (int q = 0; q < r; ++q) {
                T tmp = CAP_A(q * r + jc);
                CAP_A(q * r + jc) = CAP_A(q * r + ip);
                CAP_A(q * r + ip) = tmp;
            }
        const T rc = cn::recip(CAP_A(jc * r + jc));
        for (int i = jc + 1; i < r; ++i) CAP_A(jc * r + i) = cn::mul(CAP_A(jc * r + i), rc);
        for (int q = jc + 1; q < r; ++q)
            for (int i = jc + 1; i < r; ++i) CAP_A(q * r + i) = cn::fnms(CAP_A(jc * r + i), CAP_A(q * r + jc), CAP_A(q * r + i));
    }
    for (int q = 0; q < r * r; ++q) Sm[q] = CAP_A(q);
    for (int q = 0; q < r; ++q) spiv[(size_t)m * r + q] = CAP_P(q);
    double c = 1.0;
    if (bad) {
        if (info[m] == 0) info[m] = -1;
        c = INFINITY;
    } else {
        // inverse column by column: solve S x = e_b with the factors
        double ninv = 0.0;
        for (int b = 0; b < r; ++b) {
            for (int i = 0; i < r; ++i) CAP_X(i) = (i == b) ? cn::one<T>() : cn::zero<T>();
            for (int i = 0; i < r; ++i) {
                const int pi = CAP_P(i);
                if (pi != i) { T tmp = CAP_X(i); CAP_X(i) = CAP_X(pi); CAP_X(pi) = tmp; }
            }
            for (int i = 0; i < r; ++i)
                for (int q = 0; q < i; ++q) CAP_X(i) = cn::fnms(CAP_A(q * r + i), CAP_X(q), CAP_X(i));
            for (int i = r - 1; i >= 0; --i) {
                for (int q = i + 1; q < r; ++q) CAP_X(i) = cn::fnms(CAP_A(q * r + i), CAP_X(q), CAP_X(i));
                CAP_X(i) = cn::mul(CAP_X(i), cn::recip(CAP_A(i * r + i)));
            }
            double s = 0.0;
            for (int i = 0; i < r; ++i) s += absv(CAP_X(i));
            if (!(s <= ninv)) ninv = s;
        }
        c = nS * ninv;
    }
    cond[m] = c;
#undef CAP_A
#undef CAP_X
#undef CAP_P
}

// acc0 - sum_{d=1..WU} u(d) xw[d-1], the row step of the back substitution with U, subtracted
// term by term from d = WU down to d = 1 (the row order of a banded triangular solve).  The terms
// d >= 2 use solutions of earlier steps, so only the last complex fnms waits for the previous
// step's result; the others run while it is computed.  Every lane-mode back substitution sums in
// this order, so that all of them agree bit for bit.
template <class T, int WU, class F>
__device__ __forceinline__ T band_dot(T acc0, const T (&xw)[WU], F u) {
    T s = acc0;
#pragma unroll
    for (int d = WU; d >= 1; --d) s = cn::fnms(u(d), xw[d - 1], s);
    return s;
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

// cap_solve with the factors already in registers (r <= RB): Sr[q][i] = S(i, q), P the pivots
template <class T, int RB>
__device__ __forceinline__ void cap_solve_reg(T (&t)[MAXR], const T (&Sr)[RB][RB], const int (&P)[RB], int r) {
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i >= r) break;
        const int pi = P[i];
#pragma unroll
        for (int q = i + 1; q < RB; ++q) {
            const bool c = (q == pi);
            const T ti = t[i], tq = t[q];
            t[i] = cn::sel(c, tq, ti);
            t[q] = cn::sel(c, ti, tq);
        }
    }
#pragma unroll
    for (int i = 0; i < RB; ++i) {
        if (i >= r) break;
#pragma unroll
        for (int q = 0; q < i; ++q) t[i] = cn::fnms(Sr[q][i], t[q], t[i]);
    }
#pragma unroll
    for (int i = RB - 1; i >= 0; --i) {
        if (i < r) {
#pragma unroll
            for (int q = i + 1; q < RB; ++q)
                if (q < r) t[i] = cn::fnms(Sr[q][i], t[q], t[i]);
            t[i] = cn::mul(t[i], cn::recip(Sr[i][i]));
        }
    }
}

// Direct solve (any shape and IO mode), the Woodbury order of banded.hpp:312-341:
//   c = P b;  y = U^-1 L^-1 P_B c (forward sweep, then backward sweep);  t = V^T y accumulated
//   as the backward sweep produces y;  z = S^-1 t;  x = y - W z (a third pass over x).
// z is formed from the y that x is built from, so the residual of x is that of the two band
// sweeps and of W alone (no cond(B) amplification through the border rows).
template <class T, int KL, int WU>
__global__ void __launch_bounds__(256)
k_solve(const T* __restrict__ Uf, const T* __restrict__ Lf, const int* __restrict__ piv,
        const T* __restrict__ Vt, const T* __restrict__ W, const T* __restrict__ S,
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
    const T* V0 = Vt + ch * ((size_t)r * n * IL) + ml;
    const T* W0 = W + ch * ((size_t)r * n * IL) + ml;
    const size_t mn = (size_t)m * n;
    T* w0 = ws + gid;

    // forward: y = L^-1 P_B c, window of KL+1 rows
    T y[KL + 1];
#pragma unroll
    for (int i = 0; i <= KL; ++i) y[i] = (i < n) ? io_load_b<T>(io, mn, k, perm[i]) : cn::zero<T>();
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
        y[KL] = (jn < n) ? io_load_b<T>(io, mn, k, perm[jn]) : cn::zero<T>();
    }
    // backward: y = U^-1 y_c, window of the last WU solutions; t = V^T y
    T t[MAXR];
#pragma unroll
    for (int a = 0; a < MAXR; ++a) t[a] = cn::zero<T>();
    T xw[WU];
#pragma unroll
    for (int d = 0; d < WU; ++d) xw[d] = cn::zero<T>();
    for (int j = n - 1; j >= 0; --j) {
        const T* Ur = U0 + (size_t)j * US * IL;
        const T acc = band_dot<T, WU>(w0[(size_t)j * NL], xw, [&](int d) { return Ur[d * IL]; });
        const T x = cn::mul(acc, Ur[0]);
        io_store<T>(io, mn, k, j, x);
#pragma unroll
        for (int a = 0; a < MAXR; ++a)
            if (a < r) t[a] = cn::fmac(V0[((size_t)a * n + j) * IL], x, t[a]);
#pragma unroll
        for (int d = WU - 1; d > 0; --d) xw[d] = xw[d - 1];
        xw[0] = x;
    }
    if (r == 0) return;
    cap_solve<T>(t, S + (size_t)m * r * r, spiv + (size_t)m * r, r);
    // x = y - W z
    for (int j = 0; j < n; ++j) {
        T v = io_load_x<T>(io, mn, k, j);
#pragma unroll
        for (int a = 0; a < MAXR; ++a)
            if (a < r) v = cn::fnms(W0[((size_t)a * n + j) * IL], t[a], v);
        io_store<T>(io, mn, k, j, v);
    }
}

// Direct completion of W = B^-1 U from its forward half (in place) and S = I + V^T W:
// one lane per (system, border column a), back substitution with U.
template <class T, int WU>
__global__ void __launch_bounds__(256)
k_border(const T* __restrict__ Uf, const T* __restrict__ Vt, T* __restrict__ W, T* __restrict__ S, int B,
         int n, int r, int US) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = gid / r;
    const int a = gid - m * r;
    if (m >= B) return;
    const size_t ch = m / IL;
    const int ml = m % IL;
    const T* U0 = Uf + ch * ((size_t)n * US * IL) + ml;
    const T* V0 = Vt + ch * ((size_t)r * n * IL) + ml;
    T* W0 = W + ch * ((size_t)r * n * IL) + ml + (size_t)a * n * IL;
    T t[MAXR];
#pragma unroll
    for (int q = 0; q < MAXR; ++q) t[q] = cn::zero<T>();
    T xw[WU];
#pragma unroll
    for (int d = 0; d < WU; ++d) xw[d] = cn::zero<T>();
    for (int j = n - 1; j >= 0; --j) {
        const T* Ur = U0 + (size_t)j * US * IL;
        const T acc = band_dot<T, WU>(W0[(size_t)j * IL], xw, [&](int d) { return Ur[d * IL]; });
        const T x = cn::mul(acc, Ur[0]);
        W0[(size_t)j * IL] = x;
#pragma unroll
        for (int q = 0; q < MAXR; ++q)
            if (q < r) t[q] = cn::fmac(V0[((size_t)q * n + j) * IL], x, t[q]);
#pragma unroll
        for (int d = WU - 1; d > 0; --d) xw[d] = xw[d - 1];
        xw[0] = x;
    }
    // column a of S (column-major): S(q, a) = delta_qa + (V^T W)(q, a)
    T* Sm = S + (size_t)m * r * r + (size_t)a * r;
#pragma unroll
    for (int q = 0; q < MAXR; ++q)
        if (q < r) Sm[q] = (q == a) ? cn::fmac(cn::one<T>(), cn::one<T>(), t[q]) : t[q];
}

// ------------------------------------------------------------------------------------------
// Streamed sweeps (the hot kernels for many narrow-band systems).  A workgroup is one wavefront
// and one interleave chunk of 16 systems, with NR lanes per system (NR in {1, 2, 4} right-hand
// sides or border columns).  Every sweep runs in blocks of SPX steps: the factor rows of block
// b+1 (chunk-contiguous, copied with global_load_lds_dwordx4) and the per-lane operands of block
// b+1 (gathered by the same instruction, one slot per lane) travel from HBM into the other half
// of a double-buffered LDS area while block b is computed from LDS; the results of block b are
// held in registers and stored at the start of block b+1, so that the one wait per block
// (s_waitcnt vmcnt(0)) covers memory operations issued a whole block earlier.
// ------------------------------------------------------------------------------------------

using lds_ptr = __attribute__((address_space(3))) void*;

// contiguous copy of nbytes (multiple of 16) from global memory into LDS, 1 KiB per instruction
__device__ __forceinline__ void dma16(const void* src_, char* lds_dst, int nbytes, int lane) {
    const char* src = static_cast<const char*>(src_);
    for (int off = 0; off < nbytes; off += 1024) {
        const int my = off + lane * 16;
        if (my < nbytes)
            __builtin_amdgcn_global_load_lds(static_cast<const void*>(src + my), (lds_ptr)(lds_dst + off), 16, 0, 0);
    }
}

__device__ __forceinline__ void dma_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// vector memory instructions dma16 issues for nbytes (one per started KiB)
__device__ __forceinline__ int dma16_count(int nbytes) { return (nbytes + 1023) >> 10; }

// workgroup barrier of the two-wave sweeps: the wave's LDS operations are complete, its vector
// memory operations are not waited for (the memory wave waits for exactly the copies the next
// block needs); the memory clobber keeps the compiler from moving memory accesses across it
__device__ __forceinline__ void wg_sync() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// s_waitcnt vmcnt(k) for a wave-uniform k in [0, 63]: wait until at most k vector memory
// operations of this wave are outstanding.  Loads complete in issue order, so after the wait every
// load older than the newest k memory operations has completed (stores may complete out of
// order; callers that need them done wait for 0).
__device__ __forceinline__ void vm_wait_le(int k) {
    switch (k) {
#define CN_VMW(K) case K: asm volatile("s_waitcnt vmcnt(" #K ")" ::: "memory"); break;
        CN_VMW(0) CN_VMW(1) CN_VMW(2) CN_VMW(3) CN_VMW(4) CN_VMW(5) CN_VMW(6) CN_VMW(7) CN_VMW(8) CN_VMW(9)
        CN_VMW(10) CN_VMW(11) CN_VMW(12) CN_VMW(13) CN_VMW(14) CN_VMW(15) CN_VMW(16) CN_VMW(17) CN_VMW(18)
        CN_VMW(19) CN_VMW(20) CN_VMW(21) CN_VMW(22) CN_VMW(23) CN_VMW(24) CN_VMW(25) CN_VMW(26) CN_VMW(27)
        CN_VMW(28) CN_VMW(29) CN_VMW(30) CN_VMW(31) CN_VMW(32) CN_VMW(33) CN_VMW(34) CN_VMW(35) CN_VMW(36)
        CN_VMW(37) CN_VMW(38) CN_VMW(39) CN_VMW(40) CN_VMW(41) CN_VMW(42) CN_VMW(43) CN_VMW(44) CN_VMW(45)
        CN_VMW(46) CN_VMW(47) CN_VMW(48) CN_VMW(49) CN_VMW(50) CN_VMW(51) CN_VMW(52) CN_VMW(53) CN_VMW(54)
        CN_VMW(55) CN_VMW(56) CN_VMW(57) CN_VMW(58) CN_VMW(59) CN_VMW(60) CN_VMW(61) CN_VMW(62)
#undef CN_VMW
        default: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    }
}

template <class T>
__device__ __forceinline__ T lds_at(const char* p) { return *reinterpret_cast<const T*>(p); }

// One value per lane gathered from an arbitrary global address into LDS slot `slot` (lane l's
// value at slot + l * sizeof(T); a double travels as two dwords into two 256-byte halves).
template <class T>
struct LaneSlot;
template <>
struct LaneSlot<zc> {
    static constexpr int BYTES = 64 * 16;
    static constexpr int LOADS = 1;            // vector memory instructions per load()
    __device__ static void load(const zc* src, char* slot) {
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(src), (lds_ptr)slot, 16, 0, 0);
    }
    __device__ static zc read(const char* slot, int lane) { return lds_at<zc>(slot + lane * 16); }
    __device__ static void zero(char* slot, int lane) { *reinterpret_cast<zc*>(slot + lane * 16) = zc{0.0, 0.0}; }
    __device__ static void write(char* slot, int lane, zc v) { *reinterpret_cast<zc*>(slot + lane * 16) = v; }
};
template <>
struct LaneSlot<double> {
    static constexpr int BYTES = 64 * 8;
    static constexpr int LOADS = 2;
    __device__ static void load(const double* src, char* slot) {
        const char* s = reinterpret_cast<const char*>(src);
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(s), (lds_ptr)slot, 4, 0, 0);
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(s + 4), (lds_ptr)(slot + 256), 4, 0, 0);
    }
    __device__ static double read(const char* slot, int lane) {
        const unsigned lo = lds_at<unsigned>(slot + lane * 4);
        const unsigned hi = lds_at<unsigned>(slot + 256 + lane * 4);
        return __hiloint2double((int)hi, (int)lo);
    }
    __device__ static void zero(char* slot, int lane) {
        *reinterpret_cast<unsigned*>(slot + lane * 4) = 0u;
        *reinterpret_cast<unsigned*>(slot + 256 + lane * 4) = 0u;
    }
    __device__ static void write(char* slot, int lane, double v) {
        *reinterpret_cast<unsigned*>(slot + lane * 4) = (unsigned)__double2loint(v);
        *reinterpret_cast<unsigned*>(slot + 256 + lane * 4) = (unsigned)__double2hiint(v);
    }
};

// Lane segments: a block of SPX consecutive entries of each lane's own vector (its segment) moves
// between HBM and LDS as whole segments.  DMA instruction i carries the segments q = SPI*i + g
// (g = L / SPX, SPI = 64 / SPX segments per instruction): lane L carries entry
// e = (L mod SPX - g) mod SPX of segment q, so that one instruction touches SPI runs of SPX entries
// instead of 64 scattered entries.  In LDS instruction i's slot sits at i * PITCH and entry e of
// segment q at lane position g * SPX + (e + g) mod SPX; the rotation by g and the padding of the
// pitch spread the simultaneous reads of one entry by all 64 lanes over the banks.
template <class T, int SPX>
struct Seg {
    static constexpr int SPI = 64 / SPX;
    static constexpr int PITCH = LaneSlot<T>::BYTES + 64 * (int)sizeof(T) / SPX;
    static constexpr int BYTES = SPX * PITCH;                  // 64 segments
    // entry of its segment that lane L carries in a DMA instruction
    __device__ static int entry(int L) { return (L - L / SPX) & (SPX - 1); }
    // slot offset and lane position of entry e of segment q
    __device__ static int slot(int q) { return (q / SPI) * PITCH; }
    __device__ static int pos(int q, int e) {
        const int g = q % SPI;
        return g * SPX + ((e + g) & (SPX - 1));
    }
};
__host__ __device__ constexpr int seg_bytes(int SPX, int esz) {
    return SPX * (64 * esz + 64 * esz / SPX);
}

// compiler ordering point for LDS data exchanged between the lanes of one wavefront (the LDS
// operations of a wavefront execute in program order)
__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// LDS bytes of one buffer of each streamed sweep (blocks of SPX steps)
__host__ __device__ constexpr int fwd_bytes(int SPX, int KL, int esz) {
    return SPX * KL * IL * esz + SPX * IL * 4 + seg_bytes(SPX, esz);
}
// (rb: the border rows the kernel is compiled for, rb >= r; rows r .. rb-1 of V^T and W are
// zero pads in LDS, so that the steps carry no test on r)
__host__ __device__ constexpr int bwd_bytes(int SPX, int US, int rb, int esz) {
    return (SPX * US + rb * SPX) * IL * esz + SPX * 64 * esz;
}
__host__ __device__ constexpr int cor_bytes(int SPX, int rb, int esz) { return rb * SPX * IL * esz + SPX * 64 * esz; }
// border rows a kernel is compiled for: 2 (r <= 2) or MAXR
__host__ __device__ constexpr int rb_for(int r) { return r <= 2 ? 2 : MAXR; }

// Backward block blk covers the rows of forward block nblk-1-blk: rows jhi down to jlo with
// jlo = (nblk-1-blk) * SPX and jhi = min(n, jlo + SPX) - 1 (the first backward block is the short
// one when SPX does not divide n), so that a block's lane operands are one forward block's results.
__device__ __forceinline__ void bwd_rows(int n, int SPX, int blk, int& jlo, int& jhi) {
    const int nblk = (n + SPX - 1) / SPX;
    jlo = (nblk - 1 - blk) * SPX;
    jhi = min(n, jlo + SPX) - 1;
}
// Backward block blk into buf:
// [U: SPX rows x US][V^T: RB x SPX (rows >= r: pads)][in: SPX lane slots, slot s = row jhi - s],
// in two parts: the factor rows (V^T rows, then U; returns the number of instructions issued, all
// whole-wave) and the lane operands (in slots; lane-dependent).
template <class T, int SPX, int RB>
__device__ __forceinline__ int bwd_dma_f(const T* __restrict__ Uf, const T* __restrict__ Vt, char* buf, size_t ch, int n,
                                         int r, int US, int blk, int lane) {
    constexpr int CE = IL * (int)sizeof(T);
    int jlo, jhi;
    bwd_rows(n, SPX, blk, jlo, jhi);
    const int cnt = jhi - jlo + 1;
    const int bU = SPX * US * CE;
    for (int a = 0; a < r; ++a) dma16(Vt + (ch * r * n + (size_t)a * n + jlo) * IL, buf + bU + a * SPX * CE, cnt * CE, lane);
    dma16(Uf + (ch * n * US + (size_t)jlo * US) * IL, buf, cnt * US * CE, lane);
    return r * dma16_count(cnt * CE) + dma16_count(cnt * US * CE);
}
// (returns the vector memory instructions issued: none when no lane of the wave is live)
template <class T, int SPX, int RB>
__device__ __forceinline__ int bwd_dma_l(char* buf, int n, int US, int blk, const T* in, size_t in_st, bool live) {
    constexpr int CE = IL * (int)sizeof(T);
    int jlo, jhi;
    bwd_rows(n, SPX, blk, jlo, jhi);
    const int bU = SPX * US * CE, bV = RB * SPX * CE;
    if (live) {
#pragma unroll
        for (int s = 0; s < SPX; ++s)
            if (jhi - s >= jlo) LaneSlot<T>::load(in + (size_t)(jhi - s) * in_st, buf + bU + bV + s * LaneSlot<T>::BYTES);
    }
    return __ballot(live) != 0 ? (jhi - jlo + 1) * LaneSlot<T>::LOADS : 0;
}

// One backward step of a lane: out_j = (in_j - sum_d U(j, j+d) out_{j+d}) / U(j, j), t += V^T(:, j) out_j.
// bb: the block buffer, row: j - jlo, sidx: slot of in_j.
template <class T, int WU, int SPX, int RB>
__device__ __forceinline__ T bwd_lane_step(const char* bb, int row, int sidx, int US, int mlx, int lane,
                                           T (&xw)[WU], T (&t)[RB]) {
    constexpr int CE = IL * (int)sizeof(T);
    const int bU = SPX * US * CE, bV = RB * SPX * CE;
    const char* ur = bb + row * US * CE + mlx * (int)sizeof(T);
    const T acc = band_dot<T, WU>(LaneSlot<T>::read(bb + bU + bV + sidx * LaneSlot<T>::BYTES, lane), xw,
                                  [&](int d) { return lds_at<T>(ur + d * CE); });
    const T x = cn::mul(acc, lds_at<T>(ur));
#pragma unroll
    for (int a = 0; a < RB; ++a) t[a] = cn::fmac(lds_at<T>(bb + bU + (a * SPX + row) * CE + mlx * (int)sizeof(T)), x, t[a]);
#pragma unroll
    for (int d = WU - 1; d > 0; --d) xw[d] = xw[d - 1];
    xw[0] = x;
    return x;
}

// LDS operands of one backward step in registers: U row (u[0] = 1 / U(j, j)), V^T column, in_j
template <class T, int WU, int RB>
struct BwdOps {
    T u[WU + 1];
    T v[RB];
    T in;
};
template <class T, int WU, int SPX, int RB>
__device__ __forceinline__ void bwd_ops_load(BwdOps<T, WU, RB>& op, const char* bb, int row, int sidx, int US, int mlx,
                                             int lane) {
    constexpr int CE = IL * (int)sizeof(T);
    const int bU = SPX * US * CE, bV = RB * SPX * CE;
    const char* ur = bb + row * US * CE + mlx * (int)sizeof(T);
#pragma unroll
    for (int d = 0; d <= WU; ++d) op.u[d] = lds_at<T>(ur + d * CE);
#pragma unroll
    for (int a = 0; a < RB; ++a) op.v[a] = lds_at<T>(bb + bU + (a * SPX + row) * CE + mlx * (int)sizeof(T));
    op.in = LaneSlot<T>::read(bb + bU + bV + sidx * LaneSlot<T>::BYTES, lane);
}
// the arithmetic of bwd_lane_step on operands in registers
template <class T, int WU, int RB>
__device__ __forceinline__ T bwd_ops_step(const BwdOps<T, WU, RB>& op, T (&xw)[WU], T (&t)[RB]) {
    const T acc = band_dot<T, WU>(op.in, xw, [&](int d) { return op.u[d]; });
    const T x = cn::mul(acc, op.u[0]);
#pragma unroll
    for (int a = 0; a < RB; ++a) t[a] = cn::fmac(op.v[a], x, t[a]);
#pragma unroll
    for (int d = WU - 1; d > 0; --d) xw[d] = xw[d - 1];
    xw[0] = x;
    return x;
}

// scheduling fence: the LDS operands of the next step are read before this step's arithmetic
__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

// ------------------------------------------------------------------------------------------
// Four-wave workgroups of the streamed sweeps.  Wave 0 (compute) runs the recurrences from LDS
// and leaves its results in an LDS stage; waves 1, 2 and 3 issue every HBM access while block b
// is computed: wave 1 the factor rows of the next blocks (whole-wave global_load_lds copies),
// wave 2 the lane operands of the next blocks (loads only), wave 3 the stores of block b-1 from
// the other stage.  A wave that issues memory instructions stalls while its own copies drain (at
// most 63 in flight per wave); with the roles split, that stall holds up neither the arithmetic
// nor the other memory waves, and one workgroup barrier per block hands buffers and stages over.
// LDS: [buffers][stage 0][stage 1], stage = stage_bytes (the larger of SPX lane slots and the
// segment layout).  Where three buffers of a sweep fit, waves 1 and 2 keep two blocks in flight
// and, before each barrier, wait only until the block needed next has landed (counted waits:
// loads complete in issue order, and each wave's count covers its own loads only).  Wave 3 waits
// for its stores once, at the end of the sweep, before the next sweep may read them.
constexpr int SWEEP_THREADS = 256;
// ------------------------------------------------------------------------------------------
__host__ __device__ constexpr int stage_bytes(int SPX, int esz) { return seg_bytes(SPX, esz); }
// LDS buffers of a sweep: nb = 3 when three of the sweep's buffers fit the launch's buffer region
// (block b is computed from buffer b % 3 while blocks b+1 and b+2 are in flight), else 2
__host__ __device__ constexpr int sweep_nbuf(int sweep_bytes, int region) { return 3 * sweep_bytes <= region ? 3 : 2; }

__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Streamed backward sweep: out_j = (in_j - sum_d U(j, j+d) out_{j+d}) / U(j, j), j = n-1 .. 0, and
// t[a] += V^T(a, j) out_j (a < RB; the entries a >= r accumulate the LDS pads and are not used).
// `in` and `out` may be the same array (in place).  Every wave calls it (t is the compute wave's).
// Full blocks run without per-step tests; a short first block step by step.
// Hand-over with the sweeps on either side (the solve; the border completion passes 0, false and
// a no-op): the first `stage_in` blocks (0..2) take their operands from the LDS stages, where the
// forward sweep left its last two blocks (forward block f in stage f & 1, slot = row - f * SPX),
// instead of from `in`; with `keep_out` the last two blocks' results stay in their stages (block
// b in stage b & 1, slot s = row jhi - s) and are not stored; `tail()` is called by wave 1 in the
// last block, when no buffer but the last block's is in use, to start the next sweep's copies.
// Block b uses buffer (b + rot) mod nb; with pre_f the factor rows of block 0 are already in
// flight (issued by wave 1 before the sweep, into buffer rot).
template <class T, int WU, int SPX, int RB, class Tail>
__device__ __forceinline__ void bwd_sweep(const T* __restrict__ Uf, const T* __restrict__ Vt, char* lds, int region,
                                          size_t ch, int n, int r, int US, const T* in, size_t in_st, T* out,
                                          size_t out_st, bool live, int wave, int lane, int mlx, T (&t)[RB],
                                          int stage_in, bool keep_out, Tail tail, int rot = 0, bool pre_f = false) {
    constexpr int GB = LaneSlot<T>::BYTES;
    const int nblk = (n + SPX - 1) / SPX;
    char* stage = lds + region;
    const int SB = stage_bytes(SPX, (int)sizeof(T));
    const int bufbytes = bwd_bytes(SPX, US, RB, (int)sizeof(T));
    const int nb = sweep_nbuf(bufbytes, region), ahead = nb - 1;
    const int bLS = (SPX * US + RB * SPX) * IL * (int)sizeof(T);   // lane slots within a buffer
    const int nin = min(stage_in, nblk);
    T xw[WU];
#pragma unroll
    for (int d = 0; d < WU; ++d) xw[d] = cn::zero<T>();
    auto buf = [&](int bk) { return lds + ((bk + rot) % nb) * bufbytes; };
    if (wave == 1) {   // factor rows of blocks 0 .. ahead-1 in flight, block 0 complete
        if (!pre_f) bwd_dma_f<T, SPX, RB>(Uf, Vt, buf(0), ch, n, r, US, 0, lane);
        vm_wait_le(ahead > 1 && nblk > 1 ? bwd_dma_f<T, SPX, RB>(Uf, Vt, buf(1), ch, n, r, US, 1, lane) : 0);
    } else if (wave == 2) {   // lane operands likewise: blocks 0 .. nin-1 out of the stages, the rest from `in`
        for (int bk = 0; bk < nin; ++bk) {
            int jlo, jhi;
            bwd_rows(n, SPX, bk, jlo, jhi);
            const int cnt = jhi - jlo + 1;
            const char* sf = stage + ((jlo / SPX) & 1) * SB;
            char* dst = buf(bk) + bLS;
#pragma unroll
            for (int s = 0; s < SPX; ++s)
                if (s < cnt) LaneSlot<T>::write(dst + s * GB, lane, LaneSlot<T>::read(sf + (cnt - 1 - s) * GB, lane));
        }
        const int c0 = nin > 0 ? 0 : bwd_dma_l<T, SPX, RB>(buf(0), n, US, 0, in, in_st, live);
        (void)c0;
        vm_wait_le(ahead > 1 && nblk > 1 && nin < 2 ? bwd_dma_l<T, SPX, RB>(buf(1), n, US, 1, in, in_st, live) : 0);
    }
    wg_sync();
    for (int b = 0; b < nblk; ++b) {
        int jlo, jhi;
        bwd_rows(n, SPX, b, jlo, jhi);
        const int cnt = jhi - jlo + 1;
        if (wave == 1) {
            // factor rows of block b+ahead into the buffer block b-1 used; then block b+1 must be complete
            const int c = b + ahead < nblk ? bwd_dma_f<T, SPX, RB>(Uf, Vt, buf(b + ahead), ch, n, r, US, b + ahead, lane)
                                           : 0;
            vm_wait_le(ahead > 1 ? c : 0);
            if (b == nblk - 1) tail();
        } else if (wave == 2) {
            // lane operands of block b+ahead into the buffer block b-1 used; then block b+1 must be complete
            const int c = b + ahead < nblk && b + ahead >= nin
                              ? bwd_dma_l<T, SPX, RB>(buf(b + ahead), n, US, b + ahead, in, in_st, live)
                              : 0;
            vm_wait_le(ahead > 1 ? c : 0);
        } else if (wave == 3) {
            if (b > 0 && live && !(keep_out && b - 1 >= nblk - 2)) {   // results of block b-1
                int jlo1, jhi1;
                bwd_rows(n, SPX, b - 1, jlo1, jhi1);
                const char* sb = stage + ((b - 1) & 1) * SB;
#pragma unroll
                for (int s = 0; s < SPX; ++s)
                    if (jhi1 - s >= jlo1) out[(size_t)(jhi1 - s) * out_st] = LaneSlot<T>::read(sb + s * GB, lane);
            }
        } else {
            const char* bb = buf(b);
            char* sb = stage + (b & 1) * SB;
            if (cnt == SPX) {   // operands of step s+1 in flight during step s
                BwdOps<T, WU, RB> op[2];
                bwd_ops_load<T, WU, SPX, RB>(op[0], bb, SPX - 1, 0, US, mlx, lane);
#pragma unroll
                for (int s = 0; s < SPX; ++s) {
                    if (s + 1 < SPX) bwd_ops_load<T, WU, SPX, RB>(op[(s + 1) & 1], bb, SPX - 2 - s, s + 1, US, mlx, lane);
                    sched_fence();
                    LaneSlot<T>::write(sb + s * GB, lane, bwd_ops_step<T, WU, RB>(op[s & 1], xw, t));
                    sched_fence();
                }
            } else {
#pragma unroll
                for (int s = 0; s < SPX; ++s)
                    if (s < cnt) LaneSlot<T>::write(sb + s * GB, lane, bwd_lane_step<T, WU, SPX, RB>(bb, cnt - 1 - s, s, US, mlx, lane, xw, t));
            }
        }
        wg_sync();
    }
    if (wave == 3) {
        if (live && !keep_out) {   // the last block: rows jhi .. 0
            int jlo, jhi;
            bwd_rows(n, SPX, nblk - 1, jlo, jhi);
            const char* sb = stage + ((nblk - 1) & 1) * SB;
#pragma unroll
            for (int s = 0; s < SPX; ++s)
                if (jhi - s >= 0) out[(size_t)(jhi - s) * out_st] = LaneSlot<T>::read(sb + s * GB, lane);
        }
        dma_wait();
    }
    wg_sync();
}

// One forward step of a lane (window y[0..KL] = rows j .. j+KL): optional row interchange,
// elimination below the pivot, shift, incoming row c.
template <class T, int KL, bool PIV>
__device__ __forceinline__ T fwd_lane_step(T (&y)[KL + 1], const char* Lrow, int p, const T& c, int mlx) {
    constexpr int CE = IL * (int)sizeof(T);
    if (PIV) cn::swap_first<T, KL>(y, p);
#pragma unroll
    for (int i = 1; i <= KL; ++i) y[i] = cn::fnms(lds_at<T>(Lrow + (i - 1) * CE + mlx * (int)sizeof(T)), y[0], y[i]);
    const T out = y[0];
#pragma unroll
    for (int i = 0; i < KL; ++i) y[i] = y[i + 1];
    y[KL] = c;
    return out;
}

// LDS operands of one forward step in registers: multipliers, pivot offset, incoming row
template <class T, int KL>
struct FwdOps {
    T l[KL];
    T c;
    int p;
};
template <class T, int KL, bool PIV>
__device__ __forceinline__ void fwd_ops_load(FwdOps<T, KL>& op, const char* Lrow, const char* prow, const char* cslot,
                                             int cpos, int mlx) {
    constexpr int CE = IL * (int)sizeof(T);
#pragma unroll
    for (int i = 0; i < KL; ++i) op.l[i] = lds_at<T>(Lrow + i * CE + mlx * (int)sizeof(T));
    op.c = LaneSlot<T>::read(cslot, cpos);
    op.p = PIV ? lds_at<int>(prow + mlx * 4) : 0;
}
template <class T, int KL, bool PIV>
__device__ __forceinline__ T fwd_ops_step(T (&y)[KL + 1], const FwdOps<T, KL>& op) {
    if (PIV) cn::swap_first<T, KL>(y, op.p);
#pragma unroll
    for (int i = 1; i <= KL; ++i) y[i] = cn::fnms(op.l[i - 1], y[0], y[i]);
    const T out = y[0];
#pragma unroll
    for (int i = 0; i < KL; ++i) y[i] = y[i + 1];
    y[KL] = op.c;
    return out;
}
// a full forward block, operands of step s+1 in flight during step s; result s into stage slot s
template <class T, int KL, int SPX, bool PIV, class CPos>
__device__ __forceinline__ void fwd_block(T (&y)[KL + 1], char* sb, const char* fb, int fL, const char* cb, CPos cpos,
                                          int mlx, int lane) {
    constexpr int CE = IL * (int)sizeof(T);
    FwdOps<T, KL> op[2];
    fwd_ops_load<T, KL, PIV>(op[0], fb, fb + fL, cb, cpos(0), mlx);
#pragma unroll
    for (int s = 0; s < SPX; ++s) {
        if (s + 1 < SPX)
            fwd_ops_load<T, KL, PIV>(op[(s + 1) & 1], fb + (s + 1) * KL * CE, fb + fL + (s + 1) * IL * 4, cb, cpos(s + 1), mlx);
        sched_fence();
        LaneSlot<T>::write(sb + s * LaneSlot<T>::BYTES, lane, fwd_ops_step<T, KL, PIV>(y, op[s & 1]));
        sched_fence();
    }
}

// Streamed solve: forward sweep (pivots, L) into the work space, backward sweep (U) in place in
// the work space with t = V^T y, z = S^-1 t, then the streamed pass x = y - W z into x.
// Arithmetic as k_solve.  IO: in place (bin == xout, right-hand side k at k * rhs_stride) or
// contiguous out of place.  The work space holds entry j of every (system, rhs) lane of the
// chunk in one contiguous row (coalesced); the right-hand side enters and x leaves by lane
// segments (Seg): per block, each lane's SPX entries travel as one run.  Forward blocks without
// any row interchange in the chunk skip the (select-based) interchanges.  Four waves: lanes
// 0..63 compute, 64..127 copy factor rows, 128..191 load the lane data, 192..255 store it (same
// lane mapping).
template <class T, int KL, int WU, int NR, int SPX, int RB>
__global__ void __launch_bounds__(SWEEP_THREADS) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_solve_s(const T* __restrict__ Uf, const T* __restrict__ Lf, const int* __restrict__ piv,
          const T* __restrict__ Vt, const T* __restrict__ W, const T* __restrict__ S,
          const int* __restrict__ spiv, const int* __restrict__ perm, T* __restrict__ ws,
          const T* bin, T* xout, int64_t rhs_stride, int B, int n, int r, int US, int bufbytes) {
    static_assert(KL + 1 <= SPX, "the window lookahead must stay within the next block");
    using SG = Seg<T, SPX>;
    constexpr int CE = IL * (int)sizeof(T);
    constexpr int GB = LaneSlot<T>::BYTES;
    constexpr int NI = IL * NR / SG::SPI;       // DMA instructions per block of segments
    static_assert(IL * NR % SG::SPI == 0, "segments fill whole DMA instructions");
    extern __shared__ __align__(16) char lds[];
    const int wave = wave_id();
    const int lane = threadIdx.x & 63;
    const int mloc = lane / NR, k = lane - mloc * NR;
    const int mlx = mloc & (IL - 1);            // LDS column of this lane's system (dead lanes alias a live one)
    const size_t ch = blockIdx.x;
    const int m = (int)ch * IL + mloc;
    const bool live = mloc < IL && m < B;
    const size_t NL = (size_t)B * NR;
    T* wl = ws + (size_t)m * NR + k;
    const int fL = SPX * KL * CE, fP = SPX * IL * 4;
    const int nblk = (n + SPX - 1) / SPX;
    const int region = bufbytes;                // the buffers of every sweep, then the two stages
    char* stage = lds + region;
    constexpr int SB = stage_bytes(SPX, (int)sizeof(T));
    // the memory wave's part of the segment DMA: entry se of segment SPI*i + sg in instruction i,
    // i.e. system ch*IL + (SPI/NR)*i + sg/NR, right-hand side sg % NR
    const int sg = lane / SPX, se = SG::entry(lane);
    const size_t seg_off = (size_t)(sg % NR) * rhs_stride + ((size_t)ch * IL + sg / NR) * n;
    const size_t seg_step = (size_t)(SG::SPI / NR) * n;
    const int seg_live = B - (int)ch * IL - sg / NR;   // instruction i is live if (SPI/NR)*i < seg_live
    const int my_slot = SG::slot(lane);                // where this lane's own segment is read/written

    // the capacitance LU of this lane's system, read at the start (RB <= 2) so that the z solve
    // between the backward sweep and the correction waits for no memory
    T Sr[RB <= 2 ? RB : 1][RB <= 2 ? RB : 1];
    int Pr[RB <= 2 ? RB : 1];
    if constexpr (RB <= 2) {
        if (wave == 0 && live) {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                Pr[q] = q < r ? spiv[(size_t)m * r + q] : q;
#pragma unroll
                for (int i = 0; i < RB; ++i) Sr[q][i] = (q < r && i < r) ? S[(size_t)m * r * r + q * r + i] : cn::zero<T>();
            }
        }
    }

    // ---- forward sweep: y_c = L^-1 P_B P b into the work space
    const int fbuf = fwd_bytes(SPX, KL, (int)sizeof(T));
    const int fnb = sweep_nbuf(fbuf, region), fahead = fnb - 1;
    // The backward sweep's first factor rows are copied during the last forward block, into a
    // backward buffer (the sweep's buffer rotation brot) that does not overlap that block's buffer.
    const int bbuf = bwd_bytes(SPX, US, RB, (int)sizeof(T));
    const int bnb = sweep_nbuf(bbuf, region);
    const int flo = ((nblk - 1) % fnb) * fbuf, fhi = flo + fbuf;
    int brot = -1;
    for (int q = 0; q < bnb && brot < 0; ++q)
        if (q * bbuf >= fhi || (q + 1) * bbuf <= flo) brot = q;
    const bool pre_u = brot >= 0;
    if (brot < 0) brot = 0;
    T y[KL + 1];
    // wave 2: right-hand-side segments of forward block bn into buffer nb.  Entry e of the block's
    // segments is right-hand-side row perm[bn * SPX + KL + 1 + e] (or zero beyond n); the SPX
    // entries of perm are read at wave-uniform addresses (scalar loads), so that no vector load
    // result is waited for between the counted waits.  Returns the segment loads issued (a lower
    // bound on the wave's vector memory instructions of the call).
    auto fwd_copy_l = [&](int bn, char* nb) -> int {
        const int base = bn * SPX + KL + 1;
        int pc = -1;
#pragma unroll
        for (int e = 0; e < SPX; ++e) {
            const int q = base + e < n ? perm[base + e] : -1;
            pc = se == e ? q : pc;
        }
        char* cs = nb + fL + fP;
        int issued = 0;
        if (__ballot(pc >= 0) != 0) {
#pragma unroll
            for (int i = 0; i < NI; ++i)
                if ((SG::SPI / NR) * i < seg_live) issued += LaneSlot<T>::LOADS;
        }
        if (pc >= 0) {
            const T* src = bin + seg_off + pc;
#pragma unroll
            for (int i = 0; i < NI; ++i)
                if ((SG::SPI / NR) * i < seg_live) LaneSlot<T>::load(src + i * seg_step, cs + i * SG::PITCH);
        } else {
#pragma unroll
            for (int i = 0; i < NI; ++i) LaneSlot<T>::zero(cs + i * SG::PITCH, lane);
        }
        return issued;
    };
    // wave 1: L rows and pivots of block bn (returns the number of instructions issued)
    auto fwd_copy_f = [&](int bn, char* nb) -> int {
        const int j0 = bn * SPX;
        const int cnt = min(SPX, n - j0);
        dma16(Lf + (ch * n * KL + (size_t)j0 * KL) * IL, nb, cnt * KL * CE, lane);
        dma16(piv + (ch * n + j0) * IL, nb + fL, cnt * IL * 4, lane);
        return dma16_count(cnt * KL * CE) + dma16_count(cnt * IL * 4);
    };
    auto fwd_store = [&](int bs) {                // forward results of block bs from its stage
        if (bs >= nblk - 2) return;               // the last two blocks reach the backward sweep in their stages
        const char* sb = stage + (bs & 1) * SB;
        const int j0 = bs * SPX;
#pragma unroll
        for (int s = 0; s < SPX; ++s)
            if (j0 + s < n) wl[(size_t)(j0 + s) * NL] = LaneSlot<T>::read(sb + s * GB, lane);
    };
    if (wave == 0) {
        const T* bl = bin + k * rhs_stride + (size_t)m * n;
#pragma unroll
        for (int i = 0; i <= KL; ++i) y[i] = (live && i < n) ? bl[perm[i]] : cn::zero<T>();
    } else if (wave == 1) {   // blocks 0 .. ahead-1 in flight, block 0 complete
        fwd_copy_f(0, lds);
        vm_wait_le(fahead > 1 && nblk > 1 ? fwd_copy_f(1, lds + fbuf) : 0);
    } else if (wave == 2) {
        fwd_copy_l(0, lds);
        vm_wait_le(fahead > 1 && nblk > 1 ? fwd_copy_l(1, lds + fbuf) : 0);
    }
    wg_sync();
    for (int b = 0; b < nblk; ++b) {
        if (wave == 1) {
            const int c = b + fahead < nblk ? fwd_copy_f(b + fahead, lds + ((b + fahead) % fnb) * fbuf) : 0;
            vm_wait_le(fahead > 1 ? c : 0);
            if (b == nblk - 1 && pre_u) bwd_dma_f<T, SPX, RB>(Uf, Vt, lds + brot * bbuf, ch, n, r, US, 0, lane);
        } else if (wave == 2) {
            const int c = b + fahead < nblk ? fwd_copy_l(b + fahead, lds + ((b + fahead) % fnb) * fbuf) : 0;
            vm_wait_le(fahead > 1 ? c : 0);
        } else if (wave == 3) {
            if (b >= 1 && live) fwd_store(b - 1);
        } else {
            const char* fb = lds + (b % fnb) * fbuf;
            const char* cb = fb + fL + fP + my_slot;
            char* sb = stage + (b & 1) * SB;
            const int j0 = b * SPX, cnt = min(SPX, n - j0);
            if (cnt == SPX) {
                // any row interchange in this block for the chunk's systems?
                int any = 0;
                for (int q = lane; q < SPX * IL; q += 64) any |= lds_at<int>(fb + fL + q * 4);
                auto cpos = [&](int s) { return SG::pos(lane, s); };
                if (__ballot(any != 0) == 0)
                    fwd_block<T, KL, SPX, false>(y, sb, fb, fL, cb, cpos, mlx, lane);
                else
                    fwd_block<T, KL, SPX, true>(y, sb, fb, fL, cb, cpos, mlx, lane);
            } else {
#pragma unroll
                for (int s = 0; s < SPX; ++s)
                    if (s < cnt)
                        LaneSlot<T>::write(sb + s * GB, lane,
                                           fwd_lane_step<T, KL, true>(y, fb + s * KL * CE, lds_at<int>(fb + fL + (s * IL + mlx) * 4),
                                                                      LaneSlot<T>::read(cb, SG::pos(lane, s)), mlx));
            }
        }
        wg_sync();
    }
    if (wave == 3) {
        if (live) fwd_store(nblk - 1);
        dma_wait();
    }
    wg_sync();

    // ---- backward sweep: y = U^-1 y_c in place, t = V^T y; its first two blocks come from the
    // forward sweep's stages, its last two stay in its own for the correction
    T t[RB];
#pragma unroll
    for (int a = 0; a < RB; ++a) t[a] = cn::zero<T>();
    const int cW = RB * SPX * CE;
    const int cbuf = cor_bytes(SPX, RB, (int)sizeof(T));
    const int cnb = sweep_nbuf(cbuf, region), cahead = cnb - 1;
    auto cor_copy_f = [&](int bn, char* nb) -> int { // wave 1: the W rows of block bn
        const int j0 = bn * SPX;
        const int cnt = min(SPX, n - j0);
        for (int a = 0; a < r; ++a) dma16(W + (ch * r * n + (size_t)a * n + j0) * IL, nb + a * SPX * CE, cnt * CE, lane);
        return r * dma16_count(cnt * CE);
    };
    // The W rows of the correction's first blocks are copied during the backward sweep's last
    // block when their buffers lie below that block's buffer (the only one still in use).
    const int npre = min(cahead, nblk);
    const bool early_w = npre * cbuf <= ((nblk - 1 + brot) % bnb) * bbuf;
    int early_c = 0;   // (wave 1) instructions of the second early block
    bwd_sweep<T, WU, SPX, RB>(Uf, Vt, lds, region, ch, n, r, US, wl, NL, wl, NL, live, wave, lane, mlx, t, 2, true,
                              [&] {
                                  if (!early_w) return;
                                  cor_copy_f(0, lds);
                                  if (npre > 1) early_c = cor_copy_f(1, lds + cbuf);
                              },
                              brot, pre_u);

    // ---- z = S^-1 t, then x = y - W z: block [W: RB x SPX (rows >= r: zero pads)][y: SPX lane
    // slots]; x leaves by segments through the stages
    if (wave == 0) {
        T z[MAXR];   // the pad entries t[a >= r] are dropped: W z must not see them
#pragma unroll
        for (int a = 0; a < MAXR; ++a) z[a] = cn::zero<T>();
#pragma unroll
        for (int a = 0; a < RB; ++a) z[a] = cn::sel(a < r, t[a], cn::zero<T>());
        if (live && r > 0) {
            if constexpr (RB <= 2) cap_solve_reg<T, RB>(z, Sr, Pr, r);
            else cap_solve<T>(z, S + (size_t)m * r * r, spiv + (size_t)m * r, r);
        }
#pragma unroll
        for (int a = 0; a < RB; ++a) t[a] = z[a];
        for (int q = r * SPX * IL + lane; q < RB * SPX * IL; q += 64) {   // zero pads of W in every buffer
            for (int u = 0; u < cnb; ++u) *reinterpret_cast<T*>(lds + u * cbuf + q * (int)sizeof(T)) = cn::zero<T>();
        }
    }
    T* xseg = xout + seg_off + se;
    auto cor_copy_l = [&](int bn, char* nb) -> int { // wave 2: the work-space rows y of block bn
        const int j0 = bn * SPX;
        if (live) {
#pragma unroll
            for (int s = 0; s < SPX; ++s)
                if (j0 + s < n) LaneSlot<T>::load(wl + (size_t)(j0 + s) * NL, nb + cW + s * GB);
        }
        return __ballot(live) != 0 ? min(SPX, n - j0) * LaneSlot<T>::LOADS : 0;
    };
    auto cor_store = [&](int bs) {                // x entries of block bs from its stage, whole segment runs
        const char* sb = stage + (bs & 1) * SB;
        const int j0 = bs * SPX;
        if (j0 + se < n) {
#pragma unroll
            for (int i = 0; i < NI; ++i)
                if ((SG::SPI / NR) * i < seg_live) xseg[i * seg_step + j0] = LaneSlot<T>::read(sb + i * SG::PITCH, lane);
        }
    };
    if (wave == 1) {
        if (early_w) {
            vm_wait_le(early_c);   // block 0 complete
        } else {
            cor_copy_f(0, lds);
            vm_wait_le(cahead > 1 && nblk > 1 ? cor_copy_f(1, lds + cbuf) : 0);
        }
    } else if (wave == 2) {
        // y of blocks 0 and 1: the backward sweep's last two blocks, from its stages (backward
        // block nblk-1-f in stage (nblk-1-f) & 1, slot s = row jhi - s)
        for (int f = 0; f < min(2, nblk); ++f) {
            const int j0 = f * SPX, cnt = min(SPX, n - j0);
            const char* sf = stage + ((nblk - 1 - f) & 1) * SB;
            char* dst = lds + (f % cnb) * cbuf + cW;
#pragma unroll
            for (int s = 0; s < SPX; ++s)
                if (s < cnt) LaneSlot<T>::write(dst + s * GB, lane, LaneSlot<T>::read(sf + (cnt - 1 - s) * GB, lane));
        }
    }
    wg_sync();
    for (int b = 0; b < nblk; ++b) {
        if (wave == 1) {
            const int c = b + cahead < nblk ? cor_copy_f(b + cahead, lds + ((b + cahead) % cnb) * cbuf) : 0;
            vm_wait_le(cahead > 1 ? c : 0);
        } else if (wave == 2) {
            const int c = b + cahead < nblk && b + cahead >= 2 ? cor_copy_l(b + cahead, lds + ((b + cahead) % cnb) * cbuf) : 0;
            vm_wait_le(cahead > 1 ? c : 0);
        } else if (wave == 3) {
            if (b >= 1) cor_store(b - 1);
        } else {
            const char* cb = lds + (b % cnb) * cbuf;
            char* sb = stage + (b & 1) * SB + my_slot;
#pragma unroll
            for (int s = 0; s < SPX; ++s) {
                T v = LaneSlot<T>::read(cb + cW + s * GB, lane);
#pragma unroll
                for (int a = 0; a < RB; ++a) v = cn::fnms(lds_at<T>(cb + (a * SPX + s) * CE + mlx * (int)sizeof(T)), t[a], v);
                LaneSlot<T>::write(sb, SG::pos(lane, s), v);
            }
        }
        wg_sync();
    }
    if (wave == 3) {
        cor_store(nblk - 1);
        dma_wait();
    }
}

// Streamed completion of W = B^-1 U (in place over its forward half) and S = I + V^T W: lane =
// (system, border column a = blockIdx.y * NR + k), the backward sweep of k_solve_s (four waves).
template <class T, int WU, int NR, int SPX, int RB>
__global__ void __launch_bounds__(SWEEP_THREADS) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_border_s(const T* __restrict__ Uf, const T* __restrict__ Vt, T* W, T* __restrict__ S, int B, int n, int r, int US,
           int bufbytes) {
    extern __shared__ __align__(16) char lds[];
    const int wave = wave_id();
    const int lane = threadIdx.x & 63;
    const int mloc = lane / NR, k = lane - mloc * NR;
    const int mlx = mloc & (IL - 1);
    const size_t ch = blockIdx.x;
    const int m = (int)ch * IL + mloc;
    const int a = blockIdx.y * NR + k;
    const bool live = mloc < IL && m < B && a < r;
    T* wa = W + (ch * r * n + (size_t)(live ? a : 0) * n) * IL + mlx;
    T t[RB];
#pragma unroll
    for (int q = 0; q < RB; ++q) t[q] = cn::zero<T>();
    bwd_sweep<T, WU, SPX, RB>(Uf, Vt, lds, bufbytes, ch, n, r, US, wa, IL, wa, IL, live, wave, lane, mlx, t, 0, false,
                              [] {});
    if (wave != 0 || !live) return;
    T* Sm = S + (size_t)m * r * r + (size_t)a * r;
#pragma unroll
    for (int q = 0; q < RB; ++q)
        if (q < r) Sm[q] = (q == a) ? cn::fmac(cn::one<T>(), cn::one<T>(), t[q]) : t[q];
}

// ------------------------------------------------------------------------------------------
// Wave mode: one wavefront per (system, right-hand side), for few or wide-band systems.
// Factors are stored per system contiguously (no interleave) and U by columns, scaled by the
// column's pivot after the factorization: Uc[j][0] = 1 / U(j, j), Uc[j][d] = U(j-d, j) / U(j, j)
// (d >= 1).  The right-hand side, the forward result and the solution stay in LDS (n entries).
// The sweeps keep the band window in the low lanes (lane 0 = the row being finished) and the
// next rows in the lanes above; a step is one broadcast of lane 0 (v_readlane), one
// multiply-add per window lane and one wave rotation (DPP wave_rol:1), which brings the next row
// into the window and parks the finished row in lane 63:
//   forward (w = kl+1 lanes): the gbtrs column order, pivot exchange of lanes 0 and p;
//   backward (w = kl+ku+1 lanes): the column-oriented order of LAPACK's banded triangular solve,
//   with the division by U(j, j) applied when the finished rows are flushed.
// The L rows + pivots (forward) and U columns (backward) of the next block of SP steps are
// copied HBM -> LDS while the current block is computed.  t = V^T y, z = S^-1 t and
// x = y - W z are parallel over the entries of x.
// ------------------------------------------------------------------------------------------
// Steps per staged block in wave mode, a template parameter SP of the sweeps: 32 for few systems
// (fewer block hand-overs per sweep), 8 when the batch fills the chip (three buffers of 8 steps
// leave LDS for 6 one-wavefront workgroups per CU at kl = 12, ku = 18, n = 395, instead of 2).
constexpr int SPW_FEW = 32, SPW_MANY = 8;
// Staging buffers of the wave-mode sweeps: three (copies issued two blocks ahead) for few
// wavefronts, two (one block ahead) when many share a CU, where the other wavefronts hide the
// copies' latency and the smaller LDS footprint lets more of them reside.
__host__ __device__ constexpr int wave_nbuf(int SP) { return SP <= SPW_MANY ? 2 : 3; }

// Ordering point of a one-wavefront workgroup: LDS operations of a wavefront execute in order,
// and the staging copies are waited for explicitly (vm_wait_le), so only the compiler has to be
// kept from moving memory operations across it (a workgroup barrier would also drain the copies
// in flight).
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// yv[j] = src[j * st] for j < n (batched loads)
template <class T>
__device__ __forceinline__ void wave_load(T* yv, const T* __restrict__ src, size_t st, int n, int lane) {
    for (int j0 = 0; j0 < n; j0 += 256) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = src[(size_t)min(j0 + 64 * u + lane, n - 1) * st];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (j0 + 64 * u + lane < n) yv[j0 + 64 * u + lane] = v[u];
    }
}


__device__ __forceinline__ double bcast(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ zc bcast(zc v, int src) { return zc{bcast(v.re, src), bcast(v.im, src)}; }

__device__ __forceinline__ double xor_shfl(double v, int mask) {
    const int lo = __shfl_xor(__double2loint(v), mask, 64);
    const int hi = __shfl_xor(__double2hiint(v), mask, 64);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ zc xor_shfl(zc v, int mask) { return zc{xor_shfl(v.re, mask), xor_shfl(v.im, mask)}; }
__device__ __forceinline__ double addv(double a, double b) { return a + b; }
__device__ __forceinline__ zc addv(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }

// DPP move of a double (both halves); lanes the control does not write (row_mask) receive 0
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}
// Sum over the 64 lanes, in every lane: pairs, quads (quad_perm), half rows (row_half_mirror),
// rows of 16 (row_mirror), then row 0 into row 1 and row 2 into row 3 (row_bcast:15), rows 0-1
// into rows 2-3 (row_bcast:31), read from lane 63.  Moves within the VALU, no LDS round trips.
__device__ __forceinline__ double wave_total(double v) {
    v += dpp_d<0xB1, 0xf>(v);    // quad_perm [1,0,3,2]
    v += dpp_d<0x4E, 0xf>(v);    // quad_perm [2,3,0,1]
    v += dpp_d<0x141, 0xf>(v);   // row_half_mirror
    v += dpp_d<0x140, 0xf>(v);   // row_mirror
    v += dpp_d<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_d<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ zc wave_total(zc v) { return zc{wave_total(v.re), wave_total(v.im)}; }

// contiguous copy of nbytes (multiple of 4) into LDS, 256 bytes per instruction
__device__ __forceinline__ void dma4(const void* src_, char* lds_dst, int nbytes, int lane) {
    const char* src = static_cast<const char*>(src_);
    for (int off = 0; off < nbytes; off += 256) {
        const int my = off + lane * 4;
        if (my < nbytes)
            __builtin_amdgcn_global_load_lds(static_cast<const void*>(src + my), (lds_ptr)(lds_dst + off), 4, 0, 0);
    }
}
// entries of T: 16-byte copies for complex (always 16-byte aligned), dwords for real
template <class T>
__device__ __forceinline__ void dmaT(const T* src, char* dst, int count, int lane) {
    if constexpr (sizeof(T) == 16) dma16(src, dst, count * 16, lane);
    else dma4(src, dst, count * (int)sizeof(T), lane);
}

__host__ __device__ constexpr int align16(int b) { return (b + 15) / 16 * 16; }
// LDS after the two vectors: zero slot, capacitance LU (MAXR^2 entries) and its pivots
constexpr int WAVE_SMALL = 16 + MAXR * MAXR * 16 + MAXR * 4 + 16;
// LDS bytes of one staging buffer of the wave-mode sweeps
__host__ __device__ constexpr int wave_buf_bytes(int KL, int US, int esz, int SP) {
    return align16(SP * KL * esz + SP * 4) > align16(SP * US * esz) ? align16(SP * KL * esz + SP * 4)
                                                                    : align16(SP * US * esz);
}

// number of copy instructions dmaT / dma4 issue for `count` entries / bytes
template <class T>
__device__ __forceinline__ int dmaT_count(int count) {
    return sizeof(T) == 16 ? (count * 16 + 1023) / 1024 : (count * (int)sizeof(T) + 255) / 256;
}
__device__ __forceinline__ int dma4_count(int nbytes) { return (nbytes + 255) / 256; }


// Staging copies of one block (returns the number of copy instructions issued)
template <class T, int SP>
__device__ __forceinline__ int wave_stage_fwd(const T* __restrict__ L0, const int* __restrict__ P0, int n, int KL,
                                              int blk, char* buf, int lane) {
    const int j0 = blk * SP, cnt = min(SP, n - j0);
    dmaT<T>(L0 + (size_t)j0 * KL, buf, cnt * KL, lane);
    dma4(P0 + j0, buf + SP * KL * (int)sizeof(T), cnt * 4, lane);
    return dmaT_count<T>(cnt * KL) + dma4_count(cnt * 4);
}
template <class T, int SP>
__device__ __forceinline__ int wave_stage_bwd(const T* __restrict__ Uc0, int n, int US, int blk, char* buf, int lane) {
    const int jhi = n - 1 - blk * SP, jlo = max(0, jhi - SP + 1);
    dmaT<T>(Uc0 + (size_t)jlo * US, buf, (jhi - jlo + 1) * US, lane);
    return dmaT_count<T>((jhi - jlo + 1) * US);
}

// One-lane rotation toward lane 0 across the whole wavefront (DPP wave_rol:1): lane i receives
// lane i+1, lane 63 receives lane 0.
__device__ __forceinline__ int rol1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x134, 0xf, 0xf, false); }
__device__ __forceinline__ double rol1(double v) {
    return __hiloint2double(rol1(__double2hiint(v)), rol1(__double2loint(v)));
}
__device__ __forceinline__ zc rol1(zc v) { return zc{rol1(v.re), rol1(v.im)}; }

// Forward sweep in place on yv (n entries): y_c = L^-1 P_B c.
// Lanes 0 .. kl hold the window rows j .. j+kl (lane 0: pivot row), the lanes above hold the next
// rows of c in order.  A step: pivot exchange of lanes 0 and p (selects, a no-op for p = 0),
// broadcast of the pivot row, one multiply-add per lane (multiplier 0 outside the window: those
// lanes read a zero slot), and one wave rotation, which moves the next row into the window and the
// finished row to lane 63.  Steps run in branch-free groups of SB <= 64 - w; after a group the
// lanes 64-SB .. 63 hold its finished rows: they are stored, and the lanes above the window are
// reloaded with the next rows of c.  The multipliers of a block of SP steps are read from LDS
// into registers before the block.  A partial last block runs step by step (SB = 0: every block).
// A chunk of a partitioned sweep passes nrow = n + KL (the KL rows after its n steps are in yv)
// and keep = true: those rows, still in the window at the end, are stored back.
template <class T>
__device__ __forceinline__ void fwd_step(T& y, const T& l, int p, int lane) {
    const T v0 = bcast(y, 0), vp = bcast(y, p);
    y = cn::sel(lane == p, v0, y);
    y = cn::sel(lane == 0, vp, y);
    y = cn::fnms(l, vp, y);
    y = rol1(y);
}

template <class T, int SB, int SP>
__device__ __forceinline__ void wave_forward(const T* __restrict__ L0, const int* __restrict__ P0, T* yv,
                                             const char* zslot, char* bufs, int bufbytes, int n, int KL, int lane,
                                             int nrow = -1, bool keep = false) {
    constexpr int ES = (int)sizeof(T);
    const int w = KL + 1;
    const int R = 64 - w;
    const int nblk = (n + SP - 1) / SP;
    const int nr = nrow < 0 ? n : nrow;   // rows present in yv (n steps; a chunk also has KL rows after them)
    const bool upd = lane >= 1 && lane <= KL;
    const bool above = lane >= w;
    T y = lane < nr ? yv[lane] : cn::zero<T>();
    int t = 0, jf = 0;            // steps since the last flush, first row finished since then
    const int Poff = SP * KL * ES;
    // NB staging buffers, copies issued NB - 1 blocks ahead
    constexpr int NB = wave_nbuf(SP), AH = NB - 1;
    wave_stage_fwd<T, SP>(L0, P0, n, KL, 0, bufs, lane);
    const int c1 = AH > 1 && nblk > 1 ? wave_stage_fwd<T, SP>(L0, P0, n, KL, 1, bufs + bufbytes, lane) : 0;
    vm_wait_le(c1);
    wave_sync();
    for (int b = 0; b < nblk; ++b) {
        char* cur = bufs + (b % NB) * bufbytes;
        const int c2 = b + AH < nblk ? wave_stage_fwd<T, SP>(L0, P0, n, KL, b + AH, bufs + ((b + AH) % NB) * bufbytes, lane)
                                     : 0;
        const int j0 = b * SP, cnt = min(SP, n - j0);
        const int pv = lane < cnt ? lds_at<int>(cur + Poff + lane * 4) : 0;   // pivot of step lane
        const char* lb = upd ? cur + (lane - 1) * ES : zslot;
        const int lst = upd ? KL * ES : 0;
        if (SB > 0 && cnt == SP) {
            T lv[SP];
#pragma unroll
            for (int sidx = 0; sidx < SP; ++sidx) lv[sidx] = lds_at<T>(lb + sidx * lst);
#pragma unroll
            for (int q = 0; q < (SB > 0 ? SP / SB : 1); ++q) {
#pragma unroll
                for (int u = 0; u < SB; ++u) fwd_step<T>(y, lv[q * SB + u], __builtin_amdgcn_readlane(pv, q * SB + u), lane);
                const int jn = j0 + (q + 1) * SB;     // first row of the new window
                if (lane >= 64 - SB) yv[jf + lane - (64 - SB)] = y;
                if (above) y = (jn + lane < nr) ? yv[jn + lane] : cn::zero<T>();
                jf += SB;
            }
        } else {
            for (int sidx = 0; sidx < cnt; ++sidx) {
                fwd_step<T>(y, lds_at<T>(lb + sidx * lst), __builtin_amdgcn_readlane(pv, sidx), lane);
                if (++t == R) {       // lane w + i: finished row jf + i
                    const int jn = j0 + sidx + 1;
                    if (above) {
                        yv[jf + lane - w] = y;
                        y = (jn + lane < nr) ? yv[jn + lane] : cn::zero<T>();
                    }
                    jf += R;
                    t = 0;
                }
            }
        }
        vm_wait_le(AH > 1 ? c2 : 0);   // block b+1 has landed
        wave_sync();
    }
    if (lane >= 64 - t) yv[jf + lane - (64 - t)] = y;   // the last t finished rows
    if (keep && lane < KL && n + lane < nr) yv[n + lane] = y;   // the rows still in the window
    wave_sync();
}

// Backward sweep in place on yv: y = U^-1 y_c with U stored by columns and scaled by the pivot of
// the column, Uc[j][d] = U(j-d, j) / U(j, j) (d >= 1), Uc[j][0] = 1 / U(j, j).  Lanes 0 .. w-1
// hold rows j, j-1, .., j-w+1, the lanes above the next rows of y_c.  A step: broadcast of the
// finished row value y_j (lane 0), every window lane subtracts Uc[j][d] y_j, one wave rotation
// (y_j moves to lane 63).  Groups, flushes and the partial last block as in the forward sweep;
// the division x_j = y_j / U(j, j) is applied to the finished rows when they are stored, from the
// column's entry 0 (1 / U(j, j)), read from the staged block when the row is in it (always for the
// groups of full blocks) and from the factors otherwise.  A chunk of a partitioned sweep passes
// rlo = -(w - 1) (the rows before its first row are in yv at negative indices) and keep = true:
// those rows, still in the window at the end, are stored back.
template <class T>
__device__ __forceinline__ void bwd_step(T& y, const T& u, int lane) {
    const T yj = bcast(y, 0);
    y = cn::fnms(u, yj, y);
    y = rol1(y);
}

template <class T, int SB, int SP>
__device__ __forceinline__ void wave_backward(const T* __restrict__ Uc0, T* yv, const char* zslot, char* bufs,
                                              int bufbytes, int n, int w, int US, int lane, int rlo = 0,
                                              bool keep = false) {
    constexpr int ES = (int)sizeof(T);
    const int R = 64 - w;
    const int nblk = (n + SP - 1) / SP;
    const bool upd = lane >= 1 && lane < w;
    const bool above = lane >= w;
    // 1 / U(row, row): entry 0 of column `row`, staged at (row - jlo) * US when jlo <= row <= jhi
    auto diag = [&](const char* cur, int jlo, int jhi, int row) -> T {
        return (row >= jlo && row <= jhi) ? lds_at<T>(cur + (row - jlo) * US * ES) : Uc0[(size_t)row * US];
    };
    T y = (n - 1 - lane >= rlo) ? yv[n - 1 - lane] : cn::zero<T>();
    int t = 0, jf = n - 1;        // steps since the last flush, first row finished since then
    constexpr int NB = wave_nbuf(SP), AH = NB - 1;
    wave_stage_bwd<T, SP>(Uc0, n, US, 0, bufs, lane);
    const int c1 = AH > 1 && nblk > 1 ? wave_stage_bwd<T, SP>(Uc0, n, US, 1, bufs + bufbytes, lane) : 0;
    vm_wait_le(c1);
    wave_sync();
    for (int b = 0; b < nblk; ++b) {
        char* cur = bufs + (b % NB) * bufbytes;
        const int c2 = b + AH < nblk ? wave_stage_bwd<T, SP>(Uc0, n, US, b + AH, bufs + ((b + AH) % NB) * bufbytes, lane)
                                     : 0;
        const int jhi = n - 1 - b * SP, jlo = max(0, jhi - SP + 1);
        const int cnt = jhi - jlo + 1;
        const char* ub = upd ? cur + lane * ES : zslot;
        const int ust = upd ? US * ES : 0;
        if (SB > 0 && cnt == SP) {
            T uv[SP];
#pragma unroll
            for (int sidx = 0; sidx < SP; ++sidx) uv[sidx] = lds_at<T>(ub + (SP - 1 - sidx) * ust);
#pragma unroll
            for (int q = 0; q < (SB > 0 ? SP / SB : 1); ++q) {
#pragma unroll
                for (int u = 0; u < SB; ++u) bwd_step<T>(y, uv[q * SB + u], lane);
                const int jn = jhi - (q + 1) * SB;    // row of lane 0 in the new window
                if (lane >= 64 - SB) {   // rows of this block: their columns are staged in cur
                    const int row = jf - (lane - (64 - SB));
                    yv[row] = cn::mul(y, lds_at<T>(cur + (row - jlo) * US * ES));
                }
                if (above) y = (jn - lane >= rlo) ? yv[jn - lane] : cn::zero<T>();
                jf -= SB;
            }
        } else {
            for (int sidx = 0; sidx < cnt; ++sidx) {
                bwd_step<T>(y, lds_at<T>(ub + (cnt - 1 - sidx) * ust), lane);
                if (++t == R) {       // lane w + i: finished row jf - i
                    const int jn = jhi - sidx - 1;
                    if (above) {
                        const int row = jf - (lane - w);
                        yv[row] = cn::mul(y, diag(cur, jlo, jhi, row));
                        y = (jn - lane >= rlo) ? yv[jn - lane] : cn::zero<T>();
                    }
                    jf -= R;
                    t = 0;
                }
            }
        }
        vm_wait_le(AH > 1 ? c2 : 0);   // block b+1 has landed
        wave_sync();
    }
    if (lane >= 64 - t) {   // the last block's buffer is still staged
        const int jhi = n - 1 - (nblk - 1) * SP, jlo = max(0, jhi - SP + 1);
        const int row = jf - (lane - (64 - t));
        yv[row] = cn::mul(y, diag(bufs + ((nblk - 1) % NB) * bufbytes, jlo, jhi, row));
    }
    if (keep && lane < w - 1 && -1 - lane >= rlo) yv[-1 - lane] = y;   // rows -1, -2, .. still in the window
    wave_sync();
}

// group size of the backward sweep for a window of w lanes (0: step by step)
__host__ __device__ constexpr int bwd_group(int w, int SP) {
    return (SP >= 32 && w <= 32) ? 32 : (SP >= 16 && w <= 48) ? 16 : w <= 56 ? 8 : 0;
}
static_assert(SPW_FEW % 32 == 0 && SPW_MANY % 8 == 0, "the step groups divide a staged block");

template <class T, int SP>
__device__ __forceinline__ void wave_backward_any(const T* __restrict__ Uc0, T* yv, const char* zslot,
                                                  char* bufs, int bufbytes, int n, int w, int US, int lane,
                                                  int rlo = 0, bool keep = false) {
    switch (bwd_group(w, SP)) {
        case 32:
            if constexpr (SP >= 32) wave_backward<T, 32, SP>(Uc0, yv, zslot, bufs, bufbytes, n, w, US, lane, rlo, keep);
            break;
        case 16:
            if constexpr (SP >= 16) wave_backward<T, 16, SP>(Uc0, yv, zslot, bufs, bufbytes, n, w, US, lane, rlo, keep);
            break;
        case 8: wave_backward<T, 8, SP>(Uc0, yv, zslot, bufs, bufbytes, n, w, US, lane, rlo, keep); break;
        default: wave_backward<T, 0, SP>(Uc0, yv, zslot, bufs, bufbytes, n, w, US, lane, rlo, keep); break;
    }
}

// Loads of the parallel phases are batched (every load of a chunk of 128 entries issued before
// the first use), so that a phase costs a few memory latencies, not one per entry.
constexpr int WCH = 2;   // 64-entry groups per chunk

// t[a] = sum_j V^T(a, j) y_j over the wave (every lane ends with the totals)
template <class T>
__device__ __forceinline__ void wave_vty(const T* __restrict__ V0, const T* yv, int n, int r, int lane, T (&t)[MAXR]) {
#pragma unroll
    for (int a = 0; a < MAXR; ++a) t[a] = cn::zero<T>();
    for (int j0 = 0; j0 < n; j0 += 64 * WCH) {
        T vv[WCH][MAXR];
#pragma unroll
        for (int u = 0; u < WCH; ++u) {
            const int j = min(j0 + 64 * u + lane, n - 1);
#pragma unroll
            for (int a = 0; a < MAXR; ++a) vv[u][a] = V0[(size_t)min(a, max(r - 1, 0)) * n + j];
        }
#pragma unroll
        for (int u = 0; u < WCH; ++u) {
            const int j = j0 + 64 * u + lane;
            const T yj = j < n ? yv[j] : cn::zero<T>();
#pragma unroll
            for (int a = 0; a < MAXR; ++a)
                if (a < r) t[a] = cn::fmac(vv[u][a], yj, t[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < MAXR; ++a)
        if (a < r) t[a] = wave_total(t[a]);
}

// Solve in wave mode: grid (B, nrhs), one wavefront each.  LDS: [yv: n entries][zero slot, S,
// pivots of S][3 staging buffers].
template <class T, int SP>
__global__ void __launch_bounds__(64)
k_solve_w(const T* __restrict__ Uc, const T* __restrict__ Lf, const int* __restrict__ piv,
          const T* __restrict__ Vt, const T* __restrict__ W, const T* __restrict__ S,
          const int* __restrict__ spiv, const int* __restrict__ perm, SolveIO io, int n, int r, int KL, int w2,
          int US, int bufbytes) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x;
    const int m = blockIdx.x, k = blockIdx.y;
    const size_t mn = (size_t)m * n;
    T* yv = reinterpret_cast<T*>(lds);
    char* zslot = lds + align16(n * (int)sizeof(T));
    T* Sl = reinterpret_cast<T*>(zslot + 16);                  // capacitance LU (r x r) and pivots
    int* spl = reinterpret_cast<int*>(Sl + MAXR * MAXR);
    char* bufs = zslot + WAVE_SMALL;
    if (lane < r * r) Sl[lane] = S[(size_t)m * r * r + lane];
    if (lane < r) spl[lane] = spiv[(size_t)m * r + lane];
    for (int j0 = 0; j0 < n; j0 += 256) {   // c = P b (batched loads)
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = io_load_b<T>(io, mn, k, perm[min(j0 + 64 * u + lane, n - 1)]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (j0 + 64 * u + lane < n) yv[j0 + 64 * u + lane] = v[u];
    }
    if (lane == 0) *reinterpret_cast<zc*>(zslot) = cn::zero<zc>();
    wave_sync();
    wave_forward<T, (SP < 32 ? SP : 32), SP>(Lf + mn * KL, piv + mn, yv, zslot, bufs, bufbytes, n, KL, lane);   // kl + 1 <= 17 lanes
    wave_backward_any<T, SP>(Uc + mn * US, yv, zslot, bufs, bufbytes, n, w2, US, lane);
    if (r == 0) {
        for (int j = lane; j < n; j += 64) io_store<T>(io, mn, k, j, yv[j]);
        return;
    }
    const T* V0 = Vt + (size_t)m * r * n;
    const T* W0 = W + (size_t)m * r * n;
    T t[MAXR];
    wave_vty<T>(V0, yv, n, r, lane, t);
    cap_solve<T>(t, Sl, spl, r);
    for (int j0 = 0; j0 < n; j0 += 64 * WCH) {   // x = y - W z (batched loads)
        T wv[WCH][MAXR];
#pragma unroll
        for (int u = 0; u < WCH; ++u) {
            const int j = min(j0 + 64 * u + lane, n - 1);
#pragma unroll
            for (int a = 0; a < MAXR; ++a) wv[u][a] = W0[(size_t)min(a, max(r - 1, 0)) * n + j];
        }
#pragma unroll
        for (int u = 0; u < WCH; ++u) {
            const int j = j0 + 64 * u + lane;
            if (j < n) {
                T v = yv[j];
#pragma unroll
                for (int a = 0; a < MAXR; ++a)
                    if (a < r) v = cn::fnms(wv[u][a], t[a], v);
                io_store<T>(io, mn, k, j, v);
            }
        }
    }
}

// Border completion in wave mode: grid (B, r); W row a = U^-1 (its forward half), in place, and
// column a of S = I + V^T W.
template <class T, int SP>
__global__ void __launch_bounds__(64)
k_border_w(const T* __restrict__ Uc, const T* __restrict__ Vt, T* __restrict__ W, T* __restrict__ S, int n, int r,
           int w2, int US, int bufbytes) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x;
    const int m = blockIdx.x, a = blockIdx.y;
    T* yv = reinterpret_cast<T*>(lds);
    char* zslot = lds + align16(n * (int)sizeof(T));
    char* bufs = zslot + WAVE_SMALL;
    T* Wa = W + ((size_t)m * r + a) * n;
    wave_load<T>(yv, Wa, 1, n, lane);
    if (lane == 0) *reinterpret_cast<zc*>(zslot) = cn::zero<zc>();
    wave_sync();
    wave_backward_any<T, SP>(Uc + (size_t)m * n * US, yv, zslot, bufs, bufbytes, n, w2, US, lane);
    for (int j = lane; j < n; j += 64) Wa[j] = yv[j];
    T t[MAXR];
    wave_vty<T>(Vt + (size_t)m * r * n, yv, n, r, lane, t);
    if (lane == 0) {
        T* Sm = S + (size_t)m * r * r + (size_t)a * r;
#pragma unroll
        for (int q = 0; q < MAXR; ++q)
            if (q < r) Sm[q] = (q == a) ? cn::fmac(cn::one<T>(), cn::one<T>(), t[q]) : t[q];
    }
}

// ------------------------------------------------------------------------------------------
// Partitioned wave mode: the sweeps of one system split over P wavefronts of one workgroup, for
// launches with fewer systems than the chip has SIMDs, where one wavefront per system leaves the
// sweep a chain of n dependent steps on a single SIMD.  Chunk c covers rows [a_c, a_{c+1}) with
// a_c = c n / P; every chunk is at least as long as the coupling widths (KL rows ahead of the
// forward sweep, WB = M - 1 rows behind the backward sweep).  Each wavefront sweeps its chunk
// with the coupling rows it would receive from its neighbour set to zero; the sweeps are linear,
// so the true result is that local result plus the chunk's response to the coupling values,
// taken from the factor-time spikes:
//   SpL[m][c][u][i], i < LS:  the forward sweep of chunk c from a unit value at head row u
//                             (rows 0 .. Lc-1 of the chunk, then the KL rows after it);
//   SpU[m][c][u][i], i < LSU: the backward sweep of chunk c from a unit value at row Lc - WB + u
//                             (i = row + WB: the WB rows before the chunk, then its Lc rows).
// The coupling values come from a short chain over the chunks (q_{c+1} = out_c + T_c q_c with
// T_c the outgoing part of the spikes, KL x KL forward and WB x WB backward), then every chunk
// adds its spikes times q_c to its rows.  The arithmetic of a system depends on (n, kl, ku, r)
// only, never on the batch, so a system solves bit for bit alike in any batch of this mode.
__host__ __device__ constexpr int part_lo(int c, int n, int P) { return (int)(((long long)c * n) / P); }
constexpr int PART_W = 32;   // largest coupling width (KL or WB) of the partitioned mode

// Spikes of one (system, chunk, unit): grid (B, P, max(KL, WB)), one wavefront each.  LDS: [yw:
// ywbytes][zero slot][NB staging buffers].
template <class T, int SP>
__global__ void __launch_bounds__(64)
k_spikes_w(const T* __restrict__ Uc, const T* __restrict__ Lf, const int* __restrict__ piv, T* __restrict__ SpL,
           T* __restrict__ SpU, int n, int P, int KL, int M, int US, int LS, int LSU, int ywbytes, int bufbytes) {
    extern __shared__ __align__(16) char lds[];
    const int lane = threadIdx.x;
    const int m = blockIdx.x, c = blockIdx.y, u = blockIdx.z;
    const int WB = M - 1;
    const int a = part_lo(c, n, P), Lc = part_lo(c + 1, n, P) - a;
    const size_t mn = (size_t)m * n;
    T* yw = reinterpret_cast<T*>(lds);
    char* zslot = lds + ywbytes;
    char* bufs = zslot + 16;
    if (lane == 0) *reinterpret_cast<zc*>(zslot) = cn::zero<zc>();
    if (c >= 1 && u < KL) {
        for (int i = lane; i < Lc + KL; i += 64) yw[i] = (i == u) ? cn::one<T>() : cn::zero<T>();
        wave_sync();
        wave_forward<T, (SP < 32 ? SP : 32), SP>(Lf + (mn + a) * KL, piv + mn + a, yw, zslot, bufs, bufbytes, Lc, KL,
                                               lane, Lc + KL, true);
        T* dst = SpL + (((size_t)m * P + c) * KL + u) * LS;
        for (int i = lane; i < Lc + KL; i += 64) dst[i] = yw[i];
        wave_sync();
    }
    if (c <= P - 2 && u < WB) {
        for (int i = lane; i < WB + Lc; i += 64) yw[i] = (i == Lc + u) ? cn::one<T>() : cn::zero<T>();
        wave_sync();
        wave_backward_any<T, SP>(Uc + (mn + a) * US, yw + WB, zslot, bufs, bufbytes, Lc, M, US, lane, -WB, true);
        T* dst = SpU + (((size_t)m * P + c) * WB + u) * LSU;
        for (int i = lane; i < WB + Lc; i += 64) dst[i] = yw[i];
    }
}

// acc + sum_{u < wdt} m[u] q[u]: the q values (one LDS slot of PART_W entries, zero beyond wdt, as
// m is) read eight at a time, so that a group costs one LDS latency, into four partial sums (term u
// into sum u mod 4, joined as acc + ((s0 + s1) + (s2 + s3))), so that the multiply-adds do not form
// one dependent chain
template <class T>
__device__ __forceinline__ T part_dot(const T (&m)[PART_W], const T* q, int wdt, T acc) {
    T ps[4] = {cn::zero<T>(), cn::zero<T>(), cn::zero<T>(), cn::zero<T>()};
#pragma unroll
    for (int u0 = 0; u0 < PART_W; u0 += 8) {
        if (u0 < wdt) {
            T qq[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) qq[j] = q[u0 + j];
#pragma unroll
            for (int j = 0; j < 8; ++j) ps[j & 3] = cn::fmac(m[u0 + j], qq[j], ps[j & 3]);
        }
    }
    return addv(acc, addv(addv(ps[0], ps[1]), addv(ps[2], ps[3])));
}

// z = S^-1 t with t across the lanes (lane i holds t_i, i < r): the capacitance LU (column-major
// in LDS, pivots sp) read once, row `lane` per lane; row interchanges in order, then the unit lower
// and the upper triangle by columns, the finished entry broadcast from its lane at each step
template <class T>
__device__ __forceinline__ T cap_solve_lanes(T ti, const T* Sm, const int* sp, int r, int lane) {
    T row[MAXR];
#pragma unroll
    for (int q = 0; q < MAXR; ++q) row[q] = (q < r && lane < r) ? Sm[q * r + lane] : cn::zero<T>();
    int pv[MAXR];
#pragma unroll
    for (int i = 0; i < MAXR; ++i) pv[i] = i < r ? sp[i] : i;
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
        if (i < r && pv[i] != i) {
            const T vi = bcast(ti, i), vp = bcast(ti, pv[i]);
            ti = cn::sel(lane == i, vp, cn::sel(lane == pv[i], vi, ti));
        }
    }
#pragma unroll
    for (int q = 0; q < MAXR; ++q) {
        if (q < r) {
            const T v = bcast(ti, q);
            if (lane > q) ti = cn::fnms(row[q], v, ti);
        }
    }
#pragma unroll
    for (int q = MAXR - 1; q >= 0; --q) {
        if (q < r) {
            if (lane == q) ti = cn::mul(ti, cn::recip(row[q]));
            const T v = bcast(ti, q);
            if (lane < q) ti = cn::fnms(row[q], v, ti);
        }
    }
    return ti;
}

// One step of a coupling chain, its w outputs over lane pairs: lane i and lane i + 32 (i < w) hold
// the terms [0, 16) and [16, 32) of output i (m: that half of row i of T_c, zero beyond w; q: the
// LDS slot of the incoming coupling values, zero beyond w); lane i returns base_i + (its half) +
// (the other half).
constexpr int PART_H = PART_W / 2;
template <class T>
__device__ __forceinline__ T part_chain_step(const T (&m)[PART_H], const T* q, T base, int lane) {
    const int half = lane >> 5;
    T ps[4] = {half ? cn::zero<T>() : base, cn::zero<T>(), cn::zero<T>(), cn::zero<T>()};
    T qq[PART_H];
#pragma unroll
    for (int j = 0; j < PART_H; ++j) qq[j] = q[half * PART_H + j];
#pragma unroll
    for (int j = 0; j < PART_H; ++j) ps[j & 3] = cn::fmac(m[j], qq[j], ps[j & 3]);
    const T mine = addv(addv(ps[0], ps[1]), addv(ps[2], ps[3]));
    const T other = xor_shfl(mine, 32);
    return half ? addv(other, mine) : addv(mine, other);
}

// rows [0, Lc) of a chunk: v_i += sum_u Sp[u][off + i] q[u] (u < wdt), written to dst[i]
template <class T>
__device__ __forceinline__ void part_fixup(const T* __restrict__ Sp, size_t ustride, int off, int wdt, const T* q,
                                           const T* src, T* dst, int Lc, bool spike, int lane) {
    for (int i0 = 0; i0 < Lc; i0 += 64) {
        const int i = i0 + lane, ic = min(i, Lc - 1);
        T v = src[ic];
        if (spike) {
            T sv[PART_W];
#pragma unroll
            for (int u = 0; u < PART_W; ++u) sv[u] = u < wdt ? Sp[(size_t)u * ustride + off + ic] : cn::zero<T>();
            v = part_dot<T>(sv, q, wdt, v);
        }
        if (i < Lc) dst[i] = v;
    }
}

// Partitioned solve: grid (B, nrhs), P wavefronts per workgroup (wavefront c: chunk c).  LDS:
// [y: n entries][zero slot, S, pivots of S][q: P x PART_W][partial V^T y: P x MAXR]
// [per wavefront: yw (ywbytes), NB staging buffers].
template <class T, int SP>
__global__ void __launch_bounds__(512)
k_solve_p(const T* __restrict__ Uc, const T* __restrict__ Lf, const int* __restrict__ piv,
          const T* __restrict__ Vt, const T* __restrict__ W, const T* __restrict__ S,
          const int* __restrict__ spiv, const int* __restrict__ perm, const T* __restrict__ SpL,
          const T* __restrict__ SpU, SolveIO io, int n, int r, int KL, int M, int US, int P, int LS, int LSU,
          int ywbytes, int bufbytes) {
    extern __shared__ __align__(16) char lds[];
    constexpr int ES = (int)sizeof(T);
    constexpr int NB = wave_nbuf(SP);
    const int lane = threadIdx.x & 63, c = threadIdx.x >> 6;
    const int m = blockIdx.x, k = blockIdx.y;
    const size_t mn = (size_t)m * n;
    const int WB = M - 1;
    const int a = part_lo(c, n, P), e = part_lo(c + 1, n, P), Lc = e - a;
    T* yv = reinterpret_cast<T*>(lds);
    char* zslot = lds + align16(n * ES);
    T* Sl = reinterpret_cast<T*>(zslot + 16);
    int* spl = reinterpret_cast<int*>(Sl + MAXR * MAXR);
    T* qv = reinterpret_cast<T*>(zslot + WAVE_SMALL);
    T* part = qv + P * PART_W;
    char* mine = reinterpret_cast<char*>(part + P * MAXR) + (size_t)c * (ywbytes + NB * bufbytes);
    T* yw = reinterpret_cast<T*>(mine);
    char* bufs = mine + ywbytes;
    const T* SpLc = SpL + ((size_t)m * P + c) * KL * LS;
    const T* SpUc = SpU + ((size_t)m * P + c) * WB * LSU;
    if (c == 0) {
        if (lane < r * r) Sl[lane] = S[(size_t)m * r * r + lane];
        if (lane < r) spl[lane] = spiv[(size_t)m * r + lane];
        if (lane == 0) *reinterpret_cast<zc*>(zslot) = cn::zero<zc>();
    }
    for (int i = lane; i < PART_W; i += 64) qv[c * PART_W + i] = cn::zero<T>();   // padding of the q slots
    const T* V0 = Vt + (size_t)m * r * n;
    const T* W0 = W + (size_t)m * r * n;
    for (int j0 = a; j0 < e; j0 += 256) {   // c = P b, this chunk's rows (batched loads)
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = io_load_b<T>(io, mn, k, perm[min(j0 + 64 * u + lane, e - 1)]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (j0 + 64 * u + lane < e) yv[j0 + 64 * u + lane] = v[u];
    }
    __syncthreads();
    // forward: the chunk and the KL rows after it, head rows zero (they arrive through q)
    for (int i = lane; i < Lc + KL; i += 64) {
        const int row = a + i;
        yw[i] = ((c > 0 && i < KL) || row >= n) ? cn::zero<T>() : yv[row];
    }
    wave_sync();
    wave_forward<T, (SP < 32 ? SP : 32), SP>(Lf + (mn + a) * KL, piv + mn + a, yw, zslot, bufs, bufbytes, Lc, KL,
                                           lane, Lc + KL, true);
    __syncthreads();
    const bool chain = c >= 1 && c <= P - 2;
    const int crow = lane & 31, chalf = lane >> 5;   // chain output and half of its terms
    T tr[PART_H];   // half `chalf` of row `crow` of T_c (forward, then backward)
#pragma unroll
    for (int i = 0; i < PART_H; ++i) {
        const int u = chalf * PART_H + i;
        tr[i] = (chain && u < KL && crow < KL) ? SpLc[(size_t)u * LS + Lc + crow] : cn::zero<T>();
    }
    if (c == 0 && lane < KL) qv[PART_W + lane] = yw[Lc + lane];
    __syncthreads();
    for (int st = 1; st <= P - 2; ++st) {   // q_{st+1} = out_st + T_st q_st
        if (c == st) {
            const T v = part_chain_step<T>(tr, qv + st * PART_W, crow < KL ? yw[Lc + crow] : cn::zero<T>(), lane);
            if (lane < KL) qv[(st + 1) * PART_W + lane] = v;
        }
        __syncthreads();
    }
    part_fixup<T>(SpLc, LS, 0, KL, qv + c * PART_W, yw, yv + a, Lc, c >= 1, lane);
    __syncthreads();
    // backward: the WB rows before the chunk (zero: they leave through q) and the chunk
    for (int i = lane; i < WB + Lc; i += 64) yw[i] = i < WB ? cn::zero<T>() : yv[a + i - WB];
    wave_sync();
    wave_backward_any<T, SP>(Uc + (mn + a) * US, yw + WB, zslot, bufs, bufbytes, Lc, M, US, lane, -WB, true);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PART_H; ++i) {
        const int u = chalf * PART_H + i;
        tr[i] = (chain && u < WB && crow < WB) ? SpUc[(size_t)u * LSU + crow] : cn::zero<T>();
    }
    if (c == P - 1 && lane < WB) qv[(P - 2) * PART_W + lane] = yw[lane];
    __syncthreads();
    for (int st = P - 2; st >= 1; --st) {   // q_{st-1} = out_st + T_st q_st
        if (c == st) {
            const T v = part_chain_step<T>(tr, qv + st * PART_W, crow < WB ? yw[crow] : cn::zero<T>(), lane);
            if (lane < WB) qv[(st - 1) * PART_W + lane] = v;
        }
        __syncthreads();
    }
    part_fixup<T>(SpUc, LSU, WB, WB, qv + c * PART_W, yw + WB, yv + a, Lc, c <= P - 2, lane);
    wave_sync();
    if (r == 0) {
        for (int i = lane; i < Lc; i += 64) io_store<T>(io, mn, k, a + i, yv[a + i]);
        return;
    }
    // Woodbury correction over this chunk's rows: partial V^T y, summed over the chunks in order
    T t[MAXR];
#pragma unroll
    for (int q = 0; q < MAXR; ++q) t[q] = cn::zero<T>();
    for (int i0 = 0; i0 < Lc; i0 += 64) {
        const int i = i0 + lane, j = a + min(i, Lc - 1);
        T vv[MAXR];
#pragma unroll
        for (int q = 0; q < MAXR; ++q) vv[q] = V0[(size_t)min(q, r - 1) * n + j];
        const T yj = i < Lc ? yv[j] : cn::zero<T>();
#pragma unroll
        for (int q = 0; q < MAXR; ++q)
            if (q < r) t[q] = cn::fmac(vv[q], yj, t[q]);
    }
#pragma unroll
    for (int q = 0; q < MAXR; ++q)
        if (q < r) t[q] = wave_total(t[q]);
    // W rows of the chunk's first 64 rows, loaded now: their latency passes during the
    // capacitance solve
    T wpre[MAXR];
    {
        const int j = a + min(lane, Lc - 1);
#pragma unroll
        for (int q = 0; q < MAXR; ++q) wpre[q] = W0[(size_t)min(q, r - 1) * n + j];
    }
    if (lane < r) {
#pragma unroll
        for (int q = 0; q < MAXR; ++q)
            if (q == lane) part[c * MAXR + q] = t[q];
    }
    __syncthreads();
    // lane q < r: t_q summed over the chunks in order (all partials read at once), then z = S^-1 t
    // across the lanes and broadcast back
    T tq = cn::zero<T>();
    if (lane < r) {
        T pp[8];
#pragma unroll
        for (int cc = 0; cc < 8; ++cc) pp[cc] = cc < P ? part[cc * MAXR + lane] : cn::zero<T>();
        tq = pp[0];
#pragma unroll
        for (int cc = 1; cc < 8; ++cc)
            if (cc < P) tq = addv(tq, pp[cc]);
    }
    tq = cap_solve_lanes<T>(tq, Sl, spl, r, lane);
#pragma unroll
    for (int q = 0; q < MAXR; ++q) t[q] = q < r ? bcast(tq, q) : cn::zero<T>();
    for (int i0 = 0; i0 < Lc; i0 += 64) {   // x = y - W z
        const int i = i0 + lane, j = a + min(i, Lc - 1);
        T wv[MAXR];
#pragma unroll
        for (int q = 0; q < MAXR; ++q) wv[q] = i0 == 0 ? wpre[q] : W0[(size_t)min(q, r - 1) * n + j];
        if (i < Lc) {
            T v = yv[j];
#pragma unroll
            for (int q = 0; q < MAXR; ++q)
                if (q < r) v = cn::fnms(wv[q], t[q], v);
            io_store<T>(io, mn, k, j, v);
        }
    }
}

// Wave mode, after the band LU: the columns of U scaled by their pivots, Uc[j][d] *= Uc[j][0]
// (= 1 / U(j, j)) for d >= 1.
template <class T>
__global__ void __launch_bounds__(256)
k_scale_ucol(T* __restrict__ Uc, size_t total, int US) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const int d = (int)(e % US);
    if (d == 0) return;
    Uc[e] = cn::mul(Uc[e], Uc[e - d]);
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
          T* __restrict__ x, int64_t rhs_stride, int B, int n, int r, int US, int nrhs, int prep, int il,
          int ucol) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = gid / nrhs;
    const int k = gid - m * nrhs;
    if (m >= B) return;
    const size_t ch = m / il;
    const int ml = m % il;
    const T* U0 = Uf + ch * ((size_t)n * US * il) + ml;
    const T* L0 = Lf + ch * ((size_t)n * KL * il) + ml;
    const int* P0 = piv + ch * ((size_t)n * il) + ml;
    // input row, work row (stride between consecutive entries)
    const T* in;
    size_t in_st;
    T* w;
    size_t w_st;
    if (prep) {
        in = Vt + ch * ((size_t)r * n * il) + ml + (size_t)k * n * il;
        in_st = il;
        w = Wt + ch * ((size_t)r * n * il) + ml + (size_t)k * n * il;
        w_st = il;
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
        T wj;
        if (!ucol) {               // row j-d entry d = U(j-d, j)
            T acc = in[(size_t)j * in_st];
#pragma unroll
            for (int d = 1; d <= WU; ++d)
                if (j - d >= 0) acc = cn::fnms(U0[((size_t)(j - d) * US + d) * il], uw[d - 1], acc);
            wj = cn::mul(acc, U0[(size_t)j * US * il]);
        } else {                   // column j entry d = U(j-d, j) / U(j, j)
            T acc = cn::mul(in[(size_t)j * in_st], U0[(size_t)j * US * il]);
#pragma unroll
            for (int d = 1; d <= WU; ++d)
                if (j - d >= 0) acc = cn::fnms(U0[((size_t)j * US + d) * il], uw[d - 1], acc);
            wj = acc;
        }
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
        const T* Lr = L0 + (size_t)j * KL * il;
#pragma unroll
        for (int i = 1; i <= KL; ++i) v = cn::fnms(Lr[(i - 1) * il], y[i], v);
        y[0] = v;
        const int p = P0[(size_t)j * il];
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
    const T* W0 = Wt + ch * ((size_t)r * n * il) + ml;
    T* xo = x + k * rhs_stride + (size_t)m * n;
    for (int i = 0; i < n; ++i) {
        T v = w[(size_t)i * w_st];
#pragma unroll
        for (int a = 0; a < MAXR; ++a)
            if (a < r) v = cn::fnms(W0[((size_t)a * n + i) * il], z[a], v);
        xo[perm[i]] = v;
    }
}

// St = I + U^T Wt (column-major r x r per system): St(a, b) = delta_ab + Wt[b][a].
template <class T>
__global__ void __launch_bounds__(256)
k_build_st(const T* __restrict__ Wt, T* __restrict__ St, int* __restrict__ info_t, int B, int n, int r, int il) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = gid / (r * r);
    const int q = gid - m * r * r;
    if (m >= B) return;
    if (q == 0) info_t[m] = 0;
    const int b = q / r, a = q - b * r;
    const size_t ch = m / il;
    const int ml = m % il;
    T v = Wt[(ch * ((size_t)r * n) + (size_t)b * n + a) * il + ml];
    if (a == b) v = cn::fmac(cn::one<T>(), cn::one<T>(), v);
    St[(size_t)m * r * r + (size_t)b * r + a] = v;
}


}  // namespace cnk
