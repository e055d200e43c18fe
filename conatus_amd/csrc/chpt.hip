// libcnhip: explicit term and SBDF stage of the 3D plane channel in poloidal-toroidal form.
//
// State (float64): the phi bucket [4][P][nP] then the psi bucket [4][P][nS], nP = N + 4,
// nS = N + 2, slot s = kx * Kzh + |kz|, columns 0/1 = (Re, Im) of mode (kx, +|kz|) and 2/3 = of
// (kx, -|kz|); slot 0 of psi holds the mean flow U (column 0) and W (column 1).  Modes are
// indexed (kx, iz), iz < Kz = 2 Kzh - 1 in FFT order (kz = iz, or iz - Kz for iz >= Kzh).
//
// Explicit term F(X) (the reference's channel_3d_pt right-hand sides,
// tests/channel3d/ch3d_cases.py:176-191: the six products u_i u_j on the 3/2 grid, N = -div(u u),
// the wall-normal curl curl and curl of N, the x-z means of Nx and Nz), six launches:
//   A  y-backward per 16 modes:  H1[kx][j][3][iz]  = u, v, w at the Gauss node y_j, from
//      u = i kx D phi + i kz psi, v = k^2 phi, w = i kz D phi - i kx psi (mean: U, 0, W)
//   B  z-backward, one (kx, y_j) per workgroup: H2[j][3][kx][m]  (inverse FFT of length Mz)
//   C  x-lines, 8 (y_j, z_m) per workgroup:   H3[j][6][kx][m]  = uu, uv, uw, vv, vw, ww
//      (two real lines per complex inverse FFT of length Mx, products, two real lines per
//      forward FFT, kx < Kx kept)
//   D  z-forward, one (kx, y_j) per workgroup: H4[kx][j][6][iz] (forward FFT, kz represented)
//   E  y-forward + the curl forms per 16 modes, with the dense matrices (N x My) of
//      A1 = S14 D01 Fy, A2 = S24 D02 Fy, A0 = S04 Fy, B0 = S02 Fy, B1 = S12 D01 Fy (Fy the
//      Gauss-Chebyshev forward transform), from the node combinations
//        a = kx^2 P_uu + 2 kx kz P_uw + kz^2 P_ww - k^2 P_vv,  q = i (kx P_uv + kz P_vw),
//        b = kx kz (P_uu - P_ww) + (kz^2 - kx^2) P_uw,        s = i (kz P_uv - kx P_vw):
//        F_phi = A1 a - A2 q - A0 (k^2 q)   (C^(4) rows),   F_psi = B0 b - B1 s   (C^(2) rows)
//      (wavenumbers times kx0, kz0); the kx = 0 plane is written as conjugate pairs.
//   M  the mean flow: F_U = -B1 P_uv + G S02 e0, F_W = -B1 P_vw (real parts), one workgroup.
// Every F entry is written (boundary rows and the unused columns of kz = 0 slots are zero).
// FFTs are Stockham autosort passes of radix 4, 2 and 3 in LDS (lengths 2^a 3^b), twiddles
// W_k = exp(-2 pi i k / M) from a table; inverse transforms carry no 1/M, forward ones do.
//
// SBDF stage: one workgroup per system line (bucket, column, slot): the line of X in LDS,
// M X from the banded stencils of the operators (phi: M = -k^2 S24 D02 + k^4 S04, psi: -k^2 S02,
// the mean: S02; the (0, 0) phi pin: 0), the right-hand side sum_{j=1..q} (c_j F_j - (a_j/dt)
// M X_j) rounded as the reference's (imex.hpp:158-176), written over X for the in-place solve.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cnhip.h"
#include "cn_common.h"

using cn::zc;

namespace {

constexpr int MODES = 16;   // modes per workgroup of the y transforms (A, E)
constexpr int XL = 8;       // (y_j, z_m) lines per workgroup of the x-line kernel (C)
constexpr int JC = 16;      // nodes per chunk of the y-forward (E)
constexpr int MAXST = 16;   // FFT passes at most

struct Plan {
    int n, r[MAXST];
};

struct Dims {
    int Kx, Kzh, Kz, N, nP, nS, P, Mx, My, Mz;
    double kx0, kz0, G;
    size_t h1() const { return (size_t)Kx * My * 3 * Kz; }
    size_t h2() const { return (size_t)My * 3 * Kx * Mz; }
    size_t h3() const { return (size_t)My * 6 * Kx * Mz; }
    size_t h4() const { return (size_t)Kx * My * 6 * Kz; }
    size_t region1() const { return h1() > h3() ? h1() : h3(); }
    size_t region2() const { return h2() > h4() ? h2() : h4(); }
    size_t bytes() const { return (region1() + region2()) * sizeof(zc); }
    size_t bytes_separate() const { return (h1() + h2() + h3() + h4()) * sizeof(zc); }
    __host__ __device__ size_t size_phi() const { return (size_t)4 * P * nP; }
};

__device__ __forceinline__ zc zadd(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ zc zsub(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ zc zscale(double s, zc a) { return zc{s * a.re, s * a.im}; }
__device__ __forceinline__ zc zmuli(zc a) { return zc{-a.im, a.re}; }            // i a
__device__ __forceinline__ zc zconj(zc a) { return zc{a.re, -a.im}; }
__device__ __forceinline__ zc zmul(zc a, zc b) { return zc{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// mode index iz -> signed kz
__device__ __forceinline__ int kz_of(int iz, int Kzh, int Kz) { return iz < Kzh ? iz : iz - Kz; }
// FFT bin of mode kz on a grid of M points
__device__ __forceinline__ int bin_of(int kz, int M) { return kz >= 0 ? kz : M + kz; }

// ---- Stockham FFT in LDS ------------------------------------------------------------------
// One radix-R pass over nl lines of length M (a -> b); Ns = product of the radices of the passes
// before; forward: y_k = sum_r v_r exp(-2 pi i r k / R) after the twiddles exp(-2 pi i r (j mod Ns) / (Ns R)).
template <int R, bool INV>
__device__ void fft_pass(const zc* __restrict__ a, zc* __restrict__ b, int M, int nl, int Ns,
                         const zc* __restrict__ tw) {
    const int per = M / R;
    const int step = M / (Ns * R);
    for (int q = threadIdx.x; q < nl * per; q += blockDim.x) {
        const int line = q / per, j = q - line * per;
        const zc* src = a + (size_t)line * M;
        zc* dst = b + (size_t)line * M;
        const int k = j % Ns;
        zc v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src[j + r * per];
#pragma unroll
        for (int r = 1; r < R; ++r) {
            const zc w = tw[k * r * step];
            v[r] = zmul(v[r], INV ? zconj(w) : w);
        }
        if constexpr (R == 2) {
            const zc t = v[1];
            v[1] = zsub(v[0], t);
            v[0] = zadd(v[0], t);
        } else if constexpr (R == 4) {
            const zc s02 = zadd(v[0], v[2]), d02 = zsub(v[0], v[2]);
            const zc s13 = zadd(v[1], v[3]), d13 = zsub(v[1], v[3]);
            const zc id13 = INV ? zmuli(d13) : zscale(-1.0, zmuli(d13));   // -i d13 (forward), +i d13 (inverse)
            v[0] = zadd(s02, s13);
            v[2] = zsub(s02, s13);
            v[1] = zadd(d02, id13);
            v[3] = zsub(d02, id13);
        } else {   // R == 3
            constexpr double H = 0.86602540378443864676;                  // sqrt(3) / 2
            const zc t1 = zadd(v[1], v[2]);
            const zc t2 = zsub(v[0], zscale(0.5, t1));
            const zc d = zscale(H, zsub(v[1], v[2]));
            const zc t3 = INV ? zmuli(d) : zscale(-1.0, zmuli(d));
            v[0] = zadd(v[0], t1);
            v[1] = zadd(t2, t3);
            v[2] = zsub(t2, t3);
        }
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
    }
}

// The whole transform of nl lines held in a (b: the second buffer); returns the buffer holding
// the result.  Every thread of the workgroup calls it; a __syncthreads() follows every pass.
template <bool INV>
__device__ zc* fft_lines(zc* a, zc* b, int M, int nl, const zc* tw, const Plan& p) {
    int Ns = 1;
    for (int s = 0; s < p.n; ++s) {
        const int R = p.r[s];
        if (R == 4) fft_pass<4, INV>(a, b, M, nl, Ns, tw);
        else if (R == 2) fft_pass<2, INV>(a, b, M, nl, Ns, tw);
        else fft_pass<3, INV>(a, b, M, nl, Ns, tw);
        __syncthreads();
        Ns *= R;
        zc* t = a;
        a = b;
        b = t;
    }
    return a;
}

// ---- A: y-backward ------------------------------------------------------------------------
// LDS: phi and psi coefficients of 16 modes, [2][N][16].  Thread: modes (2 mp, 2 mp + 1),
// nodes j = g, g + 32, ...
__global__ void __launch_bounds__(256)
k_pt_yback(const double* __restrict__ X, const double* __restrict__ Tb, zc* __restrict__ H1, Dims d) {
    extern __shared__ __align__(16) char lds[];
    zc* cf = reinterpret_cast<zc*>(lds);                  // [2][N][MODES]
    const int kx = blockIdx.y, iz0 = blockIdx.x * MODES, tid = threadIdx.x, N = d.N;
    const double* Xpsi = X + d.size_phi();
    for (int q = tid; q < 2 * MODES * N; q += blockDim.x) {
        const int n = q % N, rest = q / N, l = rest % MODES, fld = rest / MODES;
        const int iz = iz0 + l;
        zc v{0.0, 0.0};
        if (iz < d.Kz) {
            const int kz = kz_of(iz, d.Kzh, d.Kz);
            const int s = kx * d.Kzh + (kz >= 0 ? kz : -kz), c = kz >= 0 ? 0 : 2;
            const bool mean = (kx == 0 && iz == 0);
            if (fld == 0) {
                if (!mean) {
                    const double* p = X + ((size_t)c * d.P + s) * d.nP + n;
                    v = zc{p[0], p[(size_t)d.P * d.nP]};
                }
            } else {
                const double* p = Xpsi + ((size_t)c * d.P + s) * d.nS + n;
                v = zc{p[0], p[(size_t)d.P * d.nS]};
            }
        }
        cf[((size_t)fld * N + n) * MODES + l] = v;
    }
    __syncthreads();
    const int mp = tid & 7, g = tid >> 3;
    const int l0 = 2 * mp;
    double ax[2], az[2], k2[2];
    bool mean[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int iz = iz0 + l0 + u;
        const int kz = kz_of(iz < d.Kz ? iz : 0, d.Kzh, d.Kz);
        ax[u] = kx * d.kx0;
        az[u] = kz * d.kz0;
        k2[u] = ax[u] * ax[u] + az[u] * az[u];
        mean[u] = (kx == 0 && iz == 0);
    }
    for (int j = g; j < d.My; j += 32) {
        const double* tb = Tb + (size_t)j * N * 2;
        zc f[2] = {{0.0, 0.0}, {0.0, 0.0}}, df[2] = {{0.0, 0.0}, {0.0, 0.0}}, ps[2] = {{0.0, 0.0}, {0.0, 0.0}};
        for (int n = 0; n < N; ++n) {
            const double t = tb[2 * n], tp = tb[2 * n + 1];
            const zc* c0 = cf + (size_t)n * MODES + l0;
            const zc* c1 = cf + ((size_t)N + n) * MODES + l0;
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const zc a = c0[u], b = c1[u];
                f[u].re = fma(t, a.re, f[u].re);
                f[u].im = fma(t, a.im, f[u].im);
                df[u].re = fma(tp, a.re, df[u].re);
                df[u].im = fma(tp, a.im, df[u].im);
                ps[u].re = fma(t, b.re, ps[u].re);
                ps[u].im = fma(t, b.im, ps[u].im);
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int iz = iz0 + l0 + u;
            if (iz >= d.Kz) continue;
            zc uu, vv, ww;
            if (mean[u]) {
                uu = zc{ps[u].re, 0.0};
                vv = zc{0.0, 0.0};
                ww = zc{ps[u].im, 0.0};
            } else {
                // u = i (ax Df + az ps), v = k2 f, w = i (az Df - ax ps)
                const zc su = zc{ax[u] * df[u].re + az[u] * ps[u].re, ax[u] * df[u].im + az[u] * ps[u].im};
                const zc sw = zc{az[u] * df[u].re - ax[u] * ps[u].re, az[u] * df[u].im - ax[u] * ps[u].im};
                uu = zmuli(su);
                vv = zscale(k2[u], f[u]);
                ww = zmuli(sw);
            }
            zc* o = H1 + (((size_t)kx * d.My + j) * 3) * d.Kz + iz;
            o[0] = uu;
            o[d.Kz] = vv;
            o[2 * (size_t)d.Kz] = ww;
        }
    }
}

// ---- B: z-backward ------------------------------------------------------------------------
__global__ void __launch_bounds__(64)
k_pt_zback(const zc* __restrict__ H1, const zc* __restrict__ Wz, zc* __restrict__ H2, Dims d, Plan pz) {
    extern __shared__ __align__(16) char lds[];
    const int Mz = d.Mz;
    zc* tw = reinterpret_cast<zc*>(lds);
    zc* a = tw + Mz;
    zc* b = a + 3 * Mz;
    const int j = blockIdx.x, kx = blockIdx.y, tid = threadIdx.x;
    for (int q = tid; q < Mz; q += blockDim.x) tw[q] = Wz[q];
    for (int q = tid; q < 3 * Mz; q += blockDim.x) a[q] = zc{0.0, 0.0};
    __syncthreads();
    const zc* in = H1 + (((size_t)kx * d.My + j) * 3) * d.Kz;
    for (int q = tid; q < 3 * d.Kz; q += blockDim.x) {
        const int f = q / d.Kz, iz = q - f * d.Kz;
        a[f * Mz + bin_of(kz_of(iz, d.Kzh, d.Kz), Mz)] = in[q];
    }
    __syncthreads();
    const zc* res = fft_lines<true>(a, b, Mz, 3, tw, pz);
    for (int q = tid; q < 3 * Mz; q += blockDim.x) {
        const int f = q / Mz, m = q - f * Mz;
        H2[(((size_t)j * 3 + f) * d.Kx + kx) * Mz + m] = res[q];
    }
}

// ---- C: x-lines ---------------------------------------------------------------------------
// LDS: tw[Mx] | in[3][Kx][XL] | phys[3][XL][Mx] (doubles) | two line buffers [XL][Mx]
__global__ void __launch_bounds__(256)
k_pt_xline(const zc* __restrict__ H2, const zc* __restrict__ Wx, zc* __restrict__ H3, Dims d, Plan px) {
    extern __shared__ __align__(16) char lds[];
    const int Mx = d.Mx, Kx = d.Kx, Mz = d.Mz;
    zc* tw = reinterpret_cast<zc*>(lds);
    zc* in = tw + Mx;
    double* ph = reinterpret_cast<double*>(in + 3 * Kx * XL);
    zc* la = reinterpret_cast<zc*>(ph + 3 * XL * Mx);
    zc* lb = la + XL * Mx;
    const int m0 = blockIdx.x * XL, j = blockIdx.y, tid = threadIdx.x;
    const int nz = (Mz - m0) < XL ? (Mz - m0) : XL;
    for (int q = tid; q < Mx; q += blockDim.x) tw[q] = Wx[q];
    for (int q = tid; q < 3 * Kx * XL; q += blockDim.x) {
        const int mm = q % XL, rest = q / XL, k = rest % Kx, f = rest / Kx;
        in[q] = mm < nz ? H2[(((size_t)j * 3 + f) * Kx + k) * Mz + m0 + mm] : zc{0.0, 0.0};
    }
    __syncthreads();
    // phase 1: u + i v per line; phase 2: w of line pairs (2p, 2p + 1)
    for (int phase = 0; phase < 2; ++phase) {
        const int nl = phase == 0 ? XL : XL / 2;
        for (int q = tid; q < nl * Mx; q += blockDim.x) {
            const int line = q / Mx, k = q - line * Mx;
            zc A{0.0, 0.0}, Bv{0.0, 0.0};
            bool conj = false;
            int kk = -1;
            if (k < Kx) kk = k;
            else if (k > Mx - Kx) { kk = Mx - k; conj = true; }
            if (kk >= 0) {
                const zc* sa = phase == 0 ? in + (size_t)(0 * Kx + kk) * XL + line : in + (size_t)(2 * Kx + kk) * XL + 2 * line;
                const zc* sb = phase == 0 ? in + (size_t)(1 * Kx + kk) * XL + line : sa + 1;
                A = sa[0];
                Bv = sb[0];
                if (kk == 0) { A.im = 0.0; Bv.im = 0.0; }
                if (conj) { A = zconj(A); Bv = zconj(Bv); }
            }
            la[q] = zadd(A, zmuli(Bv));
        }
        __syncthreads();
        const zc* res = fft_lines<true>(la, lb, Mx, nl, tw, px);
        for (int q = tid; q < nl * Mx; q += blockDim.x) {
            const int line = q / Mx, x = q - line * Mx;
            if (phase == 0) {
                ph[(0 * XL + line) * Mx + x] = res[q].re;
                ph[(1 * XL + line) * Mx + x] = res[q].im;
            } else {
                ph[(2 * XL + 2 * line) * Mx + x] = res[q].re;
                ph[(2 * XL + 2 * line + 1) * Mx + x] = res[q].im;
            }
        }
        __syncthreads();
    }
    // products, two per forward transform: (uu, uv), (uw, vv), (vw, ww)
    const double inv = 1.0 / Mx;
    for (int pair = 0; pair < 3; ++pair) {
        const int f1 = 2 * pair, f2 = 2 * pair + 1;          // product indices 0 uu 1 uv 2 uw 3 vv 4 vw 5 ww
        for (int q = tid; q < XL * Mx; q += blockDim.x) {
            const int line = q / Mx, x = q - line * Mx;
            const double u = ph[(0 * XL + line) * Mx + x], v = ph[(1 * XL + line) * Mx + x],
                         w = ph[(2 * XL + line) * Mx + x];
            la[q] = pair == 0 ? zc{u * u, u * v} : pair == 1 ? zc{u * w, v * v} : zc{v * w, w * w};
        }
        __syncthreads();
        const zc* res = fft_lines<false>(la, lb, Mx, XL, tw, px);
        for (int q = tid; q < XL * Kx; q += blockDim.x) {
            const int mm = q % XL, k = q / XL;
            const zc Z = res[(size_t)mm * Mx + k];
            const zc Zc = zconj(res[(size_t)mm * Mx + (k == 0 ? 0 : Mx - k)]);
            const zc Af = zscale(0.5 * inv, zadd(Z, Zc));
            const zc Bf = zscale(-0.5 * inv, zmuli(zsub(Z, Zc)));   // (Z - conj Z_-k) / (2 i)
            if (mm < nz) {
                H3[(((size_t)j * 6 + f1) * Kx + k) * Mz + m0 + mm] = Af;
                H3[(((size_t)j * 6 + f2) * Kx + k) * Mz + m0 + mm] = Bf;
            }
        }
        __syncthreads();
    }
}

// ---- D: z-forward -------------------------------------------------------------------------
__global__ void __launch_bounds__(128)
k_pt_zfwd(const zc* __restrict__ H3, const zc* __restrict__ Wz, zc* __restrict__ H4, Dims d, Plan pz) {
    extern __shared__ __align__(16) char lds[];
    const int Mz = d.Mz;
    zc* tw = reinterpret_cast<zc*>(lds);
    zc* a = tw + Mz;
    zc* b = a + 6 * Mz;
    const int j = blockIdx.x, kx = blockIdx.y, tid = threadIdx.x;
    for (int q = tid; q < Mz; q += blockDim.x) tw[q] = Wz[q];
    for (int q = tid; q < 6 * Mz; q += blockDim.x) {
        const int f = q / Mz, m = q - f * Mz;
        a[q] = H3[(((size_t)j * 6 + f) * d.Kx + kx) * Mz + m];
    }
    __syncthreads();
    const zc* res = fft_lines<false>(a, b, Mz, 6, tw, pz);
    const double inv = 1.0 / Mz;
    zc* out = H4 + (((size_t)kx * d.My + j) * 6) * d.Kz;
    for (int q = tid; q < 6 * d.Kz; q += blockDim.x) {
        const int f = q / d.Kz, iz = q - f * d.Kz;
        out[q] = zscale(inv, res[f * Mz + bin_of(kz_of(iz, d.Kzh, d.Kz), Mz)]);
    }
}

// ---- E: y-forward and the curl forms ------------------------------------------------------
// LDS: combinations [JC][MODES][5] (a, q, k^2 q, b, s) | output rows [MODES][2][nP + nS] doubles
__global__ void __launch_bounds__(256)
k_pt_yfwd(const zc* __restrict__ H4, const double* __restrict__ Am, double* __restrict__ F, Dims d) {
    extern __shared__ __align__(16) char lds[];
    zc* cmb = reinterpret_cast<zc*>(lds);                        // [JC][MODES][5]
    double* orow = reinterpret_cast<double*>(cmb + JC * MODES * 5);   // [MODES][2 (re, im)][nP + nS]
    const int kx = blockIdx.y, iz0 = blockIdx.x * MODES, tid = threadIdx.x, N = d.N, My = d.My;
    const int nrow = d.nP + d.nS;
    const int l = tid & (MODES - 1), g = tid >> 4;               // 16 row groups
    const int iz = iz0 + l;
    const bool live = iz < d.Kz && !(kx == 0 && iz >= d.Kzh);
    const int kz = kz_of(live ? iz : 0, d.Kzh, d.Kz);
    const double ax = kx * d.kx0, az = kz * d.kz0, k2 = ax * ax + az * az;
    constexpr int RPT = 7;                                        // rows per thread (N <= 112)
    zc fphi[RPT], fpsi[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) { fphi[r] = zc{0.0, 0.0}; fpsi[r] = zc{0.0, 0.0}; }
    const size_t msz = (size_t)N * My;
    for (int j0 = 0; j0 < My; j0 += JC) {
        const int nj = (My - j0) < JC ? (My - j0) : JC;
        __syncthreads();
        for (int q = tid; q < JC * MODES; q += blockDim.x) {
            const int ll = q % MODES, jj = q / MODES, izz = iz0 + ll;
            zc* c = cmb + ((size_t)jj * MODES + ll) * 5;
            if (jj >= nj || izz >= d.Kz) {
                for (int t = 0; t < 5; ++t) c[t] = zc{0.0, 0.0};
                continue;
            }
            const int kzz = kz_of(izz, d.Kzh, d.Kz);
            const double bx = kx * d.kx0, bz = kzz * d.kz0, kk = bx * bx + bz * bz;
            const zc* p = H4 + (((size_t)kx * My + j0 + jj) * 6) * d.Kz + izz;
            const zc uu = p[0], uv = p[d.Kz], uw = p[2 * (size_t)d.Kz], vv = p[3 * (size_t)d.Kz],
                     vw = p[4 * (size_t)d.Kz], ww = p[5 * (size_t)d.Kz];
            // a = bx^2 uu + 2 bx bz uw + bz^2 ww - k^2 vv ; q = i (bx uv + bz vw)
            const zc a = zsub(zadd(zadd(zscale(bx * bx, uu), zscale(2.0 * bx * bz, uw)), zscale(bz * bz, ww)),
                              zscale(kk, vv));
            const zc qv = zmuli(zadd(zscale(bx, uv), zscale(bz, vw)));
            const zc bb = zadd(zscale(bx * bz, zsub(uu, ww)), zscale(bz * bz - bx * bx, uw));
            const zc sv = zmuli(zsub(zscale(bz, uv), zscale(bx, vw)));
            c[0] = a;
            c[1] = qv;
            c[2] = zscale(kk, qv);
            c[3] = bb;
            c[4] = sv;
        }
        __syncthreads();
        for (int jj = 0; jj < nj; ++jj) {
            const zc* c = cmb + ((size_t)jj * MODES + l) * 5;
            const zc a = c[0], qv = c[1], kq = c[2], bb = c[3], sv = c[4];
            const int j = j0 + jj;
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                const int m = g + 16 * r;
                if (m >= N) break;
                const double* am = Am + (size_t)m * My + j;
                const double a1 = am[0], a2 = am[msz], a0 = am[2 * msz], b0 = am[3 * msz], b1 = am[4 * msz];
                fphi[r].re = fma(a1, a.re, fphi[r].re);
                fphi[r].im = fma(a1, a.im, fphi[r].im);
                fphi[r].re = fma(-a2, qv.re, fphi[r].re);
                fphi[r].im = fma(-a2, qv.im, fphi[r].im);
                fphi[r].re = fma(-a0, kq.re, fphi[r].re);
                fphi[r].im = fma(-a0, kq.im, fphi[r].im);
                fpsi[r].re = fma(b0, bb.re, fpsi[r].re);
                fpsi[r].im = fma(b0, bb.im, fpsi[r].im);
                fpsi[r].re = fma(-b1, sv.re, fpsi[r].re);
                fpsi[r].im = fma(-b1, sv.im, fpsi[r].im);
            }
        }
    }
    __syncthreads();
    // rows into LDS: [mode][re/im][phi rows (nP) | psi rows (nS)]
    double* ore = orow + (size_t)l * 2 * nrow;
    double* oim = ore + nrow;
    if (g == 0) {
        for (int r = 0; r < 4; ++r) ore[r] = oim[r] = 0.0;
        ore[d.nP] = oim[d.nP] = ore[d.nP + 1] = oim[d.nP + 1] = 0.0;
    }
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int m = g + 16 * r;
        if (m >= N) break;
        ore[4 + m] = fphi[r].re;
        oim[4 + m] = fphi[r].im;
        ore[d.nP + 2 + m] = fpsi[r].re;
        oim[d.nP + 2 + m] = fpsi[r].im;
    }
    __syncthreads();
    // write-out: each mode's rows are contiguous per column; 16 threads per mode
    double* Fphi = F;
    double* Fpsi = F + d.size_phi();
    for (int ll = 0; ll < MODES; ++ll) {
        const int izz = iz0 + ll;
        if (izz >= d.Kz || (kx == 0 && izz >= d.Kzh)) continue;
        const int kzz = kz_of(izz, d.Kzh, d.Kz);
        const int s = kx * d.Kzh + (kzz >= 0 ? kzz : -kzz), c = kzz >= 0 ? 0 : 2;
        const double* sre = orow + (size_t)ll * 2 * nrow;
        const double* sim = sre + nrow;
        const bool mean = (kx == 0 && izz == 0);
        const bool mirror = (kx == 0 && kzz > 0);     // (0, -kz) = conj (0, kz)
        const bool dead = (kzz == 0);                 // no (kx, -0) mode: columns 2, 3 zero
        for (int q = tid; q < nrow; q += blockDim.x) {
            const bool isphi = q < d.nP;
            double* base = isphi ? Fphi + (size_t)s * d.nP + q : Fpsi + (size_t)s * d.nS + (q - d.nP);
            const size_t cs = isphi ? (size_t)d.P * d.nP : (size_t)d.P * d.nS;   // column stride
            double re = sre[q], im = sim[q];
            if (mean) {
                if (isphi) { base[0] = 0.0; base[cs] = 0.0; }   // the (0, 0) pin; psi columns 0, 1: k_pt_mean
                base[2 * cs] = 0.0;
                base[3 * cs] = 0.0;
                continue;
            }
            base[c * cs] = re;
            base[(c + 1) * cs] = im;
            if (mirror) {
                base[2 * cs] = re;
                base[3 * cs] = -im;
            } else if (dead) {
                base[2 * cs] = 0.0;
                base[3 * cs] = 0.0;
            }
        }
    }
}

// ---- M: the mean flow ---------------------------------------------------------------------
__global__ void __launch_bounds__(128)
k_pt_mean(const zc* __restrict__ H4, const double* __restrict__ Am, const double* __restrict__ s02c0,
          double* __restrict__ F, Dims d) {
    const int m = threadIdx.x, N = d.N, My = d.My;
    double* Fpsi = F + d.size_phi();
    const size_t cs = (size_t)d.P * d.nS;
    if (m < 2) { Fpsi[m] = 0.0; Fpsi[cs + m] = 0.0; }
    if (m >= N) return;
    const double* b1 = Am + 4 * (size_t)N * My + (size_t)m * My;
    double fu = 0.0, fw = 0.0;
    for (int j = 0; j < My; ++j) {
        const zc* p = H4 + ((size_t)j * 6) * d.Kz;        // kx = 0, iz = 0
        fu = fma(-b1[j], p[1 * (size_t)d.Kz].re, fu);     // -B1 P_uv
        fw = fma(-b1[j], p[4 * (size_t)d.Kz].re, fw);     // -B1 P_vw
    }
    Fpsi[2 + m] = fu + d.G * s02c0[m];
    Fpsi[cs + 2 + m] = fw;
}

// ---- SBDF stage ---------------------------------------------------------------------------
struct StageArgs {
    const double* f[4];    // F levels, newest first (f[0]: F of the current state)
    double* mx0;           // M X of the current state (written)
    const double* mx[3];   // older M X levels (newest first)
    double c[5], adt[5];   // c_j, a_j / dt for j = 1..q
    int q;
};

// stencils: s04[m][5], s2d[m][5] (offsets 0, 2, .., 8), s02[m][3] (offsets 0, 2, 4)
__global__ void __launch_bounds__(128)
k_pt_sbdf(double* __restrict__ X, const double* __restrict__ s04, const double* __restrict__ s2d,
          const double* __restrict__ s02, StageArgs a, Dims d) {
    __shared__ double xl[128];
    const int s = blockIdx.x, c = blockIdx.y, bucket = blockIdx.z, row = threadIdx.x;
    const int n = bucket == 0 ? d.nP : d.nS;
    const size_t off = bucket == 0 ? ((size_t)c * d.P + s) * d.nP : d.size_phi() + ((size_t)c * d.P + s) * d.nS;
    double* x = X + off;
    if (row < n) xl[row] = x[row];
    __syncthreads();
    if (row >= n) return;
    const int kx = s / d.Kzh, kzv = s - kx * d.Kzh;
    const double ax = kx * d.kx0, az = kzv * d.kz0, k2 = ax * ax + az * az;
    double mx = 0.0;
    const int nb = bucket == 0 ? 4 : 2;
    if (row >= nb) {
        const int m = row - nb, N = d.N;
        if (bucket == 0) {
            if (s != 0) {
                const double k4 = k2 * k2;
                for (int t = 0; t < 5; ++t) {
                    const int col = m + 2 * t;
                    if (col >= N) break;
                    const double coef = (-k2) * s2d[m * 5 + t] + k4 * s04[m * 5 + t];
                    mx = mx + coef * xl[col];
                }
            }
        } else {
            for (int t = 0; t < 3; ++t) {
                const int col = m + 2 * t;
                if (col >= N) break;
                const double coef = s == 0 ? s02[m * 3 + t] : (-k2) * s02[m * 3 + t];
                mx = mx + coef * xl[col];
            }
        }
    }
    a.mx0[off + row] = mx;
    double rhs = 0.0;
    for (int j = 1; j <= a.q; ++j) {
        const double fj = a.f[j - 1][off + row];
        const double mj = j == 1 ? mx : a.mx[j - 2][off + row];
        const double t1 = a.c[j] * fj, t2 = a.adt[j] * mj;
        rhs = rhs + (t1 - t2);
    }
    x[row] = rhs;
}

Dims make_dims(int Nx, int Ny, int Nz, int Mx, int My, int Mz, double kx0, double kz0, double G) {
    Dims d{};
    d.Kx = Nx / 2;
    d.Kzh = Nz / 2;
    d.Kz = Nz - 1;
    d.N = Ny;
    d.nP = Ny + 4;
    d.nS = Ny + 2;
    d.P = d.Kx * d.Kzh;
    d.Mx = Mx;
    d.My = My;
    d.Mz = Mz;
    d.kx0 = kx0;
    d.kz0 = kz0;
    d.G = G;
    return d;
}

// radix plan of M = 2^a 3^b: fours, then a two, then threes (false if M has another factor)
bool make_plan(int M, Plan& p) {
    p.n = 0;
    int m = M;
    while (m % 4 == 0 && p.n < MAXST) { p.r[p.n++] = 4; m /= 4; }
    if (m % 2 == 0 && p.n < MAXST) { p.r[p.n++] = 2; m /= 2; }
    while (m % 3 == 0 && p.n < MAXST) { p.r[p.n++] = 3; m /= 3; }
    return m == 1 && M >= 2;
}

size_t lds_a(const Dims& d) { return (size_t)2 * d.N * MODES * sizeof(zc); }
size_t lds_b(const Dims& d) { return (size_t)7 * d.Mz * sizeof(zc); }
size_t lds_c(const Dims& d) {
    return ((size_t)d.Mx + 3 * d.Kx * XL + 2 * XL * d.Mx) * sizeof(zc) + (size_t)3 * XL * d.Mx * sizeof(double);
}
size_t lds_d(const Dims& d) { return (size_t)13 * d.Mz * sizeof(zc); }
size_t lds_e(const Dims& d) {
    return (size_t)JC * MODES * 5 * sizeof(zc) + (size_t)MODES * 2 * (d.nP + d.nS) * sizeof(double);
}

}  // namespace

extern "C" {

size_t cn_chpt_workspace_bytes(int Nx, int Ny, int Nz, int Mx, int My, int Mz) {
    if (Nx < 2 || Nz < 2 || Ny < 6 || Mx < 1 || My < 1 || Mz < 1) return 0;
    return make_dims(Nx, Ny, Nz, Mx, My, Mz, 1.0, 1.0, 0.0).bytes();
}

int cn_chpt_explicit(const double* X, double* F, void* ws, size_t ws_bytes, const double* Tb, const double* Am,
                     const double* s02c0, const void* Wx, const void* Wz, int Nx, int Ny, int Nz, int Mx, int My,
                     int Mz, double kx0, double kz0, double G, void* stream) {
    if (!X || !F || !ws || !Tb || !Am || !s02c0 || !Wx || !Wz) {
        cn::set_error("cn_chpt_explicit: null pointer");
        return CN_EINVAL;
    }
    if (Nx < 2 || Nx % 2 || Nz < 2 || Nz % 2 || Ny < 6 || Mx < Nx || My < Ny || Mz < Nz) {
        cn::set_error("cn_chpt_explicit: inconsistent sizes");
        return CN_EINVAL;
    }
    const Dims d = make_dims(Nx, Ny, Nz, Mx, My, Mz, kx0, kz0, G);
    Plan px{}, pz{};
    if (!make_plan(Mx, px) || !make_plan(Mz, pz)) {
        cn::set_error("cn_chpt_explicit: the dealiased Fourier grids must have 2^a 3^b points (got " +
                      std::to_string(Mx) + ", " + std::to_string(Mz) + ")");
        return CN_EUNSUPPORTED;
    }
    if (Ny > 16 * 7) {
        cn::set_error("cn_chpt_explicit: Ny > 112 not supported by the y-forward kernel");
        return CN_EUNSUPPORTED;
    }
    if (d.nP > 128) {
        cn::set_error("cn_chpt_explicit: Ny + 4 > 128 not supported by the stage kernel");
        return CN_EUNSUPPORTED;
    }
    const size_t cap = 160 * 1024;
    if (lds_a(d) > cap || lds_b(d) > cap || lds_c(d) > cap || lds_d(d) > cap || lds_e(d) > cap) {
        cn::set_error("cn_chpt_explicit: grid too large for the transform kernels (LDS)");
        return CN_EUNSUPPORTED;
    }
    if (ws_bytes < d.bytes()) {
        cn::set_error("cn_chpt_explicit: workspace smaller than cn_chpt_workspace_bytes");
        return CN_EINVAL;
    }
    zc* w = static_cast<zc*>(ws);
    zc *h1, *h2, *h3, *h4;
    if (ws_bytes >= d.bytes_separate()) {       // every intermediate in its own region (inspectable)
        h1 = w;
        h2 = h1 + d.h1();
        h3 = h2 + d.h2();
        h4 = h3 + d.h3();
    } else {
        h1 = w;
        h3 = w;
        h2 = w + d.region1();
        h4 = h2;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 gy((d.Kz + MODES - 1) / MODES, d.Kx);
    hipLaunchKernelGGL(k_pt_yback, gy, dim3(256), lds_a(d), s, X, Tb, h1, d);
    hipLaunchKernelGGL(k_pt_zback, dim3(My, d.Kx), dim3(64), lds_b(d), s, h1, static_cast<const zc*>(Wz), h2, d, pz);
    hipLaunchKernelGGL(k_pt_xline, dim3((Mz + XL - 1) / XL, My), dim3(256), lds_c(d), s, h2,
                       static_cast<const zc*>(Wx), h3, d, px);
    hipLaunchKernelGGL(k_pt_zfwd, dim3(My, d.Kx), dim3(128), lds_d(d), s, h3, static_cast<const zc*>(Wz), h4, d, pz);
    hipLaunchKernelGGL(k_pt_yfwd, gy, dim3(256), lds_e(d), s, h4, Am, F, d);
    hipLaunchKernelGGL(k_pt_mean, dim3(1), dim3(128), 0, s, h4, Am, s02c0, F, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        cn::set_error(std::string("cn_chpt_explicit: ") + hipGetErrorString(e));
        return CN_EHIP;
    }
    return CN_OK;
}

int cn_chpt_sbdf_stage(double* X, const double* const* f_levels, double* mx0, const double* const* mx_hist, int q,
                       const double* c, const double* a_over_dt, const double* s04, const double* s2d,
                       const double* s02, int Nx, int Ny, int Nz, double kx0, double kz0, void* stream) {
    if (!X || !f_levels || !mx0 || !c || !a_over_dt || !s04 || !s2d || !s02 || q < 1 || q > 4 ||
        (q > 1 && !mx_hist)) {
        cn::set_error("cn_chpt_sbdf_stage: null pointer or q outside 1..4");
        return CN_EINVAL;
    }
    if (Nx < 2 || Nx % 2 || Nz < 2 || Nz % 2 || Ny < 6 || Ny + 4 > 128) {
        cn::set_error("cn_chpt_sbdf_stage: inconsistent sizes (Ny + 4 <= 128 required)");
        return CN_EINVAL;
    }
    const Dims d = make_dims(Nx, Ny, Nz, 1, 1, 1, kx0, kz0, 0.0);
    StageArgs a{};
    for (int j = 0; j < q; ++j) {
        if (!f_levels[j] || (j > 0 && !mx_hist[j - 1])) {
            cn::set_error("cn_chpt_sbdf_stage: null history level");
            return CN_EINVAL;
        }
        a.f[j] = f_levels[j];
        if (j > 0) a.mx[j - 1] = mx_hist[j - 1];
    }
    a.mx0 = mx0;
    a.q = q;
    for (int j = 1; j <= q; ++j) {
        a.c[j] = c[j];
        a.adt[j] = a_over_dt[j];
    }
    hipLaunchKernelGGL(k_pt_sbdf, dim3(d.P, 4, 2), dim3(128), 0, static_cast<hipStream_t>(stream), X, s04, s2d, s02,
                       a, d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        cn::set_error(std::string("cn_chpt_sbdf_stage: ") + hipGetErrorString(e));
        return CN_EHIP;
    }
    return CN_OK;
}

}  // extern "C"
