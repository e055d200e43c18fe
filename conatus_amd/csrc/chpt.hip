// libcnhip: explicit term and SBDF stage of the 3D plane channel in poloidal-toroidal form.
//
// State (float64): the phi bucket [4][P][nP] then the psi bucket [4][P][nS], nP = N + 4,
// nS = N + 2, slot s = kx * Kzh + |kz|, columns 0/1 = (Re, Im) of mode (kx, +|kz|) and 2/3 = of
// (kx, -|kz|); slot 0 of psi holds the mean flow U (column 0) and W (column 1).  Modes are
// indexed (kx, iz), iz < Kz = 2 Kzh - 1 in FFT order (kz = iz, or iz - Kz for iz >= Kzh).
//
// Explicit term F(X) (the reference's channel_3d_pt right-hand sides,
// tests/channel3d/ch3d_cases.py:176-191: the six products u_i u_j on the 3/2 grid, N = -div(u u),
// the wall-normal curl curl and curl of N, the x-z means of Nx and Nz), five launches:
//   A  y-backward per 8 modes on the FP64 matrix cores (parity-folded Chebyshev tables):
//      H1[kx][j][3][iz] = u, v, w at the Gauss node y_j, from
//      u = i kx D phi + i kz psi, v = k^2 phi, w = i kz D phi - i kx psi (mean: U, 0, W)
//   B  z-backward, 8 kx of one y_j per workgroup: H2[j][m][3][kx]  (inverse FFTs of length Mz,
//      written through LDS in runs of 8 kx)
//   C  x-lines, two (y_j, z_m) per workgroup: H3[j][m][6][kx] = uu, uv, uw, vv, vw, ww (two real
//      lines per complex inverse FFT of length Mx, the products, two real lines per forward
//      FFT, kx < Kx kept); input and output lines are contiguous
//   D  z-forward, 8 kx of one y_j per workgroup: H4[kx][j][6][iz] (forward FFTs, kz represented)
//   E  y-forward + the curl forms per 8 modes on the FP64 matrix cores, with the parity-folded
//      matrices A1 = S14 D01 Fy, A2 = S24 D02 Fy, A0 = S04 Fy, B0 = S02 Fy, B1 = S12 D01 Fy (Fy the
//      Gauss-Chebyshev forward transform), from the node combinations
//        a = kx^2 P_uu + 2 kx kz P_uw + kz^2 P_ww - k^2 P_vv,  q = i (kx P_uv + kz P_vw),
//        b = kx kz (P_uu - P_ww) + (kz^2 - kx^2) P_uw,        s = i (kz P_uv - kx P_vw):
//        F_phi = A1 a - A2 q - A0 (k^2 q)   (C^(4) rows),   F_psi = B0 b - B1 s   (C^(2) rows)
//      (wavenumbers times kx0, kz0); the kx = 0 plane is written as conjugate pairs, and the
//      mean flow F_U = -B1 Re P_uv + G S02 e0, F_W = -B1 Re P_vw in psi slot 0.
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
#include "fft_reg.h"

using cn::zc;

namespace {

constexpr int ZK = 8;       // kx per workgroup of the z transforms (B, D): runs of 8 entries
constexpr int MAXST = 16;   // FFT passes at most

struct Plan {
    int n, r[MAXST];
};

struct Dims {
    int Kx, Kzh, Kz, N, nP, nS, P, Mx, My, Mz;
    double kx0, kz0, G;
    size_t h1() const { return (size_t)Kx * My * 3 * Kz; }
    size_t h2() const { return (size_t)My * Mz * 3 * Kx; }
    size_t h3() const { return (size_t)My * Mz * 6 * Kx; }
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

// ---- y transforms: parity folding on the FP64 matrix cores -----------------------------------
// The Gauss nodes are symmetric, y_{My-1-j} = -y_j, and T_n(-y) = (-1)^n T_n(y).  Both y transforms
// therefore run over the first Mh = ceil(My / 2) nodes only:
//  * backward: with E = the sum over even n and O = over odd n of T_n(y_j) c_n, f(y_j) = E + O and
//    f(y_{My-1-j}) = E - O (for T'_n the roles of even and odd n swap);
//  * forward: row m of S04 Fy, S24 D02 Fy and S02 Fy has A[m][My-1-j] = (-1)^m A[m][j], of S14 D01 Fy
//    and S12 D01 Fy (-1)^(m+1) (the operators keep or flip the parity of n), so with
//    c+(j) = c(j) + c(My-1-j) and c-(j) = c(j) - c(My-1-j) (the middle node of an odd My: c and 0)
//    a row of parity sign +1 sums A[m][j] c+(j), one of sign -1 sums A[m][j] c-(j).
// The products are v_mfma_f64_16x16x4f64: A (16 x 4) lane l holds A[l % 16][l / 16], B (4 x 16)
// lane l holds B[l / 16][l % 16], D (16 x 16) lane l holds D[4 r + l / 16][l % 16] in register r.
// The 16 columns are 8 modes x (re, im): column 2 l + p is part p of mode iz0 + l.
typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int YM = 8;   // modes per workgroup of the y transforms

__host__ __device__ constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// Table geometry, a function of (N, My) only: Mh folded nodes (MHP: padded to 16, Mhs: to 8), KS
// coefficient pairs (padded to 8), RE / RO rows of the even / odd row group (padded to 16).  The
// loops over node steps (E) and coefficient steps (A) run in pairs of 4 with the operands of the
// next step loaded ahead, so the tables carry one more zero step: KT = KS + 4, MhT = Mhs + 4 rows.
struct Fold {
    int Mh, MHP, Mhs, MhT, KS, KT, RE, RO;
    __host__ __device__ Fold(int N, int My)
        : Mh((My + 1) / 2), MHP(16 * cdiv((My + 1) / 2, 16)), Mhs(8 * cdiv((My + 1) / 2, 8)),
          MhT(8 * cdiv((My + 1) / 2, 8) + 4), KS(8 * cdiv((N + 1) / 2, 8)), KT(8 * cdiv((N + 1) / 2, 8) + 4),
          RE(16 * cdiv((N + 1) / 2, 16)), RO(16 * cdiv(N / 2, 16)) {}
};

// ---- A: y-backward ------------------------------------------------------------------------
// Tables Tb[4][KT][MHP]: T_{2k}, T_{2k+1}, T'_{2k}, T'_{2k+1} at the first Mh nodes (zero padded).
// Workgroup: 8 modes, one wave per tile of 16 folded nodes.  The phi and psi coefficients of the 8
// modes are staged in LDS as [2 KS][17] (column 2 l + p, zero beyond N); B operands
// i c = (-Im c, Re c) are read from the partner column.  Per tile: f = T phi, i Df = T' (i phi), i ps = T (i psi), each as E +- O.
// u = kx (i Df) + kz (i ps), v = k^2 f, w = kz (i Df) - kx (i ps); the mean (kx = kz = 0): u = U,
// v = 0, w = W (psi column 0 = U, column 1 = W).
__global__ void __launch_bounds__(512)
k_pt_yback(const double* __restrict__ X, const double* __restrict__ Tb, zc* __restrict__ H1, Dims d) {
    extern __shared__ __align__(16) char lds[];
    const Fold fo(d.N, d.My);
    const int NK = 2 * fo.KS;
    double* ca = reinterpret_cast<double*>(lds);             // phi coefficients [NK][17]
    double* cb = ca + (size_t)NK * 17;                        // psi coefficients [NK][17]
    const int kx = blockIdx.y, iz0 = blockIdx.x * YM, tid = threadIdx.x, N = d.N;
    const double* Xpsi = X + d.size_phi();
    for (int q = tid; q < 2 * 16 * NK; q += blockDim.x) {
        const int n = q % NK, rest = q / NK, col = rest & 15, fld = rest >> 4;
        const int iz = iz0 + (col >> 1), part = col & 1;
        double v = 0.0;
        if (n < N && iz < d.Kz) {
            const int kz = kz_of(iz, d.Kzh, d.Kz);
            const int s = kx * d.Kzh + (kz >= 0 ? kz : -kz), c = (kz >= 0 ? 0 : 2) + part;
            if (fld == 0) {
                if (!(kx == 0 && iz == 0)) v = X[((size_t)c * d.P + s) * d.nP + n];
            } else {
                v = Xpsi[((size_t)c * d.P + s) * d.nS + n];
            }
        }
        (fld ? cb : ca)[(size_t)n * 17 + col] = v;
    }
    __syncthreads();
    const int wave = tid >> 6, lane = tid & 63, col = lane & 15, kq = lane >> 4;
    const double sg = (col & 1) ? 1.0 : -1.0;                 // (i c)_p = sg * c_(p ^ 1)
    const size_t plane = (size_t)fo.KT * fo.MHP;
    const double* tb = Tb + 16 * wave + (lane & 15);
    d4 fE = {0.0, 0.0, 0.0, 0.0}, fO = fE, dE = fE, dO = fE, sE = fE, sO = fE;
    for (int k0 = 0; k0 < fo.KS; k0 += 4) {
        const int k = k0 + kq;
        const double* t = tb + (size_t)k * fo.MHP;
        const double Te = t[0], To = t[plane], De = t[2 * plane], Do = t[3 * plane];
        const double* a0 = ca + (size_t)(2 * k) * 17;
        const double* b0 = cb + (size_t)(2 * k) * 17;
        const double ae = a0[col], ao = a0[17 + col];
        const double iae = sg * a0[col ^ 1], iao = sg * a0[17 + (col ^ 1)];
        const double ibe = sg * b0[col ^ 1], ibo = sg * b0[17 + (col ^ 1)];
        fE = __builtin_amdgcn_mfma_f64_16x16x4f64(Te, ae, fE, 0, 0, 0);
        fO = __builtin_amdgcn_mfma_f64_16x16x4f64(To, ao, fO, 0, 0, 0);
        dE = __builtin_amdgcn_mfma_f64_16x16x4f64(Do, iao, dE, 0, 0, 0);
        dO = __builtin_amdgcn_mfma_f64_16x16x4f64(De, iae, dO, 0, 0, 0);
        sE = __builtin_amdgcn_mfma_f64_16x16x4f64(Te, ibe, sE, 0, 0, 0);
        sO = __builtin_amdgcn_mfma_f64_16x16x4f64(To, ibo, sO, 0, 0, 0);
    }
    const int iz = iz0 + (col >> 1), part = col & 1;
    if (iz >= d.Kz) return;
    const int kz = kz_of(iz, d.Kzh, d.Kz);
    const double ax = kx * d.kx0, az = kz * d.kz0, k2 = ax * ax + az * az;
    const bool mean = (kx == 0 && iz == 0);
    double* h = reinterpret_cast<double*>(H1);
    const size_t fs = 2 * (size_t)d.Kz;                       // field stride in doubles
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = 16 * wave + 4 * r + kq;
        if (j >= fo.Mh) continue;
        const int jm = d.My - 1 - j;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            if (side == 1 && jm == j) break;
            const double sn = side ? -1.0 : 1.0;
            const double f = fE[r] + sn * fO[r], idf = dE[r] + sn * dO[r], ips = sE[r] + sn * sO[r];
            double* o = h + 2 * ((((size_t)kx * d.My + (side ? jm : j)) * 3) * d.Kz + iz);
            if (mean) {                       // i ps = (-W, U): part 1 holds U, part 0 holds -W
                if (part) { o[0] = ips; o[fs + 1] = 0.0; o[2 * fs + 1] = 0.0; }
                else { o[1] = 0.0; o[fs] = 0.0; o[2 * fs] = -ips; }
                continue;
            }
            o[part] = ax * idf + az * ips;
            o[fs + part] = k2 * f;
            o[2 * fs + part] = az * idf - ax * ips;
        }
    }
}

// ---- B: z-backward ------------------------------------------------------------------------
// Workgroup: ZK consecutive kx of one y_j; per field the ZK lines are transformed together and
// written as H2[j][m][f][kx0 .. kx0 + ZK) (runs of ZK entries).  LDS: tw[Mz] | two [ZK][Mz] buffers.
__global__ void __launch_bounds__(512)
k_pt_zback(const zc* __restrict__ H1, const zc* __restrict__ Wz, zc* __restrict__ H2, Dims d, Plan pz) {
    extern __shared__ __align__(16) char lds[];
    const int Mz = d.Mz, Kx = d.Kx;
    zc* tw = reinterpret_cast<zc*>(lds);
    zc* a = tw + Mz;
    zc* b = a + ZK * Mz;
    const int kx0 = blockIdx.x * ZK, j = blockIdx.y, tid = threadIdx.x;
    const int nk = (Kx - kx0) < ZK ? (Kx - kx0) : ZK;
    for (int q = tid; q < Mz; q += blockDim.x) tw[q] = Wz[q];
    for (int f = 0; f < 3; ++f) {
        for (int q = tid; q < ZK * Mz; q += blockDim.x) a[q] = zc{0.0, 0.0};
        __syncthreads();
        for (int q = tid; q < nk * d.Kz; q += blockDim.x) {
            const int kl = q / d.Kz, iz = q - kl * d.Kz;
            a[kl * Mz + bin_of(kz_of(iz, d.Kzh, d.Kz), Mz)] = H1[(((size_t)(kx0 + kl) * d.My + j) * 3 + f) * d.Kz + iz];
        }
        __syncthreads();
        zc* res = fft_lines<true>(a, b, Mz, ZK, tw, pz);
        for (int q = tid; q < Mz * nk; q += blockDim.x) {
            const int m = q / nk, kl = q - m * nk;
            H2[(((size_t)j * Mz + m) * 3 + f) * Kx + kx0 + kl] = res[(size_t)kl * Mz + m];
        }
        __syncthreads();
        if (res != a) { zc* t = a; a = b; b = t; }      // keep `a` as the buffer filled next
    }
}

// ---- C: x-lines ---------------------------------------------------------------------------
// Workgroup: the two lines (y_j, z_m), (y_j, z_m+1).  Inverse transforms: (u + i v) of each line
// and (w_m + i w_m+1); then per product pair (uu, uv), (uw, vv), (vw, ww) one forward transform
// per line.  LDS: tw[Mx] | X[3][Mx] (the time-domain lines) | Y[2][Mx], Z[2][Mx] (transforms).
__global__ void __launch_bounds__(128)
k_pt_xline(const zc* __restrict__ H2, const zc* __restrict__ Wx, zc* __restrict__ H3, Dims d, Plan px) {
    extern __shared__ __align__(16) char lds[];
    const int Mx = d.Mx, Kx = d.Kx, Mz = d.Mz;
    zc* tw = reinterpret_cast<zc*>(lds);
    zc* X = tw + Mx;
    zc* Y = X + 3 * Mx;
    zc* Z = Y + 2 * Mx;
    const int m0 = 2 * blockIdx.x, j = blockIdx.y, tid = threadIdx.x;
    const int two = (Mz - m0) < 2 ? 1 : 2;
    for (int q = tid; q < Mx; q += blockDim.x) tw[q] = Wx[q];
    const zc* in = H2 + ((size_t)j * Mz + m0) * 3 * Kx;    // [line][f][kx]
    // spectra: line 0: u0 + i v0, line 1: u1 + i v1, line 2: w0 + i w1 (Hermitian completions)
    for (int q = tid; q < 3 * Mx; q += blockDim.x) {
        const int line = q / Mx, k = q - line * Mx;
        int kk = -1;
        bool conj = false;
        if (k < Kx) kk = k;
        else if (k > Mx - Kx) { kk = Mx - k; conj = true; }
        zc A{0.0, 0.0}, Bv{0.0, 0.0};
        if (kk >= 0) {
            if (line < 2) {
                if (line < two) {
                    A = in[((size_t)line * 3 + 0) * Kx + kk];
                    Bv = in[((size_t)line * 3 + 1) * Kx + kk];
                }
            } else {
                A = in[(size_t)2 * Kx + kk];
                if (two == 2) Bv = in[((size_t)3 + 2) * Kx + kk];
            }
            if (kk == 0) { A.im = 0.0; Bv.im = 0.0; }
            if (conj) { A = zconj(A); Bv = zconj(Bv); }
        }
        X[q] = zadd(A, zmuli(Bv));
    }
    __syncthreads();
    {
        const zc* res = fft_lines<true>(X, Y, Mx, 2, tw, px);      // Y, Z as the scratch of 2 lines
        if (res != X) {
            for (int q = tid; q < 2 * Mx; q += blockDim.x) X[q] = res[q];
            __syncthreads();
        }
        const zc* rw = fft_lines<true>(X + 2 * Mx, Z, Mx, 1, tw, px);
        if (rw != X + 2 * Mx) {
            for (int q = tid; q < Mx; q += blockDim.x) X[2 * Mx + q] = rw[q];
            __syncthreads();
        }
    }
    const double inv = 1.0 / Mx;
    zc* out = H3 + ((size_t)j * Mz + m0) * 6 * Kx;          // [line][f6][kx]
    for (int pair = 0; pair < 3; ++pair) {
        for (int q = tid; q < 2 * Mx; q += blockDim.x) {
            const int line = q / Mx, x = q - line * Mx;
            const double u = X[(size_t)line * Mx + x].re, v = X[(size_t)line * Mx + x].im;
            const double w = line == 0 ? X[2 * Mx + x].re : X[2 * Mx + x].im;
            Y[q] = pair == 0 ? zc{u * u, u * v} : pair == 1 ? zc{u * w, v * v} : zc{v * w, w * w};
        }
        __syncthreads();
        const zc* res = fft_lines<false>(Y, Z, Mx, 2, tw, px);
        for (int q = tid; q < two * Kx; q += blockDim.x) {
            const int line = q / Kx, k = q - line * Kx;
            const zc Zk = res[(size_t)line * Mx + k];
            const zc Zc = zconj(res[(size_t)line * Mx + (k == 0 ? 0 : Mx - k)]);
            out[((size_t)line * 6 + 2 * pair) * Kx + k] = zscale(0.5 * inv, zadd(Zk, Zc));
            out[((size_t)line * 6 + 2 * pair + 1) * Kx + k] = zscale(-0.5 * inv, zmuli(zsub(Zk, Zc)));
        }
        __syncthreads();
    }
}

// ---- D: z-forward -------------------------------------------------------------------------
// Workgroup: ZK consecutive kx of one y_j; per product the lines are gathered from
// H3[j][m][f][kx0 ..) (runs of ZK entries), transformed and written as H4[kx][j][f][iz].
__global__ void __launch_bounds__(512)
k_pt_zfwd(const zc* __restrict__ H3, const zc* __restrict__ Wz, zc* __restrict__ H4, Dims d, Plan pz) {
    extern __shared__ __align__(16) char lds[];
    const int Mz = d.Mz, Kx = d.Kx;
    zc* tw = reinterpret_cast<zc*>(lds);
    zc* a = tw + Mz;
    zc* b = a + ZK * Mz;
    const int kx0 = blockIdx.x * ZK, j = blockIdx.y, tid = threadIdx.x;
    const int nk = (Kx - kx0) < ZK ? (Kx - kx0) : ZK;
    for (int q = tid; q < Mz; q += blockDim.x) tw[q] = Wz[q];
    const double inv = 1.0 / Mz;
    for (int f = 0; f < 6; ++f) {
        for (int q = tid; q < Mz * ZK; q += blockDim.x) {
            const int m = q / ZK, kl = q - m * ZK;
            a[(size_t)kl * Mz + m] = kl < nk ? H3[(((size_t)j * Mz + m) * 6 + f) * Kx + kx0 + kl] : zc{0.0, 0.0};
        }
        __syncthreads();
        zc* res = fft_lines<false>(a, b, Mz, ZK, tw, pz);
        for (int q = tid; q < nk * d.Kz; q += blockDim.x) {
            const int kl = q / d.Kz, iz = q - kl * d.Kz;
            H4[(((size_t)(kx0 + kl) * d.My + j) * 6 + f) * d.Kz + iz] =
                zscale(inv, res[(size_t)kl * Mz + bin_of(kz_of(iz, d.Kzh, d.Kz), Mz)]);
        }
        __syncthreads();
        if (res != a) { zc* t = a; a = b; b = t; }
    }
}

// ---- register-split FFTs for the z and x transforms -------------------------------------------
// A line of M = FA FB points is transformed by a group of 16 lanes (t = lane & 15): lanes t < FB
// take the samples n = FB n1 + t (n1 < FA), do an FA-point DFT in registers, multiply by
// exp(-+2 pi i t k1 / M) (the table W, forward signs) and write row k1 of the line's exchange
// tile at column t (rows padded to FB + 1); lanes t < FA then read row t and do an FB-point DFT,
// which leaves output k = t + FA k2 (k2 < FB) in register k2.  One LDS transpose per transform.
template <int FA, int FB>
struct Split {
    static constexpr int M = FA * FB, RS = FB + 1, XS = FA * RS, NV = FA > FB ? FA : FB;
};

template <int FA, int FB, bool INV>
__device__ __forceinline__ void split_first(zc* v, int t, const zc* __restrict__ W, zc* xt) {
    using S = Split<FA, FB>;
    cnf::dft<FA, INV>(v);
#pragma unroll
    for (int k1 = 0; k1 < FA; ++k1) {
        const int e = (t * k1) % S::M;
        const zc w = W[e];
        const zc x = e == 0 ? v[k1] : zmul(v[k1], INV ? zconj(w) : w);
        xt[k1 * S::RS + t] = x;
    }
}

template <int FA, int FB, bool INV>
__device__ __forceinline__ void split_second(zc* v, int t, const zc* xt) {
    using S = Split<FA, FB>;
#pragma unroll
    for (int n2 = 0; n2 < FB; ++n2) v[n2] = xt[t * S::RS + n2];
    cnf::dft<FB, INV>(v);
}

// bin b of an M-point grid -> mode index iz (kz = b or b - M kept when |kz| < Kzh), or -1
__device__ __forceinline__ int iz_of_bin(int b, int M, int Kzh, int Kz) {
    return b < Kzh ? b : (b > M - Kzh ? b - M + Kz : -1);
}

// B (register split): workgroup = 16 consecutive kx of one (y_j, field), one line per 16 lanes;
// the outputs are staged as [m][17] and written in runs of 16 kx.
template <int FA, int FB>
__global__ void __launch_bounds__(256)
k_pt_zback_s(const zc* __restrict__ H1, const zc* __restrict__ Wz, zc* __restrict__ H2, Dims d) {
    using S = Split<FA, FB>;
    constexpr int M = S::M;
    extern __shared__ __align__(16) char lds[];
    zc* sh = reinterpret_cast<zc*>(lds);                    // 16 exchange tiles, then [M][17]
    const int tid = threadIdx.x, L = tid >> 4, t = tid & 15;
    const int kx0 = blockIdx.x * 16, j = blockIdx.y, f = blockIdx.z, kx = kx0 + L;
    const bool on = kx < d.Kx;
    zc v[S::NV];
    if (t < FB) {
        const zc* src = H1 + (((size_t)(on ? kx : 0) * d.My + j) * 3 + f) * d.Kz;
#pragma unroll
        for (int n1 = 0; n1 < FA; ++n1) {
            const int iz = iz_of_bin(FB * n1 + t, M, d.Kzh, d.Kz);
            v[n1] = (on && iz >= 0) ? src[iz] : zc{0.0, 0.0};
        }
        split_first<FA, FB, true>(v, t, Wz, sh + L * S::XS);
    }
    __syncthreads();
    if (t < FA) split_second<FA, FB, true>(v, t, sh + L * S::XS);
    __syncthreads();
    if (t < FA) {
#pragma unroll
        for (int k2 = 0; k2 < FB; ++k2) sh[(t + FA * k2) * 17 + L] = v[k2];
    }
    __syncthreads();
    const int nk = (d.Kx - kx0) < 16 ? (d.Kx - kx0) : 16;
    for (int q = tid; q < M * 16; q += blockDim.x) {
        const int m = q >> 4, l = q & 15;
        if (l < nk) H2[(((size_t)j * M + m) * 3 + f) * d.Kx + kx0 + l] = sh[m * 17 + l];
    }
}

// D (register split): workgroup = 16 consecutive kx of one (y_j, product); the input lines are
// gathered from H3 in runs of 16 kx into [m][17], transformed, and each lane writes its kept bins.
template <int FA, int FB>
__global__ void __launch_bounds__(256)
k_pt_zfwd_s(const zc* __restrict__ H3, const zc* __restrict__ Wz, zc* __restrict__ H4, Dims d) {
    using S = Split<FA, FB>;
    constexpr int M = S::M;
    extern __shared__ __align__(16) char lds[];
    zc* sh = reinterpret_cast<zc*>(lds);                    // [M][17], then 16 exchange tiles
    const int tid = threadIdx.x, L = tid >> 4, t = tid & 15;
    const int kx0 = blockIdx.x * 16, j = blockIdx.y, f = blockIdx.z, kx = kx0 + L;
    const int nk = (d.Kx - kx0) < 16 ? (d.Kx - kx0) : 16;
    for (int q = tid; q < M * 16; q += blockDim.x) {
        const int m = q >> 4, l = q & 15;
        sh[m * 17 + l] = l < nk ? H3[(((size_t)j * M + m) * 6 + f) * d.Kx + kx0 + l] : zc{0.0, 0.0};
    }
    __syncthreads();
    zc v[S::NV];
    if (t < FB) {
#pragma unroll
        for (int n1 = 0; n1 < FA; ++n1) v[n1] = sh[(FB * n1 + t) * 17 + L];
    }
    __syncthreads();
    if (t < FB) split_first<FA, FB, false>(v, t, Wz, sh + L * S::XS);
    __syncthreads();
    if (t < FA) {
        split_second<FA, FB, false>(v, t, sh + L * S::XS);
        if (kx < d.Kx) {
            const double inv = 1.0 / M;
            zc* dst = H4 + (((size_t)kx * d.My + j) * 6 + f) * d.Kz;
#pragma unroll
            for (int k2 = 0; k2 < FB; ++k2) {
                const int iz = iz_of_bin(t + FA * k2, M, d.Kzh, d.Kz);
                if (iz >= 0) dst[iz] = zscale(inv, v[k2]);
            }
        }
    }
}

// C (register split): one wave per (y_j; z_m, z_m+1).  Inverse: groups 0, 1, 2 transform
// u_m + i v_m, u_m+1 + i v_m+1 and w_m + i w_m+1 (Hermitian completions) into the physical lines
// ph[3][M]; then per z line the three product pairs (uu, uv), (uw, vv), (vw, ww), one per group,
// are transformed forward, staged in natural order and split into the two real spectra.
template <int FA, int FB>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
k_pt_xline_s(const zc* __restrict__ H2, const zc* __restrict__ Wx, zc* __restrict__ H3, Dims d) {
    using S = Split<FA, FB>;
    constexpr int M = S::M, XE = S::XS > M ? S::XS : M;
    extern __shared__ __align__(16) char lds[];
    zc* ex = reinterpret_cast<zc*>(lds);                    // [3][XE] exchange tiles / spectra (groups 0..2)
    zc* ph = ex + 3 * XE;                                   // [3][M] physical lines
    const int lane = threadIdx.x, g = lane >> 4, t = lane & 15;
    const int Kx = d.Kx, Mz = d.Mz;
    const int m0 = 2 * blockIdx.x, j = blockIdx.y;
    const int two = (Mz - m0) < 2 ? 1 : 2;
    const zc* in = H2 + ((size_t)j * Mz + m0) * 3 * Kx;     // [line][f][kx]
    zc* xt = ex + (g < 3 ? g : 0) * XE;                     // group 3 idles (no tile)
    zc v[S::NV];
    if (g < 3 && t < FB) {
#pragma unroll
        for (int n1 = 0; n1 < FA; ++n1) {
            const int k = FB * n1 + t;
            int kk = -1;
            bool cj = false;
            if (k < Kx) kk = k;
            else if (k > M - Kx) { kk = M - k; cj = true; }
            zc A{0.0, 0.0}, B{0.0, 0.0};
            if (kk >= 0) {
                if (g < 2) {
                    if (g < two) {
                        A = in[((size_t)g * 3 + 0) * Kx + kk];
                        B = in[((size_t)g * 3 + 1) * Kx + kk];
                    }
                } else {
                    A = in[(size_t)2 * Kx + kk];
                    if (two == 2) B = in[((size_t)3 + 2) * Kx + kk];
                }
                if (kk == 0) { A.im = 0.0; B.im = 0.0; }
                if (cj) { A = zconj(A); B = zconj(B); }
            }
            v[n1] = zadd(A, zmuli(B));
        }
        split_first<FA, FB, true>(v, t, Wx, xt);
    }
    __syncthreads();
    if (g < 3 && t < FA) {
        split_second<FA, FB, true>(v, t, xt);
#pragma unroll
        for (int k2 = 0; k2 < FB; ++k2) ph[g * M + t + FA * k2] = v[k2];
    }
    __syncthreads();
    const double h = 0.5 / M;
    for (int zl = 0; zl < two; ++zl) {
        if (g < 3 && t < FB) {
#pragma unroll
            for (int n1 = 0; n1 < FA; ++n1) {
                const int x = FB * n1 + t;
                const zc uv = ph[zl * M + x], wp = ph[2 * M + x];
                const double u = uv.re, vv = uv.im, w = zl ? wp.im : wp.re;
                v[n1] = g == 0 ? zc{u * u, u * vv} : g == 1 ? zc{u * w, vv * vv} : zc{vv * w, w * w};
            }
            split_first<FA, FB, false>(v, t, Wx, xt);
        }
        __syncthreads();
        if (g < 3 && t < FA) split_second<FA, FB, false>(v, t, xt);
        __syncthreads();
        if (g < 3 && t < FA) {
#pragma unroll
            for (int k2 = 0; k2 < FB; ++k2) xt[t + FA * k2] = v[k2];
        }
        __syncthreads();
        if (g < 3) {
            zc* out = H3 + ((size_t)j * Mz + m0 + zl) * 6 * Kx;
            for (int k = t; k < Kx; k += 16) {
                const zc Zk = xt[k];
                const zc Zc = zconj(xt[k == 0 ? 0 : M - k]);
                out[(size_t)(2 * g) * Kx + k] = zscale(h, zadd(Zk, Zc));
                out[(size_t)(2 * g + 1) * Kx + k] = zscale(-h, zmuli(zsub(Zk, Zc)));
            }
        }
        __syncthreads();
    }
}

// ---- E: y-forward and the curl forms ----------------------------------------------------------
// Tables Am[5][MhT][RE + RO]: A1 = S14 D01 Fy, A2 = S24 D02 Fy, A0 = S04 Fy, B0 = S02 Fy,
// B1 = S12 D01 Fy at the first Mh nodes, node-major, rows permuted even first (m = 2 i at i < RE)
// then odd (m = 2 i + 1 at RE + i), zero padded.  Workgroup: one wave, 8 modes; PSI selects the
// bucket, NRT = (RE + RO) / 16 row tiles.  A lane loads the own part (p) and the partner part
// (p ^ 1) of the products at its node pair, (i z)_p = sg z_(p ^ 1), forms the folded
// combinations, and all row tiles accumulate in registers; the next node step's operands are
// loaded while the current step's products run.  The rows are staged in LDS ([col][nrow],
// boundary rows zero) and written as contiguous runs per column.
// The mean (kx = kz = 0) is the psi column pair with b = 0 and -s = -(Re P_uv, Re P_vw), plus
// G S02 e0 in column 0: F_U = -B1 Re P_uv + G S02 e0, F_W = -B1 Re P_vw; its phi rows are zero.
template <bool PSI, int NRT>
__device__ __forceinline__ void yfwd_bucket(const zc* __restrict__ H4, const double* __restrict__ Am,
                                            const double* __restrict__ s02c0, double* __restrict__ F, const Dims& d) {
    extern __shared__ __align__(16) char lds[];
    double* orow = reinterpret_cast<double*>(lds);           // [16][nrow]
    constexpr bool psi = PSI;
    constexpr int NM = PSI ? 2 : 3;                          // matrices of the bucket
    constexpr int F2 = PSI ? 1 : 3;                          // own-part product in slot 2: uv / vv
    const Fold fo(d.N, d.My);
    const int kx = blockIdx.y, iz0 = blockIdx.x * YM, lane = threadIdx.x, N = d.N;
    if (kx == 0 && iz0 >= d.Kzh) return;                     // mirrors, written with their partners
    const int col = lane & 15, part = col & 1, kq = lane >> 4;
    const int iz = iz0 + (col >> 1);
    const bool live = iz < d.Kz && !(kx == 0 && iz >= d.Kzh);
    const int kz = kz_of(live ? iz : 0, d.Kzh, d.Kz);
    const double bx = kx * d.kx0, bz = kz * d.kz0, kk = bx * bx + bz * bz;
    const bool mean = (kx == 0 && iz == 0);
    const double sg = part ? 1.0 : -1.0;
    const int RP = 16 * NRT, nre = fo.RE / 16;
    const size_t ts = (size_t)fo.MhT * RP;                    // matrix stride
    const size_t fs = 2 * (size_t)d.Kz;                       // product stride in doubles
    const double* H = reinterpret_cast<const double*>(H4) + 2 * ((size_t)kx * d.My * 6 * d.Kz + (live ? iz : iz0));
    const double* ab = Am + (PSI ? 3 : 0) * ts + (size_t)kq * RP + (lane & 15);
    d4 acc[NRT];
#pragma unroll
    for (int t = 0; t < NRT; ++t) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    // load only issues the loads (nodes past Mh and dead lanes read node 0, which mma masks), so
    // that no loaded value is consumed before the next step
    auto load = [&](int j0, double (&P)[2][6], double (&A)[NM][NRT]) {
        const int j = j0 + kq;
        const bool on = live && j < fo.Mh;
        const int ja = on ? j : 0, jb = on ? d.My - 1 - j : 0;           // in bounds either way
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const double* pp = H + (size_t)(side ? jb : ja) * 6 * fs;
            P[side][0] = pp[part];                  // uu
            P[side][1] = pp[2 * fs + part];         // uw
            P[side][2] = pp[F2 * fs + part];        // vv (phi) / uv (psi)
            P[side][3] = pp[5 * fs + part];         // ww
            P[side][4] = pp[fs + (part ^ 1)];       // uv, partner part
            P[side][5] = pp[4 * fs + (part ^ 1)];   // vw, partner part
        }
        const double* a = ab + (size_t)j0 * RP;
#pragma unroll
        for (int t = 0; t < NM; ++t)
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt) A[t][rt] = a[t * ts + 16 * rt];
    };
    auto mma = [&](int j0, const double (&Pr)[2][6], const double (&A)[NM][NRT]) {
        const bool mid = d.My - 1 - (j0 + kq) == j0 + kq;
        const bool on = live && j0 + kq < fo.Mh;
        double P[2][6];
#pragma unroll
        for (int side = 0; side < 2; ++side)
#pragma unroll
            for (int k = 0; k < 6; ++k) P[side][k] = on ? Pr[side][k] : 0.0;
        double c[2][3];
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const double* q = P[side];
            if constexpr (!PSI) {
                const double nq = -sg * (bx * q[4] + bz * q[5]);
                c[side][0] = bx * bx * q[0] + 2.0 * bx * bz * q[1] + bz * bz * q[3] - kk * q[2];
                c[side][1] = nq;
                c[side][2] = kk * nq;
            } else if (mean) {                                   // -(Re P_uv, Re P_vw)
                c[side][0] = 0.0;
                c[side][1] = -(part ? q[5] : q[2]);
                c[side][2] = 0.0;
            } else {
                c[side][0] = bx * bz * (q[0] - q[3]) + (bz * bz - bx * bx) * q[1];
                c[side][1] = -sg * (bz * q[4] - bx * q[5]);
                c[side][2] = 0.0;
            }
        }
        double cp[3], cm[3];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            cp[t] = mid ? c[0][t] : c[0][t] + c[1][t];
            cm[t] = mid ? 0.0 : c[0][t] - c[1][t];
        }
        // matrix-outer, tile-inner: consecutive products go to different accumulators (each
        // accumulator still sums its matrices in the same order)
        double be[NM], bo[NM];                  // B operand of the even / odd row tiles
        if constexpr (!PSI) {
            be[0] = cm[0]; bo[0] = cp[0];
            be[1] = cp[1]; bo[1] = cm[1];
            be[2] = cp[2]; bo[2] = cm[2];
        } else {
            be[0] = cp[0]; bo[0] = cm[0];
            be[1] = cm[1]; bo[1] = cp[1];
        }
#pragma unroll
        for (int t = 0; t < NM; ++t)
#pragma unroll
            for (int rt = 0; rt < NRT; ++rt)
                acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[t][rt], rt < nre ? be[t] : bo[t], acc[rt], 0, 0, 0);
    };
    double Pa[2][6], Pn[2][6], Aa[NM][NRT], An[NM][NRT];
    // the scheduling barriers keep each step's loads ahead of the previous step's products
    load(0, Pa, Aa);
    for (int j0 = 0; j0 < fo.Mhs; j0 += 8) {
        load(j0 + 4, Pn, An);
        __builtin_amdgcn_sched_barrier(0);
        mma(j0, Pa, Aa);
        __builtin_amdgcn_sched_barrier(0);
        load(j0 + 8, Pa, Aa);
        __builtin_amdgcn_sched_barrier(0);
        mma(j0 + 4, Pn, An);
        __builtin_amdgcn_sched_barrier(0);
    }
    const int nrow = psi ? d.nS : d.nP, off = psi ? 2 : 4;
    for (int q = lane; q < 16 * off; q += 64) orow[(size_t)(q / off) * nrow + q % off] = 0.0;
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * rt + 4 * r + kq;
            const int m = i < fo.RE ? 2 * i : 2 * (i - fo.RE) + 1;
            if (m >= N) continue;
            double v = acc[rt][r];
            if (psi && mean && part == 0) v += d.G * s02c0[m];
            orow[(size_t)col * nrow + off + m] = v;
        }
    }
    __syncthreads();
    double* Fb = psi ? F + d.size_phi() : F;
    const size_t cs = (size_t)d.P * nrow;                     // column stride
    for (int ll = 0; ll < YM; ++ll) {
        const int izz = iz0 + ll;
        if (izz >= d.Kz || (kx == 0 && izz >= d.Kzh)) continue;
        const int kzz = kz_of(izz, d.Kzh, d.Kz);
        const int s = kx * d.Kzh + (kzz >= 0 ? kzz : -kzz), c = kzz >= 0 ? 0 : 2;
        const double* sre = orow + (size_t)(2 * ll) * nrow;
        const double* sim = sre + nrow;
        const bool mirror = (kx == 0 && kzz > 0);   // (0, -kz) = conj (0, kz)
        const bool dead = (kzz == 0);               // no (kx, -0) mode: columns 2, 3 zero
        double* base = Fb + (size_t)s * nrow;
        for (int q = lane; q < nrow; q += 64) {
            const double re = sre[q], im = sim[q];
            base[c * cs + q] = re;
            base[(c + 1) * cs + q] = im;
            if (mirror) {
                base[2 * cs + q] = re;
                base[3 * cs + q] = -im;
            } else if (dead) {
                base[2 * cs + q] = 0.0;
                base[3 * cs + q] = 0.0;
            }
        }
    }
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
    const bool on = row < n;
    const int kx = s / d.Kzh, kzv = s - kx * d.Kzh;
    const double ax = kx * d.kx0, az = kzv * d.kz0, k2 = ax * ax + az * az;
    const int nb = bucket == 0 ? 4 : 2, m = row - nb, N = d.N;
    // every global operand is requested before the barrier (one memory round trip per wave)
    const double xv = on ? x[row] : 0.0;
    double fv[4], mv[3], cf[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 4; ++j) fv[j] = (on && j < a.q) ? a.f[j][off + row] : 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) mv[j] = (on && j + 1 < a.q) ? a.mx[j][off + row] : 0.0;
    if (on && row >= nb) {
        if (bucket == 0) {
            if (s != 0) {
                const double k4 = k2 * k2;
#pragma unroll
                for (int t = 0; t < 5; ++t) cf[t] = (-k2) * s2d[m * 5 + t] + k4 * s04[m * 5 + t];
            }
        } else {
#pragma unroll
            for (int t = 0; t < 3; ++t) cf[t] = s == 0 ? s02[m * 3 + t] : (-k2) * s02[m * 3 + t];
        }
    }
    xl[row] = xv;
    __syncthreads();
    if (!on) return;
    double mx = 0.0;
    if (row >= nb) {
        const int nt = bucket == 0 ? (s != 0 ? 5 : 0) : 3;
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            const int col = m + 2 * t;
            if (t >= nt || col >= N) break;
            mx = mx + cf[t] * xl[col];
        }
    }
    a.mx0[off + row] = mx;
    double rhs = 0.0;
#pragma unroll
    for (int j = 1; j <= 4; ++j) {
        if (j > a.q) break;
        const double mj = j == 1 ? mx : mv[j - 2];
        const double t1 = a.c[j] * fv[j - 1], t2 = a.adt[j] * mj;
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

// E launch: blockIdx.z = 0 the phi bucket, 1 the psi bucket (their waves share the SIMDs)
template <int NRT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_pt_yfwd(const zc* __restrict__ H4, const double* __restrict__ Am, const double* __restrict__ s02c0,
          double* __restrict__ F, Dims d) {
    if (blockIdx.z == 0) yfwd_bucket<false, NRT>(H4, Am, s02c0, F, d);
    else yfwd_bucket<true, NRT>(H4, Am, s02c0, F, d);
}


size_t lds_a(const Dims& d) { return (size_t)2 * 2 * Fold(d.N, d.My).KS * 17 * sizeof(double); }

// Register splits M = FA x FB of the z and x transforms; other 2^a 3^b sizes use the Stockham
// kernels.
#define CN_PT_SPLITS(X) X(12, 4, 3) X(18, 6, 3) X(24, 4, 6) X(36, 6, 6) X(48, 8, 6) X(72, 6, 12) \
    X(96, 6, 16) X(144, 12, 12) X(192, 12, 16)

template <int FA, int FB>
size_t lds_zs() {
    using S = Split<FA, FB>;
    return (size_t)(16 * S::XS > S::M * 17 ? 16 * S::XS : S::M * 17) * sizeof(zc);
}
template <int FA, int FB>
size_t lds_xs() {
    using S = Split<FA, FB>;
    return (size_t)(3 * (S::XS > S::M ? S::XS : S::M) + 3 * S::M) * sizeof(zc);
}

bool zback_split(const Dims& d, hipStream_t s, const zc* h1, const zc* Wz, zc* h2) {
    const dim3 g((d.Kx + 15) / 16, d.My, 3);
    switch (d.Mz) {
#define CN_X(M, FA, FB) \
    case M: hipLaunchKernelGGL((k_pt_zback_s<FA, FB>), g, dim3(256), (lds_zs<FA, FB>()), s, h1, Wz, h2, d); return true;
        CN_PT_SPLITS(CN_X)
#undef CN_X
        default: return false;
    }
}
bool zfwd_split(const Dims& d, hipStream_t s, const zc* h3, const zc* Wz, zc* h4) {
    const dim3 g((d.Kx + 15) / 16, d.My, 6);
    switch (d.Mz) {
#define CN_X(M, FA, FB) \
    case M: hipLaunchKernelGGL((k_pt_zfwd_s<FA, FB>), g, dim3(256), (lds_zs<FA, FB>()), s, h3, Wz, h4, d); return true;
        CN_PT_SPLITS(CN_X)
#undef CN_X
        default: return false;
    }
}
bool xline_split(const Dims& d, hipStream_t s, const zc* h2, const zc* Wx, zc* h3) {
    const dim3 g((d.Mz + 1) / 2, d.My);
    switch (d.Mx) {
#define CN_X(M, FA, FB) \
    case M: hipLaunchKernelGGL((k_pt_xline_s<FA, FB>), g, dim3(64), (lds_xs<FA, FB>()), s, h2, Wx, h3, d); return true;
        CN_PT_SPLITS(CN_X)
#undef CN_X
        default: return false;
    }
}

template <int NRT>
void launch_yfwd_n(dim3 g, size_t lds, hipStream_t s, const zc* h4, const double* Am, const double* s02c0, double* F,
                   const Dims& d) {
    hipLaunchKernelGGL((k_pt_yfwd<NRT>), dim3(g.x, g.y, 2), dim3(64), lds, s, h4, Am, s02c0, F, d);
}

// the y-forward launch, instantiated per row-tile count (N <= 128: at most 8 tiles)
void launch_yfwd(dim3 g, size_t lds, hipStream_t s, const zc* h4, const double* Am, const double* s02c0, double* F,
                 const Dims& d) {
    const Fold fo(d.N, d.My);
    switch ((fo.RE + fo.RO) / 16) {
        case 2: launch_yfwd_n<2>(g, lds, s, h4, Am, s02c0, F, d); break;
        case 3: launch_yfwd_n<3>(g, lds, s, h4, Am, s02c0, F, d); break;
        case 4: launch_yfwd_n<4>(g, lds, s, h4, Am, s02c0, F, d); break;
        case 5: launch_yfwd_n<5>(g, lds, s, h4, Am, s02c0, F, d); break;
        case 6: launch_yfwd_n<6>(g, lds, s, h4, Am, s02c0, F, d); break;
        case 7: launch_yfwd_n<7>(g, lds, s, h4, Am, s02c0, F, d); break;
        default: launch_yfwd_n<8>(g, lds, s, h4, Am, s02c0, F, d); break;
    }
}
size_t lds_b(const Dims& d) { return ((size_t)1 + 2 * ZK) * d.Mz * sizeof(zc); }
size_t lds_c(const Dims& d) { return (size_t)8 * d.Mx * sizeof(zc); }
size_t lds_d(const Dims& d) { return lds_b(d); }
size_t lds_e(const Dims& d) { return (size_t)16 * (d.nP > d.nS ? d.nP : d.nS) * sizeof(double); }

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
    if (Ny > 128 || My > 256) {
        cn::set_error("cn_chpt_explicit: Ny > 128 or My > 256 not supported by the y transforms");
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
    const dim3 gy((d.Kz + YM - 1) / YM, d.Kx);
    hipLaunchKernelGGL(k_pt_yback, gy, dim3(64 * (Fold(d.N, d.My).MHP / 16)), lds_a(d), s, X, Tb, h1, d);
    const dim3 gz((d.Kx + ZK - 1) / ZK, My);
    const zc* wz = static_cast<const zc*>(Wz);
    const zc* wx = static_cast<const zc*>(Wx);
    if (!zback_split(d, s, h1, wz, h2))
        hipLaunchKernelGGL(k_pt_zback, gz, dim3(512), lds_b(d), s, h1, wz, h2, d, pz);
    if (!xline_split(d, s, h2, wx, h3))
        hipLaunchKernelGGL(k_pt_xline, dim3((Mz + 1) / 2, My), dim3(128), lds_c(d), s, h2, wx, h3, d, px);
    if (!zfwd_split(d, s, h3, wz, h4))
        hipLaunchKernelGGL(k_pt_zfwd, gz, dim3(512), lds_d(d), s, h3, wz, h4, d, pz);
    launch_yfwd(gy, lds_e(d), s, h4, Am, s02c0, F, d);
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
