// libcnhip explicit term of the 3D plane channel (primitive variables):
//   F = base + S_{0->2} coefficients of -u.grad(u) on the 3/2-dealiased grid, in the x, y and z
// momentum rows of every Fourier mode (kx, kz) (the reference's Fourier x Chebyshev x Fourier
// explicit pipeline: python/conatus/timestep/fused.py:994-1085 with the 3D transforms of
// python/conatus/timestep/xp_transforms.py:197-548 — factors, derivative chains, backward
// transforms, products, forward transforms, C^(lam) conversion into the stacked rows).
//
// State X: (Kx * Kz, n) complex, mode m = kx * Kz + iz, kz = iz (iz < Nz/2) or iz - Kz, Kz = Nz - 1;
// fidx[f][nn] = column of Chebyshev coefficient nn of velocity component f (u, v, w).
// Five launches, each a transform along one axis with its operands staged in LDS (sums over the
// modes, split by parity: symmetric Chebyshev nodes, and for even M the Fourier pairs q, q + M/2):
//   A  y-backward (per 8 modes):    H1[kx][j][6][iz]  = (u, u_y, v, v_y, w, w_y) at y_j
//   B  z-backward (per 4 kx, y_j):  H2[j][zm][9][kx]  = per component: value, d/dz, d/dy at z_m
//   C  x-line (per y_j, z_m):       x-backward of the 12 factors (d/dx = i kx kx0 on the modes),
//                                   the products -u.grad(u), x-forward (kx < Kx): H3[j][zm][3][kx]
//   D  z-forward (per 4 kx, y_j):   H4[j][kx][3][iz]  (modes kz of the represented range)
//   E  y-forward + conversion (per 8 modes): F[m][R + 4 nn + e] = base + sum_j Ay[j][nn] H4[..]
// Fourier sums use one twiddle table per axis, W_q = exp(2 pi i q / M); the index (k q mod M) is
// carried incrementally.  H3 reuses H1's space and H4 reuses H2's (each is dead by then).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cnhip.h"
#include "cn_common.h"

using cn::zc;

namespace {

constexpr int MA = 8;   // modes per workgroup of the y transforms (A, E): 128-byte runs in iz
constexpr int KB = 4;   // kx per workgroup of the z transforms (B, D)

struct ChDims {
    int Kx, Kz, N, Mx, My, Mz;
    size_t h1() const { return (size_t)Kx * My * 6 * Kz; }
    size_t h2() const { return (size_t)My * Mz * 9 * Kx; }
    size_t h3() const { return (size_t)My * Mz * 3 * Kx; }
    size_t h4() const { return (size_t)My * Kx * 3 * Kz; }
    size_t region1() const { return h1() > h3() ? h1() : h3(); }
    size_t region2() const { return h2() > h4() ? h2() : h4(); }
    size_t bytes() const { return (region1() + region2()) * sizeof(zc); }
};

// y-backward of 8 modes.  The Gauss-Chebyshev nodes are symmetric, y_{My-1-j} = -y_j, and
// T_n(-y) = (-1)^n T_n(y), T_n'(-y) = (-1)^(n+1) T_n'(y): the even-n and odd-n sums at y_j give
// the values at both nodes of the pair (the middle node, y = 0 for odd My, on its own).
__global__ void __launch_bounds__(256)
k_ch_yback(const zc* __restrict__ X, const int* __restrict__ fidx, const double* __restrict__ By,
           zc* __restrict__ H1, int Kx, int Kz, int N, int My, int n, size_t sx, size_t sw) {
    extern __shared__ __align__(16) char lds[];
    X += blockIdx.z * sx;                            // member blockIdx.z: its state and workspace
    H1 += blockIdx.z * sw;
    zc* c = reinterpret_cast<zc*>(lds);              // [3][N][MA]
    const int kx = blockIdx.y, iz0 = blockIdx.x * MA, tid = threadIdx.x;
    for (int q = tid; q < 3 * N * MA; q += blockDim.x) {
        const int l = q % MA, fn = q / MA;           // fn = f * N + nn
        const int iz = iz0 + l;
        c[q] = iz < Kz ? X[(size_t)(kx * Kz + iz) * n + fidx[fn]] : zc{0.0, 0.0};
    }
    __syncthreads();
    const int l = tid % MA, iz = iz0 + l, P = (My + 1) / 2;
    for (int p = tid / MA; p < 3 * P; p += blockDim.x / MA) {
        const int f = p / P, j = p - f * P, jr = My - 1 - j;
        const double* b = By + (size_t)j * N * 2;   // (T_n(y_j), T_n'(y_j)) pairs
        const zc* cf = c + (size_t)f * N * MA + l;
        double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};   // even: v.re v.im d.re d.im, odd: same
        for (int nn = 0; nn < N; nn += 2) {
            const double b0 = b[2 * nn], b1 = b[2 * nn + 1];
            const zc x = cf[nn * MA];
            s[0] = fma(b0, x.re, s[0]);
            s[1] = fma(b0, x.im, s[1]);
            s[2] = fma(b1, x.re, s[2]);
            s[3] = fma(b1, x.im, s[3]);
            if (nn + 1 < N) {
                const double c0 = b[2 * nn + 2], c1 = b[2 * nn + 3];
                const zc y = cf[(nn + 1) * MA];
                s[4] = fma(c0, y.re, s[4]);
                s[5] = fma(c0, y.im, s[5]);
                s[6] = fma(c1, y.re, s[6]);
                s[7] = fma(c1, y.im, s[7]);
            }
        }
        if (iz < Kz) {
            zc* h = H1 + ((size_t)(kx * My + j) * 6 + 2 * f) * Kz + iz;
            h[0] = zc{s[0] + s[4], s[1] + s[5]};
            h[Kz] = zc{s[2] + s[6], s[3] + s[7]};
            if (jr != j) {
                zc* hr = H1 + ((size_t)(kx * My + jr) * 6 + 2 * f) * Kz + iz;
                hr[0] = zc{s[0] - s[4], s[1] - s[5]};
                hr[Kz] = zc{s[6] - s[2], s[7] - s[3]};
            }
        }
    }
}

__global__ void __launch_bounds__(256)
k_ch_zback(const zc* __restrict__ H1, const zc* __restrict__ Wz, zc* __restrict__ H2, int Kx, int Kz, int Nz,
           int My, int Mz, double kz0, size_t sw) {
    extern __shared__ __align__(16) char lds[];
    H1 += blockIdx.z * sw;
    H2 += blockIdx.z * sw;
    zc* a = reinterpret_cast<zc*>(lds);              // [KB][6][Kz]
    zc* W = a + (size_t)KB * 6 * Kz;                 // [Mz]
    const int kxb = blockIdx.x * KB, j = blockIdx.y, tid = threadIdx.x;
    for (int q = tid; q < KB * 6 * Kz; q += blockDim.x) {
        const int l = q / (6 * Kz), r = q - l * 6 * Kz, kx = kxb + l;
        a[q] = kx < Kx ? H1[(size_t)(kx * My + j) * 6 * Kz + r] : zc{0.0, 0.0};
    }
    for (int q = tid; q < Mz; q += blockDim.x) W[q] = Wz[q];
    __syncthreads();
    const int half = Nz / 2;
    const int neg0 = Mz - (Kz - half);               // (kz mod Mz) of the first negative kz
    for (int p = tid; p < Mz * KB; p += blockDim.x) {
        const int l = p % KB, zm = p / KB, kx = kxb + l;
        const zc* al = a + (size_t)l * 6 * Kz;
        double acc[18];
#pragma unroll
        for (int q = 0; q < 18; ++q) acc[q] = 0.0;
        int idx = 0;                                 // (kz mod Mz) * zm mod Mz
        for (int iz = 0; iz < Kz; ++iz) {
            if (iz == half) idx = (int)(((long long)neg0 * zm) % Mz);
            const zc e = W[idx];
            const double kz = (double)(iz < half ? iz : iz - Kz) * kz0;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const zc v = al[(2 * cc) * Kz + iz], d = al[(2 * cc + 1) * Kz + iz];
                const double pr = fma(e.re, v.re, -(e.im * v.im)), pi = fma(e.re, v.im, e.im * v.re);
                acc[6 * cc + 0] += pr;
                acc[6 * cc + 1] += pi;
                acc[6 * cc + 2] = fma(-kz, pi, acc[6 * cc + 2]);   // d/dz: i kz (e v)
                acc[6 * cc + 3] = fma(kz, pr, acc[6 * cc + 3]);
                acc[6 * cc + 4] += fma(e.re, d.re, -(e.im * d.im));
                acc[6 * cc + 5] += fma(e.re, d.im, e.im * d.re);
            }
            idx += zm;
            if (idx >= Mz) idx -= Mz;
        }
        if (kx < Kx) {
            zc* h = H2 + (size_t)(j * Mz + zm) * 9 * Kx + kx;
#pragma unroll
            for (int q = 0; q < 9; ++q) h[(size_t)q * Kx] = zc{acc[2 * q], acc[2 * q + 1]};
        }
    }
}

__global__ void __launch_bounds__(256)
k_ch_xline(const zc* __restrict__ H2, const zc* __restrict__ Wx, zc* __restrict__ H3, int Kx, int Mz, int Mx,
           double kx0, size_t sw) {
    extern __shared__ __align__(16) char lds[];
    H2 += blockIdx.z * sw;
    H3 += blockIdx.z * sw;
    zc* z = reinterpret_cast<zc*>(lds);              // [9][Kx]: (val, dz, dy) of u, v, w per kx
    zc* W = z + 9 * Kx;                              // [Mx]
    double* G = reinterpret_cast<double*>(W + Mx);   // [3][Mx] products on the line
    const int zm = blockIdx.x, j = blockIdx.y, tid = threadIdx.x;
    const zc* src = H2 + (size_t)(j * Mz + zm) * 9 * Kx;
    for (int q = tid; q < 9 * Kx; q += blockDim.x) z[q] = src[q];
    for (int q = tid; q < Mx; q += blockDim.x) W[q] = Wx[q];
    __syncthreads();
    for (int i = tid; i < Mx; i += blockDim.x) {
        double f[12];                                // per component: value, d/dx, d/dy, d/dz
#pragma unroll
        for (int q = 0; q < 12; ++q) f[q] = 0.0;
        int m = 0;                                   // kx i mod Mx
        for (int kx = 0; kx < Kx; ++kx) {
            const zc t = W[m];
            const double wk = kx == 0 ? 1.0 : 2.0;
            const zc e{wk * t.re, wk * t.im};
            const double kk = kx * kx0;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const zc v = z[(3 * cc) * Kx + kx], dz = z[(3 * cc + 1) * Kx + kx], dy = z[(3 * cc + 2) * Kx + kx];
                const double pr = fma(e.re, v.re, -(e.im * v.im)), pi = fma(e.re, v.im, e.im * v.re);
                f[4 * cc + 0] += pr;
                f[4 * cc + 1] = fma(-kk, pi, f[4 * cc + 1]);
                f[4 * cc + 2] += fma(e.re, dy.re, -(e.im * dy.im));
                f[4 * cc + 3] += fma(e.re, dz.re, -(e.im * dz.im));
            }
            m += i;
            if (m >= Mx) m -= Mx;
        }
        const double u = f[0], v = f[4], w = f[8];
#pragma unroll
        for (int o = 0; o < 3; ++o)
            G[o * Mx + i] = -fma(u, f[4 * o + 1], fma(v, f[4 * o + 2], w * f[4 * o + 3]));
    }
    __syncthreads();
    const double inv = 1.0 / Mx;
    zc* dst = H3 + (size_t)(j * Mz + zm) * 3 * Kx;
    for (int p = tid; p < 3 * Kx; p += blockDim.x) {
        const int o = p / Kx, kx = p - o * Kx;
        const double* g = G + (size_t)o * Mx;
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int i = 0; i < Mx; ++i) {
            const zc t = W[m];
            const double gi = g[i] * inv;
            re = fma(t.re, gi, re);
            im = fma(-t.im, gi, im);
            m += kx;
            if (m >= Mx) m -= Mx;
        }
        dst[p] = zc{re, im};
    }
}

// Even M: the sums of the pair (q, q + M/2) share their terms, e^{2 pi i k (q + M/2)/M} =
// (-1)^k e^{2 pi i k q / M}: a backward transform accumulates the even-k and odd-k terms apart
// (value at q: even + odd, at q + M/2: even - odd), a forward one sums g_q + g_{q+M/2} (even k)
// or g_q - g_{q+M/2} (odd k) over q < M/2.  Half the multiply-adds of the direct sums.

// one kz term of the z-backward sums of the three components (value, d/dz, d/dy), into acc
__device__ __forceinline__ void zback_term(double (&acc)[18], const zc e, const zc* al, int Kz, int iz, double kz) {
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
        const zc v = al[(2 * cc) * Kz + iz], d = al[(2 * cc + 1) * Kz + iz];
        const double pr = fma(e.re, v.re, -(e.im * v.im)), pi = fma(e.re, v.im, e.im * v.re);
        acc[6 * cc + 0] += pr;
        acc[6 * cc + 1] += pi;
        acc[6 * cc + 2] = fma(-kz, pi, acc[6 * cc + 2]);
        acc[6 * cc + 3] = fma(kz, pr, acc[6 * cc + 3]);
        acc[6 * cc + 4] += fma(e.re, d.re, -(e.im * d.im));
        acc[6 * cc + 5] += fma(e.re, d.im, e.im * d.re);
    }
}

// one kx term of the x-backward sums of the 12 factors (value, d/dx, d/dy, d/dz of u, v, w)
__device__ __forceinline__ void xback_term(double (&f)[12], const zc e, const zc* zl, int Kx, int kx, double kk) {
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
        const zc v = zl[(3 * cc) * Kx + kx], dz = zl[(3 * cc + 1) * Kx + kx], dy = zl[(3 * cc + 2) * Kx + kx];
        const double pr = fma(e.re, v.re, -(e.im * v.im)), pi = fma(e.re, v.im, e.im * v.re);
        f[4 * cc + 0] += pr;
        f[4 * cc + 1] = fma(-kk, pi, f[4 * cc + 1]);
        f[4 * cc + 2] += fma(e.re, dy.re, -(e.im * dy.im));
        f[4 * cc + 3] += fma(e.re, dz.re, -(e.im * dz.im));
    }
}

__global__ void __launch_bounds__(256)
k_ch_zback2(const zc* __restrict__ H1, const zc* __restrict__ Wz, zc* __restrict__ H2, int Kx, int Kz, int Nz,
            int My, int Mz, double kz0, size_t sw) {
    extern __shared__ __align__(16) char lds[];
    H1 += blockIdx.z * sw;
    H2 += blockIdx.z * sw;
    zc* a = reinterpret_cast<zc*>(lds);              // [KB][6][Kz]
    zc* W = a + (size_t)KB * 6 * Kz;                 // [Mz]
    const int kxb = blockIdx.x * KB, j = blockIdx.y, tid = threadIdx.x;
    for (int q = tid; q < KB * 6 * Kz; q += blockDim.x) {
        const int l = q / (6 * Kz), r = q - l * 6 * Kz, kx = kxb + l;
        a[q] = kx < Kx ? H1[(size_t)(kx * My + j) * 6 * Kz + r] : zc{0.0, 0.0};
    }
    for (int q = tid; q < Mz; q += blockDim.x) W[q] = Wz[q];
    __syncthreads();
    const int half = Nz / 2, Hz = Mz / 2;
    const int neg0 = Mz - (Kz - half);
    for (int p = tid; p < Hz * KB; p += blockDim.x) {
        const int l = p % KB, zm = p / KB, kx = kxb + l;
        const zc* al = a + (size_t)l * 6 * Kz;
        double ev[18], od[18];
#pragma unroll
        for (int q = 0; q < 18; ++q) {
            ev[q] = 0.0;
            od[q] = 0.0;
        }
        int idx = 0;
        for (int iz = 0; iz < Kz; ++iz) {
            if (iz == half) idx = (int)(((long long)neg0 * zm) % Mz);
            const zc e = W[idx];
            const int kzi = iz < half ? iz : iz - Kz;
            const double kz = (double)kzi * kz0;
            if (kzi & 1)                             // uniform across the workgroup
                zback_term(od, e, al, Kz, iz, kz);
            else
                zback_term(ev, e, al, Kz, iz, kz);
            idx += zm;
            if (idx >= Mz) idx -= Mz;
        }
        if (kx < Kx) {
            zc* h0 = H2 + (size_t)(j * Mz + zm) * 9 * Kx + kx;
            zc* h1 = H2 + (size_t)(j * Mz + zm + Hz) * 9 * Kx + kx;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                h0[(size_t)q * Kx] = zc{ev[2 * q] + od[2 * q], ev[2 * q + 1] + od[2 * q + 1]};
                h1[(size_t)q * Kx] = zc{ev[2 * q] - od[2 * q], ev[2 * q + 1] - od[2 * q + 1]};
            }
        }
    }
}

constexpr int XL = 2;   // grid lines per workgroup of k_ch_xline2

__global__ void __launch_bounds__(192)
k_ch_xline2(const zc* __restrict__ H2, const zc* __restrict__ Wx, zc* __restrict__ H3, int Kx, int Mz, int Mx,
            double kx0, size_t sw) {
    extern __shared__ __align__(16) char lds[];
    H2 += blockIdx.z * sw;
    H3 += blockIdx.z * sw;
    zc* z = reinterpret_cast<zc*>(lds);              // [XL][9][Kx]
    zc* W = z + XL * 9 * Kx;                         // [Mx]
    double* G = reinterpret_cast<double*>(W + Mx);   // [XL][3][Mx]
    const int zm0 = blockIdx.x * XL, j = blockIdx.y, tid = threadIdx.x, Hx = Mx / 2;
    const zc* src = H2 + (size_t)(j * Mz + zm0) * 9 * Kx;   // the XL lines are consecutive
    for (int q = tid; q < XL * 9 * Kx; q += blockDim.x) z[q] = src[q];
    for (int q = tid; q < Mx; q += blockDim.x) W[q] = Wx[q];
    __syncthreads();
    for (int p = tid; p < XL * Hx; p += blockDim.x) {
        const int ln = p / Hx, i = p - ln * Hx;
        const zc* zl = z + (size_t)ln * 9 * Kx;
        double ev[12], od[12];                       // per component: value, d/dx, d/dy, d/dz
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            ev[q] = 0.0;
            od[q] = 0.0;
        }
        // kx = 0 (weight 1) into the even sums, then pairs (odd, even) with weight 2
        xback_term(ev, W[0], zl, Kx, 0, 0.0);
        int m = i;                                   // kx i mod Mx at kx = 1
        for (int kx = 1; kx < Kx; kx += 2) {
            const zc t = W[m];
            xback_term(od, zc{2.0 * t.re, 2.0 * t.im}, zl, Kx, kx, kx * kx0);
            m += i;
            if (m >= Mx) m -= Mx;
            if (kx + 1 < Kx) {
                const zc t2 = W[m];
                xback_term(ev, zc{2.0 * t2.re, 2.0 * t2.im}, zl, Kx, kx + 1, (kx + 1) * kx0);
                m += i;
                if (m >= Mx) m -= Mx;
            }
        }
        double* Gl = G + (size_t)ln * 3 * Mx;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            double f[12];
#pragma unroll
            for (int q = 0; q < 12; ++q) f[q] = side ? ev[q] - od[q] : ev[q] + od[q];
            const double u = f[0], v = f[4], w = f[8];
#pragma unroll
            for (int o = 0; o < 3; ++o)
                Gl[o * Mx + i + side * Hx] = -fma(u, f[4 * o + 1], fma(v, f[4 * o + 2], w * f[4 * o + 3]));
        }
    }
    __syncthreads();
    // sums and differences of the pairs (i, i + Mx/2), in place: [0, Hx) sums, [Hx, Mx) differences
    for (int p = tid; p < XL * 3 * Hx; p += blockDim.x) {
        const int lo = p / Hx, i = p - lo * Hx;
        double* g = G + (size_t)lo * Mx;
        const double a0 = g[i], a1 = g[i + Hx];
        g[i] = a0 + a1;
        g[i + Hx] = a0 - a1;
    }
    __syncthreads();
    const double inv = 1.0 / Mx;
    for (int p = tid; p < XL * 3 * Kx; p += blockDim.x) {
        const int lo = p / Kx, kx = p - lo * Kx, ln = lo / 3;
        const double* g = G + (size_t)lo * Mx + ((kx & 1) ? Hx : 0);
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int i = 0; i < Hx; ++i) {
            const zc t = W[m];
            const double gi = g[i] * inv;
            re = fma(t.re, gi, re);
            im = fma(-t.im, gi, im);
            m += kx;
            if (m >= Mx) m -= Mx;
        }
        H3[(size_t)(j * Mz + zm0 + ln) * 3 * Kx + (lo - 3 * ln) * Kx + kx] = zc{re, im};
    }
}

__global__ void __launch_bounds__(256)
k_ch_zfwd2(const zc* __restrict__ H3, const zc* __restrict__ Wz, zc* __restrict__ H4, int Kx, int Kz, int Nz,
           int Mz, size_t sw) {
    extern __shared__ __align__(16) char lds[];
    H3 += blockIdx.z * sw;
    H4 += blockIdx.z * sw;
    zc* g = reinterpret_cast<zc*>(lds);              // [KB][3][Mz]: sums in [0, Hz), differences in [Hz, Mz)
    zc* W = g + (size_t)KB * 3 * Mz;                 // [Mz]
    const int kxb = blockIdx.x * KB, j = blockIdx.y, tid = threadIdx.x, Hz = Mz / 2;
    for (int q = tid; q < KB * 3 * Hz; q += blockDim.x) {
        const int l = q / (3 * Hz), r = q - l * 3 * Hz, o = r / Hz, zm = r - o * Hz, kx = kxb + l;
        zc a0{0.0, 0.0}, a1{0.0, 0.0};
        if (kx < Kx) {
            a0 = H3[((size_t)(j * Mz + zm) * 3 + o) * Kx + kx];
            a1 = H3[((size_t)(j * Mz + zm + Hz) * 3 + o) * Kx + kx];
        }
        zc* gl = g + (size_t)(l * 3 + o) * Mz;
        gl[zm] = zc{a0.re + a1.re, a0.im + a1.im};
        gl[zm + Hz] = zc{a0.re - a1.re, a0.im - a1.im};
    }
    for (int q = tid; q < Mz; q += blockDim.x) W[q] = Wz[q];
    __syncthreads();
    const int half = Nz / 2;
    const double inv = 1.0 / Mz;
    for (int p = tid; p < Kz * KB * 3; p += blockDim.x) {
        const int iz = p % Kz, lo = p / Kz, l = lo / 3, o = lo - 3 * l, kx = kxb + l;
        const int kzi = iz < half ? iz : iz - Kz;
        const int kzm = iz < half ? iz : iz - Kz + Mz;
        const zc* gl = g + (size_t)(l * 3 + o) * Mz + ((kzi & 1) ? Hz : 0);
        double re = 0.0, im = 0.0;
        int idx = 0;
        for (int zm = 0; zm < Hz; ++zm) {
            const zc t = W[idx];
            const zc v = gl[zm];
            re += fma(t.re, v.re, t.im * v.im);
            im += fma(t.re, v.im, -(t.im * v.re));
            idx += kzm;
            if (idx >= Mz) idx -= Mz;
        }
        if (kx < Kx) H4[((size_t)(j * Kx + kx) * 3 + o) * Kz + iz] = zc{re * inv, im * inv};
    }
}

__global__ void __launch_bounds__(256)
k_ch_zfwd(const zc* __restrict__ H3, const zc* __restrict__ Wz, zc* __restrict__ H4, int Kx, int Kz, int Nz, int Mz,
          size_t sw) {
    extern __shared__ __align__(16) char lds[];
    H3 += blockIdx.z * sw;
    H4 += blockIdx.z * sw;
    zc* g = reinterpret_cast<zc*>(lds);              // [KB][3][Mz]
    zc* W = g + (size_t)KB * 3 * Mz;                 // [Mz]
    const int kxb = blockIdx.x * KB, j = blockIdx.y, tid = threadIdx.x;
    for (int q = tid; q < KB * 3 * Mz; q += blockDim.x) {
        const int l = q / (3 * Mz), r = q - l * 3 * Mz, o = r / Mz, zm = r - o * Mz, kx = kxb + l;
        g[q] = kx < Kx ? H3[((size_t)(j * Mz + zm) * 3 + o) * Kx + kx] : zc{0.0, 0.0};
    }
    for (int q = tid; q < Mz; q += blockDim.x) W[q] = Wz[q];
    __syncthreads();
    const int half = Nz / 2;
    const double inv = 1.0 / Mz;
    for (int p = tid; p < Kz * KB * 3; p += blockDim.x) {
        const int iz = p % Kz, lo = p / Kz, l = lo / 3, o = lo - 3 * l, kx = kxb + l;
        const int kzm = iz < half ? iz : iz - Kz + Mz;   // kz mod Mz
        const zc* gl = g + (size_t)(l * 3 + o) * Mz;
        double re = 0.0, im = 0.0;
        int idx = 0;
        for (int zm = 0; zm < Mz; ++zm) {
            const zc t = W[idx];
            const zc v = gl[zm];
            re += fma(t.re, v.re, t.im * v.im);      // conj(t) v
            im += fma(t.re, v.im, -(t.im * v.re));
            idx += kzm;
            if (idx >= Mz) idx -= Mz;
        }
        if (kx < Kx) H4[((size_t)(j * Kx + kx) * 3 + o) * Kz + iz] = zc{re * inv, im * inv};
    }
}

// y-forward + conversion of 8 modes: Ay[My-1-j][n] = (-1)^n Ay[j][n] (S_{0->2} keeps the parity
// of n), so the sums run over j < My/2 on g_j + g_{My-1-j} (even n) or g_j - g_{My-1-j} (odd n),
// plus the middle node for odd My.
__global__ void __launch_bounds__(256)
k_ch_yfwd(const zc* __restrict__ H4, const double* __restrict__ Ay, const zc* __restrict__ base, zc* __restrict__ F,
          int Kx, int Kz, int N, int My, int n, int R, size_t sx, size_t sw) {
    extern __shared__ __align__(16) char lds[];
    H4 += blockIdx.z * sw;
    F += blockIdx.z * sx;
    zc* g = reinterpret_cast<zc*>(lds);              // [My][3][MA]; after the fold: sums at j < Hy, differences at My-1-j
    const int kx = blockIdx.y, iz0 = blockIdx.x * MA, tid = threadIdx.x, Hy = My / 2;
    for (int q = tid; q < My * 3 * MA; q += blockDim.x) {
        const int l = q % MA, r = q / MA, jj = r / 3, o = r - 3 * jj, iz = iz0 + l;
        g[q] = iz < Kz ? H4[((size_t)(jj * Kx + kx) * 3 + o) * Kz + iz] : zc{0.0, 0.0};
    }
    __syncthreads();
    for (int q = tid; q < Hy * 3 * MA; q += blockDim.x) {
        const int l = q % MA, r = q / MA, jj = r / 3, o = r - 3 * jj;
        zc* a = g + (size_t)(jj * 3 + o) * MA + l;
        zc* b = g + (size_t)((My - 1 - jj) * 3 + o) * MA + l;
        const zc a0 = *a, a1 = *b;
        *a = zc{a0.re + a1.re, a0.im + a1.im};
        *b = zc{a0.re - a1.re, a0.im - a1.im};
    }
    __syncthreads();
    const int l = tid % MA, iz = iz0 + l;
    for (int p = tid / MA; p < 3 * N; p += blockDim.x / MA) {
        const int e = p / N, nn = p - e * N;
        double re = 0.0, im = 0.0;
        for (int jj = 0; jj < Hy; ++jj) {
            const double a = Ay[(size_t)jj * N + nn];
            const zc v = g[(((nn & 1) ? My - 1 - jj : jj) * 3 + e) * MA + l];
            re = fma(a, v.re, re);
            im = fma(a, v.im, im);
        }
        if (My & 1) {
            const double a = Ay[(size_t)Hy * N + nn];
            const zc v = g[(Hy * 3 + e) * MA + l];
            re = fma(a, v.re, re);
            im = fma(a, v.im, im);
        }
        if (iz < Kz) {
            const size_t at = (size_t)(kx * Kz + iz) * n + R + 4 * nn + e;
            F[at] = zc{base[at].re + re, base[at].im + im};
        }
    }
}

}  // namespace

extern "C" {

size_t cn_ch3d_workspace_bytes(int Nx, int Ny, int Nz, int Mx, int My, int Mz) {
    if (Nx < 2 || Nz < 2 || Ny < 2 || Mx < 1 || My < 1 || Mz < 1) return 0;
    return ChDims{Nx / 2, Nz - 1, Ny, Mx, My, Mz}.bytes();
}

int cn_ch3d_explicit_many(const void* X, const int* fidx, const void* base, void* F, void* ws, size_t ws_bytes,
                          const double* By, const void* Wx, const void* Wz, const double* Ay, int Nx, int Ny, int Nz,
                          int Mx, int My, int Mz, int n, int R, double kx0, double kz0, int members, void* stream) {
    if (!X || !fidx || !base || !F || !ws || !By || !Wx || !Wz || !Ay) {
        cn::set_error("cn_ch3d_explicit: null pointer");
        return CN_EINVAL;
    }
    if (Nx < 2 || Nx % 2 || Nz < 2 || Nz % 2 || Ny < 2 || Mx < Nx || My < Ny || Mz < Nz || R < 0 ||
        n < R + 4 * Ny || members < 1 || members > 65535) {
        cn::set_error("cn_ch3d_explicit: inconsistent sizes");
        return CN_EINVAL;
    }
    const ChDims d{Nx / 2, Nz - 1, Ny, Mx, My, Mz};
    if (ws_bytes / (size_t)members < d.bytes()) {
        cn::set_error("cn_ch3d_explicit: workspace smaller than members x cn_ch3d_workspace_bytes");
        return CN_EINVAL;
    }
    const size_t lds_a = (size_t)3 * d.N * MA * sizeof(zc);
    const size_t lds_b = ((size_t)KB * 6 * d.Kz + Mz) * sizeof(zc);
    const size_t lds_c = ((size_t)9 * d.Kx + Mx) * sizeof(zc) + (size_t)3 * Mx * sizeof(double);
    const size_t lds_d = ((size_t)KB * 3 * Mz + Mz) * sizeof(zc);
    const size_t lds_e = (size_t)My * 3 * MA * sizeof(zc);
    const bool split_x = Mx % 2 == 0 && Mz % XL == 0, split_z = Mz % 2 == 0;
    const size_t lds_c2 = ((size_t)XL * 9 * d.Kx + Mx) * sizeof(zc) + (size_t)XL * 3 * Mx * sizeof(double);
    const size_t cap = 160 * 1024;
    if (lds_a > cap || lds_b > cap || lds_c > cap || lds_d > cap || lds_e > cap || (split_x && lds_c2 > cap)) {
        cn::set_error("cn_ch3d_explicit: grid too large for the transform kernels (LDS)");
        return CN_EUNSUPPORTED;
    }
    // member m: state and F at m * sx, workspace (both regions) at m * sw (grid z = member)
    const size_t sx = (size_t)d.Kx * d.Kz * n, sw = d.region1() + d.region2();
    zc* r1 = static_cast<zc*>(ws);
    zc* r2 = r1 + d.region1();
    hipStream_t s = static_cast<hipStream_t>(stream);
    const unsigned mz = (unsigned)members;
    const dim3 gy((d.Kz + MA - 1) / MA, d.Kx, mz), gz((d.Kx + KB - 1) / KB, My, mz), gx(Mz, My, mz);
    hipLaunchKernelGGL(k_ch_yback, gy, dim3(256), lds_a, s, static_cast<const zc*>(X), fidx, By, r1, d.Kx, d.Kz,
                       d.N, My, n, sx, sw);
    if (split_z)
        hipLaunchKernelGGL(k_ch_zback2, gz, dim3(256), lds_b, s, r1, static_cast<const zc*>(Wz), r2, d.Kx, d.Kz, Nz,
                           My, Mz, kz0, sw);
    else
        hipLaunchKernelGGL(k_ch_zback, gz, dim3(256), lds_b, s, r1, static_cast<const zc*>(Wz), r2, d.Kx, d.Kz, Nz,
                           My, Mz, kz0, sw);
    if (split_x)
        hipLaunchKernelGGL(k_ch_xline2, dim3(Mz / XL, My, mz), dim3(192), lds_c2, s, r2, static_cast<const zc*>(Wx),
                           r1, d.Kx, Mz, Mx, kx0, sw);
    else
        hipLaunchKernelGGL(k_ch_xline, gx, dim3(256), lds_c, s, r2, static_cast<const zc*>(Wx), r1, d.Kx, Mz, Mx, kx0,
                           sw);
    if (split_z)
        hipLaunchKernelGGL(k_ch_zfwd2, gz, dim3(256), lds_d, s, r1, static_cast<const zc*>(Wz), r2, d.Kx, d.Kz, Nz,
                           Mz, sw);
    else
        hipLaunchKernelGGL(k_ch_zfwd, gz, dim3(256), lds_d, s, r1, static_cast<const zc*>(Wz), r2, d.Kx, d.Kz, Nz, Mz,
                           sw);
    hipLaunchKernelGGL(k_ch_yfwd, gy, dim3(256), lds_e, s, r2, Ay, static_cast<const zc*>(base), static_cast<zc*>(F),
                       d.Kx, d.Kz, d.N, My, n, R, sx, sw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        cn::set_error(std::string("cn_ch3d_explicit: ") + hipGetErrorString(e));
        return CN_EHIP;
    }
    return CN_OK;
}

int cn_ch3d_explicit(const void* X, const int* fidx, const void* base, void* F, void* ws, size_t ws_bytes,
                     const double* By, const void* Wx, const void* Wz, const double* Ay, int Nx, int Ny, int Nz,
                     int Mx, int My, int Mz, int n, int R, double kx0, double kz0, void* stream) {
    return cn_ch3d_explicit_many(X, fidx, base, F, ws, ws_bytes, By, Wx, Wz, Ay, Nx, Ny, Nz, Mx, My, Mz, n, R, kx0,
                                 kz0, 1, stream);
}

}  // extern "C"
