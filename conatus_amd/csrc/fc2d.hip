// libcnhip Fourier x Chebyshev explicit-term pipeline of 2D Rayleigh-Benard convection:
//   F = S_{0->2} coefficients of ( -u.grad(u), -u.grad(T) ) on the 3/2-dealiased grid
// (the reference's FourierChebyshev2DRHS, python/conatus/timestep/fc2d.py:567-643: 8 factors ->
// 3 outputs, derivative chains, backward transforms, products, forward transforms, C^(lam)
// conversion into the stacked rows).
//
// Two launches (grid y: ensemble member, the members' states (K, n) one after another; for even
// Mx, k_fc_line2 is A with the Fourier sums split into even and odd modes; when members x K >= 4096
// a third launch, k_fc_zback, does A's z-backward once per mode instead of once per line, which
// would otherwise read every member's whole state once per grid line):
//   A  one workgroup per grid line z_j (Mz workgroups per member), everything in LDS:
//      z-backward of the 6 factor columns of the modes (u, u_z, w, w_z, T, T_z; the z-derivative
//      is folded into the matrix Bz_1 = T_n'(z_j)), x-backward of the 8 factors u, u_x, u_z, w,
//      w_x, w_z, T_x, T_z at the Mx points of the line (x-derivative i k kappa on the modes), the 3
//      products, and the x-forward transform of the products (modes k < K: dealiasing).  The
//      Fourier weights come from one table of Mx twiddles W_m = exp(2 pi i m / Mx)
//      (e^{i k kappa x_i} = W_{k i mod Mx}).  Output Gh[o][k][j].
//   B  one workgroup per mode k: z-forward + S_{0->2} conversion of the 3 products,
//      F[k][R + 4n + e] = sum_j Az[j][n] Gh[e][k][j], Az = (S_{0->2} Fz)^T (keeps n < Nz).
// Grids: Mx = floor(3/2 Nx), Mz = floor(3/2 Nz) (exact dealiasing of quadratic products).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cnhip.h"
#include "cn_common.h"

using cn::zc;

namespace {

__global__ void __launch_bounds__(256)
k_fc_line(const zc* __restrict__ X, const int* __restrict__ fidx, const double* __restrict__ Bz,
          const zc* __restrict__ Wt, zc* __restrict__ Gh, int K, int n, int N, int Mz, int Mx, double kappa) {
    extern __shared__ __align__(16) char lds[];
    zc* Zj = reinterpret_cast<zc*>(lds);            // [6][K]: (u, uz, w, wz, T, Tz) at z_j per mode
    zc* W = Zj + 6 * K;                              // [Mx] twiddles
    double* G = reinterpret_cast<double*>(W + Mx);   // [3][Mx] products on the line
    double* b = G + 3 * Mx;                          // [2][N]: T_n(z_j), T_n'(z_j)
    const int j = blockIdx.x, tid = threadIdx.x;
    X += (size_t)blockIdx.y * K * n;
    Gh += (size_t)blockIdx.y * 3 * K * Mz;
    for (int q = tid; q < Mx; q += blockDim.x) W[q] = Wt[q];
    for (int q = tid; q < 2 * N; q += blockDim.x) {
        const int v = q / N, nn = q - v * N;
        b[q] = Bz[((size_t)v * N + nn) * Mz + j];
    }
    __syncthreads();
    // z-backward: Zj[2f + v][k] = sum_n Bz_v[n][j] X[k][pos_f(n)]
    for (int p = tid; p < 3 * K; p += blockDim.x) {
        const int f = p / K, k = p - f * K;
        const zc* Xk = X + (size_t)k * n;
        const int* pos = fidx + (size_t)f * N;
        double r0 = 0.0, i0 = 0.0, r1 = 0.0, i1 = 0.0;
#pragma unroll 8
        for (int nn = 0; nn < N; ++nn) {
            const zc x = Xk[pos[nn]];
            const double b0 = b[nn], b1 = b[N + nn];
            r0 = fma(b0, x.re, r0);
            i0 = fma(b0, x.im, i0);
            r1 = fma(b1, x.re, r1);
            i1 = fma(b1, x.im, i1);
        }
        Zj[(2 * f) * K + k] = zc{r0, i0};
        Zj[(2 * f + 1) * K + k] = zc{r1, i1};
    }
    __syncthreads();
    // x-backward of the 8 factors at x_i and the products
    for (int i = tid; i < Mx; i += blockDim.x) {
        double u = 0, ux = 0, uz = 0, w = 0, wx = 0, wz = 0, Tx = 0, Tz = 0;
        int m = 0;                                   // k i mod Mx
        for (int k = 0; k < K; ++k) {
            const zc t = W[m];
            const double wk = k == 0 ? 1.0 : 2.0;
            const zc e{wk * t.re, wk * t.im};
            const double kk = k * kappa;
            const zc zu = Zj[k], zuz = Zj[K + k], zw = Zj[2 * K + k], zwz = Zj[3 * K + k], zT = Zj[4 * K + k],
                     zTz = Zj[5 * K + k];
            // p = e z: Re(p) is the value, -k kappa Im(p) the x-derivative
            const double pu_r = fma(e.re, zu.re, -(e.im * zu.im)), pu_i = fma(e.re, zu.im, e.im * zu.re);
            const double pw_r = fma(e.re, zw.re, -(e.im * zw.im)), pw_i = fma(e.re, zw.im, e.im * zw.re);
            const double pT_i = fma(e.re, zT.im, e.im * zT.re);
            u += pu_r;
            ux = fma(-kk, pu_i, ux);
            uz += fma(e.re, zuz.re, -(e.im * zuz.im));
            w += pw_r;
            wx = fma(-kk, pw_i, wx);
            wz += fma(e.re, zwz.re, -(e.im * zwz.im));
            Tx = fma(-kk, pT_i, Tx);
            Tz += fma(e.re, zTz.re, -(e.im * zTz.im));
            m += i;
            if (m >= Mx) m -= Mx;
        }
        G[i] = -fma(u, ux, w * uz);
        G[Mx + i] = -fma(u, wx, w * wz);
        G[2 * Mx + i] = -fma(u, Tx, w * Tz);
    }
    __syncthreads();
    // x-forward: Gh[o][k][j] = (1/Mx) sum_i conj(W_{k i mod Mx}) G[o][i].  When the 3K outputs
    // fit the workgroup, one thread per output; else the three products of mode k by one thread
    // (one twiddle read for three outputs).  Same sums in the same order either way.
    const double inv = 1.0 / Mx;
    if (3 * K <= (int)blockDim.x) {
        for (int p = tid; p < 3 * K; p += blockDim.x) {
            const int o = p / K, k = p - o * K;
            const double* g = G + (size_t)o * Mx;
            double re = 0.0, im = 0.0;
            int m = 0;
            for (int i = 0; i < Mx; ++i) {
                const zc t = W[m];
                const double gi = g[i] * inv;
                re = fma(t.re, gi, re);
                im = fma(-t.im, gi, im);
                m += k;
                if (m >= Mx) m -= Mx;
            }
            Gh[((size_t)o * K + k) * Mz + j] = zc{re, im};
        }
        return;
    }
    for (int k = tid; k < K; k += blockDim.x) {
        double re[3] = {0.0, 0.0, 0.0}, im[3] = {0.0, 0.0, 0.0};
        int m = 0;
        for (int i = 0; i < Mx; ++i) {
            const zc t = W[m];
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                const double gi = G[o * Mx + i] * inv;
                re[o] = fma(t.re, gi, re[o]);
                im[o] = fma(-t.im, gi, im[o]);
            }
            m += k;
            if (m >= Mx) m -= Mx;
        }
#pragma unroll
        for (int o = 0; o < 3; ++o) Gh[((size_t)o * K + k) * Mz + j] = zc{re[o], im[o]};
    }
}

// one mode's terms of the 8 factors at x_i (e = weight times twiddle): u, u_x, u_z, w, w_x, w_z,
// T_x, T_z; p = e z, Re(p) the value, -k kappa Im(p) the x-derivative
__device__ __forceinline__ void fc_term(double (&f)[8], const zc e, const zc* Zj, int K, int k, double kk) {
    const zc zu = Zj[k], zuz = Zj[K + k], zw = Zj[2 * K + k], zwz = Zj[3 * K + k], zT = Zj[4 * K + k],
             zTz = Zj[5 * K + k];
    const double pu_r = fma(e.re, zu.re, -(e.im * zu.im)), pu_i = fma(e.re, zu.im, e.im * zu.re);
    const double pw_r = fma(e.re, zw.re, -(e.im * zw.im)), pw_i = fma(e.re, zw.im, e.im * zw.re);
    const double pT_i = fma(e.re, zT.im, e.im * zT.re);
    f[0] += pu_r;
    f[1] = fma(-kk, pu_i, f[1]);
    f[2] += fma(e.re, zuz.re, -(e.im * zuz.im));
    f[3] += pw_r;
    f[4] = fma(-kk, pw_i, f[4]);
    f[5] += fma(e.re, zwz.re, -(e.im * zwz.im));
    f[6] = fma(-kk, pT_i, f[6]);
    f[7] += fma(e.re, zTz.re, -(e.im * zTz.im));
}

// k_fc_line for even Mx: the points x_i and x_{i+Mx/2} from one pass over the modes
// (e^{i k kappa x_{i+Mx/2}} = (-1)^k e^{i k kappa x_i}: even-k and odd-k terms summed apart), and
// the x-forward over i < Mx/2 of G_i + G_{i+Mx/2} (even k) or G_i - G_{i+Mx/2} (odd k).
// z-backward of every grid line at once for large batches: one workgroup per ZB modes of a member
// reads those modes' coefficients once, Zg[member][j][2f + v][k] = sum_n Bz_v[n][j] X[k][pos_f(n)]
// (the same sums, in the same order, as the fused form inside k_fc_line2); each Bz entry read
// from L2 serves the ZB modes (one workgroup per mode read the whole table once per mode).
constexpr int ZB = 4;
__global__ void __launch_bounds__(256)
k_fc_zback(const zc* __restrict__ X, const int* __restrict__ fidx, const double* __restrict__ Bz,
           zc* __restrict__ Zg, int K, int n, int N, int Mz) {
    extern __shared__ __align__(16) char lds[];
    zc* c = reinterpret_cast<zc*>(lds);             // [ZB][3][N]
    const int k0 = blockIdx.x * ZB, tid = threadIdx.x;
    const int nk = min(ZB, K - k0);
    const zc* X0 = X + ((size_t)blockIdx.y * K + k0) * n;
    Zg += (size_t)blockIdx.y * Mz * 6 * K;
    for (int q = tid; q < ZB * 3 * N; q += blockDim.x) {
        const int kk = q / (3 * N), r = q - kk * 3 * N;
        c[q] = kk < nk ? X0[(size_t)kk * n + fidx[r]] : zc{0.0, 0.0};
    }
    __syncthreads();
    for (int p = tid; p < 3 * Mz; p += blockDim.x) {
        const int f = p / Mz, j = p - f * Mz;
        double r0[ZB], i0[ZB], r1[ZB], i1[ZB];
#pragma unroll
        for (int kk = 0; kk < ZB; ++kk) r0[kk] = i0[kk] = r1[kk] = i1[kk] = 0.0;
#pragma unroll 4
        for (int nn = 0; nn < N; ++nn) {
            const double b0 = Bz[(size_t)nn * Mz + j], b1 = Bz[((size_t)N + nn) * Mz + j];
#pragma unroll
            for (int kk = 0; kk < ZB; ++kk) {
                const zc x = c[((size_t)kk * 3 + f) * N + nn];
                r0[kk] = fma(b0, x.re, r0[kk]);
                i0[kk] = fma(b0, x.im, i0[kk]);
                r1[kk] = fma(b1, x.re, r1[kk]);
                i1[kk] = fma(b1, x.im, i1[kk]);
            }
        }
#pragma unroll
        for (int kk = 0; kk < ZB; ++kk) {
            if (kk >= nk) break;
            zc* z = Zg + ((size_t)j * 6 + 2 * f) * K + k0 + kk;
            z[0] = zc{r0[kk], i0[kk]};
            z[K] = zc{r1[kk], i1[kk]};
        }
    }
}

template <bool PRE>
__global__ void __launch_bounds__(256)
k_fc_line2(const zc* __restrict__ X, const int* __restrict__ fidx, const double* __restrict__ Bz,
           const zc* __restrict__ Wt, zc* __restrict__ Gh, const zc* __restrict__ Zg, int K, int n, int N, int Mz,
           int Mx, double kappa) {
    extern __shared__ __align__(16) char lds[];
    zc* Zj = reinterpret_cast<zc*>(lds);            // [6][K]
    zc* W = Zj + 6 * K;                              // [Mx]
    double* G = reinterpret_cast<double*>(W + Mx);   // [3][Mx]
    double* b = G + 3 * Mx;                          // [2][N]
    const int j = blockIdx.x, tid = threadIdx.x, Hx = Mx / 2;
    X += (size_t)blockIdx.y * K * n;
    Gh += (size_t)blockIdx.y * 3 * K * Mz;
    for (int q = tid; q < Mx; q += blockDim.x) W[q] = Wt[q];
    if constexpr (PRE) {
        const zc* z = Zg + ((size_t)blockIdx.y * Mz + j) * 6 * K;
        for (int q = tid; q < 6 * K; q += blockDim.x) Zj[q] = z[q];
        __syncthreads();
    } else {
        for (int q = tid; q < 2 * N; q += blockDim.x) {
            const int v = q / N, nn = q - v * N;
            b[q] = Bz[((size_t)v * N + nn) * Mz + j];
        }
        __syncthreads();
        for (int p = tid; p < 3 * K; p += blockDim.x) {
            const int f = p / K, k = p - f * K;
            const zc* Xk = X + (size_t)k * n;
            const int* pos = fidx + (size_t)f * N;
            double r0 = 0.0, i0 = 0.0, r1 = 0.0, i1 = 0.0;
#pragma unroll 8
            for (int nn = 0; nn < N; ++nn) {
                const zc x = Xk[pos[nn]];
                const double b0 = b[nn], b1 = b[N + nn];
                r0 = fma(b0, x.re, r0);
                i0 = fma(b0, x.im, i0);
                r1 = fma(b1, x.re, r1);
                i1 = fma(b1, x.im, i1);
            }
            Zj[(2 * f) * K + k] = zc{r0, i0};
            Zj[(2 * f + 1) * K + k] = zc{r1, i1};
        }
        __syncthreads();
    }
    // x-backward: the mode sum of point i is split over PS consecutive lanes (PS = 1, 2 or 4, the
    // most that keep PS * Hx within the workgroup), each over a contiguous run of mode pairs
    // (k, k+1), k odd; the partial sums meet by lane exchanges, in a fixed order.
    const int PS = 4 * Hx <= (int)blockDim.x ? 4 : (2 * Hx <= (int)blockDim.x ? 2 : 1);
    const int npair = K / 2;                         // pairs (2q+1, 2q+2), q < npair (the last may lack k+1)
    for (int pt = tid; pt < PS * Hx; pt += blockDim.x) {
        const int i = pt / PS, part = pt - i * PS;
        const int q0 = part * npair / PS, q1 = (part + 1) * npair / PS;
        double ev[8], od[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            ev[q] = 0.0;
            od[q] = 0.0;
        }
        if (part == 0) fc_term(ev, W[0], Zj, K, 0, 0.0);
        int m = (int)(((long long)(2 * q0 + 1) * i) % Mx);   // k i mod Mx at k = 2 q0 + 1
        for (int k = 2 * q0 + 1; k < 2 * q1 + 1 && k < K; k += 2) {
            const zc t = W[m];
            fc_term(od, zc{2.0 * t.re, 2.0 * t.im}, Zj, K, k, k * kappa);
            m += i;
            if (m >= Mx) m -= Mx;
            if (k + 1 < K) {
                const zc t2 = W[m];
                fc_term(ev, zc{2.0 * t2.re, 2.0 * t2.im}, Zj, K, k + 1, (k + 1) * kappa);
                m += i;
                if (m >= Mx) m -= Mx;
            }
        }
        for (int mask = 1; mask < PS; mask <<= 1) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                ev[q] += __shfl_xor(ev[q], mask, 64);
                od[q] += __shfl_xor(od[q], mask, 64);
            }
        }
        if (part != 0) continue;
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            double f[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) f[q] = side ? ev[q] - od[q] : ev[q] + od[q];
            const int at = i + side * Hx;
            G[at] = -fma(f[0], f[1], f[3] * f[2]);
            G[Mx + at] = -fma(f[0], f[4], f[3] * f[5]);
            G[2 * Mx + at] = -fma(f[0], f[6], f[3] * f[7]);
        }
    }
    __syncthreads();
    for (int p = tid; p < 3 * Hx; p += blockDim.x) {   // sums in [0, Hx), differences in [Hx, Mx)
        const int o = p / Hx, i = p - o * Hx;
        double* g = G + (size_t)o * Mx;
        const double a0 = g[i], a1 = g[i + Hx];
        g[i] = a0 + a1;
        g[i + Hx] = a0 - a1;
    }
    __syncthreads();
    const double inv = 1.0 / Mx;
    for (int p = tid; p < 3 * K; p += blockDim.x) {
        const int o = p / K, k = p - o * K;
        const double* g = G + (size_t)o * Mx + ((k & 1) ? Hx : 0);
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int i = 0; i < Hx; ++i) {
            const zc t = W[m];
            const double gi = g[i] * inv;
            re = fma(t.re, gi, re);
            im = fma(-t.im, gi, im);
            m += k;
            if (m >= Mx) m -= Mx;
        }
        Gh[((size_t)o * K + k) * Mz + j] = zc{re, im};
    }
}

// (ZM modes of a member per workgroup: each Az entry read from L2 serves the ZM modes; the sums
// of every output keep the order of the one-mode form)
template <int ZM>
__global__ void __launch_bounds__(256)
k_fc_mode_rows(const zc* __restrict__ Gh, const double* __restrict__ Az, zc* __restrict__ F, int K, int Mz, int N,
               int n, int R) {
    extern __shared__ __align__(16) char lds[];
    zc* g = reinterpret_cast<zc*>(lds);              // [ZM][3][Mz]: Gh[e][k0 + kk][:]
    const int k0 = blockIdx.x * ZM, tid = threadIdx.x;
    const int nk = min(ZM, K - k0);
    Gh += (size_t)blockIdx.y * 3 * K * Mz;
    F += (size_t)blockIdx.y * K * n;
    for (int q = tid; q < ZM * 3 * Mz; q += blockDim.x) {
        const int kk = q / (3 * Mz), r = q - kk * 3 * Mz;
        const int e = r / Mz, jj = r - e * Mz;
        g[q] = kk < nk ? Gh[((size_t)e * K + k0 + kk) * Mz + jj] : zc{0.0, 0.0};
    }
    __syncthreads();
    for (int p = tid; p < 3 * N; p += blockDim.x) {
        const int e = p / N, nn = p - e * N;
        double re[ZM], im[ZM];
#pragma unroll
        for (int kk = 0; kk < ZM; ++kk) re[kk] = im[kk] = 0.0;
        for (int jj = 0; jj < Mz; ++jj) {
            const double a = Az[(size_t)jj * N + nn];
#pragma unroll
            for (int kk = 0; kk < ZM; ++kk) {
                const zc v = g[((size_t)kk * 3 + e) * Mz + jj];
                re[kk] = fma(a, v.re, re[kk]);
                im[kk] = fma(a, v.im, im[kk]);
            }
        }
#pragma unroll
        for (int kk = 0; kk < ZM; ++kk) {
            if (kk >= nk) break;
            F[(size_t)(k0 + kk) * n + R + 4 * nn + e] = zc{re[kk], im[kk]};
        }
    }
}

}  // namespace

extern "C" {

int cn_rb2d_explicit_many(const void* X, const int* fidx, void* F, void* Gh, const double* Bz, const void* Wx,
                          const double* Az, int K, int Nz, int Mx, int Mz, int n, int R, double kappa, int members,
                          void* stream) {
    if (!X || !fidx || !F || !Gh || !Bz || !Wx || !Az) {
        cn::set_error("cn_rb2d_explicit: null pointer");
        return CN_EINVAL;
    }
    if (K < 1 || Nz < 2 || Mx < 1 || Mz < 1 || n < R + 4 * Nz || members < 1 || members > 65535) {
        cn::set_error("cn_rb2d_explicit: inconsistent sizes");
        return CN_EINVAL;
    }
    const size_t lds_a = (size_t)(6 * K + Mx) * sizeof(zc) + (size_t)(3 * Mx + 2 * Nz) * sizeof(double);
    const size_t lds_b = (size_t)3 * Mz * sizeof(zc);
    if (lds_a > 64 * 1024 || lds_b > 64 * 1024) {
        cn::set_error("cn_rb2d_explicit: grid too large for the line kernels (LDS)");
        return CN_EUNSUPPORTED;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    // many modes: the z-backward of each mode once (k_fc_zback) into the scratch after the line
    // products, then the lines read their values; few modes: fused into the line kernel
    zc* Zg = static_cast<zc*>(Gh) + (size_t)members * 3 * K * Mz;
    const bool pre = Mx % 2 == 0 && (long long)members * K >= 4096 && (size_t)ZB * 3 * Nz * sizeof(zc) <= 64 * 1024;
    if (pre) {
        hipLaunchKernelGGL(k_fc_zback, dim3((K + ZB - 1) / ZB, members), dim3(256), (size_t)ZB * 3 * Nz * sizeof(zc), s,
                           static_cast<const zc*>(X), fidx, Bz, Zg, K, n, Nz, Mz);
        hipLaunchKernelGGL(k_fc_line2<true>, dim3(Mz, members), dim3(256), lds_a, s, static_cast<const zc*>(X), fidx,
                           Bz, static_cast<const zc*>(Wx), static_cast<zc*>(Gh), static_cast<const zc*>(Zg), K, n, Nz,
                           Mz, Mx, kappa);
    } else if (Mx % 2 == 0)
        hipLaunchKernelGGL(k_fc_line2<false>, dim3(Mz, members), dim3(256), lds_a, s, static_cast<const zc*>(X), fidx,
                           Bz, static_cast<const zc*>(Wx), static_cast<zc*>(Gh), static_cast<const zc*>(Zg), K, n, Nz,
                           Mz, Mx, kappa);
    else
        hipLaunchKernelGGL(k_fc_line, dim3(Mz, members), dim3(256), lds_a, s, static_cast<const zc*>(X), fidx, Bz,
                           static_cast<const zc*>(Wx), static_cast<zc*>(Gh), K, n, Nz, Mz, Mx, kappa);
    if (pre && (size_t)ZB * lds_b <= 64 * 1024)
        hipLaunchKernelGGL(k_fc_mode_rows<ZB>, dim3((K + ZB - 1) / ZB, members), dim3(256), ZB * lds_b, s,
                           static_cast<const zc*>(Gh), Az, static_cast<zc*>(F), K, Mz, Nz, n, R);
    else
        hipLaunchKernelGGL(k_fc_mode_rows<1>, dim3(K, members), dim3(256), lds_b, s, static_cast<const zc*>(Gh), Az,
                           static_cast<zc*>(F), K, Mz, Nz, n, R);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        cn::set_error(std::string("cn_rb2d_explicit: ") + hipGetErrorString(e));
        return CN_EHIP;
    }
    return CN_OK;
}

int cn_rb2d_explicit(const void* X, const int* fidx, void* F, void* Gh, const double* Bz, const void* Wx,
                     const double* Az, int K, int Nz, int Mx, int Mz, int n, int R, double kappa, void* stream) {
    return cn_rb2d_explicit_many(X, fidx, F, Gh, Bz, Wx, Az, K, Nz, Mx, Mz, n, R, kappa, 1, stream);
}

}  // extern "C"
