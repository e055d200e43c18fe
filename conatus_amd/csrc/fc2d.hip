// libcnhip Fourier x Chebyshev explicit-term pipeline of 2D Rayleigh-Benard convection:
//   F = S_{0->2} coefficients of ( -u.grad(u), -u.grad(T) ) on the 3/2-dealiased grid
// (the reference's FourierChebyshev2DRHS, python/conatus/timestep/fc2d.py:567-643: 8 factors ->
// 3 outputs, derivative chains, backward transforms, products, forward transforms, C^(lam)
// conversion into the stacked rows).
//
// Four launches, each an HBM/L2-streaming pass with one thread per output entry and the
// transform expressed as a small dense matrix applied along one axis (so no transpose pass):
//   A  z-backward: Z[q][k][j] = sum_n Bz_v[n][j] X[k][pos_f(n)]     (q = (f, v): u, uz, w, wz, T, Tz;
//      the z-derivative is folded into the matrix: Bz_1 = T_n'(z_j))
//   B  x-backward + products, per grid point (i, j): the 8 factors u, ux, uz, w, wx, wz, Tx, Tz
//      are accumulated over k in registers and only the 3 products are written
//   C  x-forward: Gh[o][k][j] = sum_i Fx[k][i] G[o][i][j]             (keeps k < K: dealiasing)
//   D  z-forward + conversion + scatter: F[k][R + 4n + e] = sum_j Az[j][n] Gh[e][k][j],
//      Az = (S_{0->2} Fz)^T  (keeps n < N)
// Grids: Mx = floor(3/2 Nx), Mz = floor(3/2 Nz) (exact dealiasing of quadratic products).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cnhip.h"
#include "cn_common.h"

using cn::zc;

namespace {

__global__ void __launch_bounds__(256)
k_fc_zback(const zc* __restrict__ X, const int* __restrict__ fidx, const double* __restrict__ Bz, zc* __restrict__ Z,
           int K, int n, int N, int Mz) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = 6 * K * Mz;
    if (t >= total) return;
    const int j = t % Mz;
    const int k = (t / Mz) % K;
    const int q = t / (Mz * K);
    const int f = q >> 1, v = q & 1;
    const zc* Xk = X + (size_t)k * n;
    const int* pos = fidx + (size_t)f * N;
    const double* B = Bz + (size_t)v * N * Mz + j;
    double re = 0.0, im = 0.0;
    for (int nn = 0; nn < N; ++nn) {
        const double b = B[(size_t)nn * Mz];
        const zc x = Xk[pos[nn]];
        re = fma(b, x.re, re);
        im = fma(b, x.im, im);
    }
    Z[t] = zc{re, im};
}

__global__ void __launch_bounds__(256)
k_fc_xback_products(const zc* __restrict__ Z, const zc* __restrict__ Ex, double* __restrict__ G, int K, int Mz,
                    int Mx, double kappa) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Mx * Mz) return;
    const int j = t % Mz;
    const int i = t / Mz;
    const size_t qs = (size_t)K * Mz;   // stride between the 6 z-backward arrays
    double u = 0, ux = 0, uz = 0, w = 0, wx = 0, wz = 0, Tx = 0, Tz = 0;
    const zc* Er = Ex + (size_t)i * K;
    for (int k = 0; k < K; ++k) {
        const zc e = Er[k];
        const double kk = k * kappa;
        const zc* Zk = Z + (size_t)k * Mz + j;
        const zc zu = Zk[0], zuz = Zk[qs], zw = Zk[2 * qs], zwz = Zk[3 * qs], zT = Zk[4 * qs], zTz = Zk[5 * qs];
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
    }
    const size_t P = (size_t)Mx * Mz;
    G[t] = -fma(u, ux, w * uz);
    G[P + t] = -fma(u, wx, w * wz);
    G[2 * P + t] = -fma(u, Tx, w * Tz);
}

__global__ void __launch_bounds__(256)
k_fc_xfwd(const double* __restrict__ G, const zc* __restrict__ Fx, zc* __restrict__ Gh, int K, int Mz, int Mx) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * K * Mz) return;
    const int j = t % Mz;
    const int k = (t / Mz) % K;
    const int o = t / (Mz * K);
    const double* Go = G + (size_t)o * Mx * Mz + j;
    const zc* F = Fx + (size_t)k * Mx;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < Mx; ++i) {
        const double g = Go[(size_t)i * Mz];
        const zc f = F[i];
        re = fma(f.re, g, re);
        im = fma(f.im, g, im);
    }
    Gh[t] = zc{re, im};
}

__global__ void __launch_bounds__(256)
k_fc_zfwd_rows(const zc* __restrict__ Gh, const double* __restrict__ Az, zc* __restrict__ F, int K, int Mz, int N,
               int n, int R) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * K * N) return;
    const int nn = t % N;
    const int k = (t / N) % K;
    const int e = t / (N * K);
    const zc* g = Gh + ((size_t)e * K + k) * Mz;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < Mz; ++j) {
        const double a = Az[(size_t)j * N + nn];
        const zc v = g[j];
        re = fma(a, v.re, re);
        im = fma(a, v.im, im);
    }
    F[(size_t)k * n + R + 4 * nn + e] = zc{re, im};
}

dim3 g256(size_t N) { return dim3((unsigned)((N + 255) / 256)); }

}  // namespace

extern "C" {

int cn_rb2d_explicit(const void* X, const int* fidx, void* F, void* Z, void* G, void* Gh, const double* Bz,
                     const void* Ex, const void* Fx, const double* Az, int K, int Nz, int Mx, int Mz, int n, int R,
                     double kappa, void* stream) {
    if (!X || !fidx || !F || !Z || !G || !Gh || !Bz || !Ex || !Fx || !Az) {
        cn::set_error("cn_rb2d_explicit: null pointer");
        return CN_EINVAL;
    }
    if (K < 1 || Nz < 2 || Mx < 1 || Mz < 1 || n < R + 4 * Nz) {
        cn::set_error("cn_rb2d_explicit: inconsistent sizes");
        return CN_EINVAL;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_fc_zback, g256((size_t)6 * K * Mz), dim3(256), 0, s, static_cast<const zc*>(X), fidx, Bz,
                       static_cast<zc*>(Z), K, n, Nz, Mz);
    hipLaunchKernelGGL(k_fc_xback_products, g256((size_t)Mx * Mz), dim3(256), 0, s, static_cast<const zc*>(Z),
                       static_cast<const zc*>(Ex), static_cast<double*>(G), K, Mz, Mx, kappa);
    hipLaunchKernelGGL(k_fc_xfwd, g256((size_t)3 * K * Mz), dim3(256), 0, s, static_cast<const double*>(G),
                       static_cast<const zc*>(Fx), static_cast<zc*>(Gh), K, Mz, Mx);
    hipLaunchKernelGGL(k_fc_zfwd_rows, g256((size_t)3 * K * Nz), dim3(256), 0, s, static_cast<const zc*>(Gh), Az,
                       static_cast<zc*>(F), K, Mz, Nz, n, R);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        cn::set_error(std::string("cn_rb2d_explicit: ") + hipGetErrorString(e));
        return CN_EHIP;
    }
    return CN_OK;
}

}  // extern "C"
