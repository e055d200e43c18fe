// Shared device helpers of libcnhip: the complex128 value type and the FP64 arithmetic of the
// banded kernels (explicit fma where a product feeds an accumulation; the library is built with
// -ffp-contract=off so that nothing else is contracted).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace cn {

// Message of the last failure on this thread (read by cn_last_error).
void set_error(const std::string& msg);

struct __align__(16) zc {
    double re, im;
};

template <class T> __device__ __forceinline__ T zero();
template <> __device__ __forceinline__ double zero<double>() { return 0.0; }
template <> __device__ __forceinline__ zc zero<zc>() { return zc{0.0, 0.0}; }

template <class T> __device__ __forceinline__ T one();
template <> __device__ __forceinline__ double one<double>() { return 1.0; }
template <> __device__ __forceinline__ zc one<zc>() { return zc{1.0, 0.0}; }

// |re| + |im|: the pivot measure of LAPACK's i?amax (dcabs1)
__device__ __forceinline__ double cabs1(double a) { return fabs(a); }
__device__ __forceinline__ double cabs1(zc a) { return fabs(a.re) + fabs(a.im); }

// a * b
__device__ __forceinline__ double mul(double a, double b) { return a * b; }
__device__ __forceinline__ zc mul(zc a, zc b) {
    return zc{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}

// a + b
__device__ __forceinline__ double add(double a, double b) { return a + b; }
__device__ __forceinline__ zc add(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }

// a - b
__device__ __forceinline__ double sub(double a, double b) { return a - b; }
__device__ __forceinline__ zc sub(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }

// a - l * u
__device__ __forceinline__ double fnms(double l, double u, double a) { return fma(-l, u, a); }
__device__ __forceinline__ zc fnms(zc l, zc u, zc a) {
    double re = fma(-l.re, u.re, a.re);
    re = fma(l.im, u.im, re);
    double im = fma(-l.re, u.im, a.im);
    im = fma(-l.im, u.re, im);
    return zc{re, im};
}

// acc + v * x
__device__ __forceinline__ double fmac(double v, double x, double acc) { return fma(v, x, acc); }
__device__ __forceinline__ zc fmac(zc v, zc x, zc acc) {
    double re = fma(v.re, x.re, acc.re);
    re = fma(-v.im, x.im, re);
    double im = fma(v.re, x.im, acc.im);
    im = fma(v.im, x.re, im);
    return zc{re, im};
}

// 1 / p.  Complex: Smith's range-reduced division (the rule of Fortran complex division),
// written without branches: a is the larger-modulus part.
__device__ __forceinline__ double recip(double p) { return 1.0 / p; }
__device__ __forceinline__ zc recip(zc p) {
    const bool big_re = fabs(p.re) >= fabs(p.im);
    const double a = big_re ? p.re : p.im;
    const double b = big_re ? p.im : p.re;
    const double t = b / a;
    const double rd = 1.0 / fma(b, t, a);
    const double tr = t * rd;
    return big_re ? zc{rd, -tr} : zc{tr, -rd};
}

__device__ __forceinline__ double shfl(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ zc shfl(zc v, int src) { return zc{__shfl(v.re, src, 64), __shfl(v.im, src, 64)}; }

// select without branches or memory (keeps register arrays in registers)
__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }
__device__ __forceinline__ zc sel(bool c, zc a, zc b) { return zc{c ? a.re : b.re, c ? a.im : b.im}; }

// swap v[0] and v[p] (p in 0..K) with per-slot selects: static register indices only
template <class T, int K>
__device__ __forceinline__ void swap_first(T (&v)[K + 1], int p) {
    const T v0 = v[0];
    T nv0 = v0;
#pragma unroll
    for (int i = 1; i <= K; ++i) {
        const bool c = (p == i);
        nv0 = sel(c, v[i], nv0);
        v[i] = sel(c, v0, v[i]);
    }
    v[0] = nv0;
}

__device__ __forceinline__ bool is_zero(double a) { return a == 0.0; }
__device__ __forceinline__ bool is_zero(zc a) { return a.re == 0.0 && a.im == 0.0; }

}  // namespace cn
