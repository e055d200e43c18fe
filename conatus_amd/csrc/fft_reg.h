// libcnhip: in-register DFTs of the sizes 2, 3, 4, 6, 8, 12, 16 (mixed radix, natural order in and
// out, forward exp(-2 pi i n k / R), inverse exp(+2 pi i n k / R), no 1/R), and the constants they
// use.  Everything here is fully unrolled at compile time: a zc v[R] held by one lane stays in
// registers.
#pragma once

#include "cn_common.h"

namespace cnf {

using cn::zc;

__device__ __forceinline__ zc zadd(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ zc zsub(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ zc zscale(double s, zc a) { return zc{s * a.re, s * a.im}; }
__device__ __forceinline__ zc zmuli(zc a) { return zc{-a.im, a.re}; }            // i a
__device__ __forceinline__ zc zconj(zc a) { return zc{a.re, -a.im}; }
__device__ __forceinline__ zc zmul(zc a, zc b) { return zc{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// cos and sin of 2 pi q / 16 and 2 pi q / 12 from the first-octant values and exact symmetries
constexpr double C8 = 0.9238795325112867;    // cos(pi / 8)
constexpr double S8 = 0.3826834323650898;    // sin(pi / 8)
constexpr double H2 = 0.7071067811865476;    // sqrt(2) / 2
constexpr double H3 = 0.8660254037844386;    // sqrt(3) / 2
__host__ __device__ constexpr double cos16(int q) {
    constexpr double t[16] = {1.0, C8, H2, S8, 0.0, -S8, -H2, -C8, -1.0, -C8, -H2, -S8, 0.0, S8, H2, C8};
    return t[q];
}
__host__ __device__ constexpr double sin16(int q) {
    constexpr double t[16] = {0.0, S8, H2, C8, 1.0, C8, H2, S8, 0.0, -S8, -H2, -C8, -1.0, -C8, -H2, -S8};
    return t[q];
}
__host__ __device__ constexpr double cos12(int q) {
    constexpr double t[12] = {1.0, H3, 0.5, 0.0, -0.5, -H3, -1.0, -H3, -0.5, 0.0, 0.5, H3};
    return t[q];
}
__host__ __device__ constexpr double sin12(int q) {
    constexpr double t[12] = {0.0, 0.5, H3, 1.0, H3, 0.5, 0.0, -0.5, -H3, -1.0, -H3, -0.5};
    return t[q];
}

// a * exp(-+ 2 pi i e / R) (forward: -, inverse: +) for R dividing 16 or 12; multiplications by
// 1, -1 and +-i are exact swaps and negations
template <int R, bool INV>
__device__ __forceinline__ zc twiddle(zc a, int e) {
    e %= R;
    if (e == 0) return a;
    if (2 * e == R) return zc{-a.re, -a.im};
    if (4 * e == R) return INV ? zmuli(a) : zc{a.im, -a.re};              // -i a forward, +i a inverse
    if (4 * e == 3 * R) return INV ? zc{a.im, -a.re} : zmuli(a);
    double c, s;
    if constexpr (16 % R == 0) {
        c = cos16(e * (16 / R));
        s = sin16(e * (16 / R));
    } else {
        c = cos12(e * (12 / R));
        s = sin12(e * (12 / R));
    }
    const zc w{c, INV ? s : -s};
    return zmul(a, w);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(zc* v);

template <bool INV>
__device__ __forceinline__ void dft2(zc* v) {
    const zc t = v[1];
    v[1] = zsub(v[0], t);
    v[0] = zadd(v[0], t);
}

template <bool INV>
__device__ __forceinline__ void dft3(zc* v) {
    const zc t1 = zadd(v[1], v[2]);
    const zc t2 = zsub(v[0], zscale(0.5, t1));
    const zc dd = zscale(H3, zsub(v[1], v[2]));
    const zc t3 = INV ? zmuli(dd) : zc{dd.im, -dd.re};
    v[0] = zadd(v[0], t1);
    v[1] = zadd(t2, t3);
    v[2] = zsub(t2, t3);
}

template <bool INV>
__device__ __forceinline__ void dft4(zc* v) {
    const zc s02 = zadd(v[0], v[2]), d02 = zsub(v[0], v[2]);
    const zc s13 = zadd(v[1], v[3]), d13 = zsub(v[1], v[3]);
    const zc r = INV ? zmuli(d13) : zc{d13.im, -d13.re};                 // -i d13 forward, +i inverse
    v[0] = zadd(s02, s13);
    v[2] = zsub(s02, s13);
    v[1] = zadd(d02, r);
    v[3] = zsub(d02, r);
}

// R = P Q: n = Q n1 + n2, k = k1 + P k2; P-point DFTs over n1, twiddles exp(-+2 pi i n2 k1 / R),
// Q-point DFTs over n2
template <int P, int Q, bool INV>
__device__ __forceinline__ void dft_pq(zc* v) {
    constexpr int R = P * Q;
    zc y[R];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
        zc s[P];
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) s[n1] = v[Q * n1 + n2];
        dft<P, INV>(s);
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) y[n2 * P + k1] = twiddle<R, INV>(s[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
        zc s[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; ++n2) s[n2] = y[n2 * P + k1];
        dft<Q, INV>(s);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = s[k2];
    }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(zc* v) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        dft2<INV>(v);
    } else if constexpr (R == 3) {
        dft3<INV>(v);
    } else if constexpr (R == 4) {
        dft4<INV>(v);
    } else if constexpr (R == 6) {
        dft_pq<2, 3, INV>(v);
    } else if constexpr (R == 8) {
        dft_pq<4, 2, INV>(v);
    } else if constexpr (R == 12) {
        dft_pq<4, 3, INV>(v);
    } else if constexpr (R == 16) {
        dft_pq<4, 4, INV>(v);
    } else {
        static_assert(R == 16, "unsupported in-register DFT size");
    }
}

}  // namespace cnf
