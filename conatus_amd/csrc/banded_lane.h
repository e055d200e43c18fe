// Launch tables of the lane-mode streamed kernels (k_solve_s, k_border_s) for one element type,
// one block length and one bound on the border rows; included by the banded_lane_*.hip
// translation units, which define CN_LANE_T (element type), CN_LANE_SPX (steps per block),
// CN_LANE_RB (border rows compiled for) and the two table names.
#pragma once
#include "banded_dev.h"

namespace cnk {
namespace {   // internal to each translation unit: the helpers below differ per (type, block length)

using T = CN_LANE_T;
constexpr int SPX = CN_LANE_SPX;
constexpr int RB = CN_LANE_RB;

template <int KL, int WU, int NR>
void launch_solve_s(dim3 g, dim3 b, size_t lds, hipStream_t s, const void* Uf, const void* Lf, const int* piv,
                    const void* Vt, const void* W, const void* S, const int* spiv, const int* perm, void* ws,
                    const void* bin, void* xout, int64_t rs, int B, int n, int r, int US, int bufbytes) {
    hipLaunchKernelGGL((k_solve_s<T, KL, WU, NR, SPX, RB>), g, b, lds, s, static_cast<const T*>(Uf),
                       static_cast<const T*>(Lf), piv, static_cast<const T*>(Vt), static_cast<const T*>(W),
                       static_cast<const T*>(S), spiv, perm, static_cast<T*>(ws), static_cast<const T*>(bin),
                       static_cast<T*>(xout), rs, B, n, r, US, bufbytes);
}
template <int KL, int WU>
SolveSFn solve_nr(int nr) {
    switch (nr) {
        case 1: return &launch_solve_s<KL, WU, 1>;
        case 2: return &launch_solve_s<KL, WU, 2>;
        case 4: return &launch_solve_s<KL, WU, 4>;
    }
    return nullptr;
}
template <int KL>
SolveSFn solve_wu(int WU, int nr) {
    switch (WU) {
        case 4: if constexpr (KL <= 4) return solve_nr<KL, 4>(nr); else return nullptr;
        case 8: return solve_nr<KL, 8>(nr);
        case 16: return solve_nr<KL, 16>(nr);
    }
    return nullptr;
}

template <int WU, int NR>
void launch_border_s(dim3 g, dim3 b, size_t lds, hipStream_t s, const void* Uf, const void* Vt, void* W, void* S, int B,
                     int n, int r, int US, int bufbytes) {
    hipLaunchKernelGGL((k_border_s<T, WU, NR, SPX, RB>), g, b, lds, s, static_cast<const T*>(Uf),
                       static_cast<const T*>(Vt), static_cast<T*>(W), static_cast<T*>(S), B, n, r, US, bufbytes);
}
template <int WU>
BorderSFn border_nr(int nr) {
    switch (nr) {
        case 1: return &launch_border_s<WU, 1>;
        case 2: return &launch_border_s<WU, 2>;
        case 4: return &launch_border_s<WU, 4>;
    }
    return nullptr;
}

}  // namespace

SolveSFn CN_LANE_SOLVE(int KL, int WU, int nr) {
    switch (KL) {
        case 1: return solve_wu<1>(WU, nr);
        case 2: return solve_wu<2>(WU, nr);
        case 4: return solve_wu<4>(WU, nr);
        case 6: return solve_wu<6>(WU, nr);
    }
    return nullptr;
}

BorderSFn CN_LANE_BORDER(int WU, int nr) {
    switch (WU) {
        case 4: return border_nr<4>(nr);
        case 8: return border_nr<8>(nr);
        case 16: return border_nr<16>(nr);
    }
    return nullptr;
}

}  // namespace cnk
