// Launch tables of the lane-mode streamed kernels (k_solve_s, k_border_s) for one element type,
// one block length, one bound on the border rows and one band width kl.  The build
// (conatus_amd/build.py) compiles this file once per combination, defining CN_LANE_ZC (1:
// complex128, 0: float64), CN_LANE_SPX (steps per block), CN_LANE_RB (border rows compiled for),
// CN_LANE_KL, the solve table's name CN_LANE_SOLVE and, in the kl = 1 objects, the border
// completion table's name CN_LANE_BORDER; one small object per combination keeps the build
// parallel.
#include "banded_dev.h"

#if !defined(CN_LANE_ZC) || !defined(CN_LANE_SPX) || !defined(CN_LANE_RB) || !defined(CN_LANE_KL) || \
    !defined(CN_LANE_SOLVE)
#error "banded_lane.hip is compiled by conatus_amd/build.py with the CN_LANE_* definitions of one variant"
#endif

namespace cnk {
namespace {   // internal to each translation unit: the helpers below differ per (type, block length)

#if CN_LANE_ZC
using T = cn::zc;
#else
using T = double;
#endif
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

#ifdef CN_LANE_BORDER
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
#endif

}  // namespace

SolveSFn CN_LANE_SOLVE(int WU, int nr) { return solve_wu<CN_LANE_KL>(WU, nr); }

#ifdef CN_LANE_BORDER
BorderSFn CN_LANE_BORDER(int WU, int nr) {
    switch (WU) {
        case 4: return border_nr<4>(nr);
        case 8: return border_nr<8>(nr);
        case 16: return border_nr<16>(nr);
    }
    return nullptr;
}
#endif

}  // namespace cnk
