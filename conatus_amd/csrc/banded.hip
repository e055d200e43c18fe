// libcnhip: batched bordered-banded factor and solve for MI355X (gfx950), FP64 real and complex.
//
// Algorithm (the reference's, core/include/conatus/banded.hpp:1-8,253-341): after moving the r
// boundary rows to the top, P A = B + U V^T with B banded (kl, ku) whose border rows are s_i e_i,
// U = [e_0 .. e_{r-1}], V^T = border rows - diag(s).  B = P_B L U (partial pivoting inside the
// band, LAPACK gbtf2 order), W = B^-1 U, S = I + V^T W, and A x = b is solved as
// y = B^-1 P b, z = S^-1 V^T y, x = y - W z.
//
// MI355X design (not a translation of the reference's warp kernels):
//  * Storage is interleaved by 16 systems: entry e of system m lives at
//    ((m/16)*len + e)*16 + m%16, so the solve's lanes (consecutive systems) read 256 contiguous
//    bytes per factor entry and a 16-system chunk is produced by one workgroup.
//  * Factor: one group of G (16/32/64) lanes per system, lane c owns band column q = c (mod M),
//    M = kl+ku+1, and holds the kl+1 active rows of that column in registers (a window that
//    shifts by one row per step, so that every register index is static).  The
//    owner of the pivot column finds the pivot, forms the reciprocal and the multipliers, and
//    broadcasts them over the wavefront (ds_bpermute); every other lane swaps and updates its own
//    column.  r extra lanes carry the columns of U (unit vectors) through the same elimination,
//    so the forward half of W = B^-1 U costs nothing extra.  Row j of U leaves the window as one
//    coalesced row store; the incoming band row is one coalesced row load.  U's diagonal is
//    stored as its reciprocal so that solves multiply instead of divide.
//  * Border: one lane per (system, border column) finishes W by back substitution (in place over
//    its forward half) and accumulates the column of S = I + V^T W; one lane per system LU-factors
//    S (partial pivoting) and forms cond_1(S) from the explicit inverse, as banded.hpp:403-432.
//  * Solve: one lane per (system, right-hand side); forward sweep (pivots + L), backward sweep
//    (U, window of the last wu solutions in registers) with V^T y accumulated as y is produced,
//    z = S^-1 V^T y, then x = y - W z.  For narrow bands the sweeps are streamed: the factor rows
//    of the next block of steps are copied HBM -> LDS (global_load_lds) while the current block
//    is computed.  Gather/scatter maps and the split complex mode of the reference's device
//    solver are folded into the loads and stores of the direct kernel.
//  * Transposed solves (the adjoint path) use Wt = B^-T V and St = I + U^T Wt, built on demand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cnhip.h"
#include "cn_common.h"
#include "banded_dev.h"

using cn::zc;

using namespace cnk;

// Lane work space of cn_bb_solve_noalloc (see scratch_for_device below): 64 MiB, enough for the
// four right-hand sides of 4096 complex systems of n = 256.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {

thread_local std::string g_err;
// kernel mode of the handles this thread creates next: 0 automatic, 1 lane, 2 wave
thread_local int g_mode_request = 0;

struct CnError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw CnError{code, msg}; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(CN_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

size_t checked_mul(std::initializer_list<size_t> dims) {
    size_t p = 1;
    for (size_t d : dims) {
        if (d != 0 && p > std::numeric_limits<size_t>::max() / d) fail(CN_EOVERFLOW, "array dimensions overflow size_t");
        p *= d;
    }
    return p;
}

// ------------------------------------------------------------------------------------------
// Host side
// ------------------------------------------------------------------------------------------
constexpr int KL_SET[] = {1, 2, 4, 6, 8, 10, 12, 14, 16};
constexpr int WU_SET[] = {4, 8, 16, 24, 32};

int pick_kl(int kl) {
    for (int v : KL_SET) if (v >= kl) return v;
    return -1;
}
int pick_wu(int w) {
    for (int v : WU_SET) if (v >= w) return v;
    return -1;
}

using FactorFn = void (*)(dim3, dim3, hipStream_t, const void*, void*, void*, int*, void*, int*, int, int,
                          int, int, int, int, int, int, int);
using SolveFn = void (*)(dim3, dim3, hipStream_t, const void*, const void*, const int*, const void*, const void*,
                         const void*, const int*, const int*, void*, SolveIO, int, int, int, int, int);

template <class T, int KL>
void launch_factor(dim3 g, dim3 b, hipStream_t s, const void* Ab, void* Uf, void* Lf, int* piv, void* Wf,
                   int* info, int B, int n, int M, int r, int G, int US, int pivot, int il, int ucol) {
    hipLaunchKernelGGL((k_band_lu<T, KL>), g, b, 0, s, static_cast<const T*>(Ab), static_cast<T*>(Uf),
                       static_cast<T*>(Lf), piv, static_cast<T*>(Wf), info, B, n, M, r, G, US, pivot, il, ucol);
}
template <class T, int KL, int WU>
void launch_solve(dim3 g, dim3 b, hipStream_t s, const void* Uf, const void* Lf, const int* piv, const void* Vt,
                  const void* W, const void* S, const int* spiv, const int* perm, void* ws, SolveIO io, int B,
                  int n, int r, int US, int nrhs) {
    hipLaunchKernelGGL((k_solve<T, KL, WU>), g, b, 0, s, static_cast<const T*>(Uf), static_cast<const T*>(Lf),
                       piv, static_cast<const T*>(Vt), static_cast<const T*>(W), static_cast<const T*>(S), spiv,
                       perm, static_cast<T*>(ws), io, B, n, r, US, nrhs);
}

using SolveTFn = void (*)(dim3, dim3, hipStream_t, const void*, const void*, const int*, const void*, void*,
                          const void*, const int*, const int*, void*, void*, int64_t, int, int, int, int, int, int,
                          int, int);

template <class T, int KL, int WU>
void launch_solve_t(dim3 g, dim3 b, hipStream_t s, const void* Uf, const void* Lf, const int* piv, const void* Vt,
                    void* Wt, const void* St, const int* stpiv, const int* perm, void* ws, void* x, int64_t rs, int B,
                    int n, int r, int US, int nrhs, int prep, int il, int ucol) {
    hipLaunchKernelGGL((k_solve_t<T, KL, WU>), g, b, 0, s, static_cast<const T*>(Uf), static_cast<const T*>(Lf),
                       piv, static_cast<const T*>(Vt), static_cast<T*>(Wt), static_cast<const T*>(St), stpiv, perm,
                       static_cast<T*>(ws), static_cast<T*>(x), rs, B, n, r, US, nrhs, prep, il, ucol);
}

template <class T, int KL>
SolveTFn solve_t_fn_kl(int WU) {
    if (WU < KL) return nullptr;
    switch (WU) {
        case 4: if constexpr (KL <= 4) return &launch_solve_t<T, KL, 4>; else return nullptr;
        case 8: if constexpr (KL <= 8) return &launch_solve_t<T, KL, 8>; else return nullptr;
        case 16: return &launch_solve_t<T, KL, 16>;
        case 24: return &launch_solve_t<T, KL, 24>;
        case 32: return &launch_solve_t<T, KL, 32>;
    }
    return nullptr;
}
template <class T>
SolveTFn solve_t_fn(int KL, int WU) {
    switch (KL) {
        case 1: return solve_t_fn_kl<T, 1>(WU);
        case 2: return solve_t_fn_kl<T, 2>(WU);
        case 4: return solve_t_fn_kl<T, 4>(WU);
        case 6: return solve_t_fn_kl<T, 6>(WU);
        case 8: return solve_t_fn_kl<T, 8>(WU);
        case 10: return solve_t_fn_kl<T, 10>(WU);
        case 12: return solve_t_fn_kl<T, 12>(WU);
        case 14: return solve_t_fn_kl<T, 14>(WU);
        case 16: return solve_t_fn_kl<T, 16>(WU);
    }
    return nullptr;
}

template <class T>
FactorFn factor_fn(int KL) {
    switch (KL) {
        case 1: return &launch_factor<T, 1>;
        case 2: return &launch_factor<T, 2>;
        case 4: return &launch_factor<T, 4>;
        case 6: return &launch_factor<T, 6>;
        case 8: return &launch_factor<T, 8>;
        case 10: return &launch_factor<T, 10>;
        case 12: return &launch_factor<T, 12>;
        case 14: return &launch_factor<T, 14>;
        case 16: return &launch_factor<T, 16>;
    }
    return nullptr;
}
template <class T, int KL>
SolveFn solve_fn_kl(int WU) {
    if (WU < KL) return nullptr;
    switch (WU) {
        case 4: if constexpr (KL <= 4) return &launch_solve<T, KL, 4>; else return nullptr;
        case 8: if constexpr (KL <= 8) return &launch_solve<T, KL, 8>; else return nullptr;
        case 16: return &launch_solve<T, KL, 16>;
        case 24: return &launch_solve<T, KL, 24>;
        case 32: return &launch_solve<T, KL, 32>;
    }
    return nullptr;
}
using BorderFn = void (*)(dim3, dim3, hipStream_t, const void*, const void*, void*, void*, int, int, int, int);
template <class T, int WU>
void launch_border(dim3 g, dim3 b, hipStream_t s, const void* Uf, const void* Vt, void* W, void* S, int B, int n,
                   int r, int US) {
    hipLaunchKernelGGL((k_border<T, WU>), g, b, 0, s, static_cast<const T*>(Uf), static_cast<const T*>(Vt),
                       static_cast<T*>(W), static_cast<T*>(S), B, n, r, US);
}
template <class T>
BorderFn border_fn(int WU) {
    switch (WU) {
        case 4: return &launch_border<T, 4>;
        case 8: return &launch_border<T, 8>;
        case 16: return &launch_border<T, 16>;
        case 24: return &launch_border<T, 24>;
        case 32: return &launch_border<T, 32>;
    }
    return nullptr;
}

// streamed kernels for r border rows (compiled for rb_for(r) rows) and blocks of spx steps
#define CN_LANE_BY_KL(tag)                                        \
    switch (KL) {                                                 \
        case 1: return lane_solve_##tag##_k1(WU, nr);             \
        case 2: return lane_solve_##tag##_k2(WU, nr);             \
        case 4: return lane_solve_##tag##_k4(WU, nr);             \
        case 6: return lane_solve_##tag##_k6(WU, nr);             \
    }                                                             \
    return nullptr;
template <class T>
SolveSFn solve_s_fn(int KL, int WU, int nr, int spx, int r) {
    const bool r2 = rb_for(r) == 2;
    if (r > MAXR) return nullptr;
    if constexpr (sizeof(T) == 16) {
        if (spx == 16) { if (r2) { CN_LANE_BY_KL(zc16_r2) } else { CN_LANE_BY_KL(zc16_r8) } }
        if (spx == 8) { if (r2) { CN_LANE_BY_KL(zc8_r2) } else { CN_LANE_BY_KL(zc8_r8) } }
    } else {
        if (spx == 16) { if (r2) { CN_LANE_BY_KL(d16_r2) } else { CN_LANE_BY_KL(d16_r8) } }
        if (spx == 8) { if (r2) { CN_LANE_BY_KL(d8_r2) } else { CN_LANE_BY_KL(d8_r8) } }
    }
    return nullptr;
}
#undef CN_LANE_BY_KL
template <class T>
BorderSFn border_s_fn(int WU, int nr, int spx, int r) {
    const bool r2 = rb_for(r) == 2;
    if (r > MAXR) return nullptr;
    if constexpr (sizeof(T) == 16) {
        if (spx == 16) return r2 ? lane_border_zc16_r2(WU, nr) : lane_border_zc16_r8(WU, nr);
        if (spx == 8) return r2 ? lane_border_zc8_r2(WU, nr) : lane_border_zc8_r8(WU, nr);
    } else {
        if (spx == 16) return r2 ? lane_border_d16_r2(WU, nr) : lane_border_d16_r8(WU, nr);
        if (spx == 8) return r2 ? lane_border_d8_r2(WU, nr) : lane_border_d8_r8(WU, nr);
    }
    return nullptr;
}

template <class T>
SolveFn solve_fn(int KL, int WU) {
    switch (KL) {
        case 1: return solve_fn_kl<T, 1>(WU);
        case 2: return solve_fn_kl<T, 2>(WU);
        case 4: return solve_fn_kl<T, 4>(WU);
        case 6: return solve_fn_kl<T, 6>(WU);
        case 8: return solve_fn_kl<T, 8>(WU);
        case 10: return solve_fn_kl<T, 10>(WU);
        case 12: return solve_fn_kl<T, 12>(WU);
        case 14: return solve_fn_kl<T, 14>(WU);
        case 16: return solve_fn_kl<T, 16>(WU);
    }
    return nullptr;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void alloc(size_t b) {
        release();
        bytes = b;
        if (b) hip_check(hipMalloc(&p, b), "hipMalloc");
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    ~DevBuf() { release(); }
};

}  // namespace

void cn::set_error(const std::string& msg) { g_err = msg; }

struct cn_bb {
    int dtype = CN_C128, B = 0, Bp = 0, n = 0, r = 0, kl = 0, ku = 0, KLP = 0, M = 0, WU = 0, US = 0, G = 0;
    int pivot = 1;
    int streamed = 1;   // 0: always use the direct solve kernel (CNHIP_SOLVE=direct)
    int mode = 0;       // 0: lane mode (one lane per system and rhs), 1: wave mode (one wavefront each)
    int il = 16;        // interleave of the factor arrays: 16 systems (lane mode) or 1 (wave mode)
    int parts = 1;      // wave mode: wavefronts per system of the solve (>= 2: partitioned sweeps, spikes)
    int wave_sp = 0;    // wave mode: staged block length asked for (0: by the launch size, wave_steps)
    int plen = 0, ls = 0, lsu = 0;   // longest chunk; row strides of the forward / backward spikes
    int nrhs_cap = 0;
    size_t esz = 16;
    double cond_max = 1.0, factor_ms = 0.0;
    int device = 0;
    DevBuf Ab, Uf, Lf, piv, Vt, W, S, spiv, cond, info, perm, ws;
    DevBuf Wt, St, stpiv, cond_t, info_t;   // transposed Woodbury data (built on demand)
    DevBuf SpL, SpU;                        // partitioned wave mode: the chunks' coupling responses
    bool tr_ready = false;
    std::vector<double> hcond;
    std::vector<int> hperm;   // host copy of the row permutation (border rows first)
};

namespace {

// Host layout of one system into the interleaved device arrays (banded.hpp:170-251 semantics).
template <class T>
void layout_one(const cn_bb& h, int m, const int32_t* ip, const int32_t* ix, const T* dv,
                const std::vector<int>& perm, const std::vector<int>& pos, std::vector<T>& Ab, std::vector<T>& Vt) {
    const int n = h.n, r = h.r, M = h.M, KLP = h.KLP;
    const size_t il = h.il, ch = m / il, ml = m % il;
    auto AB = [&](int i, int d) -> T& { return Ab[(ch * ((size_t)n * M) + (size_t)i * M + d) * il + ml]; };
    auto VT = [&](int a, int j) -> T& { return Vt[(ch * ((size_t)r * n) + (size_t)a * n + j) * il + ml]; };
    for (int i = 0; i < n; ++i) {
        const int o = perm[i];
        if (i < r) {
            double s = 0.0;
            for (int p = ip[o]; p < ip[o + 1]; ++p) {
                s = std::max(s, static_cast<double>(std::abs(dv[p])));
                VT(i, ix[p]) += dv[p];
            }
            if (s == 0.0) s = 1.0;
            AB(i, KLP) = T(s);
            VT(i, i) -= T(s);
        } else {
            for (int p = ip[o]; p < ip[o + 1]; ++p) {
                if (dv[p] == T(0)) continue;
                const int j = ix[p];
                AB(i, j - i + KLP) += dv[p];
            }
        }
    }
    (void)pos;
}

// wave mode: one staging buffer, and the whole LDS of a wavefront ([x: n entries][2 buffers])
// staged block of the wave-mode sweeps for a launch of `waves` wavefronts: the longer block while
// the batch leaves CUs idle at two workgroups per CU, the shorter one (more resident) beyond
int wave_steps(size_t waves) { return waves > 512 ? SPW_MANY : SPW_FEW; }
constexpr int PART_MAX_B = 256;   // automatic partitioned sweeps up to this many systems
int wave_buf(const cn_bb& h, int sp) { return wave_buf_bytes(h.KLP, h.US, (int)h.esz, sp); }
size_t wave_lds_bytes(const cn_bb& h, int sp) {
    return (size_t)align16(h.n * (int)h.esz) + WAVE_SMALL + wave_nbuf(sp) * (size_t)wave_buf(h, sp);
}

// Partitioned wave mode.  LDS of one wavefront's chunk rows (the chunk and its coupling rows)
int part_yw_bytes(const cn_bb& h) { return align16((h.plen + std::max(h.KLP, h.M - 1)) * (int)h.esz); }
// LDS of the partitioned solve with staged blocks of sp steps
size_t part_lds_bytes(const cn_bb& h, int P, int sp) {
    return (size_t)align16(h.n * (int)h.esz) + WAVE_SMALL + (size_t)P * (PART_W + MAXR) * h.esz +
           (size_t)P * (part_yw_bytes(h) + wave_nbuf(sp) * (size_t)wave_buf(h, sp));
}
// staged block length of the partitioned solve: the longest whose LDS fits (0: none)
int part_steps(const cn_bb& h, int P) {
    for (int sp : {16, SPW_MANY})
        if (part_lds_bytes(h, P, sp) <= (size_t)LDS_CAP) return sp;
    return 0;
}
// Wavefronts per system: 8, else 4, when the chunks are at least one coupling width (max(kl,
// kl + ku)) long; CNHIP_PARTS = 2, 4 or 8 asks for that count (2 splits the chain too little to pay
// for the coupling: measured no faster than one wavefront); 1 when none applies or the LDS does not
// fit.
int pick_parts(cn_bb& h) {
    const int wmax = std::max(h.KLP, h.M - 1);
    if (h.KLP < 1 || wmax > PART_W) return 1;
    const char* e = std::getenv("CNHIP_PARTS");
    const int want = e ? std::atoi(e) : 0;
    for (int P : {8, 4, 2}) {
        if (want > 0 ? P != want : P == 2) continue;
        if (h.n / P < wmax) continue;
        h.plen = (h.n + P - 1) / P;
        if (part_steps(h, P) == 0) continue;
        return P;
    }
    h.plen = 0;
    return 1;
}

// Block length of the streamed sweeps: 16 steps when two buffers of buf16 bytes (the largest sweep
// buffer) and the two stages fit in LDS, else 8 (0: none fits); CNHIP_SPX=8 forces the short
// blocks (tests of both variants).
int pick_spx(int buf16, int buf8, int esz) {
    const char* e = std::getenv("CNHIP_SPX");
    const bool force8 = e && e[0] == '8' && e[1] == 0;
    if (!force8 && 2 * buf16 + 2 * stage_bytes(16, esz) <= LDS_CAP) return 16;
    return 2 * buf8 + 2 * stage_bytes(8, esz) <= LDS_CAP ? 8 : 0;
}
// Buffer region of a streamed launch: every sweep gets three buffers when they fit beside the
// stages, else two (the kernels derive the same counts from the region, sweep_nbuf)
int buffer_region(std::initializer_list<int> sweeps, int spx, int esz) {
    const int room = LDS_CAP - 2 * stage_bytes(spx, esz);
    int region = 0;
    for (int x : sweeps) region = std::max(region, (3 * x <= room ? 3 : 2) * x);
    return region;
}
// LDS bytes of a streamed launch: the buffer region and two stages
size_t streamed_lds(int region, int spx, int esz) { return (size_t)region + 2 * (size_t)stage_bytes(spx, esz); }

// LDS buffer region of the streamed border completion (block length in spx), or -1
int border_bufbytes(const cn_bb& h, int& spx) {
    spx = 0;
    if (!h.streamed || h.WU > 16) return -1;
    const int e = (int)h.esz;
    const int rb = rb_for(h.r);
    spx = pick_spx(bwd_bytes(16, h.US, rb, e), bwd_bytes(8, h.US, rb, e), e);
    return spx ? buffer_region({bwd_bytes(spx, h.US, rb, e)}, spx, e) : -1;
}

int nr_for(int r) { return r >= 4 ? 4 : (r >= 2 ? 2 : 1); }

template <class T>
void run_factor(cn_bb& h, hipStream_t s) {
    const int spb = 256 / h.G;
    const dim3 fb(256), fg((h.Bp + spb - 1) / spb);
    FactorFn f = factor_fn<T>(h.KLP);
    if (!f) fail(CN_ELOGIC, "no factor kernel for kl " + std::to_string(h.KLP));
    // the factor kernel indexes within an interleave chunk with 32-bit offsets
    if ((int64_t)h.n * std::max({h.M, h.US, h.r, 1}) * h.il >= ((int64_t)1 << 31))
        fail(CN_EOVERFLOW, "BandedBatch: n = " + std::to_string(h.n) + " is too large for the factor kernel");
    // band LU, with the forward half L^-1 P_B U of W in the r extra lanes
    f(fg, fb, s, h.Ab.p, h.Uf.p, h.Lf.p, static_cast<int*>(h.piv.p), h.W.p, static_cast<int*>(h.info.p), h.B, h.n,
      h.M, h.r, h.G, h.US, h.pivot, h.il, h.mode);
    hip_check(hipGetLastError(), "k_band_lu launch");
    if (h.mode == 1) {
        const size_t tot = (size_t)h.B * h.n * h.US;
        hipLaunchKernelGGL((k_scale_ucol<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                           static_cast<T*>(h.Uf.p), tot, h.US);
        hip_check(hipGetLastError(), "k_scale_ucol launch");
        if (h.parts > 1) {
            const int ywb = part_yw_bytes(h), bb = wave_buf(h, SPW_FEW);
            const size_t lds = (size_t)ywb + 16 + wave_nbuf(SPW_FEW) * (size_t)bb;
            hipLaunchKernelGGL((k_spikes_w<T, SPW_FEW>), dim3(h.B, h.parts, std::max(h.KLP, h.M - 1)), dim3(64), lds,
                               s, static_cast<const T*>(h.Uf.p), static_cast<const T*>(h.Lf.p),
                               static_cast<const int*>(h.piv.p), static_cast<T*>(h.SpL.p), static_cast<T*>(h.SpU.p),
                               h.n, h.parts, h.KLP, h.M, h.US, h.ls, h.lsu, ywb, bb);
            hip_check(hipGetLastError(), "k_spikes_w launch");
        }
    }
    h.tr_ready = false;   // the transposed Woodbury data follow the factors (built on demand)
    if (h.r == 0) return;
    // W = U^-1 (L^-1 P_B U) in place and S = I + V^T W, then the LU of S
    if (h.mode == 1) {
        const int sp = h.wave_sp ? h.wave_sp : wave_steps((size_t)h.B * h.r);
        const size_t lds = wave_lds_bytes(h, sp);
        if (h.dtype == CN_C128)
            hipLaunchKernelGGL((sp == SPW_MANY ? k_border_w<zc, SPW_MANY> : k_border_w<zc, SPW_FEW>), dim3(h.B, h.r),
                               dim3(64), lds, s, static_cast<const zc*>(h.Uf.p), static_cast<const zc*>(h.Vt.p),
                               static_cast<zc*>(h.W.p), static_cast<zc*>(h.S.p), h.n, h.r, h.M, h.US,
                               wave_buf(h, sp));
        else
            hipLaunchKernelGGL((sp == SPW_MANY ? k_border_w<double, SPW_MANY> : k_border_w<double, SPW_FEW>),
                               dim3(h.B, h.r), dim3(64), lds, s, static_cast<const double*>(h.Uf.p),
                               static_cast<const double*>(h.Vt.p), static_cast<double*>(h.W.p),
                               static_cast<double*>(h.S.p), h.n, h.r, h.M, h.US, wave_buf(h, sp));
        hip_check(hipGetLastError(), "k_border_w launch");
    }
    int spx = 0;
    const int bb = h.mode == 1 ? -1 : border_bufbytes(h, spx);
    BorderSFn fs = bb > 0 ? border_s_fn<T>(h.WU, nr_for(h.r), spx, h.r) : nullptr;
    if (h.mode == 1) {
        // done above
    } else if (fs) {
        const int nr = nr_for(h.r);
        fs(dim3((h.B + IL - 1) / IL, (h.r + nr - 1) / nr), dim3(SWEEP_THREADS), streamed_lds(bb, spx, (int)h.esz), s, h.Uf.p,
           h.Vt.p, h.W.p, h.S.p,
           h.B, h.n, h.r, h.US, bb);
        hip_check(hipGetLastError(), "k_border_s launch");
    } else {
        BorderFn fd = border_fn<T>(h.WU);
        if (!fd) fail(CN_ELOGIC, "no border kernel for wu " + std::to_string(h.WU));
        const int lanes = h.B * h.r;
        fd(dim3((lanes + 255) / 256), dim3(256), s, h.Uf.p, h.Vt.p, h.W.p, h.S.p, h.B, h.n, h.r, h.US);
        hip_check(hipGetLastError(), "k_border launch");
    }
    hipLaunchKernelGGL((k_cap_lu<T>), dim3((h.B + CAP_LANES - 1) / CAP_LANES), dim3(CAP_LANES), cap_lds_bytes<T>(), s,
                       static_cast<T*>(h.S.p),
                       static_cast<int*>(h.spiv.p), static_cast<double*>(h.cond.p), static_cast<int*>(h.info.p),
                       h.B, h.r);
    hip_check(hipGetLastError(), "k_cap_lu launch");
}

// Transposed Woodbury data: Wt = B^-T V (rows B^-T v_a), St = I + U^T Wt and its LU, checked.
template <class T>
void run_prepare_t(cn_bb& h, hipStream_t s) {
    if (h.r > 0) {
        SolveTFn ft = solve_t_fn<T>(h.KLP, h.WU);
        if (!ft) fail(CN_ELOGIC, "no transposed solve kernel");
        const int lanes = h.B * h.r;
        ft(dim3((lanes + 255) / 256), dim3(256), s, h.Uf.p, h.Lf.p, static_cast<const int*>(h.piv.p), h.Vt.p, h.Wt.p,
           h.St.p, static_cast<const int*>(h.stpiv.p), static_cast<const int*>(h.perm.p), nullptr, nullptr, 0, h.B, h.n,
           h.r, h.US, h.r, 1, h.il, h.mode);
        hip_check(hipGetLastError(), "k_solve_t (Wt) launch");
        const int l2 = h.B * h.r * h.r;
        hipLaunchKernelGGL((k_build_st<T>), dim3((l2 + 255) / 256), dim3(256), 0, s, static_cast<const T*>(h.Wt.p),
                           static_cast<T*>(h.St.p), static_cast<int*>(h.info_t.p), h.B, h.n, h.r, h.il);
        hip_check(hipGetLastError(), "k_build_st launch");
        hipLaunchKernelGGL((k_cap_lu<T>), dim3((h.B + CAP_LANES - 1) / CAP_LANES), dim3(CAP_LANES), cap_lds_bytes<T>(),
                           s, static_cast<T*>(h.St.p),
                           static_cast<int*>(h.stpiv.p), static_cast<double*>(h.cond_t.p),
                           static_cast<int*>(h.info_t.p), h.B, h.r);
        hip_check(hipGetLastError(), "k_cap_lu (transposed) launch");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        std::vector<int> info_t(h.B);
        hip_check(hipMemcpy(info_t.data(), h.info_t.p, sizeof(int) * h.B, hipMemcpyDeviceToHost), "hipMemcpy info_t");
        for (int m = 0; m < h.B; ++m)
            if (info_t[m] != 0)
                fail(CN_ESINGULAR, "BandedBatch: transposed capacitance matrix of mode " + std::to_string(m) +
                                       " is singular");
    }
    h.tr_ready = true;
}

void factor_and_check(cn_bb& h, hipStream_t s, bool check) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (check) {
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        hip_check(hipEventRecord(e0, s), "hipEventRecord");
    }
    if (h.dtype == CN_C128) run_factor<zc>(h, s); else run_factor<double>(h, s);
    if (!check) return;
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h.factor_ms = ms;
    std::vector<int> info(h.B);
    hip_check(hipMemcpy(info.data(), h.info.p, sizeof(int) * h.B, hipMemcpyDeviceToHost), "hipMemcpy info");
    h.hcond.assign(h.B, 1.0);
    if (h.r > 0)
        hip_check(hipMemcpy(h.hcond.data(), h.cond.p, sizeof(double) * h.B, hipMemcpyDeviceToHost), "hipMemcpy cond");
    h.cond_max = 1.0;
    for (int m = 0; m < h.B; ++m) {
        if (info[m] > 0)
            fail(CN_ESINGULAR, "BandedBatch: banded part of mode " + std::to_string(m) + " is singular at pivot " +
                                   std::to_string(info[m]) + (h.pivot ? "" : " (no pivoting was requested)") +
                                   "; the non-border rows must form an invertible banded operator");
        if (info[m] < 0)
            fail(CN_ESINGULAR, "BandedBatch: capacitance matrix of mode " + std::to_string(m) +
                                   " is singular: the boundary rows are not independent of the banded operator");
        if (std::isnan(h.hcond[m]))
            fail(CN_ESINGULAR, "BandedBatch: capacitance condition number of mode " + std::to_string(m) + " is NaN");
        h.cond_max = std::max(h.cond_max, h.hcond[m]);
    }
}

template <class T>
void create_impl(cn_bb& h, const int32_t* const* indptr, const int32_t* const* indices, const void* const* data,
                 const int64_t* nnz, const int32_t* border) {
    const int n = h.n, B = h.B, r = h.r;
    // CSR structure and finiteness (core/bindings/module.cpp:75-95, banded.hpp:173-180)
    for (int m = 0; m < B; ++m) {
        const int32_t* ip = indptr[m];
        const int32_t* ix = indices[m];
        const T* dv = static_cast<const T*>(data[m]);
        if (!ip || (nnz[m] > 0 && (!ix || !dv))) fail(CN_EINVAL, "CSR: null array for matrix " + std::to_string(m));
        if (ip[0] != 0 || (int64_t)ip[n] != nnz[m])
            fail(CN_EINVAL, "inconsistent CSR arrays (indptr[0] = 0, indptr[n] = nnz = len(indices) = len(data) "
                            "required) for matrix " + std::to_string(m));
        for (int i = 0; i < n; ++i)
            if (ip[i + 1] < ip[i]) fail(CN_EINVAL, "CSR: indptr decreases at row " + std::to_string(i));
        for (int64_t p = 0; p < nnz[m]; ++p) {
            if (ix[p] < 0 || ix[p] >= n)
                fail(CN_EINVAL, "CSR: column index " + std::to_string(ix[p]) + " outside [0, " + std::to_string(n) + ")");
            const std::complex<double> v(dv[p]);
            if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) {
                int row = 0;
                while (ip[row + 1] <= p) ++row;
                fail(CN_EINVAL, "BorderedBanded: non-finite matrix entry at (" + std::to_string(row) + ", " +
                                    std::to_string(ix[p]) + ") of matrix " + std::to_string(m) +
                                    ": the operator must be finite");
            }
        }
    }
    std::vector<char> isb(n, 0);
    for (int a = 0; a < r; ++a) {
        const int i = border[a];
        if (i < 0 || i >= n) fail(CN_EINVAL, "border row out of range");
        if (isb[i]) fail(CN_EINVAL, "duplicate border row");
        isb[i] = 1;
    }
    std::vector<int> perm(border, border + r), pos(n);
    for (int i = 0; i < n; ++i) if (!isb[i]) perm.push_back(i);
    for (int i = 0; i < n; ++i) pos[perm[i]] = i;
    // common band of the bulk rows (banded.hpp:144-168), max over the batch
    int kl = 0, ku = 0;
    for (int m = 0; m < B; ++m) {
        const int32_t* ip = indptr[m];
        const int32_t* ix = indices[m];
        const T* dv = static_cast<const T*>(data[m]);
        for (int i = r; i < n; ++i) {
            const int o = perm[i];
            for (int p = ip[o]; p < ip[o + 1]; ++p) {
                if (dv[p] == T(0)) continue;
                kl = std::max(kl, i - ix[p]);
                ku = std::max(ku, ix[p] - i);
            }
        }
    }
    h.kl = kl;
    h.ku = ku;
    h.KLP = pick_kl(std::max(kl, 1));
    if (h.KLP < 0) fail(CN_EUNSUPPORTED, "BandedBatch: kl = " + std::to_string(kl) + " exceeds the compiled kernels (<= 16)");
    h.M = h.KLP + ku + 1;
    int G = 16;
    while (G < h.M + r) G *= 2;
    if (G > 64)
        fail(CN_EUNSUPPORTED, "BandedBatch: band kl+ku+1 = " + std::to_string(h.M) + " plus border rank " +
                                  std::to_string(r) + " exceeds one wavefront (64 lanes)");
    h.G = G;
    h.WU = pick_wu(h.KLP + ku);
    if (h.WU < 0) fail(CN_EUNSUPPORTED, "BandedBatch: kl+ku = " + std::to_string(h.KLP + ku) + " exceeds 32");
    h.US = h.WU + 1;
    h.Bp = (B + 63) / 64 * 64;   // whole 64-system workgroups of the streamed solve stay in bounds
    // kernel mode: lane mode (one lane per system and right-hand side, streamed sweeps) for many
    // narrow-band systems; wave mode (one wavefront per system and right-hand side) for few or
    // wide-band systems.  CNHIP_MODE=lane|wave forces one where it applies.
    {
        const int e = (int)sizeof(T);
        const int rb = rb_for(r);
        const int lane_buf = std::max({fwd_bytes(8, h.KLP, e), bwd_bytes(8, h.US, rb, e), cor_bytes(8, rb, e)});
        const bool lane_fast = h.KLP + 1 <= 8 && h.WU <= 16 && 2 * lane_buf + 2 * stage_bytes(8, e) <= LDS_CAP;
        const bool wave_ok =
            (size_t)align16(n * e) + WAVE_SMALL + 3 * (size_t)wave_buf_bytes(h.KLP, h.US, e, SPW_FEW) <=
                (size_t)LDS_CAP;
        const char* em = std::getenv("CNHIP_MODE");
        std::string want = g_mode_request == 1   ? "lane"
                           : g_mode_request == 2 ? "wave"
                           : g_mode_request == 3 ? "part"
                           : g_mode_request == 4 ? "wave8"
                           : g_mode_request == 5 ? "wave32"
                           : em                  ? em
                                                 : "";
        // wave8 / wave32: wave mode with that staged block length (results do not depend on it)
        if (want == "wave8" || want == "wave32") {
            h.wave_sp = want == "wave8" ? SPW_MANY : SPW_FEW;
            want = "wave";
        }
        if (want == "wave" || want == "part") h.mode = wave_ok ? 1 : 0;
        else if (want == "lane") h.mode = 0;
        else h.mode = (lane_fast && B >= 1024) || !wave_ok ? 0 : 1;
        h.il = h.mode == 1 ? 1 : IL;
        // partitioned sweeps when asked for, and by default while the batch has fewer systems than
        // a quarter of the chip's SIMDs (one wavefront per system would leave most of them idle)
        if (h.mode == 1 && (want == "part" || (want.empty() && B <= PART_MAX_B))) h.parts = pick_parts(h);
    }
    const size_t es = sizeof(T);
    const size_t nAb = checked_mul({(size_t)h.Bp, (size_t)n, (size_t)h.M});
    const size_t nU = checked_mul({(size_t)h.Bp, (size_t)n, (size_t)h.US});
    const size_t nL = checked_mul({(size_t)h.Bp, (size_t)n, (size_t)h.KLP});
    const size_t nV = checked_mul({(size_t)h.Bp, (size_t)r, (size_t)n});
    std::vector<T> Ab(nAb, T(0)), Vt(nV, T(0));
    for (int m = 0; m < B; ++m)
        layout_one<T>(h, m, indptr[m], indices[m], static_cast<const T*>(data[m]), perm, pos, Ab, Vt);
    hip_check(hipGetDevice(&h.device), "hipGetDevice");
    h.Ab.alloc(nAb * es);
    h.Uf.alloc(nU * es);
    h.Lf.alloc(nL * es);
    h.piv.alloc((size_t)h.Bp * n * sizeof(int));
    h.Vt.alloc(std::max<size_t>(nV, 1) * es);
    h.W.alloc(std::max<size_t>(nV, 1) * es);
    h.S.alloc(std::max<size_t>((size_t)h.Bp * r * r, 1) * es);
    h.spiv.alloc(std::max<size_t>((size_t)h.Bp * r, 1) * sizeof(int));
    h.cond.alloc((size_t)h.Bp * sizeof(double));
    h.info.alloc((size_t)h.Bp * sizeof(int));
    h.perm.alloc((size_t)n * sizeof(int));
    if (h.parts > 1) {
        h.ls = h.plen + h.KLP;
        h.lsu = h.plen + h.M - 1;
        h.SpL.alloc(checked_mul({(size_t)B, (size_t)h.parts, (size_t)h.KLP, (size_t)h.ls, es}));
        h.SpU.alloc(checked_mul({(size_t)B, (size_t)h.parts, (size_t)(h.M - 1), (size_t)h.lsu, es}));
    }
    h.Wt.alloc(std::max<size_t>(nV, 1) * es);
    h.St.alloc(std::max<size_t>((size_t)h.Bp * r * r, 1) * es);
    h.stpiv.alloc(std::max<size_t>((size_t)h.Bp * r, 1) * sizeof(int));
    h.cond_t.alloc((size_t)h.Bp * sizeof(double));
    h.info_t.alloc((size_t)h.Bp * sizeof(int));
    hip_check(hipMemset(h.info_t.p, 0, (size_t)h.Bp * sizeof(int)), "hipMemset info_t");
    hip_check(hipMemset(h.W.p, 0, std::max<size_t>(nV, 1) * es), "hipMemset W");
    hip_check(hipMemset(h.Wt.p, 0, std::max<size_t>(nV, 1) * es), "hipMemset Wt");
    hip_check(hipMemcpy(h.Ab.p, Ab.data(), nAb * es, hipMemcpyHostToDevice), "hipMemcpy Ab");
    if (nV) hip_check(hipMemcpy(h.Vt.p, Vt.data(), nV * es, hipMemcpyHostToDevice), "hipMemcpy Vt");
    hip_check(hipMemcpy(h.perm.p, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy perm");
    h.hperm.assign(perm.begin(), perm.begin() + n);
    hip_check(hipMemset(h.Uf.p, 0, nU * es), "hipMemset U");
    hip_check(hipMemset(h.Lf.p, 0, nL * es), "hipMemset L");
    hip_check(hipMemset(h.info.p, 0, (size_t)h.Bp * sizeof(int)), "hipMemset info");
    hip_check(hipMemset(h.piv.p, 0, (size_t)h.Bp * n * sizeof(int)), "hipMemset piv");
    hip_check(hipMemset(h.S.p, 0, std::max<size_t>((size_t)h.Bp * r * r, 1) * es), "hipMemset S");
    hip_check(hipMemset(h.cond.p, 0, (size_t)h.Bp * sizeof(double)), "hipMemset cond");
}

// bytes of the lane work space of a solve of nrhs right-hand sides (one entry per system, rhs, row)
size_t solve_ws_bytes(const cn_bb& h, int nrhs) {
    return checked_mul({(size_t)h.B, (size_t)nrhs, (size_t)h.n, h.esz});
}

void reserve_impl(cn_bb& h, int nrhs) {
    if (nrhs <= h.nrhs_cap) return;
    h.ws.alloc(solve_ws_bytes(h, nrhs));
    h.nrhs_cap = nrhs;
}

// ---- the allocation-free solve: its lane work space is a fixed device array of this library ----
// Every device the library runs on holds one copy of `scratch` for the life of the process; its
// address is looked up once per device (hipGetSymbolAddress, with that device current) and kept.
// The kernels write every work-space entry before they read it, so nothing needs clearing.
// Calls that overlap in time on one device share the array and race.
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_DEVICES] = {};

void* scratch_for_device(int dev) {
    if (dev < 0 || dev >= SCRATCH_DEVICES) fail(CN_EUNSUPPORTED, "cn_bb_solve_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) fail(CN_EHIP, "cn_bb_solve_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

void prepare_t(cn_bb& h, hipStream_t s) {
    if (h.tr_ready) return;
    if (h.dtype == CN_C128) run_prepare_t<zc>(h, s); else run_prepare_t<double>(h, s);
}

void solve_t_impl(cn_bb& h, void* x, int nrhs, int64_t rs, hipStream_t s) {
    if (!h.tr_ready || nrhs > h.nrhs_cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
        if (cs != hipStreamCaptureStatusNone)
            fail(CN_EINVAL, "BandedBatch transposed solve: prepare_transpose / reserve must run before graph capture");
        prepare_t(h, s);
        reserve_impl(h, nrhs);
    }
    SolveTFn f = (h.dtype == CN_C128) ? solve_t_fn<zc>(h.KLP, h.WU) : solve_t_fn<double>(h.KLP, h.WU);
    if (!f) fail(CN_ELOGIC, "no transposed solve kernel");
    const int lanes = h.B * nrhs;
    f(dim3((lanes + 255) / 256), dim3(256), s, h.Uf.p, h.Lf.p, static_cast<const int*>(h.piv.p), h.Vt.p, h.Wt.p,
      h.St.p, static_cast<const int*>(h.stpiv.p), static_cast<const int*>(h.perm.p), h.ws.p, x, rs, h.B, h.n, h.r,
      h.US, nrhs, 0, h.il, h.mode);
    hip_check(hipGetLastError(), "k_solve_t launch");
}

// LDS buffer region of the streamed solve (k_solve_s), or -1 if it does not apply
int streamed_bufbytes(const cn_bb& h, const SolveIO& io, int& spx) {
    spx = 0;
    if (!h.streamed) return -1;
    if (!(io.mode == 0 || (io.mode == 1 && !io.rows && !io.cols))) return -1;
    if (h.KLP + 1 > 8 || h.WU > 16) return -1;
    const int e = (int)h.esz;
    const int rb = rb_for(h.r);
    auto buf = [&](int x) { return std::max({fwd_bytes(x, h.KLP, e), bwd_bytes(x, h.US, rb, e), cor_bytes(x, rb, e)}); };
    spx = pick_spx(buf(16), buf(8), e);
    return spx ? buffer_region({fwd_bytes(spx, h.KLP, e), bwd_bytes(spx, h.US, rb, e), cor_bytes(spx, rb, e)}, spx, e)
               : -1;
}

// ws: the lane work space of the solve (nullptr: the handle's own, grown to nrhs right-hand
// sides here when needed; else caller-provided memory of at least solve_ws_bytes(h, nrhs) bytes)
void solve_impl(cn_bb& h, SolveIO io, int nrhs, hipStream_t s, void* ws = nullptr) {
    if (ws == nullptr && nrhs > h.nrhs_cap) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hip_check(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
        if (cs != hipStreamCaptureStatusNone)
            fail(CN_EINVAL, "BandedBatch solve: work space for " + std::to_string(nrhs) +
                                " right-hand sides not reserved before graph capture (call cn_bb_reserve)");
        reserve_impl(h, nrhs);
    }
    if (ws == nullptr) ws = h.ws.p;
    if (h.mode == 1 && h.parts > 1) {
        const int P = h.parts, sp = part_steps(h, P);
        const size_t lds = part_lds_bytes(h, P, sp);
        const int ywb = part_yw_bytes(h), bb = wave_buf(h, sp);
        if (h.dtype == CN_C128)
            hipLaunchKernelGGL((sp == 16 ? k_solve_p<zc, 16> : k_solve_p<zc, SPW_MANY>), dim3(h.B, nrhs), dim3(64 * P),
                               lds, s, static_cast<const zc*>(h.Uf.p), static_cast<const zc*>(h.Lf.p),
                               static_cast<const int*>(h.piv.p), static_cast<const zc*>(h.Vt.p),
                               static_cast<const zc*>(h.W.p), static_cast<const zc*>(h.S.p),
                               static_cast<const int*>(h.spiv.p), static_cast<const int*>(h.perm.p),
                               static_cast<const zc*>(h.SpL.p), static_cast<const zc*>(h.SpU.p), io, h.n, h.r, h.KLP,
                               h.M, h.US, P, h.ls, h.lsu, ywb, bb);
        else
            hipLaunchKernelGGL((sp == 16 ? k_solve_p<double, 16> : k_solve_p<double, SPW_MANY>), dim3(h.B, nrhs),
                               dim3(64 * P), lds, s, static_cast<const double*>(h.Uf.p),
                               static_cast<const double*>(h.Lf.p), static_cast<const int*>(h.piv.p),
                               static_cast<const double*>(h.Vt.p), static_cast<const double*>(h.W.p),
                               static_cast<const double*>(h.S.p), static_cast<const int*>(h.spiv.p),
                               static_cast<const int*>(h.perm.p), static_cast<const double*>(h.SpL.p),
                               static_cast<const double*>(h.SpU.p), io, h.n, h.r, h.KLP, h.M, h.US, P, h.ls, h.lsu,
                               ywb, bb);
        hip_check(hipGetLastError(), "k_solve_p launch");
        return;
    }
    if (h.mode == 1) {
        const int sp = h.wave_sp ? h.wave_sp : wave_steps((size_t)h.B * nrhs);
        const size_t lds = wave_lds_bytes(h, sp);
        if (h.dtype == CN_C128)
            hipLaunchKernelGGL((sp == SPW_MANY ? k_solve_w<zc, SPW_MANY> : k_solve_w<zc, SPW_FEW>), dim3(h.B, nrhs),
                               dim3(64), lds, s, static_cast<const zc*>(h.Uf.p), static_cast<const zc*>(h.Lf.p),
                               static_cast<const int*>(h.piv.p), static_cast<const zc*>(h.Vt.p),
                               static_cast<const zc*>(h.W.p), static_cast<const zc*>(h.S.p),
                               static_cast<const int*>(h.spiv.p), static_cast<const int*>(h.perm.p), io, h.n, h.r,
                               h.KLP, h.M, h.US, wave_buf(h, sp));
        else
            hipLaunchKernelGGL((sp == SPW_MANY ? k_solve_w<double, SPW_MANY> : k_solve_w<double, SPW_FEW>),
                               dim3(h.B, nrhs), dim3(64), lds, s, static_cast<const double*>(h.Uf.p),
                               static_cast<const double*>(h.Lf.p), static_cast<const int*>(h.piv.p),
                               static_cast<const double*>(h.Vt.p), static_cast<const double*>(h.W.p),
                               static_cast<const double*>(h.S.p), static_cast<const int*>(h.spiv.p),
                               static_cast<const int*>(h.perm.p), io, h.n, h.r, h.KLP, h.M, h.US, wave_buf(h, sp));
        hip_check(hipGetLastError(), "k_solve_w launch");
        return;
    }
    int spx = 0;
    const int bufbytes = streamed_bufbytes(h, io, spx);
    if (bufbytes > 0 && solve_s_fn<double>(h.KLP, h.WU, 1, spx, h.r)) {
        // groups of up to 4 right-hand sides, one launch each (in place: rhs k at k * rhs_stride)
        char* xb = static_cast<char*>(io.x);
        const char* bb = static_cast<const char*>(io.mode == 0 ? io.x : io.rhs);
        const int64_t rs = io.mode == 0 ? io.rhs_stride : 0;
        for (int k0 = 0; k0 < nrhs;) {
            const int nr = nr_for(nrhs - k0);
            SolveSFn fs = (h.dtype == CN_C128) ? solve_s_fn<zc>(h.KLP, h.WU, nr, spx, h.r)
                                               : solve_s_fn<double>(h.KLP, h.WU, nr, spx, h.r);
            const size_t off = (size_t)k0 * (size_t)rs * h.esz;
            fs(dim3((h.B + IL - 1) / IL), dim3(SWEEP_THREADS), streamed_lds(bufbytes, spx, (int)h.esz), s, h.Uf.p, h.Lf.p,
               static_cast<const int*>(h.piv.p), h.Vt.p, h.W.p, h.S.p, static_cast<const int*>(h.spiv.p),
               static_cast<const int*>(h.perm.p), ws, bb + off, xb + off, rs, h.B, h.n, h.r, h.US, bufbytes);
            hip_check(hipGetLastError(), "k_solve_s launch");
            k0 += nr;
        }
        return;
    }
    SolveFn f = (h.dtype == CN_C128) ? solve_fn<zc>(h.KLP, h.WU) : solve_fn<double>(h.KLP, h.WU);
    if (!f) fail(CN_ELOGIC, "no solve kernel for (kl, wu) = (" + std::to_string(h.KLP) + ", " + std::to_string(h.WU) + ")");
    const int lanes = h.B * nrhs;
    f(dim3((lanes + 255) / 256), dim3(256), s, h.Uf.p, h.Lf.p, static_cast<const int*>(h.piv.p), h.Vt.p, h.W.p,
      h.S.p, static_cast<const int*>(h.spiv.p), static_cast<const int*>(h.perm.p), ws, io, h.B, h.n, h.r, h.US,
      nrhs);
    hip_check(hipGetLastError(), "k_solve launch");
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        g_err.clear();
        return CN_OK;
    } catch (const CnError& e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return CN_EOVERFLOW;
    } catch (const std::exception& e) {
        g_err = e.what();
        return CN_ELOGIC;
    }
}

}  // namespace

extern "C" {

const char* cn_last_error(void) { return g_err.c_str(); }

const char* cn_version(void) { return "cnhip 0.1.0 gfx950"; }

int cn_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int cn_bb_create(cn_bb_t* out, int dtype, int batch, int n, const int32_t* const* indptr,
                 const int32_t* const* indices, const void* const* data, const int64_t* nnz,
                 const int32_t* border, int r, int pivot, int col_border) {
    return guarded([&] {
        if (!out) fail(CN_EINVAL, "cn_bb_create: null output handle");
        *out = nullptr;
        if (dtype != CN_F64 && dtype != CN_C128) fail(CN_EINVAL, "cn_bb_create: dtype must be CN_F64 or CN_C128");
        if (batch < 1) fail(CN_EINVAL, "BandedBatch: empty batch");
        if (n < 1) fail(CN_EINVAL, "BandedBatch: n must be >= 1, got " + std::to_string(n));
        if (r < 0 || r > n) fail(CN_EINVAL, "BandedBatch: border rank must be in [0, n]");
        if (r > MAXR)
            fail(CN_EUNSUPPORTED, "BandedBatch: border rank " + std::to_string(r) + " > " + std::to_string(MAXR) +
                                      " not supported on the device");
        if (col_border) fail(CN_EUNSUPPORTED, "BandedBatch: the column border is not supported by this library");
        if (!indptr || !indices || !data || !nnz || (r > 0 && !border)) fail(CN_EINVAL, "cn_bb_create: null array");
        auto h = std::make_unique<cn_bb>();
        h->dtype = dtype;
        h->B = batch;
        h->n = n;
        h->r = r;
        h->pivot = pivot ? 1 : 0;
        h->esz = dtype == CN_C128 ? 16 : 8;
        if (const char* e = std::getenv("CNHIP_SOLVE")) h->streamed = std::string(e) != "direct";
        if (dtype == CN_C128) create_impl<std::complex<double>>(*h, indptr, indices, data, nnz, border);
        else create_impl<double>(*h, indptr, indices, data, nnz, border);
        factor_and_check(*h, nullptr, true);
        reserve_impl(*h, 1);
        *out = h.release();
    });
}

int cn_bb_factor(cn_bb_t h, void* stream, int check) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_factor: null handle");
        factor_and_check(*h, static_cast<hipStream_t>(stream), check != 0);
    });
}

int cn_bb_reserve(cn_bb_t h, int nrhs) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_reserve: null handle");
        if (nrhs < 1) fail(CN_EINVAL, "cn_bb_reserve: nrhs must be >= 1");
        reserve_impl(*h, nrhs);
    });
}

int cn_bb_solve(cn_bb_t h, void* x, int nrhs, int64_t rhs_stride, int transpose, void* stream) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_solve: null handle");
        if (!x) fail(CN_EINVAL, "BandedBatch.solve_ptr: null device pointer");
        if (nrhs < 1) fail(CN_EINVAL, "cn_bb_solve: nrhs must be >= 1");
        if (nrhs > 1 && rhs_stride < (int64_t)h->B * h->n)
            fail(CN_EINVAL, "cn_bb_solve: rhs_stride must be >= batch * n");
        if (transpose) {
            solve_t_impl(*h, x, nrhs, rhs_stride, static_cast<hipStream_t>(stream));
            return;
        }
        SolveIO io{};
        io.mode = 0;
        io.x = x;
        io.rhs_stride = rhs_stride;
        solve_impl(*h, io, nrhs, static_cast<hipStream_t>(stream));
    });
}

int cn_bb_solve_noalloc(cn_bb_t h, void* x, int nrhs, int64_t rhs_stride, void* stream) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_solve_noalloc: null handle");
        if (!x) fail(CN_EINVAL, "cn_bb_solve_noalloc: null device pointer");
        if (nrhs < 1) fail(CN_EINVAL, "cn_bb_solve_noalloc: nrhs must be >= 1");
        if (nrhs > 1 && rhs_stride < (int64_t)h->B * h->n)
            fail(CN_EINVAL, "cn_bb_solve_noalloc: rhs_stride must be >= batch * n");
        const size_t need = solve_ws_bytes(*h, nrhs);
        if (need > SCRATCH_BYTES)
            fail(CN_EINVAL, "cn_bb_solve_noalloc: the work space of " + std::to_string(nrhs) + " right-hand side(s) of " +
                                std::to_string(h->B) + " systems of n = " + std::to_string(h->n) + " needs " +
                                std::to_string(need) + " bytes, more than the library's fixed " +
                                std::to_string(SCRATCH_BYTES) + " (use cn_bb_solve, or fewer right-hand sides per call)");
        int dev = -1;
        hip_check(hipGetDevice(&dev), "hipGetDevice");
        if (dev != h->device)
            fail(CN_EINVAL, "cn_bb_solve_noalloc: called with device " + std::to_string(dev) +
                                " current, the handle lives on device " + std::to_string(h->device));
        SolveIO io{};
        io.mode = 0;
        io.x = x;
        io.rhs_stride = rhs_stride;
        solve_impl(*h, io, nrhs, static_cast<hipStream_t>(stream), scratch_for_device(dev));
    });
}

int cn_bb_solve_mapped(cn_bb_t h, const void* rhs, const int64_t* rows, void* x, const int64_t* cols, void* stream) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_solve_mapped: null handle");
        if (!rhs || !x) fail(CN_EINVAL, "BandedBatch.solve_mapped_ptr: null device pointer");
        SolveIO io{};
        io.mode = 1;
        io.x = x;
        io.rhs = rhs;
        io.rows = rows;
        io.cols = cols;
        solve_impl(*h, io, 1, static_cast<hipStream_t>(stream));
    });
}

int cn_bb_solve_split(cn_bb_t h, const double* re, const int64_t* rows_c, const int64_t* rows_s, double* x,
                      const int64_t* cols_c, const int64_t* cols_s, void* stream) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_solve_split: null handle");
        if (h->dtype != CN_C128) fail(CN_EINVAL, "BandedBatch.solve_split_ptr: complex128 factors required");
        if (!re || !x || !rows_c || !rows_s || !cols_c || !cols_s)
            fail(CN_EINVAL, "BandedBatch.solve_split_ptr: null device pointer");
        SolveIO io{};
        io.mode = 2;
        io.x = x;
        io.rhs = re;
        io.rows = rows_c;
        io.rows2 = rows_s;
        io.cols = cols_c;
        io.cols2 = cols_s;
        solve_impl(*h, io, 1, static_cast<hipStream_t>(stream));
    });
}

int cn_bb_prepare_transpose(cn_bb_t h) {
    return guarded([&] {
        if (!h) fail(CN_EINVAL, "cn_bb_prepare_transpose: null handle");
        prepare_t(*h, nullptr);
    });
}

int cn_bb_props_get(cn_bb_t h, cn_bb_props* p) {
    return guarded([&] {
        if (!h || !p) fail(CN_EINVAL, "cn_bb_props_get: null argument");
        p->dtype = h->dtype;
        p->batch = h->B;
        p->n = h->n;
        p->kl = h->kl;
        p->ku = h->ku;
        p->rank = h->r;
        p->pivot = h->pivot;
        p->kl_kernel = h->KLP;
        p->wu_kernel = h->WU;
        p->lanes = h->G;
        p->nrhs_reserved = h->nrhs_cap;
        p->transpose_ready = h->tr_ready ? 1 : 0;
        p->cond_max = h->cond_max;
        p->factor_ms = h->factor_ms;
        p->solve_mode = h->mode == 1 && h->parts > 1 ? 2 : h->mode;
        p->factor_bytes = (int64_t)(h->Uf.bytes + h->Lf.bytes + h->piv.bytes + h->Vt.bytes + h->W.bytes + h->S.bytes +
                                    h->spiv.bytes + h->SpL.bytes + h->SpU.bytes);
    });
}

int cn_bb_cond(cn_bb_t h, double* out) {
    return guarded([&] {
        if (!h || !out) fail(CN_EINVAL, "cn_bb_cond: null argument");
        for (int m = 0; m < h->B; ++m) out[m] = m < (int)h->hcond.size() ? h->hcond[m] : 1.0;
    });
}

void cn_bb_destroy(cn_bb_t h) { delete h; }

int cn_bb_copy_systems(cn_bb_t dst, int dst_first, cn_bb_t src, int src_first, int count, void* stream) {
    return guarded([&] {
        if (!dst || !src) fail(CN_EINVAL, "cn_bb_copy_systems: null handle");
        const cn_bb& S_ = *src;
        cn_bb& D = *dst;
        if (S_.dtype != D.dtype || S_.n != D.n || S_.r != D.r || S_.KLP != D.KLP || S_.M != D.M || S_.US != D.US ||
            S_.WU != D.WU || S_.il != D.il || S_.mode != D.mode || S_.parts != D.parts || S_.pivot != D.pivot ||
            S_.device != D.device)
            fail(CN_EINVAL, "cn_bb_copy_systems: the handles differ in type, size, band, border rank, kernel mode or device");
        if (count < 0 || src_first < 0 || dst_first < 0 || src_first + count > S_.B || dst_first + count > D.B)
            fail(CN_EINVAL, "cn_bb_copy_systems: system range out of bounds");
        if (count == 0) return;
        // the row permutation (border rows first) must be the same: compared on the host copies, so
        // that the call stays asynchronous
        if (S_.hperm != D.hperm) fail(CN_EINVAL, "cn_bb_copy_systems: the handles have different border rows");
        hipStream_t st = static_cast<hipStream_t>(stream);
        const size_t il = D.il, es = D.esz;
        // interleaved arrays: entry e of system m at ((m / il) * len + e) * il + m % il
        auto inter = [&](const DevBuf& from, DevBuf& to, size_t len, size_t esz) {
            for (int q = 0; q < count;) {
                const int ms = src_first + q, md = dst_first + q;
                if (il == 1) {
                    hip_check(hipMemcpyAsync(static_cast<char*>(to.p) + (size_t)md * len * esz,
                                             static_cast<const char*>(from.p) + (size_t)ms * len * esz,
                                             (size_t)count * len * esz, hipMemcpyDeviceToDevice, st),
                              "cn_bb_copy_systems");
                    return;
                }
                const char* sp = static_cast<const char*>(from.p) + (((size_t)ms / il) * len * il + ms % il) * esz;
                char* dp = static_cast<char*>(to.p) + (((size_t)md / il) * len * il + md % il) * esz;
                hip_check(hipMemcpy2DAsync(dp, il * esz, sp, il * esz, esz, len, hipMemcpyDeviceToDevice, st),
                          "cn_bb_copy_systems");
                ++q;
            }
        };
        // system-major arrays: len entries per system, contiguous
        auto major = [&](const DevBuf& from, DevBuf& to, size_t len, size_t esz) {
            if (len == 0) return;
            hip_check(hipMemcpyAsync(static_cast<char*>(to.p) + (size_t)dst_first * len * esz,
                                     static_cast<const char*>(from.p) + (size_t)src_first * len * esz,
                                     (size_t)count * len * esz, hipMemcpyDeviceToDevice, st),
                      "cn_bb_copy_systems");
        };
        const size_t n = D.n, r = D.r;
        inter(S_.Ab, D.Ab, n * D.M, es);
        inter(S_.Uf, D.Uf, n * D.US, es);
        inter(S_.Lf, D.Lf, n * D.KLP, es);
        inter(S_.piv, D.piv, n, sizeof(int));
        if (r) {
            inter(S_.Vt, D.Vt, r * n, es);
            inter(S_.W, D.W, r * n, es);
        }
        major(S_.S, D.S, r * r, es);
        major(S_.spiv, D.spiv, r, sizeof(int));
        if (D.parts > 1) {
            major(S_.SpL, D.SpL, (size_t)D.parts * D.KLP * D.ls, es);
            major(S_.SpU, D.SpU, (size_t)D.parts * (D.M - 1) * D.lsu, es);
        }
        major(S_.cond, D.cond, 1, sizeof(double));
        major(S_.info, D.info, 1, sizeof(int));
        if ((int)D.hcond.size() < D.B) D.hcond.assign(D.B, 1.0);
        for (int q = 0; q < count; ++q)
            D.hcond[dst_first + q] = src_first + q < (int)S_.hcond.size() ? S_.hcond[src_first + q] : 1.0;
        D.cond_max = 1.0;
        for (double c : D.hcond) D.cond_max = std::max(D.cond_max, c);
        D.tr_ready = false;   // the transposed Woodbury data follow the factors (built on demand)
    });
}

int cn_bb_request_mode(int mode) {
    if (mode < 0 || mode > 5) {
        g_err = "cn_bb_request_mode: mode must be 0 (automatic), 1 (lane), 2 (wave), 3 (partitioned wave), "
                "4 (wave, blocks of 8 steps) or 5 (wave, blocks of 32 steps)";
        return CN_EINVAL;
    }
    g_mode_request = mode;
    return CN_OK;
}

}  // extern "C"
