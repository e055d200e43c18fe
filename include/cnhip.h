/* cnhip.h — C ABI of libcnhip.so, the MI355X (gfx950) bordered-banded batch solver.
 *
 * Plain pointers, sizes and a hipStream_t passed as void*; no torch or C++ types cross this
 * boundary.  Every entry point returns a status (CN_OK or a CN_E* code); the message of the
 * last failure on the calling thread is cn_last_error().  Status codes map to the exception
 * types of the reference binding (core/bindings/module.cpp:40-95, banded.hpp:45-56):
 *   CN_EINVAL    -> ValueError     (shapes, CSR structure, non-finite entries, border rows)
 *   CN_ESINGULAR -> RuntimeError   (message contains "singular", the mode index and the pivot)
 *   CN_EOVERFLOW -> OverflowError  (size products)
 *   CN_ELOGIC    -> internal error (a layout bug; never expected)
 *   CN_EHIP      -> RuntimeError   (a HIP runtime failure)
 *   CN_EUNSUPPORTED -> NotImplementedError (a shape outside the compiled kernel set)
 *
 * Device memory: the handle owns its factor storage (allocated at create).  Right-hand sides
 * and solutions live in caller-owned device memory (e.g. torch tensors) and are solved in place
 * on the caller's stream; solve entry points neither allocate nor synchronise the host once
 * cn_bb_reserve() has sized the work space, so they can be captured in a HIP graph.
 */
#ifndef CNHIP_H
#define CNHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CN_OK 0
#define CN_EINVAL 1
#define CN_ESINGULAR 2
#define CN_EOVERFLOW 3
#define CN_ELOGIC 4
#define CN_EHIP 5
#define CN_EUNSUPPORTED 6

#define CN_F64 0  /* float64 factors and data */
#define CN_C128 1 /* complex128 (interleaved re, im) */

typedef struct cn_bb* cn_bb_t;

typedef struct cn_bb_props {
    int32_t dtype;      /* CN_F64 / CN_C128 */
    int32_t batch;      /* number of systems */
    int32_t n;          /* size of every system */
    int32_t kl, ku;     /* common band of the bulk rows (max over the batch) */
    int32_t rank;       /* border rows r */
    int32_t pivot;      /* 1: partial pivoting inside the band */
    int32_t kl_kernel;  /* band widths of the compiled kernel chosen (kl_kernel >= kl, ...) */
    int32_t wu_kernel;  /* ... wu_kernel >= kl_kernel + ku) */
    int32_t lanes;      /* lanes per system in the factor kernel */
    int32_t nrhs_reserved;
    int32_t transpose_ready;
    double cond_max;    /* max over the batch of cond_1 of the capacitance matrix */
    double factor_ms;   /* device time of the last factorization */
    int64_t factor_bytes; /* device bytes held by the factors */
    int32_t solve_mode;   /* 0: one lane per (system, rhs), streamed sweeps; 1: one wavefront per (system, rhs);
                             2: one workgroup per (system, rhs), its sweeps partitioned over 2-8 wavefronts */
} cn_bb_props;

/* Message of the last failed call on this thread ("" if none). */
const char* cn_last_error(void);

/* Number of visible HIP devices (0 without a GPU); never fails. */
int cn_device_count(void);

/* Library version string. */
const char* cn_version(void);

/* Create and factor a batch of bordered-banded systems.
 * Replaces _core.BorderedBandedBatch{D,Z}.__init__ (core/bindings/module.cpp:164-169) and
 * cuda.BandedBatch{D,Z}.__init__ (core/bindings/module.cpp:326-332, core/src/cuda/banded.cu:1138-1237).
 * mats: `batch` square CSR matrices of size n (int32 indptr[n+1], int32 indices[nnz], data
 *   float64 or complex128 [nnz]); border: r distinct rows moved to the top (boundary rows);
 * pivot: 1 = partial pivoting in the band (default of the reference), 0 = none (declared
 *   diagonally dominant).  col_border must be 0 (column border not supported by this library).
 * The layout (border rows first, common band, s-scaled border, V^T) follows
 * core/include/conatus/banded.hpp:170-251; the factorization runs on the current device and the
 * call synchronises once to check singular pivots (CN_ESINGULAR names the mode and pivot). */
int cn_bb_create(cn_bb_t* out, int dtype, int batch, int n,
                 const int32_t* const* indptr, const int32_t* const* indices,
                 const void* const* data, const int64_t* nnz,
                 const int32_t* border, int r, int pivot, int col_border);

/* Re-run the device factorization from the stored unfactored layout (the factor kernels alone,
 * for timing them; the stored matrices do not change).  Asynchronous on `stream` when check = 0;
 * check = 1 synchronises and reports singular systems like cn_bb_create. */
int cn_bb_factor(cn_bb_t h, void* stream, int check);

/* Reserve the device work space for solves of up to nrhs right-hand sides per system (the
 * solve entry points then allocate nothing: graph-capturable). */
int cn_bb_reserve(cn_bb_t h, int nrhs);

/* In-place batched solve A_m x = b_m on caller device memory.
 * x: nrhs blocks of (batch, n) C-contiguous values; block k starts at element k * rhs_stride.
 * transpose = 1 solves A_m^T x = b_m (plain transpose).
 * Replaces BandedBatch.solve_ptr / solve_transpose_ptr (core/bindings/module.cpp:379-401) and,
 * with host staging in the Python layer, _core.BorderedBandedBatch{D,Z}.solve
 * (core/bindings/module.cpp:181-206). */
int cn_bb_solve(cn_bb_t h, void* x, int nrhs, int64_t rhs_stride, int transpose, void* stream);

/* In-place batched solve that allocates nothing: cn_bb_solve (transpose = 0) with the lane work
 * space taken from a fixed 64 MiB device array of the library instead of handle memory, so that
 * no hipMalloc can happen in the call even when the handle never reserved nrhs right-hand sides.
 * No counterpart in the reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  Limit: batch * nrhs * n * element size
 * <= 64 MiB (CN_EINVAL otherwise); the current device must be the handle's.  Every call on one
 * device shares that array, so calls that overlap in time on one device race. */
int cn_bb_solve_noalloc(cn_bb_t h, void* x, int nrhs, int64_t rhs_stride, void* stream);

/* Out-of-place solve with the gather/scatter fused: b_m[i] = rhs[rows[m*n+i]],
 * x[cols[m*n+i]] = x_m[i]; rows/cols int64 (batch, n) device arrays or NULL (contiguous).
 * Replaces BandedBatch.solve_mapped_ptr (core/bindings/module.cpp:335-343). */
int cn_bb_solve_mapped(cn_bb_t h, const void* rhs, const int64_t* rows, void* x,
                       const int64_t* cols, void* stream);

/* Complex systems of real data: b = re[rows_c] + i re[rows_s]; x[cols_c] = Re, x[cols_s] = Im
 * (complex128 handles only).  Replaces BandedBatch.solve_split_ptr
 * (core/bindings/module.cpp:344-353). */
int cn_bb_solve_split(cn_bb_t h, const double* re, const int64_t* rows_c, const int64_t* rows_s,
                      double* x, const int64_t* cols_c, const int64_t* cols_s, void* stream);

/* Build the transposed Woodbury data now (before graph capture of transposed solves).
 * Replaces BandedBatch.prepare_transpose (core/bindings/module.cpp:392-394). */
int cn_bb_prepare_transpose(cn_bb_t h);

/* Properties (n, kl, ku, rank, batch, cond_max, factor_ms, ...).  Replaces the read-only
 * properties of cuda.BandedBatch (core/bindings/module.cpp:402-409). */
int cn_bb_props_get(cn_bb_t h, cn_bb_props* p);

/* Copy the per-system capacitance condition numbers (batch doubles) to host memory. */
int cn_bb_cond(cn_bb_t h, double* out);

/* Release the handle and its device memory. */
void cn_bb_destroy(cn_bb_t h);

/* Copy the factored data of systems src_first .. src_first+count-1 of `src` into systems
 * dst_first .. of `dst` (asynchronous device copies on `stream`; no refactorization): both handles
 * must have the same type, n, band, border rows, kernel mode and device.  For servers that move a
 * run's factored mode systems into a free slot of a resident batch; no counterpart in the
 * reference (its batches are built once, core/include/conatus/cuda.hpp:57-120). */
int cn_bb_copy_systems(cn_bb_t dst, int dst_first, cn_bb_t src, int src_first, int count, void* stream);

/* Kernel mode of the handles the calling thread creates from now on: 0 automatic (lane mode for
 * many narrow-band systems, partitioned wave mode for up to 256 systems long enough to split,
 * wave mode otherwise), 1 lane mode (one lane per system and right-hand side, streamed sweeps),
 * 2 wave mode (one wavefront per system and right-hand side), 3 partitioned wave mode (the sweeps
 * of a system split over 2-8 wavefronts, coupled through spikes computed at factor time), 4 and 5
 * wave mode with staged blocks of 8 and of 32 steps (else chosen by the launch size; the results
 * do not depend on it); a mode that does not fit the shape falls back to the next one.  For the
 * per-size-class autotuner;
 * no counterpart in the reference (its CUDA batch picks band vs level-scheduled solves by flags,
 * core/include/conatus/cuda.hpp:57-120). */
int cn_bb_request_mode(int mode);

/* ---- IMEX stage (core/include/conatus/imex.hpp) -------------------------------------------
 * Batched block-diagonal CSR operators on the device and the fused right-hand-side assembly of
 * M dX/dt + L X = F.  Every product and sum is rounded separately, in the reference's order, so
 * the assembled right-hand sides equal the host integrator's bit for bit.  N = batch * n. */
typedef struct cn_csr* cn_csr_t;

/* Upload `batch` CSR matrices of size n (same dtype) as one block-diagonal device operator.
 * Replaces the device copy of M / L in cuda.Imex (core/src/cuda/imex.cu:56-86). */
int cn_csr_create(cn_csr_t* out, int dtype, int batch, int n, const int32_t* const* indptr,
                  const int32_t* const* indices, const void* const* data, const int64_t* nnz);
void cn_csr_destroy(cn_csr_t A);

/* y = alpha * A x (+ y if accumulate), one row per thread, storage order (banded.hpp:70-76;
 * DeviceCsr.__matmul__, python/conatus/timestep/device.py:78-105). */
int cn_csr_matvec(cn_csr_t A, const void* x, void* y, double alpha, int accumulate, void* stream);

/* One SBDF stage (imex.hpp:158-176, lincomb.py:146-162): mx0 = M x; f0 = fext, or g x + h
 * (g, h may be NULL); rhs = sum_{j=1..q} c[j] F^{n-j} - a_over_dt[j] M X^{n-j}, level 1 being
 * (f0, mx0) and levels 2..q read from f_hist[j-2], mx_hist[j-2].  f0 may be NULL (not stored). */
int cn_sbdf_stage(cn_csr_t M, const void* x, const void* g, const void* h, const void* fext, void* mx0,
                  void* f0, const void* const* mx_hist, const void* const* f_hist, int q, const double* c,
                  const double* a_over_dt, void* rhs, void* stream);

/* RK stage right-hand side (imex.hpp:193-199, lincomb.py:73-139):
 * rhs = mx0 + sum_{k<stages} dtA[k] F[k] - dtH[k] LX[k] (zero coefficients skipped). */
int cn_rk_rhs(int dtype, int64_t N, const void* mx0, const void* const* F, const void* const* LX,
              const double* dtA, const double* dtH, int stages, void* rhs, void* stream);

/* out = g * x + h (pointwise explicit term of cuda.Imex, core/src/cuda/imex.cu:47-55). */
int cn_pointwise(int dtype, int64_t N, void* out, const void* x, const void* g, const void* h, void* stream);

/* ---- Fourier x Chebyshev explicit term (2D Rayleigh-Benard) -------------------------------
 * F rows of the momentum and temperature equations from the state X (K modes x n, complex):
 * S_{0->2} forward(-u.grad u, -u.grad T) on the 3/2 grid.  Replaces the reference's GPU 2D
 * pipeline FourierChebyshev2DRHS.evaluate (python/conatus/timestep/fc2d.py:567-643) and the fused
 * product evaluation (python/conatus/timestep/fused.py:994-1085).  Layout: coefficient m of field f (u, w, T) of
 * mode k at X[k*n + fidx[f*Nz + m]] (fidx: 3 x Nz device ints), equation e rows at F[k*n + R + 4*m + e]; other rows of F are left
 * untouched.  Scratch: Gh (9 K Mz complex: 3 K Mz line products, 6 K Mz line values).  Tables: Bz (2 x Nz x Mz: T_n, T_n' at the nodes),
 * Wx (Mx complex twiddles exp(2 pi i m / Mx)), Az (Mz x Nz: (S02 Fz)^T).  Needs
 * (6 K + Mx) 16 + (3 Mx + 2 Nz) 8 <= 65536 bytes of LDS (CN_EUNSUPPORTED otherwise). */
int cn_rb2d_explicit(const void* X, const int* fidx, void* F, void* Gh, const double* Bz, const void* Wx,
                     const double* Az, int K, int Nz, int Mx, int Mz, int n, int R, double kappa, void* stream);

/* The same for `members` states of one discretization stored one after another (X and F:
 * members x K x n, Gh: members x 9 K Mz): the explicit term of an ensemble in one launch pair
 * (the reference evaluates an ensemble's F through BatchedRHS, python/conatus/ensembles/batched.py:116-440).
 * 1 <= members <= 65535. */
int cn_rb2d_explicit_many(const void* X, const int* fidx, void* F, void* Gh, const double* Bz, const void* Wx,
                          const double* Az, int K, int Nz, int Mx, int Mz, int n, int R, double kappa, int members,
                          void* stream);

/* ---- Fourier x Chebyshev x Fourier explicit term (3D plane channel) -------------------------
 * F rows of the x, y and z momentum equations from the state X (Kx*Kz modes x n, complex;
 * Kx = Nx/2, Kz = Nz-1, mode kx*Kz + iz with kz = iz for iz < Nz/2 else iz - Kz):
 * F = base + S_{0->2} forward(-u.grad u) on the Mx x My x Mz grid.  Replaces the reference's 3D
 * explicit evaluation (python/conatus/timestep/fused.py:994-1085 with the 3D transforms of
 * python/conatus/timestep/xp_transforms.py:197-548).  Layout: coefficient m of velocity
 * component f (u, v, w) of mode k at X[k*n + fidx[f*Ny + m]]; equation e rows at
 * F[k*n + R + 4*m + e] (base: the constant part, same layout as F); other rows of F are left
 * untouched.  Tables: By (My x Ny x 2: T_n(y_j), T_n'(y_j)), Wx, Wz (Mx, Mz complex twiddles
 * exp(2 pi i q / M)), Ay (My x Ny: (S02 Fy)^T); kx0 = 2 pi / Lx, kz0 = 2 pi / Lz.  Workspace:
 * ws of at least cn_ch3d_workspace_bytes(...) device bytes.  Five launches on `stream`. */
size_t cn_ch3d_workspace_bytes(int Nx, int Ny, int Nz, int Mx, int My, int Mz);
int cn_ch3d_explicit(const void* X, const int* fidx, const void* base, void* F, void* ws, size_t ws_bytes,
                     const double* By, const void* Wx, const void* Wz, const double* Ay, int Nx, int Ny, int Nz,
                     int Mx, int My, int Mz, int n, int R, double kx0, double kz0, void* stream);
/* The same for `members` (1 .. 65535) states of one grid stored one after another (X and F:
 * members x Kx*Kz x n; base: one member's, shared), each with its own workspace region of
 * cn_ch3d_workspace_bytes(...) (ws_bytes >= members times that): one launch sequence for all
 * members, member m's arithmetic that of its own cn_ch3d_explicit call.  No counterpart in the
 * reference, whose ensembles batch equal right-hand sides only (ensembles/batched.py:55-613);
 * for ensembles and server pages of channel runs. */
int cn_ch3d_explicit_many(const void* X, const int* fidx, const void* base, void* F, void* ws, size_t ws_bytes,
                          const double* By, const void* Wx, const void* Wz, const double* Ay, int Nx, int Ny, int Nz,
                          int Mx, int My, int Mz, int n, int R, double kx0, double kz0, int members, void* stream);

/* ---- 3D plane channel, poloidal-toroidal (velocity-vorticity) form ---------------------------
 * The reference's channel_3d_pt (tests/channel3d/ch3d_cases.py:90-217).  State X (float64): the
 * phi systems [4][P][Ny + 4] then the psi systems [4][P][Ny + 2], P = (Nx/2)(Nz/2) slots
 * s = kx (Nz/2) + |kz|, columns (Re, Im) of mode (kx, +|kz|) then of (kx, -|kz|); phi rows
 * [phi(-1), phi(1), phi'(-1), phi'(1), C^(4) equation rows], psi rows [psi(-1), psi(1), C^(2)
 * rows]; psi slot 0 holds the mean flow U (column 0) and W (column 1).
 *
 * F = the explicit right-hand sides of every row (the curl curl and curl of N = -div(u u) for the
 * phi and psi rows, the x-z means of Nx + G and Nz for U and W, zero elsewhere) from the six
 * products on the Mx x My x Mz grid (Mx, Mz of the form 2^a 3^b).  Replaces the reference's
 * evaluation of those right-hand sides (python/conatus/timestep/spectral.py:94-122 with the 3D
 * transforms of python/conatus/timestep/xp_transforms.py:197-548).  The y tables are folded by
 * the parity of the Chebyshev polynomials over the symmetric Gauss nodes (y_{My-1-j} = -y_j),
 * with Mh = ceil(My/2), KS = 8 ceil(ceil(Ny/2)/8), Mhs = 8 ceil(Mh/8), MHP = 16 ceil(Mh/16),
 * RE = 16 ceil(ceil(Ny/2)/16), RO = 16 ceil(floor(Ny/2)/16), zero padded:
 *   Tb [4][KS + 4][MHP]: T_{2k}(y_j), T_{2k+1}(y_j), T'_{2k}(y_j), T'_{2k+1}(y_j), j < Mh;
 *   Am [5][Mhs + 4][RE + RO]: S14 D01 Fy, S24 D02 Fy, S04 Fy, S02 Fy, S12 D01 Fy (Fy the
 *      Gauss-Chebyshev forward transform) at node j < Mh, row m at index m/2 (even m) or
 *      RE + (m-1)/2 (odd m);
 * s02c0 (column 0 of S02), Wx, Wz (Mx, Mz complex twiddles exp(-2 pi i k / M)).  Workspace: ws
 * of at least cn_chpt_workspace_bytes(...) bytes.  Six launches on `stream`. */
size_t cn_chpt_workspace_bytes(int Nx, int Ny, int Nz, int Mx, int My, int Mz);
int cn_chpt_explicit(const double* X, double* F, void* ws, size_t ws_bytes, const double* Tb, const double* Am,
                     const double* s02c0, const void* Wx, const void* Wz, int Nx, int Ny, int Nz, int Mx, int My,
                     int Mz, double kx0, double kz0, double G, void* stream);

/* One SBDF stage of that state (core/include/conatus/imex.hpp:158-176): mx0 = M X (M from the
 * banded stencils s04, s2d (Ny x 5: S04 and S24 D02 at offsets 0, 2, .., 8) and s02 (Ny x 3)
 * and k^2 of each slot), then X = sum_{j=1..q} c[j] F_j - a_over_dt[j] M X_j in place (level 1:
 * f_levels[0] and mx0; levels j >= 2: f_levels[j-1], mx_hist[j-2]), ready for the in-place solve
 * of each bucket.  Replaces the stage of the reference's IVPSolver step
 * (python/conatus/timestep/integrator.py:361-373, lincomb.py:146-162) for this problem. */
int cn_chpt_sbdf_stage(double* X, const double* const* f_levels, double* mx0, const double* const* mx_hist, int q,
                       const double* c, const double* a_over_dt, const double* s04, const double* s2d,
                       const double* s02, int Nx, int Ny, int Nz, double kx0, double kz0, void* stream);

/* ---- IMEX integrator handle ---------------------------------------------------------------------
 * The reference's device integrator cuda.Imex{D,Z} (core/src/cuda/imex.cu:56-244, bindings
 * core/bindings/module.cpp:411-442; host twin ImexIntegrator, core/include/conatus/imex.hpp:73-219)
 * for callers that bind the library directly: batch systems M_k dX/dt + L_k X = F of size n (CSR
 * per system, as cn_bb_create takes them), border rows, scheme "SBDF1".."SBDF4", "RK111", "RK222"
 * or "RK443", constant dt.  F is zero or pointwise (cn_imex_set_pointwise: F = g .* X + h, device
 * arrays of batch * n entries, copied).  cn_imex_run advances the (batch, n) device state x in
 * place by nsteps on `stream` (the factorization of a new implicit operator synchronises once);
 * *t_out = t0 plus one addition of dt per step.  A run that holds an SBDF history refuses an
 * (x, t0) other than the one the previous run returned (CN_EINVAL; cn_imex_reset starts a new
 * trajectory, factorizations kept).  The arithmetic is that of conatus_amd.runtime.imex.Imex,
 * bit for bit. */
typedef struct cn_imex* cn_imex_t;
int cn_imex_create(cn_imex_t* out, int dtype, int batch, int n, const int32_t* const* M_indptr,
                   const int32_t* const* M_indices, const void* const* M_data, const int64_t* M_nnz,
                   const int32_t* const* L_indptr, const int32_t* const* L_indices, const void* const* L_data,
                   const int64_t* L_nnz, const int32_t* border, int r, const char* scheme, double dt, int pivot);
int cn_imex_set_pointwise(cn_imex_t h, const void* g, const void* hp, void* stream);
int cn_imex_run(cn_imex_t h, void* x, double t0, int nsteps, double* t_out, void* stream);
int cn_imex_reset(cn_imex_t h);
int cn_imex_info(cn_imex_t h, int* history_length, int* n_factorizations, int64_t* iteration);
void cn_imex_destroy(cn_imex_t h);

/* ---- Batched grid <-> coefficient transforms along the last axis ------------------------------
 * out[r][k] = sum_m A[k][m] in[r][m] for every row r < rows of a C-contiguous (rows, nin) array,
 * out (rows, nout): parts = 1 for real rows, 2 for complex rows ((re, im) interleaved, the same
 * real map A = Ar applied to both parts), or a complex map A = Ar + i Ai (Ai non-null, parts = 2).
 * Ar and Ai are (nout, nin) row-major device matrices: the transform of one basis and direction,
 * built by the caller with the reference's conventions (Gauss-Chebyshev nodes and the (-1)^n
 * DCT-II / M scaling with c_0 halved; RealFourier [cos_k, msin_k] pairs; ComplexFourier in fftfreq
 * order with the Nyquist entry zero).  Replaces the device transforms behind
 * python/conatus/runtime/cuda.py:93-125 (forward, backward) and their bindings cheb_forward,
 * cheb_backward, realfourier_forward, realfourier_backward, complexfourier_forward,
 * complexfourier_backward (core/bindings/module.cpp:300-495).  Out of place (in != out); one
 * launch on `stream`, no allocation, no host synchronisation. */
int cn_tr_rows(const double* in, double* out, int64_t rows, int nin, int nout, int parts, const double* Ar,
               const double* Ai, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CNHIP_H */
