// The batched IMEX integrator behind the C ABI (cn_imex_*): the reference's device integrator
// cuda.Imex{D,Z} (core/src/cuda/imex.cu:56-244, core/bindings/module.cpp:411-442) for callers that
// bind the library directly.  Host orchestration only: every device operation is one of the
// library's own entry points (cn_csr_matvec, cn_sbdf_stage, cn_rk_rhs, cn_pointwise, cn_bb_*),
// issued in the order conatus_amd.runtime.imex.Imex issues them, so that the two agree bit for bit.
//
// Schemes: SBDF1..4 (constant step, startup orders 1..q: level j of the history ring is the state
// j steps back), and the IMEX Runge-Kutta tableaux RK111, RK222, RK443 (cumulative form
// M X_i = M X_0 + dt sum_j (A_ij F_j - H_ij L X_j), one factorization per distinct H_ii).  F is
// zero or pointwise, F = g .* X + h.  The time advances by one addition of dt per step
// (imex.hpp:112-116).  A run with a history refuses an (x, t0) that is not what the previous run
// returned (the continuation rule of imex.cu:145-151).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "cn_common.h"
#include "cnhip.h"

namespace {

struct ImexFail {
    int code;
    std::string msg;
};

[[noreturn]] void ifail(int code, const std::string& m) { throw ImexFail{code, m}; }

void hcheck(hipError_t e, const char* what) {
    if (e != hipSuccess) ifail(CN_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

// the status of a library call, with its message kept
void lcheck(int st, const char* what) {
    if (st != CN_OK) ifail(st, std::string(what) + ": " + cn_last_error());
}

struct DevMem {
    void* p = nullptr;
    void alloc(size_t b) {
        if (b) hcheck(hipMalloc(&p, b), "hipMalloc");
    }
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

// host CSR of one system
struct Csr {
    std::vector<int32_t> ip, ix;
    std::vector<double> re, im;   // im empty for real systems
};

// the RK tableaux of imex.hpp:31-51 (cumulative form), the SBDF constants of imex.hpp:63-71
struct Tableau {
    std::vector<std::vector<double>> A, H;
    std::vector<double> c;
};

bool rk_tableau(const std::string& name, Tableau& t) {
    if (name == "RK111") {
        t.A = {{0, 0}, {1, 0}};
        t.H = {{0, 0}, {0, 1}};
        t.c = {0, 1};
        return true;
    }
    if (name == "RK222") {
        const double g = (2.0 - std::sqrt(2.0)) / 2.0, d = 1.0 - 1.0 / g / 2.0;
        t.A = {{0, 0, 0}, {g, 0, 0}, {d, 1 - d, 0}};
        t.H = {{0, 0, 0}, {0, g, 0}, {0, 1 - g, g}};
        t.c = {0, g, 1};
        return true;
    }
    if (name == "RK443") {
        t.A = {{0, 0, 0, 0, 0}, {1.0 / 2, 0, 0, 0, 0}, {11.0 / 18, 1.0 / 18, 0, 0, 0}, {5.0 / 6, -5.0 / 6, 1.0 / 2, 0, 0},
               {1.0 / 4, 7.0 / 4, 3.0 / 4, -7.0 / 4, 0}};
        t.H = {{0, 0, 0, 0, 0}, {0, 1.0 / 2, 0, 0, 0}, {0, 1.0 / 6, 1.0 / 2, 0, 0}, {0, -1.0 / 2, 1.0 / 2, 1.0 / 2, 0},
               {0, 3.0 / 2, -3.0 / 2, 1.0 / 2, 1.0 / 2}};
        t.c = {0, 1.0 / 2, 2.0 / 3, 1.0 / 2, 1};
        return true;
    }
    return false;
}

void sbdf_coefficients(int q, std::vector<double>& a, std::vector<double>& c) {
    switch (q) {
        case 1: a = {1.0, -1.0}; c = {0.0, 1.0}; break;
        case 2: a = {1.5, -2.0, 0.5}; c = {0.0, 2.0, -1.0}; break;
        case 3: a = {11.0 / 6, -3.0, 1.5, -1.0 / 3}; c = {0.0, 3.0, -3.0, 1.0}; break;
        default: a = {25.0 / 12, -4.0, 3.0, -4.0 / 3, 0.25}; c = {0.0, 4.0, -6.0, 4.0, -1.0}; break;
    }
}

}  // namespace

struct cn_imex {
    int dtype = CN_F64, batch = 0, n = 0, order = 0, pivot = 1;
    size_t esz = 8;
    std::string scheme;
    double dt = 0.0;
    std::vector<int32_t> border;
    std::vector<Csr> M, L;
    cn_csr_t Md = nullptr, Ld = nullptr;
    Tableau tab;
    std::map<std::pair<double, double>, cn_bb_t> factors;
    DevMem rhs, g, h, last;
    std::vector<std::unique_ptr<DevMem>> mx, f;            // SBDF ring: (M X, F) per level slot
    std::vector<std::unique_ptr<DevMem>> Xs, Fs, LXs;      // RK stages
    DevMem mx0;
    bool has_g = false, has_h = false;
    std::vector<int> levels;                              // ring slots, newest first
    bool has_last = false;
    double last_t = 0.0;
    long long iteration = 0;

    size_t bytes() const { return (size_t)batch * n * esz; }
    ~cn_imex() {
        for (auto& kv : factors) cn_bb_destroy(kv.second);
        if (Md) cn_csr_destroy(Md);
        if (Ld) cn_csr_destroy(Ld);
    }

    // alpha M_k + beta L_k of every system, built as scipy builds it in factor_operator (the
    // scaled matrices added entry by entry, a sum that is exactly zero dropped), then factored
    cn_bb_t factor(double alpha, double beta) {
        const auto key = std::make_pair(alpha, beta);
        auto it = factors.find(key);
        if (it != factors.end()) return it->second;
        const bool cplx = dtype == CN_C128;
        std::vector<std::vector<int32_t>> ips(batch), ixs(batch);
        std::vector<std::vector<double>> vals(batch);
        std::vector<const int32_t*> pip(batch), pix(batch);
        std::vector<const void*> pdv(batch);
        std::vector<int64_t> nnz(batch);
        for (int m = 0; m < batch; ++m) {
            const Csr& A = M[m];
            const Csr& B = L[m];
            auto& ip = ips[m];
            auto& ix = ixs[m];
            auto& v = vals[m];
            ip.assign(1, 0);
            for (int i = 0; i < n; ++i) {
                int p = A.ip[i], q = B.ip[i];
                const int pe = A.ip[i + 1], qe = B.ip[i + 1];
                while (p < pe || q < qe) {
                    const int ca = p < pe ? A.ix[p] : n, cb = q < qe ? B.ix[q] : n;
                    const int col = ca < cb ? ca : cb;
                    double sr = 0.0, si = 0.0;
                    if (ca == col && cb == col) {
                        sr = alpha * A.re[p] + beta * B.re[q];
                        if (cplx) si = alpha * A.im[p] + beta * B.im[q];
                        ++p;
                        ++q;
                    } else if (ca == col) {
                        sr = alpha * A.re[p];
                        if (cplx) si = alpha * A.im[p];
                        ++p;
                    } else {
                        sr = beta * B.re[q];
                        if (cplx) si = beta * B.im[q];
                        ++q;
                    }
                    if (sr == 0.0 && si == 0.0) continue;
                    ix.push_back(col);
                    v.push_back(sr);
                    if (cplx) v.push_back(si);
                }
                ip.push_back((int32_t)ix.size());
            }
            pip[m] = ip.data();
            pix[m] = ix.data();
            pdv[m] = v.data();
            nnz[m] = (int64_t)ix.size();
        }
        cn_bb_t hb = nullptr;
        lcheck(cn_bb_create(&hb, dtype, batch, n, pip.data(), pix.data(), pdv.data(), nnz.data(), border.data(),
                            (int)border.size(), pivot, 0),
               "cn_imex: factor");
        const int st = cn_bb_factor(hb, nullptr, 1);
        if (st != CN_OK) {
            const std::string msg = cn_last_error();
            cn_bb_destroy(hb);
            ifail(st, "cn_imex: factor: " + msg);
        }
        const int rs = cn_bb_reserve(hb, 1);
        if (rs != CN_OK) {
            cn_bb_destroy(hb);
            ifail(rs, "cn_imex: reserve");
        }
        factors[key] = hb;
        return hb;
    }

    void sbdf_step(void* x, hipStream_t s) {
        std::vector<int> lv;
        if ((int)levels.size() < order) {
            int slot = 0;
            while (std::find(levels.begin(), levels.end(), slot) != levels.end()) ++slot;
            lv.push_back(slot);
            lv.insert(lv.end(), levels.begin(), levels.end());
        } else {
            lv.push_back(levels.back());
            lv.insert(lv.end(), levels.begin(), levels.end() - 1);
        }
        const int q = (int)lv.size();
        std::vector<double> a, c;
        sbdf_coefficients(q, a, c);
        const void* mxh[4] = {nullptr, nullptr, nullptr, nullptr};
        const void* fh[4] = {nullptr, nullptr, nullptr, nullptr};
        for (int j = 1; j < q; ++j) {
            mxh[j - 1] = mx[lv[j]]->p;
            fh[j - 1] = f[lv[j]]->p;
        }
        double cc[5] = {0, 0, 0, 0, 0}, adt[5] = {0, 0, 0, 0, 0};
        for (int j = 1; j <= q; ++j) {
            cc[j] = c[j];
            adt[j] = a[j] / dt;
        }
        lcheck(cn_sbdf_stage(Md, x, has_g ? g.p : nullptr, has_h ? h.p : nullptr, nullptr, mx[lv[0]]->p, f[lv[0]]->p,
                             mxh, fh, q, cc, adt, rhs.p, s),
               "cn_imex: SBDF stage");
        cn_bb_t fac = factor(a[0] / dt, 1.0);
        lcheck(cn_bb_solve_mapped(fac, rhs.p, nullptr, x, nullptr, s), "cn_imex: solve");
        levels = lv;
    }

    void rk_step(void* x, hipStream_t s) {
        const auto& A = tab.A;
        const auto& H = tab.H;
        const int st = (int)tab.c.size() - 1;
        const int64_t N = (int64_t)batch * n;
        std::vector<void*> X(st + 1);
        X[0] = x;
        for (int j = 1; j < st; ++j) X[j] = Xs[j - 1]->p;
        X[st] = x;
        lcheck(cn_csr_matvec(Md, x, mx0.p, 1.0, 0, s), "cn_imex: M X0");
        auto needF = [&](int j) {
            for (int i = j + 1; i <= st; ++i)
                if (A[i][j] != 0) return true;
            return false;
        };
        auto needLX = [&](int j) {
            for (int i = j + 1; i <= st; ++i)
                if (H[i][j] != 0) return true;
            return false;
        };
        for (int i = 1; i <= st; ++i) {
            const int j = i - 1;
            if (needF(j))
                lcheck(cn_pointwise(dtype, N, Fs[j]->p, X[j], has_g ? g.p : nullptr, has_h ? h.p : nullptr, s),
                       "cn_imex: F");
            if (needLX(j)) lcheck(cn_csr_matvec(Ld, X[j], LXs[j]->p, 1.0, 0, s), "cn_imex: L X");
            const void* Fp[4] = {nullptr, nullptr, nullptr, nullptr};
            const void* Lp[4] = {nullptr, nullptr, nullptr, nullptr};
            double dA[4] = {0, 0, 0, 0}, dH[4] = {0, 0, 0, 0};
            for (int k = 0; k < i; ++k) {
                dA[k] = dt * A[i][k];
                dH[k] = dt * H[i][k];
                if (dA[k] != 0) Fp[k] = Fs[k]->p;
                if (dH[k] != 0) Lp[k] = LXs[k]->p;
            }
            lcheck(cn_rk_rhs(dtype, N, mx0.p, Fp, Lp, dA, dH, i, rhs.p, s), "cn_imex: RK stage");
            cn_bb_t fac = factor(1.0, dt * H[i][i]);
            lcheck(cn_bb_solve_mapped(fac, rhs.p, nullptr, X[i], nullptr, s), "cn_imex: solve");
        }
    }
};

namespace {

template <class F>
int iguard(F&& fn) {
    try {
        fn();
        cn::set_error("");
        return CN_OK;
    } catch (const ImexFail& e) {
        cn::set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        cn::set_error("cn_imex: host allocation failed");
        return CN_EOVERFLOW;
    } catch (const std::exception& e) {
        cn::set_error(e.what());
        return CN_ELOGIC;
    }
}

void read_csr(int dtype, int batch, int n, const int32_t* const* ip, const int32_t* const* ix, const void* const* dv,
              const int64_t* nnz, std::vector<Csr>& out, const char* what) {
    if (!ip || !ix || !dv || !nnz) ifail(CN_EINVAL, std::string("cn_imex_create: null ") + what + " array");
    out.resize(batch);
    for (int m = 0; m < batch; ++m) {
        if (!ip[m] || (nnz[m] > 0 && (!ix[m] || !dv[m])))
            ifail(CN_EINVAL, std::string("cn_imex_create: null ") + what + " array of system " + std::to_string(m));
        if (ip[m][0] != 0 || ip[m][n] != nnz[m])
            ifail(CN_EINVAL, std::string("cn_imex_create: ") + what + " of system " + std::to_string(m) +
                                 ": indptr must start at 0 and end at nnz");
        Csr& c = out[m];
        c.ip.assign(ip[m], ip[m] + n + 1);
        c.ix.assign(ix[m], ix[m] + nnz[m]);
        c.re.resize(nnz[m]);
        if (dtype == CN_C128) c.im.resize(nnz[m]);
        const double* d = static_cast<const double*>(dv[m]);
        for (int64_t p = 0; p < nnz[m]; ++p) {
            if (dtype == CN_C128) {
                c.re[p] = d[2 * p];
                c.im[p] = d[2 * p + 1];
            } else {
                c.re[p] = d[p];
            }
            if (!std::isfinite(c.re[p]) || (dtype == CN_C128 && !std::isfinite(c.im[p])))
                ifail(CN_EINVAL, std::string("cn_imex_create: non-finite entry in ") + what);
        }
        for (int i = 0; i < n; ++i) {
            if (c.ip[i + 1] < c.ip[i]) ifail(CN_EINVAL, std::string("cn_imex_create: ") + what + ": indptr decreases");
            for (int p = c.ip[i]; p < c.ip[i + 1]; ++p) {
                if (c.ix[p] < 0 || c.ix[p] >= n) ifail(CN_EINVAL, std::string("cn_imex_create: ") + what + ": column out of range");
                if (p > c.ip[i] && c.ix[p] <= c.ix[p - 1])
                    ifail(CN_EINVAL, std::string("cn_imex_create: ") + what + ": columns of a row must increase");
            }
        }
    }
}

}  // namespace

extern "C" {

int cn_imex_create(cn_imex_t* out, int dtype, int batch, int n, const int32_t* const* M_indptr,
                   const int32_t* const* M_indices, const void* const* M_data, const int64_t* M_nnz,
                   const int32_t* const* L_indptr, const int32_t* const* L_indices, const void* const* L_data,
                   const int64_t* L_nnz, const int32_t* border, int r, const char* scheme, double dt, int pivot) {
    return iguard([&] {
        if (!out) ifail(CN_EINVAL, "cn_imex_create: null output handle");
        *out = nullptr;
        if (dtype != CN_F64 && dtype != CN_C128) ifail(CN_EINVAL, "cn_imex_create: dtype must be CN_F64 or CN_C128");
        if (batch < 1 || n < 1) ifail(CN_EINVAL, "cn_imex_create: batch and n must be >= 1");
        if (!scheme) ifail(CN_EINVAL, "cn_imex_create: null scheme");
        if (!(dt > 0.0) || !std::isfinite(dt)) ifail(CN_EINVAL, "cn_imex_create: dt must be > 0");
        if (r < 0 || r > n || (r > 0 && !border)) ifail(CN_EINVAL, "cn_imex_create: border rows");
        auto h = std::make_unique<cn_imex>();
        h->dtype = dtype;
        h->batch = batch;
        h->n = n;
        h->esz = dtype == CN_C128 ? 16 : 8;
        h->scheme = scheme;
        h->dt = dt;
        h->pivot = pivot ? 1 : 0;
        h->border.assign(border, border + r);
        const std::string sc = scheme;
        if (sc == "SBDF1" || sc == "SBDF2" || sc == "SBDF3" || sc == "SBDF4") {
            h->order = sc[4] - '0';
        } else if (!rk_tableau(sc, h->tab)) {
            ifail(CN_EINVAL, "cn_imex_create: unknown scheme " + sc + " (SBDF1..SBDF4, RK111, RK222, RK443)");
        }
        read_csr(dtype, batch, n, M_indptr, M_indices, M_data, M_nnz, h->M, "M");
        read_csr(dtype, batch, n, L_indptr, L_indices, L_data, L_nnz, h->L, "L");
        lcheck(cn_csr_create(&h->Md, dtype, batch, n, M_indptr, M_indices, M_data, M_nnz), "cn_imex_create: M");
        lcheck(cn_csr_create(&h->Ld, dtype, batch, n, L_indptr, L_indices, L_data, L_nnz), "cn_imex_create: L");
        const size_t B = h->bytes();
        h->rhs.alloc(B);
        if (h->order) {
            for (int i = 0; i < h->order; ++i) {
                h->mx.push_back(std::make_unique<DevMem>());
                h->mx.back()->alloc(B);
                h->f.push_back(std::make_unique<DevMem>());
                h->f.back()->alloc(B);
                hcheck(hipMemset(h->f.back()->p, 0, B), "hipMemset");
            }
        } else {
            const int st = (int)h->tab.c.size() - 1;
            h->mx0.alloc(B);
            for (int i = 0; i < st; ++i) {
                h->Xs.push_back(std::make_unique<DevMem>());
                h->Xs.back()->alloc(B);
                h->Fs.push_back(std::make_unique<DevMem>());
                h->Fs.back()->alloc(B);
                hcheck(hipMemset(h->Fs.back()->p, 0, B), "hipMemset");
                h->LXs.push_back(std::make_unique<DevMem>());
                h->LXs.back()->alloc(B);
            }
        }
        h->last.alloc(B);
        *out = h.release();
    });
}

int cn_imex_set_pointwise(cn_imex_t h, const void* g, const void* hh, void* stream) {
    return iguard([&] {
        if (!h) ifail(CN_EINVAL, "cn_imex_set_pointwise: null handle");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const size_t B = h->bytes();
        h->has_g = g != nullptr;
        h->has_h = hh != nullptr;
        if (g) {
            if (!h->g.p) h->g.alloc(B);
            hcheck(hipMemcpyAsync(h->g.p, g, B, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync g");
        }
        if (hh) {
            if (!h->h.p) h->h.alloc(B);
            hcheck(hipMemcpyAsync(h->h.p, hh, B, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync h");
        }
    });
}

int cn_imex_run(cn_imex_t h, void* x, double t0, int nsteps, double* t_out, void* stream) {
    return iguard([&] {
        if (!h || !x) ifail(CN_EINVAL, "cn_imex_run: null handle or state");
        if (nsteps < 0) ifail(CN_EINVAL, "cn_imex_run: nsteps must be >= 0");
        hipStream_t s = static_cast<hipStream_t>(stream);
        const size_t B = h->bytes();
        if (!h->levels.empty()) {   // continuation: (x, t0) must be what the previous run returned
            bool same = h->has_last && h->last_t == t0;
            if (same) {
                std::vector<unsigned char> a(B), b(B);
                hcheck(hipMemcpyAsync(a.data(), x, B, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
                hcheck(hipMemcpyAsync(b.data(), h->last.p, B, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
                hcheck(hipStreamSynchronize(s), "hipStreamSynchronize");
                same = std::memcmp(a.data(), b.data(), B) == 0;
            }
            if (!same)
                ifail(CN_EINVAL, "cn_imex_run: " + h->scheme + " holds the history of a previous run, and (x0, t0) is "
                                     "not the state and time that run returned; call cn_imex_reset to start a new "
                                     "trajectory, or pass the returned (x, t) to continue");
        }
        double t = t0;
        for (int k = 0; k < nsteps; ++k) {
            if (h->order) h->sbdf_step(x, s);
            else h->rk_step(x, s);
            ++h->iteration;
            t += h->dt;
        }
        if (h->order) {
            hcheck(hipMemcpyAsync(h->last.p, x, B, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
            h->has_last = true;
            h->last_t = t;
        }
        if (t_out) *t_out = t;
    });
}

int cn_imex_reset(cn_imex_t h) {
    return iguard([&] {
        if (!h) ifail(CN_EINVAL, "cn_imex_reset: null handle");
        h->levels.clear();
        h->has_last = false;
        h->iteration = 0;
    });
}

int cn_imex_info(cn_imex_t h, int* history_length, int* n_factorizations, int64_t* iteration) {
    return iguard([&] {
        if (!h) ifail(CN_EINVAL, "cn_imex_info: null handle");
        if (history_length) *history_length = (int)h->levels.size();
        if (n_factorizations) *n_factorizations = (int)h->factors.size();
        if (iteration) *iteration = h->iteration;
    });
}

void cn_imex_destroy(cn_imex_t h) { delete h; }

}  // extern "C"
