// extern "C" entry points declared in include/idrs_capi.h.
#include "idrs_capi.h"

#include "idrs.hpp"
#include "krylov.hpp"
#include "batch.hpp"
#include "mixed.hpp"
#include "mm.hpp"
#include "ritz.hpp"
#include "general.hpp"
#include "lu.hpp"

#include <cstdarg>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

struct idrs_user_matrix_s {
    idrs::DevMatrix *M;
};
struct idrs_matrix_s {
    idrs::Operator op;
};

namespace idrs {

static thread_local std::string g_err;

void set_last_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
}

template <class F>
static int guarded(F &&f) {
    try {
        g_err.clear();
        f();
        return IDRS_OK;
    } catch (const HipError &) {
        return IDRS_ERR_HIP;
    } catch (const std::invalid_argument &e) {
        g_err = e.what();
        return IDRS_ERR_ARG;
    } catch (const std::bad_alloc &) {
        g_err = "host allocation failed";
        return IDRS_ERR_INTERNAL;
    } catch (const std::exception &e) {
        g_err = e.what();
        return IDRS_ERR_INTERNAL;
    }
}

// device copies of host vectors for the host-pointer entry points
template <class V>
struct DVec {
    V *p = nullptr;
    size_t n = 0;
    DVec() = default;
    explicit DVec(size_t cnt) : n(cnt) { HIP_TRY(hipMalloc(&p, sizeof(V) * (cnt ? cnt : 1))); }
    DVec(const V *h, size_t cnt, hipStream_t st) : DVec(cnt) {
        if (h && cnt) HIP_TRY(hipMemcpyAsync(p, h, sizeof(V) * cnt, hipMemcpyHostToDevice, st));
    }
    ~DVec() { (void)hipFree(p); }
    DVec(const DVec &) = delete;
    DVec &operator=(const DVec &) = delete;
    DVec &operator=(DVec &&o) noexcept {
        std::swap(p, o.p);
        std::swap(n, o.n);
        return *this;
    }
    void download(V *h, hipStream_t st) const {
        HIP_TRY(hipMemcpyAsync(h, p, sizeof(V) * n, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
};

static void need(bool c, const char *msg) {
    if (!c) throw std::invalid_argument(msg);
}

template <class T>
static int idrs_entry(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                      idrs_stats_t *stats) {
    return guarded([&] {
        need(A && b && x, "null argument");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        idrs_opts_t o{};
        if (opts) o = *opts;
        SolveOpts<T> so;
        so.has_m1 = o.has_M1 != 0;
        so.M1 = (o.has_M1 && o.M1) ? &o.M1->op : nullptr;
        if (o.has_tolerance) so.tol = o.tolerance;
        if (o.has_maximum_iterations) {
            so.has_maxit = true;
            so.maxit = o.maximum_iterations;
        }
        if (o.has_variant) so.variant = o.variant;
        so.omega = (const T *)o.omega;
        so.n_omega = o.omega ? o.n_omega : 0;
        so.resvec = o.resvec;
        so.resvec_len = o.resvec_len;
        so.p_seed = o.p_seed ? o.p_seed : 1;
        need(o.graph_mode >= 0 && o.graph_mode <= 2, "graph_mode must be 0, 1 or 2");
        so.graph_mode = o.graph_mode;
        so.H = (T *)o.H;
        so.nritz = o.H ? o.H_nritz : 0;
        const int prow = (so.variant == 1) ? s : 1;
        SolveStats ss;
        SpmvTimer &tm = spmv_timer();
        tm.enabled = o.time_spmv != 0;
        if (o.device_pointers) {
            so.x0 = (const T *)o.x0;
            so.U0 = (const T *)o.U0;
            so.PT_raw = (const T *)o.PT_raw;
            idrs_solve<T>(A->op, n, s, (const T *)b, (T *)x, so, ss, st);
        } else {
            DVec<T> db((const T *)b, n, st), dx(n);
            DVec<T> dx0, dU0, dP;
            if (o.x0) {
                dx0 = DVec<T>((const T *)o.x0, n, st);
                so.x0 = dx0.p;
            }
            if (o.U0) {
                dU0 = DVec<T>((const T *)o.U0, (size_t)n * s, st);
                so.U0 = dU0.p;
            }
            if (o.PT_raw) {
                dP = DVec<T>((const T *)o.PT_raw, (size_t)n * prow, st);
                so.PT_raw = dP.p;
            }
            idrs_solve<T>(A->op, n, s, db.p, dx.p, so, ss, st);
            dx.download((T *)x, st);
        }
        tm.enabled = false;
        if (stats) {
            stats->flag = ss.flag;
            stats->iterations = ss.iterations;
            stats->relres = ss.relres;
            stats->solve_ms = ss.solve_ms;
            stats->spmv_ms = ss.spmv_ms;
            stats->spmv_launches = ss.spmv_launches;
            stats->reserved = 0;
        }
    });
}

// IDRS_RITZ: the solve with its Hessenberg matrix, then omega, U0 and H
// (ritz_module.f90:225-556).  opts->has_M1/M1 is the M argument; the
// shift-and-invert start block continues the shadow space's stream.
template <class T>
static int ritz_entry(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                      idrs_stats_t *stats, double *omega, int n_omega, double *U0, int dim_u0, double *H, int h_rows,
                      int h_cols, int *nritz_out, int *converged) {
    return guarded([&] {
        need(A && b && x, "null argument");
        need(s >= 1 && s <= SMAX, "s must be in 1..64");
        need(n_omega >= 0 && dim_u0 >= 0 && h_rows >= 0 && h_cols >= 0, "negative output size");
        need((n_omega == 0 || omega) && (dim_u0 == 0 || U0) && ((h_rows == 0 && h_cols == 0) || H),
             "output array missing");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        idrs_opts_t o{};
        if (opts) o = *opts;
        need(!o.U0 && !o.omega && !o.H, "IDRS_RITZ returns U0, omega and H; it takes none of them");
        need(!o.time_spmv, "SpMV timing is not offered by IDRS_RITZ");
        SolveOpts<T> so;
        if (o.has_tolerance) so.tol = o.tolerance;
        if (o.has_maximum_iterations) {
            so.has_maxit = true;
            so.maxit = o.maximum_iterations;
        }
        if (o.has_variant) so.variant = o.variant;
        so.resvec = o.resvec;
        so.resvec_len = o.resvec_len;
        so.p_seed = o.p_seed ? o.p_seed : 1;
        need(o.graph_mode >= 0 && o.graph_mode <= 2, "graph_mode must be 0, 1 or 2");
        so.graph_mode = o.graph_mode;
        const int prow = (so.variant == 1) ? s : 1;
        RitzRequest<T> rq;
        if (o.has_M1) {
            rq.has_M = true;
            rq.M = o.M1 ? &o.M1->op : nullptr;
        }
        rq.omega = (T *)omega;
        rq.n_omega = n_omega;
        rq.dimU0 = dim_u0;
        rq.H = (T *)H;
        rq.h_rows = h_rows;
        rq.h_cols = h_cols;
        rq.seed = so.p_seed;
        rq.offset = (unsigned long long)prow * (unsigned long long)n;
        SolveStats ss;
        int nritz = 0;
        if (o.device_pointers) {
            so.x0 = (const T *)o.x0;
            so.PT_raw = (const T *)o.PT_raw;
            rq.U0 = (T *)U0;
            nritz = idrs_ritz_solve<T>(A->op, n, s, (const T *)b, (T *)x, so, rq, ss, st);
        } else {
            DVec<T> db((const T *)b, n, st), dx(n), dx0, dP, dU((size_t)n * dim_u0);
            if (o.x0) {
                dx0 = DVec<T>((const T *)o.x0, n, st);
                so.x0 = dx0.p;
            }
            if (o.PT_raw) {
                dP = DVec<T>((const T *)o.PT_raw, (size_t)n * prow, st);
                so.PT_raw = dP.p;
            }
            rq.U0 = dim_u0 ? dU.p : nullptr;
            nritz = idrs_ritz_solve<T>(A->op, n, s, db.p, dx.p, so, rq, ss, st);
            dx.download((T *)x, st);
            if (dim_u0) dU.download((T *)U0, st);
        }
        if (nritz_out) *nritz_out = nritz;
        if (converged) *converged = rq.converged;
        if (stats) {
            stats->flag = ss.flag;
            stats->iterations = ss.iterations;
            stats->relres = ss.relres;
            stats->solve_ms = ss.solve_ms;
            stats->spmv_ms = 0.0;
            stats->spmv_launches = 0;
            stats->reserved = 0;
        }
    });
}

// ---------------------------------------------------------------------------
// allocation-free IDRS: every workspace buffer of the solve is carved from
// `scratch`, a region of NOALLOC_SCRATCH_BYTES in the code object's device
// data, so the call makes no hipMalloc/hipFree, no pinned host allocation,
// no event and no graph.  Each GPU holds its own copy of the array (until
// the process ends); its address on the current device is looked up once
// per device with hipGetSymbolAddress and kept, and only that address ever
// reaches a kernel or a copy.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on the current device (looked up on the
// first call on that device)
static void *scratch_on_current_device() {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    need(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

template <class T>
static int idrs_noalloc_entry(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                              idrs_stats_t *stats) {
    return guarded([&] {
        need(A && b && x, "null argument");
        idrs_opts_t o{};
        if (opts) o = *opts;
        need(!o.time_spmv, "SpMV timing is not offered by the allocation-free solve");
        need(!o.H, "the Hessenberg output is not offered by the allocation-free solve");
        need(op_prepared(A->op),
             "the operator's sliced copy is not built yet: run one ordinary product or solve on it first");
        hipStream_t st = lib_stream();
        Arena ar;
        ar.base = (char *)scratch_on_current_device();
        ar.cap = NOALLOC_SCRATCH_BYTES;
        SolveOpts<T> so;
        so.arena = &ar;
        so.has_m1 = o.has_M1 != 0;
        so.M1 = (o.has_M1 && o.M1) ? &o.M1->op : nullptr;
        if (o.has_tolerance) so.tol = o.tolerance;
        if (o.has_maximum_iterations) {
            so.has_maxit = true;
            so.maxit = o.maximum_iterations;
        }
        if (o.has_variant) so.variant = o.variant;
        so.omega = (const T *)o.omega;
        so.n_omega = o.omega ? o.n_omega : 0;
        so.resvec = o.resvec;
        so.resvec_len = o.resvec_len;
        so.p_seed = o.p_seed ? o.p_seed : 1;
        so.graph_mode = 2;
        so.x0 = (const T *)o.x0;
        so.U0 = (const T *)o.U0;
        so.PT_raw = (const T *)o.PT_raw;
        SolveStats ss;
        idrs_solve<T>(A->op, n, s, (const T *)b, (T *)x, so, ss, st);
        if (stats) {
            stats->flag = ss.flag;
            stats->iterations = ss.iterations;
            stats->relres = ss.relres;
            stats->solve_ms = ss.solve_ms;
            stats->spmv_ms = 0.0;
            stats->spmv_launches = 0;
            stats->reserved = 0;
        }
    });
}

enum class Kry { QMRIDR, MSIDRS, MSQMRIDR };

template <class T>
static int krylov_entry(Kry kind, idrs_matrix_t A, int n, const double *b, int ns, const double *sigma, int s,
                        const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats) {
    return guarded([&] {
        need(A && b && X, "null argument");
        need(kind == Kry::QMRIDR || (sigma && ns >= 1), "sigma must hold at least one shift");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        idrs_krylov_opts_t o{};
        if (opts) o = *opts;
        KrylovOpts<T> ko;
        ko.has_m1 = o.has_M1 != 0;
        ko.M1 = (o.has_M1 && o.M1) ? &o.M1->op : nullptr;
        if (o.has_tolerance) ko.tol = o.tolerance;
        if (o.has_maximum_iterations) {
            ko.has_maxit = true;
            ko.maxit = o.maximum_iterations;
        }
        if (o.has_variant) ko.variant = o.variant;
        ko.has_inner_s = o.has_inner_s != 0;
        ko.inner_s = o.inner_s;
        ko.has_inner_tol = o.has_inner_tolerance != 0;
        ko.inner_tol = o.inner_tolerance;
        ko.has_inner_it = o.has_inner_iterations != 0;
        ko.inner_it = o.inner_iterations;
        ko.omega = (const T *)o.omega;
        ko.n_omega = o.omega ? o.n_omega : 0;
        ko.resvec = o.resvec;
        ko.resvec_len = o.resvec_len;
        ko.p_seed = o.p_seed ? o.p_seed : 1;
        ko.colfac = (T *)o.colfac;
        ko.want_colfac = kind == Kry::MSIDRS && o.colfac != nullptr;
        const int nsig = kind == Kry::QMRIDR ? 1 : ns;
        const bool pt_rows_s = !(kind == Kry::MSIDRS && ko.variant == 2);
        KrylovStats ks;
        auto run = [&](const T *db, T *dX) {
            switch (kind) {
            case Kry::QMRIDR: qmridr_solve<T>(A->op, n, s, db, dX, ko, ks, st); break;
            case Kry::MSIDRS: msidrs_solve<T>(A->op, n, ns, (const T *)sigma, s, db, dX, ko, ks, st); break;
            case Kry::MSQMRIDR: msqmridr_solve<T>(A->op, n, ns, (const T *)sigma, s, db, dX, ko, ks, st); break;
            }
        };
        if (o.device_pointers) {
            ko.x0 = (const T *)o.x0;
            ko.PT_raw = (const T *)o.PT_raw;
            run((const T *)b, (T *)X);
        } else {
            need(n >= 1, "empty system");
            DVec<T> db((const T *)b, n, st), dX((size_t)n * nsig), dx0, dP;
            if (o.x0 && kind == Kry::QMRIDR) {
                dx0 = DVec<T>((const T *)o.x0, n, st);
                ko.x0 = dx0.p;
            }
            if (o.PT_raw && pt_rows_s) {
                need(s >= 1 && s <= SMAX, "s must be in 1..64");
                dP = DVec<T>((const T *)o.PT_raw, (size_t)n * s, st);
                ko.PT_raw = dP.p;
            }
            run(db.p, dX.p);
            dX.download((T *)X, st);
        }
        if (relres && kind != Kry::QMRIDR)
            for (int k = 0; k < ns; k++) relres[k] = ks.relres_vec[k];
        if (stats) {
            stats->flag = ks.flag;
            stats->iterations = ks.iterations;
            stats->relres = ks.relres;
            stats->solve_ms = ks.solve_ms;
            stats->spmv_ms = 0.0;
            stats->spmv_launches = 0;
            stats->reserved = 0;
        }
    });
}

template <class T>
static int batched_entry(idrs_matrix_t A, const double *b, int s, const idrs_opts_t *opts, double *x,
                         idrs_stats_t *stats) {
    return guarded([&] {
        need(A && b && x && stats, "null argument");
        const Partition *part = A->op.K->part.get();
        need(part != nullptr, "not a batch matrix (idrs_batch_format)");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        idrs_opts_t o{};
        if (opts) o = *opts;
        need(!o.has_M1 && !o.U0 && !o.resvec && !o.H, "M1, U0, resvec and H are not offered for batches");
        SolveOpts<T> so;
        if (o.has_tolerance) so.tol = o.tolerance;
        if (o.has_maximum_iterations) {
            so.has_maxit = true;
            so.maxit = o.maximum_iterations;
        }
        if (o.has_variant) so.variant = o.variant;
        so.omega = (const T *)o.omega;
        so.n_omega = o.omega ? o.n_omega : 0;
        so.p_seed = o.p_seed ? o.p_seed : 1;
        need(o.graph_mode >= 0 && o.graph_mode <= 2, "graph_mode must be 0, 1 or 2");
        so.graph_mode = o.graph_mode;
        const int nsys = part->nsys;
        const long ntot = part->h_row0[nsys];
        const int prow = (so.variant == 1) ? s : 1;
        std::vector<SolveStats> ss;
        double ms = 0.0;
        if (o.device_pointers) {
            so.x0 = (const T *)o.x0;
            so.PT_raw = (const T *)o.PT_raw;
            batch_idrs_solve<T>(A->op, s, (const T *)b, (T *)x, so, ss, ms, st);
        } else {
            DVec<T> db((const T *)b, ntot, st), dx(ntot), dx0, dP;
            if (o.x0) {
                dx0 = DVec<T>((const T *)o.x0, ntot, st);
                so.x0 = dx0.p;
            }
            if (o.PT_raw) {
                need(s >= 1 && s <= PDOT_MAX, "batched solves take s in 1..16");
                dP = DVec<T>((const T *)o.PT_raw, (size_t)ntot * prow, st);
                so.PT_raw = dP.p;
            }
            batch_idrs_solve<T>(A->op, s, db.p, dx.p, so, ss, ms, st);
            dx.download((T *)x, st);
        }
        for (int k = 0; k < nsys; k++) {
            stats[k].flag = ss[k].flag;
            stats[k].iterations = ss[k].iterations;
            stats[k].relres = ss[k].relres;
            stats[k].solve_ms = ms;
            stats[k].spmv_ms = 0.0;
            stats[k].spmv_launches = 0;
            stats[k].reserved = 0;
        }
    });
}

static int batched_mixed_entry(idrs_matrix_t A, const double *b, int s, const idrs_opts_t *opts, double inner_tol,
                               int max_outer, double *x, idrs_stats_t *stats, int *outer_steps) {
    return guarded([&] {
        need(A && b && x && stats && outer_steps, "null argument");
        const Partition *part = A->op.K->part.get();
        need(part != nullptr, "not a batch matrix (idrs_batch_format)");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        idrs_opts_t o{};
        if (opts) o = *opts;
        need(!o.has_M1 && !o.U0 && !o.resvec && !o.omega && !o.x0 && !o.PT_raw && !o.H,
             "M1, U0, omega, resvec, x0, PT_raw and H are not offered here");
        need(o.graph_mode >= 0 && o.graph_mode <= 2, "graph_mode must be 0, 1 or 2");
        SolveOpts<double> so;
        if (o.has_tolerance) so.tol = o.tolerance;
        if (o.has_maximum_iterations) {
            so.has_maxit = true;
            so.maxit = o.maximum_iterations;
        }
        if (o.has_variant) so.variant = o.variant;
        so.p_seed = o.p_seed ? o.p_seed : 1;
        so.graph_mode = o.graph_mode;
        const int nsys = part->nsys;
        const long ntot = part->h_row0[nsys];
        std::vector<MixedStats> ms;
        double solve_ms = 0.0;
        if (o.device_pointers) {
            batch_mixed_solve(A->op, s, b, x, so, inner_tol, max_outer, ms, solve_ms, st);
        } else {
            DVec<double> db(b, ntot, st), dx(ntot);
            batch_mixed_solve(A->op, s, db.p, dx.p, so, inner_tol, max_outer, ms, solve_ms, st);
            dx.download(x, st);
        }
        for (int k = 0; k < nsys; k++) {
            stats[k].flag = ms[k].flag;
            stats[k].iterations = ms[k].iterations;
            stats[k].relres = ms[k].relres;
            stats[k].solve_ms = solve_ms;
            stats[k].spmv_ms = 0.0;
            stats[k].spmv_launches = 0;
            stats[k].reserved = 0;
            outer_steps[k] = ms[k].outer;
        }
    });
}

template <class T>
static int mixed_entry(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double inner_tol,
                       int max_outer, int fp32_vectors, double *x, idrs_stats_t *stats, int *outer_steps) {
    return guarded([&] {
        need(A && b && x, "null argument");
        need(s >= 1 && s <= SMAX, "s must be in 1..64");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        idrs_opts_t o{};
        if (opts) o = *opts;
        need(!o.has_M1 && !o.U0 && !o.resvec && !o.omega && !o.H, "M1, U0, omega, resvec and H are not offered here");
        SolveOpts<T> so;
        if (o.has_tolerance) so.tol = o.tolerance;
        if (o.has_maximum_iterations) {
            so.has_maxit = true;
            so.maxit = o.maximum_iterations;
        }
        if (o.has_variant) so.variant = o.variant;
        so.p_seed = o.p_seed ? o.p_seed : 1;
        const int prow = so.variant == 1 ? s : 1;
        MixedStats ms;
        if (o.device_pointers) {
            so.x0 = (const T *)o.x0;
            so.PT_raw = (const T *)o.PT_raw;
            mixed_solve<T>(A->op, n, s, (const T *)b, (T *)x, so, inner_tol, max_outer, fp32_vectors != 0, ms, st);
        } else {
            DVec<T> db((const T *)b, n, st), dx(n), dx0, dP;
            if (o.x0) {
                dx0 = DVec<T>((const T *)o.x0, n, st);
                so.x0 = dx0.p;
            }
            if (o.PT_raw) {
                dP = DVec<T>((const T *)o.PT_raw, (size_t)n * prow, st);
                so.PT_raw = dP.p;
            }
            mixed_solve<T>(A->op, n, s, db.p, dx.p, so, inner_tol, max_outer, fp32_vectors != 0, ms, st);
            dx.download((T *)x, st);
        }
        if (stats) {
            stats->flag = ms.flag;
            stats->iterations = ms.iterations;
            stats->relres = ms.relres;
            stats->solve_ms = ms.solve_ms;
            stats->spmv_ms = 0.0;
            stats->spmv_launches = 0;
            stats->reserved = 0;
        }
        if (outer_steps) *outer_steps = ms.outer;
    });
}

template <class T>
static int matvec_entry(idrs_matrix_t A, const double *v, double *w, int what) {
    return guarded([&] {
        need(A && v && w, "null argument");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        long n = A->op.rows();
        DVec<T> dv((const T *)v, n, st), dw(n), w1(n), w2(n), w3(n);
        EpiArgs<T> none{};
        if (what == 0) op_apply<T>(A->op, dv.p, dw.p, w1.p, w2.p, EPI_NONE, none, nullptr, st);
        else if (what == 1) op_precon<T>(A->op, dv.p, dw.p, st);
        else op_scaleback<T>(A->op, dv.p, dw.p, w1.p, w2.p, w3.p, st);
        dw.download((T *)w, st);
    });
}

template <class T>
static int matvec_block_entry(idrs_matrix_t A, int ncols, const double *V, double *W) {
    return guarded([&] {
        need(A && V && W, "null argument");
        need(ncols >= 1, "ncols must be at least 1");
        need(A->op.K->nrows == A->op.K->ncols, "block products need a square matrix");
        const long rows = A->op.rows();
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        const size_t cnt = (size_t)rows * ncols;
        DVec<T> dv((const T *)V, cnt, st), dw(cnt), w1(cnt), w2(cnt);
        op_apply_cols<T>(A->op, ncols, rows, dv.p, dw.p, w1.p, w2.p, st);
        dw.download((T *)W, st);
    });
}

template <class T>
static int ms_scaleback_entry(idrs_matrix_t A, int ns, int degree, const double *alpha_s, const double *Y, double *X,
                              int device) {
    return guarded([&] {
        need(A && alpha_s && Y && X, "null argument");
        need(ns >= 1 && degree >= 0, "bad arguments");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        const long n = A->op.rows();
        // blocks of up to 8 shifts share each product's read of the matrix
        const int nb = std::min(ns, SPMV_COLS);
        DVec<T> dy(device ? 0 : n * nb), dx(device ? 0 : n * nb), w1(n * nb), w2(n * nb);
        std::vector<std::vector<std::complex<double>>> a;
        for (int k0 = 0; k0 < ns; k0 += nb) {
            const int m = std::min(nb, ns - k0);
            a.assign(m, std::vector<std::complex<double>>(degree + 1));
            for (int j = 0; j < m; j++)
                for (int i = 0; i <= degree; i++) {
                    const long q = 2 * ((long)(k0 + j) * (degree + 1) + i);
                    a[j][i] = {alpha_s[q], alpha_s[q + 1]};
                }
            const T *yk = (const T *)Y + (long)k0 * n;
            T *xk = (T *)X + (long)k0 * n;
            if (device) {
                op_scaleback_cols<T>(A->op, m, n, a, yk, xk, w1.p, w2.p, st);
            } else {
                HIP_TRY(hipMemcpyAsync(dy.p, yk, sizeof(T) * n * m, hipMemcpyHostToDevice, st));
                op_scaleback_cols<T>(A->op, m, n, a, dy.p, dx.p, w1.p, w2.p, st);
                HIP_TRY(hipMemcpyAsync(xk, dx.p, sizeof(T) * n * m, hipMemcpyDeviceToHost, st));
            }
        }
        HIP_TRY(hipStreamSynchronize(st));
    });
}

template <class T>
static int user_mv_entry(idrs_user_matrix_t K, const double *v, double *w) {
    return guarded([&] {
        need(K && v && w, "null argument");
        hipStream_t st = lib_stream();
        DVec<T> dv((const T *)v, K->M->ncols, st), dw(K->M->nrows);
        user_spmv<T>(*K->M, dv.p, dw.p, st);
        dw.download((T *)w, st);
    });
}

template <class T>
static int make_p_entry(int n, int s, int method, const double *PT_raw, const double *r, uint64_t seed, double *PT) {
    return guarded([&] {
        need(n >= 1 && s >= 1 && s <= SMAX && PT, "bad arguments");
        need(method >= 1 && method <= 4, "method must be 1..4");
        need(method != 2 || r, "method 2 needs r");
        hipStream_t st = lib_stream();
        int rows = method == 1 ? s : 1;
        DVec<T> draw, dr, dPT((size_t)n * rows);
        if (PT_raw) draw = DVec<T>((const T *)PT_raw, (size_t)n * rows, st);
        if (r) dr = DVec<T>((const T *)r, n, st);
        make_p_device<T>(n, s, method, PT_raw ? draw.p : nullptr, dr.p, dPT.p, seed ? seed : 1, st);
        dPT.download((T *)PT, st);
    });
}

static ValKind kind_of(const void *vals, bool cplx) {
    if (!vals) return ValKind::Pattern;
    return cplx ? ValKind::Complex : ValKind::Real;
}

}  // namespace idrs

using namespace idrs;

extern "C" {

const char *idrs_last_error(void) { return g_err.c_str(); }
int idrs_version(void) { return 100; }

int idrs_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

#define MAKE_UM(expr)                                             \
    guarded([&] {                                                 \
        need(A != nullptr, "null output handle");                 \
        *A = nullptr;                                             \
        DevMatrix *M = (expr);                                    \
        *A = new idrs_user_matrix_s{M};                           \
    })

int idrs_crs_format_f64(int nrows, int ncols, const int *row_ptr, const int *col_ind, const double *rval,
                        idrs_user_matrix_t *A) {
    return MAKE_UM((need(row_ptr && (col_ind || nrows == 0), "null index array"),
                    make_crs(nrows, ncols, row_ptr, col_ind, rval, kind_of(rval, false), 2)));
}
int idrs_crs_format_c128(int nrows, int ncols, const int *row_ptr, const int *col_ind, const double *cval,
                         idrs_user_matrix_t *A) {
    return MAKE_UM((need(row_ptr && cval, "null array"),
                    make_crs(nrows, ncols, row_ptr, col_ind, cval, ValKind::Complex, 2)));
}
int idrs_coo_format_f64(int nrows, int ncols, int nnz, const int *indx, const int *jndx, const double *rval,
                        idrs_user_matrix_t *A) {
    return MAKE_UM((need((indx && jndx) || nnz == 0, "null index array"),
                    make_coo(nrows, ncols, nnz, indx, jndx, rval, kind_of(rval, false))));
}
int idrs_coo_format_c128(int nrows, int ncols, int nnz, const int *indx, const int *jndx, const double *cval,
                         idrs_user_matrix_t *A) {
    return MAKE_UM((need(((indx && jndx) || nnz == 0) && cval, "null array"),
                    make_coo(nrows, ncols, nnz, indx, jndx, cval, ValKind::Complex)));
}
int idrs_mm_info(const char *path, int *rep, int *field, int *symm, int *nrows, int *ncols, int64_t *entries) {
    return guarded([&] {
        need(path != nullptr, "null path");
        const MMInfo in = mm_info(path);
        if (rep) *rep = in.rep;
        if (field) *field = in.field;
        if (symm) *symm = in.symm;
        if (nrows) *nrows = in.nrows;
        if (ncols) *ncols = in.ncols;
        if (entries) *entries = in.entries;
    });
}
int idrs_mm_matrix(const char *path, int format, idrs_user_matrix_t *A) {
    return MAKE_UM((need(path != nullptr, "null path"),
                    need(format == 2 || format == 3, "mm_matrix format must be 2 (crs) or 3 (coo)"), [&] {
                        const MMCoo c = mm_coo(path);
                        const ValKind k = c.info.field == MM_COMPLEX   ? ValKind::Complex
                                          : c.info.field == MM_PATTERN ? ValKind::Pattern
                                                                       : ValKind::Real;
                        DevMatrix *M = make_coo(c.info.nrows, c.info.ncols, (long)c.indx.size(), c.indx.data(),
                                                c.jndx.data(), c.val.data(), k);
                        M->orig_format = format;
                        return M;
                    }()));
}
int idrs_mm_dense(const char *path, double *vals) {
    return guarded([&] {
        need(path && vals, "null argument");
        MMInfo in;
        const std::vector<double> a = mm_dense(path, in);
        std::memcpy(vals, a.data(), sizeof(double) * a.size());
    });
}
int idrs_cds_format_f64(int nrows, int ncols, int half_band, const double *rband, idrs_user_matrix_t *A) {
    return MAKE_UM(make_cds(nrows, ncols, half_band, rband, kind_of(rband, false)));
}
int idrs_cds_format_c128(int nrows, int ncols, int half_band, const double *cband, idrs_user_matrix_t *A) {
    return MAKE_UM((need(cband != nullptr, "null band"), make_cds(nrows, ncols, half_band, cband, ValKind::Complex)));
}
int idrs_dia_format_f64(int nrows, int ncols, int ndiag, const int *offsets, const double *vals,
                        idrs_user_matrix_t *A) {
    return MAKE_UM((need(offsets != nullptr, "null offsets"),
                    make_dia(nrows, ncols, ndiag, offsets, vals, kind_of(vals, false))));
}
int idrs_dia_format_c128(int nrows, int ncols, int ndiag, const int *offsets, const double *vals,
                         idrs_user_matrix_t *A) {
    return MAKE_UM((need(offsets && vals, "null array"), make_dia(nrows, ncols, ndiag, offsets, vals, ValKind::Complex)));
}
#undef MAKE_UM

int idrs_user_matrix_destroy(idrs_user_matrix_t A) {
    return guarded([&] {
        if (!A) return;
        delete A->M;
        delete A;
    });
}

int idrs_user_matrix_autotune(idrs_user_matrix_t K, int reps, int *group) {
    return guarded([&] {
        need(K != nullptr, "null handle");
        need(reps >= 1, "reps must be positive");
        const int g = autotune_group(*K->M, reps, lib_stream(), nullptr);
        if (group) *group = g;
    });
}

int idrs_user_matrix_layout(idrs_user_matrix_t K, int *group, int *row_slabs) {
    return guarded([&] {
        need(K != nullptr, "null handle");
        const DevMatrix &M = *K->M;
        if (group) *group = M.fmt == DevFmt::CRS ? M.group : 0;
        if (row_slabs) *row_slabs = M.row_slabs;
    });
}
int idrs_user_matrix_set_row_slabs(idrs_user_matrix_t K, int row_slabs) {
    return guarded([&] {
        need(K != nullptr, "null handle");
        need(row_slabs == 4096 || row_slabs == 8192 || row_slabs == 16384 || row_slabs == 32768,
             "row_slabs must be 4096, 8192, 16384 or 32768");
        K->M->row_slabs = row_slabs;
    });
}
int idrs_user_matrix_set_group(idrs_user_matrix_t K, int group) {
    return guarded([&] {
        need(K != nullptr, "null handle");
        need(group == 1 || group == 2 || group == 4 || group == 8 || group == 16 || group == 32 || group == 64,
             "group must be 1, 2, 4, 8, 16, 32 or 64");
        DevMatrix &M = *K->M;
        if (M.fmt != DevFmt::CRS) return;
        if (group == 1 && !M.sell) build_sell(M, lib_stream());
        if (group != 1) M.sell.reset();
        M.group = group;
    });
}

int idrs_user_matrix_info(idrs_user_matrix_t A, int *nrows, int *ncols, int64_t *nnz, int *format) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        if (nrows) *nrows = A->M->nrows;
        if (ncols) *ncols = A->M->ncols;
        if (nnz) *nnz = A->M->nnz;
        if (format) *format = A->M->orig_format;
    });
}

int idrs_user_mv_f64(idrs_user_matrix_t K, const double *v, double *w) { return user_mv_entry<double>(K, v, w); }
int idrs_user_mv_c128(idrs_user_matrix_t K, const double *v, double *w) { return user_mv_entry<c128>(K, v, w); }

int idrs_real_diagonal(idrs_user_matrix_t K, double *D) {
    return guarded([&] {
        need(K && D, "null argument");
        need(K->M->kind != ValKind::Complex, "real_diagonal of a complex matrix");
        HIP_TRY(hipMemcpy(D, K->M->diag, sizeof(double) * K->M->nrows, hipMemcpyDeviceToHost));
    });
}
int idrs_complex_diagonal(idrs_user_matrix_t K, double *D) {
    return guarded([&] {
        need(K && D, "null argument");
        int n = K->M->nrows;
        if (K->M->kind == ValKind::Complex) {
            HIP_TRY(hipMemcpy(D, K->M->diag, sizeof(c128) * n, hipMemcpyDeviceToHost));
        } else {
            std::vector<double> r(n);
            HIP_TRY(hipMemcpy(r.data(), K->M->diag, sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int i = 0; i < n; i++) {
                D[2 * i] = r[i];
                D[2 * i + 1] = 0.0;
            }
        }
    });
}

int idrs_matrix_create(idrs_user_matrix_t K, idrs_matrix_t *A) {
    return guarded([&] {
        need(K && A, "null argument");
        need(K->M->nrows == K->M->ncols, "system matrix must be square");
        *A = new idrs_matrix_s;
        (*A)->op.K = K->M;
        (*A)->op.alpha.assign(1, std::complex<double>(1.0));
    });
}
int idrs_matrix_destroy(idrs_matrix_t A) {
    return guarded([&] { delete A; });
}
int idrs_matrix_set_precon_f64(idrs_matrix_t A, int which_precon, const double *shift) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        need(which_precon == 0 || which_precon == 1,
             "which_precon must be 0 (none) or 1 (diagonal); dense/band LU (2, 3): idrs_matrix_set_precon_lu");
        A->op.lu.reset();
        std::complex<double> sh = shift ? std::complex<double>(*shift, 0.0) : 0.0;
        op_set_diag_precon(A->op, which_precon, shift ? &sh : nullptr);
    });
}
int idrs_matrix_set_precon_c128(idrs_matrix_t A, int which_precon, const double *shift) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        need(which_precon == 0 || which_precon == 1,
             "which_precon must be 0 (none) or 1 (diagonal); dense/band LU (2, 3): idrs_matrix_set_precon_lu");
        A->op.lu.reset();
        std::complex<double> sh = shift ? std::complex<double>(shift[0], shift[1]) : 0.0;
        op_set_diag_precon(A->op, which_precon, shift ? &sh : nullptr);
    });
}
int idrs_matrix_stored_values(idrs_matrix_t A, int64_t *count, int *unit_diagonal) {
    return guarded([&] {
        need(A && count, "null argument");
        *count = A->op.stored_values();
        if (unit_diagonal) *unit_diagonal = (A->op.sval && A->op.unit_diag >= 0) ? 1 : 0;
    });
}
int idrs_matrix_set_seed_f64(idrs_matrix_t A, double seed) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        A->op.seed = std::complex<double>(seed, 0.0);
    });
}
int idrs_matrix_set_seed_c128(idrs_matrix_t A, const double *seed) {
    return guarded([&] {
        need(A && seed, "null argument");
        A->op.seed = std::complex<double>(seed[0], seed[1]);
    });
}

static int set_poly(idrs_matrix_t A, int preconditioner, int degree, const double *foci, const double *center,
                    const double *H, int ldh, int ncol, bool cplx) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        need(preconditioner >= 0 && preconditioner <= 3, "preconditioner must be 0..3");
        std::complex<double> f[2], c;
        std::vector<std::complex<double>> h;
        if (foci) {
            if (cplx) {
                f[0] = {foci[0], foci[1]};
                f[1] = {foci[2], foci[3]};
            } else {
                f[0] = foci[0];
                f[1] = foci[1];
            }
        }
        if (center) c = cplx ? std::complex<double>(center[0], center[1]) : std::complex<double>(center[0], 0.0);
        if (H) {
            need(ldh > 0 && ncol > 0, "bad H dimensions");
            h.resize((size_t)ldh * ncol);
            for (size_t q = 0; q < h.size(); q++)
                h[q] = cplx ? std::complex<double>(H[2 * q], H[2 * q + 1]) : std::complex<double>(H[q], 0.0);
        }
        if (preconditioner == 3) need(H != nullptr, "Hessenberg preconditioner needs H");
        op_set_poly(A->op, preconditioner, preconditioner ? degree : 0, foci ? f : nullptr,
                    center ? &c : nullptr, H ? h.data() : nullptr, ldh, ncol, cplx);
    });
}

int idrs_real_preconditioner(idrs_matrix_t A, int preconditioner, int degree, const double *foci,
                             const double *center, const double *H, int ldh, int ncol) {
    return set_poly(A, preconditioner, degree, foci, center, H, ldh, ncol, false);
}
int idrs_complex_preconditioner(idrs_matrix_t A, int preconditioner, int degree, const double *foci,
                                const double *center, const double *H, int ldh, int ncol) {
    return set_poly(A, preconditioner, degree, foci, center, H, ldh, ncol, true);
}
int idrs_matrix_set_alpha(idrs_matrix_t A, int degree, const double *alpha) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        need(degree >= 0 && (degree == 0 || alpha), "bad polynomial");
        A->op.degree = degree;
        A->op.poly_kind = degree ? POLY_HESSENBERG : POLY_NONE;
        A->op.alpha.assign((size_t)degree + 1, std::complex<double>(1.0));
        if (degree)
            for (int i = 0; i <= degree; i++) A->op.alpha[i] = {alpha[2 * i], alpha[2 * i + 1]};
    });
}

static int coeffs_entry(int kind, int degree, const double *foci, const double *center, const double *H, int ldh,
                        int ncol, int is_complex, double *alpha) {
    return guarded([&] {
        need(degree >= 0 && alpha, "bad arguments");
        std::complex<double> f[2], c;
        std::vector<std::complex<double>> h, a((size_t)degree + 1);
        if (foci) {
            f[0] = {foci[0], foci[1]};
            f[1] = {foci[2], foci[3]};
        }
        if (center) c = {center[0], center[1]};
        if (H) {
            need(ldh > 0 && ncol > 0, "bad H dimensions");
            h.resize((size_t)ldh * ncol);
            for (size_t q = 0; q < h.size(); q++) h[q] = {H[2 * q], H[2 * q + 1]};
        }
        if (degree == 0) a[0] = 1.0;
        else poly_coeffs(kind, degree, foci ? f : nullptr, center ? &c : nullptr, H ? h.data() : nullptr, ldh, ncol,
                         is_complex != 0, a.data());
        for (int i = 0; i <= degree; i++) {
            alpha[2 * i] = a[i].real();
            alpha[2 * i + 1] = is_complex ? a[i].imag() : 0.0;
        }
    });
}
int idrs_chebyshev_coeffs(int degree, const double *foci, int is_complex, double *alpha) {
    return coeffs_entry(POLY_CHEBYSHEV, degree, foci, nullptr, nullptr, 0, 0, is_complex, alpha);
}
int idrs_neumann_coeffs(int degree, const double *center, int is_complex, double *alpha) {
    return coeffs_entry(POLY_NEUMANN, degree, nullptr, center, nullptr, 0, 0, is_complex, alpha);
}
int idrs_compute_alpha_coeffs(int degree, const double *H, int ldh, int ncol, int is_complex, double *alpha) {
    return coeffs_entry(POLY_HESSENBERG, degree, nullptr, nullptr, H, ldh, ncol, is_complex, alpha);
}

int idrs_matrix_get_alpha(idrs_matrix_t A, double *alpha, int maxlen, int *degree) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        if (degree) *degree = A->op.degree;
        if (alpha)
            for (int i = 0; i <= A->op.degree && i < maxlen; i++) {
                alpha[2 * i] = A->op.alpha[i].real();
                alpha[2 * i + 1] = A->op.alpha[i].imag();
            }
    });
}

int idrs_matvec_f64(idrs_matrix_t A, const double *v, double *w) { return matvec_entry<double>(A, v, w, 0); }
int idrs_matvec_c128(idrs_matrix_t A, const double *v, double *w) { return matvec_entry<c128>(A, v, w, 0); }
int idrs_matvec_block_f64(idrs_matrix_t A, int ncols, const double *V, double *W) {
    return matvec_block_entry<double>(A, ncols, V, W);
}
int idrs_matvec_block_c128(idrs_matrix_t A, int ncols, const double *V, double *W) {
    return matvec_block_entry<c128>(A, ncols, V, W);
}
// ---------------------------------------------------------------------------
// generalized operators: mass and damping matrices, block systems,
// deflation (matrix_module.f90:169-264, 306-400, 1284-1623)
// ---------------------------------------------------------------------------
int idrs_matrix_set_mass(idrs_matrix_t A, idrs_user_matrix_t M) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        if (M) need(M->M->nrows == A->op.K->nrows && M->M->ncols == A->op.K->ncols, "M must have K's shape");
        A->op.Mm = M ? M->M : nullptr;
    });
}
int idrs_matrix_set_damping(idrs_matrix_t A, idrs_user_matrix_t C) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        if (C) need(C->M->nrows == A->op.K->nrows && C->M->ncols == A->op.K->ncols, "C must have K's shape");
        A->op.Cm = C ? C->M : nullptr;
    });
}
int idrs_matrix_set_flags(idrs_matrix_t A, int block_system, int multishift) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        if (block_system) need(A->op.ndefl == 0, "set the block flag before make_deflation");
        A->op.block = block_system != 0;
        A->op.multishift = multishift != 0;
    });
}
int idrs_matrix_rows(idrs_matrix_t A, int *rows) {
    return guarded([&] {
        need(A && rows, "null argument");
        *rows = (int)A->op.rows();
    });
}
int idrs_matrix_set_precon_diag(idrs_matrix_t A, idrs_user_matrix_t PK, idrs_user_matrix_t PM, idrs_user_matrix_t PC,
                                const double *shift, int is_complex) {
    return guarded([&] {
        need(A && PK, "null argument");
        const long n = A->op.K->nrows;
        for (idrs_user_matrix_t U : {PK, PM, PC})
            if (U) need(U->M->nrows == n, "preconditioner matrices must have K's rows");
        hipStream_t st = lib_stream();
        // real_diagonal / complex_diagonal of each matrix (zeros -> 1), host
        auto diag = [&](const DevMatrix &D) {
            std::vector<std::complex<double>> h(n);
            if (D.kind == ValKind::Complex) {
                std::vector<c128> c(n);
                HIP_TRY(hipMemcpyAsync(c.data(), D.diag, sizeof(c128) * n, hipMemcpyDeviceToHost, st));
                HIP_TRY(hipStreamSynchronize(st));
                for (long i = 0; i < n; i++) h[i] = {c[i].re, c[i].im};
            } else {
                std::vector<double> r(n);
                HIP_TRY(hipMemcpyAsync(r.data(), D.diag, sizeof(double) * n, hipMemcpyDeviceToHost, st));
                HIP_TRY(hipStreamSynchronize(st));
                for (long i = 0; i < n; i++) h[i] = r[i];
            }
            return h;
        };
        std::vector<std::complex<double>> d = diag(*PK->M);
        bool cdiag = is_complex != 0 || PK->M->kind == ValKind::Complex;
        if (shift) {
            const std::complex<double> s = is_complex ? std::complex<double>(shift[0], shift[1]) : shift[0];
            if (s.imag() != 0.0) cdiag = true;
            const std::complex<double> s2 = s * s;
            if (PM && PC) {                       // user_module.f90:1361-1364
                const auto dc = diag(*PC->M), dm = diag(*PM->M);
                for (long i = 0; i < n; i++) d[i] = (d[i] + s * dc[i]) + s2 * dm[i];
            } else if (PC) {                      // :1365-1369
                const auto dc = diag(*PC->M);
                for (long i = 0; i < n; i++) d[i] = d[i] + s * dc[i];
                for (long i = 0; i < n; i++) d[i] = d[i] + s2;
            } else if (PM) {                      // :1370-1372
                const auto dm = diag(*PM->M);
                for (long i = 0; i < n; i++) d[i] = d[i] - s * dm[i];
            } else {                              // :1373-1376
                for (long i = 0; i < n; i++) d[i] = d[i] - s;
            }
            for (idrs_user_matrix_t U : {PM, PC})
                if (U && U->M->kind == ValKind::Complex) cdiag = true;
        }
        if (!cdiag)
            for (long i = 0; i < n; i++) d[i] = d[i].real();
        A->op.lu.reset();
        op_set_diag_values(A->op, d, cdiag);
    });
}

int idrs_matrix_set_precon_lu(idrs_matrix_t A, int which, idrs_user_matrix_t PK, idrs_user_matrix_t PM,
                              idrs_user_matrix_t PC, const double *shift, int is_complex) {
    return guarded([&] {
        need(A && PK, "null argument");
        need(which == 2 || which == 3, "which_precon must be 2 (dense LU) or 3 (band LU)");
        need(PK->M->nrows == A->op.K->nrows, "preconditioner matrices must have K's rows");
        std::complex<double> s;
        if (shift) s = is_complex ? std::complex<double>(shift[0], shift[1]) : std::complex<double>(shift[0], 0.0);
        const bool cplx = is_complex != 0 || PK->M->kind == ValKind::Complex ||
                          (shift && PM && PM->M->kind == ValKind::Complex) ||
                          (shift && PC && PC->M->kind == ValKind::Complex);
        hipStream_t st = lib_stream();
        auto F = lu_build(which, *PK->M, PM ? PM->M : nullptr, PC ? PC->M : nullptr, shift ? &s : nullptr, cplx, st);
        A->op.lu = F;
        A->op.which_precon = which;
    });
}

int idrs_matrix_share_lu(idrs_matrix_t A, idrs_matrix_t B) {
    return guarded([&] {
        need(A && B, "null argument");
        need(B->op.lu != nullptr, "the source operator has no LU preconditioner");
        need(B->op.lu->n == A->op.K->nrows, "the LU preconditioner has another number of rows");
        A->op.lu = B->op.lu;
        A->op.which_precon = B->op.which_precon;
    });
}

int idrs_matrix_set_shift_invert(idrs_matrix_t A, int shift_invert, const double *shift, int is_complex) {
    return guarded([&] {
        need(A != nullptr, "null handle");
        need(!shift_invert || shift, "shift-and-invert needs the shift");
        A->op.shift_invert = shift_invert != 0;
        A->op.si_shift = !shift ? std::complex<double>(0.0, 0.0)
                                : is_complex ? std::complex<double>(shift[0], shift[1])
                                             : std::complex<double>(shift[0], 0.0);
    });
}

extern "C++" {
template <class T>
static int make_deflation_entry(idrs_matrix_t A, int nd, const double *V) {
    return guarded([&] {
        need(A && V, "null argument");
        need(nd >= 1 && nd <= MAX_DEFL, "deflation takes 1..16 vectors");
        hipStream_t st = lib_stream();
        const long n = A->op.rows();
        DVec<T> dv((const T *)V, (size_t)n * nd, st);
        gen_make_deflation<T>(A->op, dv.p, nd, st);
    });
}
}  // extern "C++"
int idrs_make_deflation_f64(idrs_matrix_t A, int nd, const double *V) { return make_deflation_entry<double>(A, nd, V); }
int idrs_make_deflation_c128(idrs_matrix_t A, int nd, const double *V) { return make_deflation_entry<c128>(A, nd, V); }

// what: 0 Pmul, 1 real_mv, 2 add_deflation(b, y), 3 scaleback(b, y)
extern "C++" {
template <class T>
static int general_entry(idrs_matrix_t A, int what, const double *b, const double *v, double *w) {
    return guarded([&] {
        need(A && v && w && (what < 2 || b), "null argument");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        gen_prepare(A->op);
        const long n = A->op.rows();
        DVec<T> dv((const T *)v, n, st), dw(n), db, t(n), u(n);
        if (b) db = DVec<T>((const T *)b, n, st);
        const bool gen = A->op.general();
        if (what == 0) {
            HIP_TRY(hipMemcpyAsync(dw.p, dv.p, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
            if (gen) gen_pmul<T>(A->op, dw.p, st);
        } else if (what == 1) {
            if (gen) {
                gen_mv<T>(A->op, dv.p, dw.p, t.p, u.p, st);
            } else {
                EpiArgs<T> none{};
                op_spmv<T>(A->op, dv.p, T(1), nullptr, T(0), dw.p, EPI_NONE, none, nullptr, st);
            }
        } else if (what == 2) {
            if (gen) gen_add_deflation<T>(A->op, db.p, dv.p, dw.p, st);
            else HIP_TRY(hipMemcpyAsync(dw.p, dv.p, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
        } else {
            if (gen) gen_scaleback<T>(A->op, db.p, dv.p, dw.p, st);
            else op_scaleback<T>(A->op, dv.p, dw.p, t.p, u.p, nullptr, st);
        }
        dw.download((T *)w, st);
    });
}
}  // extern "C++"
int idrs_pmul_f64(idrs_matrix_t A, const double *v, double *w) { return general_entry<double>(A, 0, nullptr, v, w); }
int idrs_pmul_c128(idrs_matrix_t A, const double *v, double *w) { return general_entry<c128>(A, 0, nullptr, v, w); }
int idrs_system_mv_f64(idrs_matrix_t A, const double *v, double *w) { return general_entry<double>(A, 1, nullptr, v, w); }
int idrs_system_mv_c128(idrs_matrix_t A, const double *v, double *w) { return general_entry<c128>(A, 1, nullptr, v, w); }
int idrs_add_deflation_f64(idrs_matrix_t A, const double *b, const double *y, double *x) {
    return general_entry<double>(A, 2, b, y, x);
}
int idrs_add_deflation_c128(idrs_matrix_t A, const double *b, const double *y, double *x) {
    return general_entry<c128>(A, 2, b, y, x);
}
int idrs_scaleback_b_f64(idrs_matrix_t A, const double *b, const double *y, double *x) {
    return general_entry<double>(A, 3, b, y, x);
}
int idrs_scaleback_b_c128(idrs_matrix_t A, const double *b, const double *y, double *x) {
    return general_entry<c128>(A, 3, b, y, x);
}

// ms scale-back with the deflation step (matrix_module.f90:724-806): per
// shift k, x = p_k(A) y_k, ms_add_deflation for sigma_k, x/A
extern "C++" {
template <class T>
static int ms_scaleback_b_entry(idrs_matrix_t A, int ns, const double *sigma, int degree, const double *alpha_s,
                                const double *b, const double *Y, double *X) {
    return guarded([&] {
        need(A && sigma && alpha_s && b && Y && X, "null argument");
        need(ns >= 1 && degree >= 0, "bad arguments");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        gen_prepare(A->op);
        const long n = A->op.rows();
        DVec<T> db((const T *)b, n, st), dy(n), dx(n), w1(n), w2(n);
        for (int k = 0; k < ns; k++) {
            std::vector<std::complex<double>> a(degree + 1);
            for (int i = 0; i <= degree; i++) {
                const long q = 2 * ((long)k * (degree + 1) + i);
                a[i] = {alpha_s[q], alpha_s[q + 1]};
            }
            const std::complex<double> sg = Tr<T>::is_complex ? std::complex<double>(sigma[2 * k], sigma[2 * k + 1])
                                                              : std::complex<double>(sigma[k], 0.0);
            HIP_TRY(hipMemcpyAsync(dy.p, (const T *)Y + (long)k * n, sizeof(T) * n, hipMemcpyHostToDevice, st));
            if (A->op.general()) gen_ms_scaleback_col<T>(A->op, a, sg, db.p, dy.p, dx.p, st);
            else op_scaleback_with<T>(A->op, a, dy.p, dx.p, w1.p, w2.p, st);
            HIP_TRY(hipMemcpyAsync((T *)X + (long)k * n, dx.p, sizeof(T) * n, hipMemcpyDeviceToHost, st));
        }
        HIP_TRY(hipStreamSynchronize(st));
    });
}
}  // extern "C++"
// X = add_deflation(A, sigma, b, Y) alone (rms_add_deflation /
// cms_add_deflation, matrix_module.f90:1508-1623); Y copied without deflation
extern "C++" {
template <class T>
static int ms_add_deflation_entry(idrs_matrix_t A, int ns, const double *sigma, const double *b, const double *Y,
                                  double *X) {
    return guarded([&] {
        need(A && sigma && b && Y && X, "null argument");
        need(ns >= 1, "bad arguments");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        gen_prepare(A->op);
        const long n = A->op.rows();
        std::vector<std::complex<double>> sg(ns);
        for (int k = 0; k < ns; k++)
            sg[k] = Tr<T>::is_complex ? std::complex<double>(sigma[2 * k], sigma[2 * k + 1])
                                      : std::complex<double>(sigma[k], 0.0);
        DVec<T> db((const T *)b, n, st), dy((const T *)Y, (size_t)n * ns, st), dx((size_t)n * ns);
        if (A->op.general()) gen_ms_add_deflation<T>(A->op, ns, sg.data(), db.p, dy.p, dx.p, st);
        else HIP_TRY(hipMemcpyAsync(dx.p, dy.p, sizeof(T) * n * ns, hipMemcpyDeviceToDevice, st));
        dx.download((T *)X, st);
    });
}
}  // extern "C++"
int idrs_ms_add_deflation_f64(idrs_matrix_t A, int n_sigma, const double *sigma, const double *b, const double *Y,
                              double *X) {
    return ms_add_deflation_entry<double>(A, n_sigma, sigma, b, Y, X);
}
int idrs_ms_add_deflation_c128(idrs_matrix_t A, int n_sigma, const double *sigma, const double *b, const double *Y,
                               double *X) {
    return ms_add_deflation_entry<c128>(A, n_sigma, sigma, b, Y, X);
}
int idrs_ms_scaleback_b_f64(idrs_matrix_t A, int n_sigma, const double *sigma, int degree, const double *alpha_s,
                            const double *b, const double *Y, double *X) {
    return ms_scaleback_b_entry<double>(A, n_sigma, sigma, degree, alpha_s, b, Y, X);
}
int idrs_ms_scaleback_b_c128(idrs_matrix_t A, int n_sigma, const double *sigma, int degree, const double *alpha_s,
                             const double *b, const double *Y, double *X) {
    return ms_scaleback_b_entry<c128>(A, n_sigma, sigma, degree, alpha_s, b, Y, X);
}

int idrs_precon_f64(idrs_matrix_t A, const double *v, double *w) { return matvec_entry<double>(A, v, w, 1); }
int idrs_precon_c128(idrs_matrix_t A, const double *v, double *w) { return matvec_entry<c128>(A, v, w, 1); }
int idrs_scaleback_f64(idrs_matrix_t A, const double *y, double *x) { return matvec_entry<double>(A, y, x, 2); }
int idrs_scaleback_c128(idrs_matrix_t A, const double *y, double *x) { return matvec_entry<c128>(A, y, x, 2); }

int idrs_ms_scaleback_f64(idrs_matrix_t A, int n_sigma, int degree, const double *alpha_s, const double *Y,
                          double *X, int device_pointers) {
    return ms_scaleback_entry<double>(A, n_sigma, degree, alpha_s, Y, X, device_pointers);
}
int idrs_ms_scaleback_c128(idrs_matrix_t A, int n_sigma, int degree, const double *alpha_s, const double *Y,
                           double *X, int device_pointers) {
    return ms_scaleback_entry<c128>(A, n_sigma, degree, alpha_s, Y, X, device_pointers);
}

int idrs_idrs_f64(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                  idrs_stats_t *stats) {
    return idrs_entry<double>(A, n, b, s, opts, x, stats);
}
int idrs_idrs_c128(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                   idrs_stats_t *stats) {
    return idrs_entry<c128>(A, n, b, s, opts, x, stats);
}
int idrs_idrs_ritz_f64(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                       idrs_stats_t *stats, double *omega, int n_omega, double *U0, int dim_u0, double *H, int h_rows,
                       int h_cols, int *nritz, int *converged) {
    return ritz_entry<double>(A, n, b, s, opts, x, stats, omega, n_omega, U0, dim_u0, H, h_rows, h_cols, nritz,
                              converged);
}
int idrs_idrs_ritz_c128(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                        idrs_stats_t *stats, double *omega, int n_omega, double *U0, int dim_u0, double *H,
                        int h_rows, int h_cols, int *nritz, int *converged) {
    return ritz_entry<c128>(A, n, b, s, opts, x, stats, omega, n_omega, U0, dim_u0, H, h_rows, h_cols, nritz,
                            converged);
}

// spectral routines on small dense matrices (host); complex values as
// (re, im) pairs
static std::vector<std::complex<double>> pairs_in(const double *v, size_t cnt) {
    std::vector<std::complex<double>> out(cnt);
    for (size_t q = 0; q < cnt; q++) out[q] = {v[2 * q], v[2 * q + 1]};
    return out;
}
static void pairs_out(const std::vector<std::complex<double>> &v, double *out) {
    for (size_t q = 0; q < v.size(); q++) {
        out[2 * q] = v[q].real();
        out[2 * q + 1] = v[q].imag();
    }
}
static std::vector<double> reals_in(const double *v, size_t cnt) {
    std::vector<double> out(cnt);
    for (size_t q = 0; q < cnt; q++) out[q] = v[2 * q];
    return out;
}

int idrs_qr_eig(int n, const double *H, int ldh, int is_complex, double *eig, double *U, int *converged) {
    return guarded([&] {
        need(n >= 1 && ldh >= n && H && eig, "bad arguments");
        std::vector<std::complex<double>> e, Uv;
        if (is_complex) {
            const auto h = pairs_in(H, (size_t)ldh * n);
            e = ritz::qr_eig(n, h.data(), ldh, U ? &Uv : nullptr, converged);
        } else {
            const auto h = reals_in(H, (size_t)ldh * n);
            e = ritz::qr_eig(n, h.data(), ldh, U ? &Uv : nullptr, converged);
        }
        pairs_out(e, eig);
        if (U) pairs_out(Uv, U);
    });
}

int idrs_filter_ritzval(int n, const double *ritzval, int *outlier) {
    return guarded([&] {
        need(n >= 1 && ritzval && outlier, "bad arguments");
        const auto o = ritz::filter_ritzval(pairs_in(ritzval, n));
        for (int i = 0; i < n; i++) outlier[i] = o[i] ? 1 : 0;
    });
}

int idrs_analyse_spectrum(int n, const double *H, int ldh, int is_complex, double *foci, double *center,
                          double *shift) {
    return guarded([&] {
        need(n >= 1 && ldh >= n && H, "bad arguments");
        if (is_complex) {
            const auto h = pairs_in(H, (size_t)ldh * n);
            std::complex<double> f[2], c, sh;
            ritz::analyse_spectrum(n, h.data(), ldh, f, &c, &sh);
            if (foci) pairs_out({f[0], f[1]}, foci);
            if (center) pairs_out({c}, center);
            if (shift) pairs_out({sh}, shift);
        } else {
            const auto h = reals_in(H, (size_t)ldh * n);
            double f[2], c, sh;
            ritz::analyse_spectrum(n, h.data(), ldh, f, &c, &sh);
            if (foci) pairs_out({f[0], f[1]}, foci);
            if (center) pairs_out({c}, center);
            if (shift) pairs_out({sh}, shift);
        }
    });
}

int idrs_si_power(int n, const double *H, int ldh, int is_complex, int s, const double *shift, uint64_t seed,
                  uint64_t offset, double *schurvec) {
    return guarded([&] {
        need(n >= 1 && ldh >= n && s >= 1 && s <= n && H && schurvec, "bad arguments");
        if (is_complex) {
            const auto h = pairs_in(H, (size_t)ldh * n);
            const std::complex<double> sh = shift ? std::complex<double>(shift[0], shift[1]) : 0.0;
            pairs_out(ritz::si_power(n, h.data(), ldh, s, sh, seed, offset), schurvec);
        } else {
            const auto h = reals_in(H, (size_t)ldh * n);
            const auto v = ritz::si_power(n, h.data(), ldh, s, shift ? shift[0] : 0.0, seed, offset);
            for (size_t q = 0; q < v.size(); q++) {
                schurvec[2 * q] = v[q];
                schurvec[2 * q + 1] = 0.0;
            }
        }
    });
}

int idrs_idrs_f64_noalloc(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts, double *x,
                          idrs_stats_t *stats) {
    return idrs_noalloc_entry<double>(A, n, b, s, opts, x, stats);
}

int idrs_qmridr_f64(idrs_matrix_t A, int n, const double *b, int s, const idrs_krylov_opts_t *opts, double *x,
                    idrs_stats_t *stats) {
    return krylov_entry<double>(Kry::QMRIDR, A, n, b, 1, nullptr, s, opts, x, nullptr, stats);
}
int idrs_qmridr_c128(idrs_matrix_t A, int n, const double *b, int s, const idrs_krylov_opts_t *opts, double *x,
                     idrs_stats_t *stats) {
    return krylov_entry<c128>(Kry::QMRIDR, A, n, b, 1, nullptr, s, opts, x, nullptr, stats);
}
int idrs_msidrs_f64(idrs_matrix_t A, int n, const double *b, int n_sigma, const double *sigma, int s,
                    const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats) {
    return krylov_entry<double>(Kry::MSIDRS, A, n, b, n_sigma, sigma, s, opts, X, relres, stats);
}
int idrs_msidrs_c128(idrs_matrix_t A, int n, const double *b, int n_sigma, const double *sigma, int s,
                     const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats) {
    return krylov_entry<c128>(Kry::MSIDRS, A, n, b, n_sigma, sigma, s, opts, X, relres, stats);
}
int idrs_msqmridr_f64(idrs_matrix_t A, int n, const double *b, int n_sigma, const double *sigma, int s,
                      const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats) {
    return krylov_entry<double>(Kry::MSQMRIDR, A, n, b, n_sigma, sigma, s, opts, X, relres, stats);
}
int idrs_msqmridr_c128(idrs_matrix_t A, int n, const double *b, int n_sigma, const double *sigma, int s,
                       const idrs_krylov_opts_t *opts, double *X, double *relres, idrs_stats_t *stats) {
    return krylov_entry<c128>(Kry::MSQMRIDR, A, n, b, n_sigma, sigma, s, opts, X, relres, stats);
}

int idrs_idrs_mixed_f64(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts,
                        double inner_tolerance, int max_outer, int fp32_vectors, double *x, idrs_stats_t *stats,
                        int *outer_steps) {
    return mixed_entry<double>(A, n, b, s, opts, inner_tolerance, max_outer, fp32_vectors, x, stats, outer_steps);
}
int idrs_idrs_mixed_c128(idrs_matrix_t A, int n, const double *b, int s, const idrs_opts_t *opts,
                         double inner_tolerance, int max_outer, int fp32_vectors, double *x, idrs_stats_t *stats,
                         int *outer_steps) {
    return mixed_entry<c128>(A, n, b, s, opts, inner_tolerance, max_outer, fp32_vectors, x, stats, outer_steps);
}

int idrs_ms_mixed_c128(idrs_matrix_t A, int method, int n, const double *b, int n_sigma, const double *sigma, int s,
                       const idrs_krylov_opts_t *opts, double inner_tolerance, int max_outer, double *X,
                       double *relres, idrs_stats_t *stats, int *outer_steps, int *corrections) {
    return guarded([&] {
        need(A && b && X && sigma && relres, "null argument");
        need(n_sigma >= 1, "sigma must hold at least one shift");
        need(A->op.rows() == n, "n does not match the matrix");
        hipStream_t st = lib_stream();
        op_prepare(A->op, st);
        idrs_krylov_opts_t o{};
        if (opts) o = *opts;
        KrylovOpts<c128> ko;
        ko.has_m1 = o.has_M1 != 0;
        if (o.has_tolerance) ko.tol = o.tolerance;
        if (o.has_maximum_iterations) {
            ko.has_maxit = true;
            ko.maxit = o.maximum_iterations;
        }
        if (o.has_variant) ko.variant = o.variant;
        ko.has_inner_s = o.has_inner_s != 0;
        ko.inner_s = o.inner_s;
        ko.has_inner_tol = o.has_inner_tolerance != 0;
        ko.inner_tol = o.inner_tolerance;
        ko.has_inner_it = o.has_inner_iterations != 0;
        ko.inner_it = o.inner_iterations;
        ko.omega = (const c128 *)o.omega;
        ko.resvec = o.resvec;
        ko.x0 = (const c128 *)o.x0;
        ko.want_colfac = o.colfac != nullptr;
        ko.p_seed = o.p_seed ? o.p_seed : 1;
        need(!o.PT_raw, "PT_raw is not offered by the complex64 multishift solve");
        MsMixedStats ms;
        if (o.device_pointers) {
            ms_mixed_solve(A->op, n, n_sigma, (const c128 *)sigma, s, (const c128 *)b, (c128 *)X, ko, method,
                           inner_tolerance, max_outer, relres, ms, st);
        } else {
            DVec<c128> db((const c128 *)b, n, st), dX((size_t)n * n_sigma);
            ms_mixed_solve(A->op, n, n_sigma, (const c128 *)sigma, s, db.p, dX.p, ko, method, inner_tolerance,
                           max_outer, relres, ms, st);
            dX.download((c128 *)X, st);
        }
        if (stats) {
            double mx = 0.0;
            for (int k = 0; k < n_sigma; k++) mx = std::max(mx, relres[k]);
            stats->flag = ms.flag;
            stats->iterations = ms.iterations;
            stats->relres = mx;
            stats->solve_ms = ms.solve_ms;
            stats->spmv_ms = 0.0;
            stats->spmv_launches = 0;
            stats->reserved = 0;
        }
        if (outer_steps) *outer_steps = ms.outer;
        if (corrections) *corrections = ms.corrections;
    });
}

int idrs_crs_transpose(idrs_user_matrix_t K, idrs_user_matrix_t *KT) {
    return guarded([&] {
        need(K && KT, "null argument");
        *KT = nullptr;
        std::unique_ptr<DevMatrix> T(make_transpose(*K->M));
        *KT = new idrs_user_matrix_s{T.release()};
    });
}
int idrs_crs_transpose_refresh(idrs_user_matrix_t K, idrs_user_matrix_t KT) {
    return guarded([&] {
        need(K && KT, "null argument");
        hipStream_t st = lib_stream();
        transpose_refresh(*K->M, *KT->M, st);
        HIP_TRY(hipStreamSynchronize(st));
    });
}
int idrs_user_matrix_set_values_f64(idrs_user_matrix_t K, const double *vals, int device_pointer) {
    return guarded([&] {
        need(K && (vals || K->M->nnz == 0), "null argument");
        need(K->M->kind != ValKind::Complex, "complex matrix: real values given");
        hipStream_t st = lib_stream();
        set_values(*K->M, vals, device_pointer != 0, st);
        HIP_TRY(hipStreamSynchronize(st));
    });
}
int idrs_crs_pattern_grad_f64(idrs_user_matrix_t K, const double *lambda, const double *x, double *grad,
                              int device_pointers) {
    return guarded([&] {
        need(K && lambda && x && (grad || K->M->nnz == 0), "null argument");
        hipStream_t st = lib_stream();
        const DevMatrix &M = *K->M;
        if (device_pointers) {
            pattern_grad<double>(M, lambda, x, grad, st);
            HIP_TRY(hipStreamSynchronize(st));
        } else {
            DVec<double> dl(lambda, M.nrows, st), dx(x, M.ncols, st), dg((size_t)M.nnz);
            pattern_grad<double>(M, dl.p, dx.p, dg.p, st);
            dg.download(grad, st);
        }
    });
}

int idrs_batch_format(int nsys, const idrs_user_matrix_t *K, idrs_user_matrix_t *A) {
    return guarded([&] {
        need(K && A && nsys >= 1, "bad arguments");
        std::vector<const DevMatrix *> ms(nsys);
        for (int k = 0; k < nsys; k++) {
            need(K[k] != nullptr, "null system matrix");
            ms[k] = K[k]->M;
        }
        std::unique_ptr<DevMatrix> M(make_blockdiag(nsys, ms.data()));
        *A = new idrs_user_matrix_s{M.release()};
    });
}
int idrs_batch_info(idrs_user_matrix_t A, int *nsys, int64_t *row0) {
    return guarded([&] {
        need(A && nsys, "null argument");
        const Partition *p = A->M->part.get();
        *nsys = p ? p->nsys : 0;
        if (p && row0)
            for (int k = 0; k <= p->nsys; k++) row0[k] = p->h_row0[k];
    });
}
int idrs_idrs_batched_mixed_f64(idrs_matrix_t A, const double *b, int s, const idrs_opts_t *opts,
                                double inner_tolerance, int max_outer, double *x, idrs_stats_t *stats,
                                int *outer_steps) {
    return batched_mixed_entry(A, b, s, opts, inner_tolerance, max_outer, x, stats, outer_steps);
}
int idrs_idrs_batched_f64(idrs_matrix_t A, const double *b, int s, const idrs_opts_t *opts, double *x,
                          idrs_stats_t *stats) {
    return batched_entry<double>(A, b, s, opts, x, stats);
}
int idrs_idrs_batched_c128(idrs_matrix_t A, const double *b, int s, const idrs_opts_t *opts, double *x,
                           idrs_stats_t *stats) {
    return batched_entry<c128>(A, b, s, opts, x, stats);
}

int idrs_make_p_f64(int n, int s, int method, const double *PT_raw, const double *r, uint64_t seed, double *PT) {
    return make_p_entry<double>(n, s, method, PT_raw, r, seed, PT);
}
int idrs_make_p_c128(int n, int s, int method, const double *PT_raw, const double *r, uint64_t seed, double *PT) {
    return make_p_entry<c128>(n, s, method, PT_raw, r, seed, PT);
}

}  // extern "C"
