// IDR(s) solver (RIDRS/CIDRS) on the device.
#pragma once

#include "matrix.hpp"

#include <stdexcept>

namespace idrs {

constexpr int SMAX = 64;             // largest shadow-space dimension s
// Largest number of P rows a one-pass P^T y kernel (product epilogue, the
// update kernels) holds in registers; wider shadow spaces take their P^T y
// in chunks of PDOT_MAX rows from a separate kernel (one grid row per chunk).
constexpr int PDOT_MAX = 16;
constexpr int MAX_PARTS = 32768;     // upper bound of blocks in any reduction grid
constexpr int MAX_NV = SMAX + 3;     // values reduced per block
static_assert(MAX_RED >= SMAX + 4, "one-workgroup reductions hold s + 4 values");

struct SolveStats {
    int flag = -1;
    double relres = 1.0;
    int iterations = 0;
    double solve_ms = 0.0;       // hipEvent time of the whole solve
    double spmv_ms = 0.0;        // summed SpMV launch time (when timing enabled)
    int spmv_launches = 0;
};

// Device memory carved out of one region the caller owns (the
// allocation-free solve): 256-byte aligned pieces taken in order, never
// freed one by one; asking for more than the region holds throws.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t bytes) {
        const size_t at = (used + 255) & ~(size_t)255;
        if (at + bytes > cap) throw std::invalid_argument(
            "the solve needs more workspace than the scratch region holds (a smaller n or s fits)");
        used = at + bytes;
        return base + at;
    }
};

// Solver options; pointers are DEVICE pointers unless noted.
template <class T>
struct SolveOpts {
    bool has_m1 = false;         // present(M1)
    const Operator *M1 = nullptr;  // v/M1 uses M1's diagonal if it has one
    double tol = (double)1e-6f;  // reference default is the single-precision literal 1e-6
    bool has_maxit = false;      // present(maximum_iterations); default min(2n, 1000)
    int maxit = 0;
    int variant = 1;
    const T *x0 = nullptr;       // device
    const T *U0 = nullptr;       // device, n*s column-major
    const T *omega = nullptr;    // HOST array of user omegas
    int n_omega = 0;
    double *resvec = nullptr;    // HOST output, length >= maxit+1
    int resvec_len = 0;
    const T *PT_raw = nullptr;   // device, raw shadow space before orthonormalisation
    unsigned long long p_seed = 1;   // default stream seed (reference: this_image() = 1)
    int graph_mode = 0;              // 0: HIP graph per cycle when n <= 2^21, 1: always, 2: never
    // Hessenberg matrix H(nritz+1, nritz) of the inner steps (idrs.inc:81-90),
    // HOST output, column-major; null: not wanted
    T *H = nullptr;
    int nritz = 0;
    // set: every device buffer of the solve comes from this region, and the
    // solve makes no allocation of any kind (no graph, no events, no pinned
    // host memory: the host polls the flag with a blocking copy per cycle)
    Arena *arena = nullptr;
};

template <class T>
int idrs_solve(const Operator &A, long n, int s, const T *b, T *x, const SolveOpts<T> &o, SolveStats &st,
               hipStream_t stream);

// The outputs of IDRS_RITZ (ritz_module.f90:225-556) beside the solution.
template <class T>
struct RitzRequest {
    bool has_M = false;            // present(M): IDRS's preconditioner and v/M in compute_U0
    const Operator *M = nullptr;   // its operator; null: v/M = v (a default type(matrix))
    T *omega = nullptr;            // HOST output omega(n_omega)
    int n_omega = 0;
    T *U0 = nullptr;               // DEVICE output U0(n, dimU0), orthonormal columns
    int dimU0 = 0;
    T *H = nullptr;                // HOST output H(h_rows, h_cols) = Hf(1:h_rows, 1:h_cols)
    int h_rows = 0, h_cols = 0;
    unsigned long long seed = 1;   // stream of the shift-and-invert power method's start block,
    unsigned long long offset = 0; // from this counter on
    int converged = 0;             // output: Ritz values converged for the omegas
};

// x = IDRS_RITZ(A, b, s, ...): the solve with H = Hf(maxit+1, maxit), then
// the requested outputs; returns nritz.  o.H / o.nritz are set here.
template <class T>
int idrs_ritz_solve(const Operator &A, long n, int s, const T *b, T *x, SolveOpts<T> o, RitzRequest<T> &rq,
                    SolveStats &st, hipStream_t stream);

// Standalone P generation (MAKE_P) for tests: PT(s,n) orthonormalised, device.
// caching allocators shared by the solvers (idrs.hip): device buffers and
// pinned host buffers that solves of equal size reuse
void *pool_alloc(size_t bytes);
void pool_free(void *p, size_t bytes);
void *pinned_take(size_t bytes, unsigned flags);
void pinned_give(void *p, size_t bytes, unsigned flags);

template <class T>
void make_p_device(long n, int s, int method, const T *PT_raw, const T *r, T *PT, unsigned long long seed,
                   hipStream_t stream);

}  // namespace idrs
