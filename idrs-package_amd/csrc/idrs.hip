// IDR(s) with bi-orthogonalisation (RIDRS / CIDRS) on one GPU.
//
// Restates MODULES/IDRS/idrs.inc:1-311 and the helpers of
// MODULES/IDRS/idrs_module.f90 (COMP_OM :719-785, NORM :787-811,
// MAKE_P :839-937, P_DOT :939-993); paths relative to the reference's IDRS/.
//
// Execution model: the whole iteration state (M, f, gamma, alpha, beta,
// omega, norms, counters, convergence flag) lives in device memory.  Each
// reference statement group maps onto either a streaming kernel over the N
// rows or a one-workgroup "scalar" kernel that reduces the block partials
// in a fixed order and runs the s x s logic.  The host enqueues whole cycles
// (s inner steps + the omega step) without waiting; every kernel returns
// immediately once the device flag says the solve has finished, and the
// host polls that flag one cycle behind, so there is no per-iteration host
// synchronisation.  Per inner step k the kernels are
//   SpMV chain of A*U(:,k) -> G(:,k), last SpMV with the P^T G(:,k) epilogue
//   fin_mu      : mu -> alpha (forward substitution), M(k:s,k), beta_k,
//                 f update, gamma for step k+1
//   orth_update : G(:,k), U(:,k) bi-orthogonalised; r -= beta G(:,k);
//                 x += beta U(:,k); ||r||^2 partials; U(:,k+1) for step k+1
//   fin_norm    : ||r||, iteration count, convergence test
// and the omega step is: SpMV chain of A*r with the (r,r),(t,t),(t,r)
// epilogue, fin_om, omega_update (r, x, ||r||^2 and the next P^T r), fin_cycle.
// Shadow spaces wider than PDOT_MAX rows take their P^T sums from a chunked
// kernel (pdot.hpp) instead of the product epilogue and the update kernel.
#include "idrs.hpp"
#include "idrs_state.hpp"
#include "pdot.hpp"

#include <cfloat>
#include <chrono>
#include <map>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace idrs {

// ---------------------------------------------------------------------------
// small caching allocators (solves of equal size reuse workspaces; a free
// would wait for the device): device buffers by (device, size), pinned
// host buffers by (size, flags).  Buffers stay allocated until the process
// ends.
// ---------------------------------------------------------------------------
namespace {
std::mutex g_pool_mu;
std::multimap<std::pair<int, size_t>, void *> g_pool;
std::multimap<std::pair<size_t, unsigned>, void *> g_pinned;
size_t pool_round(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    return bytes ? bytes : 256;
}
}  // namespace

void *pool_alloc(size_t bytes) {
    bytes = pool_round(bytes);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> g(g_pool_mu);
        auto it = g_pool.find({dev, bytes});
        if (it != g_pool.end()) {
            void *p = it->second;
            g_pool.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    HIP_TRY(hipMalloc(&p, bytes));
    return p;
}
void pool_free(void *p, size_t bytes) {
    if (!p) return;
    int dev = 0;
    if (hipPointerAttribute_t at{}; hipPointerGetAttributes(&at, p) == hipSuccess) dev = at.device;
    std::lock_guard<std::mutex> g(g_pool_mu);
    g_pool.emplace(std::make_pair(dev, pool_round(bytes)), p);
}
void *pinned_take(size_t bytes, unsigned flags) {
    bytes = pool_round(bytes);
    {
        std::lock_guard<std::mutex> g(g_pool_mu);
        auto it = g_pinned.find({bytes, flags});
        if (it != g_pinned.end()) {
            void *p = it->second;
            g_pinned.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    HIP_TRY(hipHostMalloc(&p, bytes, flags));
    return p;
}
void pinned_give(void *p, size_t bytes, unsigned flags) {
    if (!p) return;
    std::lock_guard<std::mutex> g(g_pool_mu);
    g_pinned.emplace(std::make_pair(pool_round(bytes), flags), p);
}

namespace {
// a device buffer from the caching pool, or (with an arena) a piece of the
// caller's region that is not freed
struct Buf {
    void *p = nullptr;
    size_t bytes = 0;
    bool pooled = false;
    Buf() = default;
    explicit Buf(size_t b, Arena *ar = nullptr) : bytes(b), pooled(ar == nullptr) {
        p = ar ? ar->take(b) : pool_alloc(b);
    }
    Buf(const Buf &) = delete;
    Buf &operator=(const Buf &) = delete;
    Buf(Buf &&o) noexcept : p(o.p), bytes(o.bytes), pooled(o.pooled) { o.p = nullptr; }
    Buf &operator=(Buf &&o) noexcept {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(pooled, o.pooled);
        return *this;
    }
    ~Buf() {
        if (pooled) pool_free(p, bytes);
    }
    template <class V> V *as() const { return (V *)p; }
};
}  // namespace

#define GRID_LOOP(q, n) for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < (n); q += (long)gridDim.x * blockDim.x)

// ---------------------------------------------------------------------------
// setup kernels
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_init(State<T> *st, int s, int method, int prow, int conjP, int maxit, double kappa, double tol,
                       int user_omega, int n_omega, int inispace, int rv_cap, T *H, int nritz) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        scalar_init(st, s, method, prow, conjP, maxit, kappa, tol, user_omega, n_omega, inispace, rv_cap);
        st->H = H;
        st->nritz = nritz;
    }
}

// partials[block] = sum |v|^2 over the block's rows
template <class T>
__global__ __launch_bounds__(256) void k_norm_partials(long n, const T *__restrict__ v, T *partials) {
    __shared__ T lds[4];
    T e[1] = {T(0)};
    GRID_LOOP(q, n) e[0] = e[0] + T(Tr<T>::abs2(v[q]));
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

// r = b - Ax (Ax may be null: x = 0), partials = |r|^2
template <class T>
__global__ __launch_bounds__(256) void k_residual(long n, const T *__restrict__ b, const T *__restrict__ ax,
                                                  T *__restrict__ r, T *partials) {
    __shared__ T lds[4];
    T e[1] = {T(0)};
    GRID_LOOP(q, n) {
        T rq = ax ? b[q] - ax[q] : b[q];
        r[q] = rq;
        e[0] = e[0] + T(Tr<T>::abs2(rq));
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_normb(State<T> *st, const T *partials, int nparts) {
    __shared__ T red[1];
    reduce_partials(partials, nparts, 1, red);
    if (threadIdx.x == 0) {
        st->normb = sqrt(sre(red[0]));
        st->tolb = st->tol * st->normb;                                    // idrs.inc:94
    }
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_init(State<T> *st, const T *partials, int nparts, double *resvec) {
    __shared__ T red[1];
    reduce_partials(partials, nparts, 1, red);
    if (threadIdx.x == 0) {
        double nr = sqrt(sre(red[0]));
        st->normr = nr;
        if (resvec) resvec[0] = nr;                                        // idrs.inc:97
        if (nr <= st->tolb) st->info = 0;                                  // idrs.inc:101
    }
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_final(State<T> *st, const T *partials, int nparts) {
    __shared__ T red[1];
    reduce_partials(partials, nparts, 1, red);
    if (threadIdx.x == 0) {
        double nr = sqrt(sre(red[0]));                                     // idrs.inc:304-307
        st->normr = nr;
        if (st->info == 0 && nr > st->tolb) st->info = 2;
    }
}

// ---------------------------------------------------------------------------
// MAKE_P (idrs_module.f90:839-937)
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_fill_uniform(long cnt, T *PT, unsigned long long seed) {
    GRID_LOOP(q, cnt) {
        if constexpr (Tr<T>::is_complex)
            PT[q] = T(uniform01(seed, (uint64_t)q), uniform01(seed, (uint64_t)(cnt + q)));
        else
            PT[q] = uniform01(seed, (uint64_t)q);
    }
}

template <class T>
__global__ void k_copy_conj(long n, const T *__restrict__ r, T *__restrict__ PT) {
    GRID_LOOP(q, n) PT[q] = Tr<T>::conj(r[q]);
}

// PT(j,:) / rho, the normalisation of a finished row
template <class T>
__device__ inline T gs_scaled(T p, double rho) {
    if constexpr (Tr<T>::is_complex)
        return T(p.re / rho, p.im / rho);
    else
        return p / rho;
}

// alpha(k) = dot_product(PT(j,:),PT(k,:)) [real] / dot_product(PT(k,:),PT(j,:)) [complex], k < j.
// SC: row j-1 is first divided by st->rho and stored (k_gs_scale's step in
// the same pass; the dot with it uses the stored value)
template <class T, int NV, bool SC>
__global__ __launch_bounds__(256) void k_gs_dots(long n, int s, int j, const State<T> *st, T *PT, T *partials,
                                                 int nparts) {
    __shared__ T lds[NV * 4];
    const double rho = SC ? st->rho : 1.0;
    T e[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) e[k] = T(0);
    GRID_LOOP(q, n) {
        const T pj = PT[j + (long)s * q];
        T pl = T(0);
        if constexpr (SC) {
            pl = gs_scaled(PT[j - 1 + (long)s * q], rho);
            PT[j - 1 + (long)s * q] = pl;
        }
#pragma unroll
        for (int k = 0; k < NV; k++)
            if (k < j) e[k] = fmaT(Tr<T>::conj((SC && k == j - 1) ? pl : PT[k + (long)s * q]), pj, e[k]);
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) partials[(long)k * nparts + blockIdx.x] = e[k];
}

// row 0 of a given P: PT = PT_raw (every row) and partials = |PT(0,:)|^2,
// the copy and k_gs_update_norm's j = 0 pass in one
template <class T>
__global__ __launch_bounds__(256) void k_gs_first(long n, int s, const T *PT_raw, T *PT, T *partials) {
    __shared__ T lds[4];
    T e[1] = {T(0)};
    GRID_LOOP(q, n) {
        for (int i = 1; i < s; i++) PT[i + (long)s * q] = PT_raw[i + (long)s * q];
        const T p = PT_raw[(long)s * q];
        PT[(long)s * q] = p;
        e[0] = e[0] + T(Tr<T>::abs2(p));
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

// k_gs_dots for rows j above PDOT_MAX: grid row y sums the coefficients
// k in [y*PDOT_MAX, min(j, (y+1)*PDOT_MAX)) with k_gs_dots' arithmetic
template <class T>
__global__ __launch_bounds__(256) void k_gs_dots_wide(long n, int s, int j, const T *__restrict__ PT, T *partials,
                                                      int nparts) {
    constexpr int NV = PDOT_MAX;
    __shared__ T lds[NV * 4];
    const int k0 = blockIdx.y * NV, cnt = min(NV, j - k0);
    T e[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) e[k] = T(0);
    GRID_LOOP(q, n) {
        const T pj = PT[j + (long)s * q];
#pragma unroll
        for (int k = 0; k < NV; k++)
            if (k < cnt) e[k] = fmaT(Tr<T>::conj(PT[k0 + k + (long)s * q]), pj, e[k]);
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < cnt; k++) partials[(long)(k0 + k) * nparts + blockIdx.x] = e[k];
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_gs(State<T> *st, const T *partials, int nparts, int nv) {
    __shared__ T red[SMAX];
    reduce_partials(partials, nparts, nv, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < nv; k++) st->gs[k] = red[k];
}

// PT(j,:) = PT(j,:) - alpha(k)*PT(k,:) for k < j; partials = |PT(j,:)|^2
template <class T>
__global__ __launch_bounds__(256) void k_gs_update_norm(long n, int s, int j, const State<T> *st, T *PT,
                                                        T *partials) {
    __shared__ T lds[4], sa[SMAX];
    if (threadIdx.x < SMAX) sa[threadIdx.x] = st->gs[threadIdx.x];
    __syncthreads();
    T e[1] = {T(0)};
    GRID_LOOP(q, n) {
        T p = PT[j + (long)s * q];
        for (int k = 0; k < j; k++) p = fmaT(-sa[k], PT[k + (long)s * q], p);
        PT[j + (long)s * q] = p;
        e[0] = e[0] + T(Tr<T>::abs2(p));
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_rho(State<T> *st, const T *partials, int nparts) {
    __shared__ T red[1];
    reduce_partials(partials, nparts, 1, red);
    if (threadIdx.x == 0) st->rho = sqrt(sre(red[0]));
}

template <class T>
__global__ void k_gs_scale(long n, int s, int j, const State<T> *st, T *PT) {
    const double rho = st->rho;
    GRID_LOOP(q, n) PT[j + (long)s * q] = gs_scaled(PT[j + (long)s * q], rho);
}

// ---------------------------------------------------------------------------
// P_DOT of r at setup: partials of PT*r (no conjugation, or conjugated PT
// for the one-row variants 3/4)
// ---------------------------------------------------------------------------
template <class T, int NV>
__global__ __launch_bounds__(256) void k_pdot(long n, const T *__restrict__ w, const T *__restrict__ P, int prow,
                                              int conjP, T *partials, int nparts) {
    __shared__ T lds[NV * 4];
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    GRID_LOOP(q, n) {
        const T wq = w[q];
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < prow) {
                T p = ldnt(&P[q * prow + i]);
                if (conjP) p = Tr<T>::conj(p);
                e[i] = fmaT(p, wq, e[i]);
            }
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; i++) partials[(long)i * nparts + blockIdx.x] = e[i];
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_f(State<T> *st, const T *partials, int nparts) {
    __shared__ T red[SMAX];
    const int prow = st->prow;
    reduce_partials(partials, nparts, prow, red);
    if (threadIdx.x == 0)
        for (int i = 0; i < st->s; i++) st->f[i] = i < prow ? red[i] : T(0);
}

// ---------------------------------------------------------------------------
// one-block reduction + scalar logic kernels
// ---------------------------------------------------------------------------
// First level of a two-level reduction of many block partials: block b sums
// its contiguous share of every value's partials ([v][nparts] in, [v][PRE]
// out), in a fixed order, so the one-workgroup kernel after it reads PRE.
constexpr int PRE = 64;
template <class T>
__global__ __launch_bounds__(256) void k_prereduce(const State<T> *st, const T *partials, int nparts, int nv,
                                                   T *out) {
    if (st->info >= 0) return;
    __shared__ T lds[4];
    const long p0 = (long)nparts * blockIdx.x / PRE, p1 = (long)nparts * (blockIdx.x + 1) / PRE;
    for (int v = 0; v < nv; v++) {
        T e[1] = {T(0)};
        for (long p = p0 + threadIdx.x; p < p1; p += blockDim.x) e[0] = e[0] + partials[(long)v * nparts + p];
        block_sum<T, 1>(e, lds);
        if (threadIdx.x == 0) out[(long)v * PRE + blockIdx.x] = e[0];
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_mu(int k, State<T> *st, const T *partials, int nparts) {
    if (st->info >= 0) return;
    __shared__ T red[SMAX];
    reduce_partials(partials, nparts, st->prow, red);
    if (threadIdx.x == 0) scalar_mu(st, k, red);
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_norm(State<T> *st, const T *partials, int nparts, double *resvec) {
    if (st->info >= 0) return;
    __shared__ T red[1];
    reduce_partials(partials, nparts, 1, red);
    if (threadIdx.x == 0) scalar_norm(st, sqrt(sre(red[0])), resvec);
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_om(State<T> *st, const T *partials, int nparts, const T *omega) {
    if (st->info >= 0) return;
    __shared__ T red[3];
    reduce_partials(partials, nparts, 3, red);
    if (threadIdx.x == 0) scalar_om(st, red, omega);
}

template <class T>
__global__ __launch_bounds__(FIN_THREADS) void k_fin_cycle(State<T> *st, const T *partials, int nparts, double *resvec) {
    if (st->info >= 0) return;
    __shared__ T red[SMAX + 1];
    reduce_partials(partials, nparts, 1 + st->prow, red);
    if (threadIdx.x == 0) scalar_cycle(st, red, resvec);
}

// ---------------------------------------------------------------------------
// streaming kernels of the iteration
// ---------------------------------------------------------------------------
template <class T>
__device__ inline T precon_div(T v, const T *m1, long q) {
    return m1 ? sdiv(v, m1[q]) : v;
}

// U(:,k) for step k (idrs.inc:155-190)
template <class T>
__global__ __launch_bounds__(256) void k_form_u(long n, int k, const State<T> *st, const T *__restrict__ r,
                                                const T *__restrict__ G, T *__restrict__ U, const T *__restrict__ m1) {
    if (st->info >= 0) return;
    __shared__ T sg[SMAX];
    const int s = st->s, jj = st->jj;
    if (threadIdx.x < SMAX) sg[threadIdx.x] = st->gamma[threadIdx.x];
    __syncthreads();
    const T om = st->om;
    if (jj > 0) {
        GRID_LOOP(q, n) {
            T v = r[q];
            for (int i = k; i < s; i++) v = fmaT(-sg[i], ldnt(&G[(long)i * n + q]), v);
            T t = om * precon_div(v, m1, q);
            for (int i = k; i < s; i++) t = fmaT(sg[i], ldnt(&U[(long)i * n + q]), t);
            U[(long)k * n + q] = t;
        }
    } else if (!st->inispace) {
        GRID_LOOP(q, n) U[(long)k * n + q] = precon_div(r[q], m1, q);
    }
}

// bi-orthogonalisation of G(:,k), U(:,k); r, x update; ||r||^2; and U(:,k+1)
// for the next inner step (idrs.inc:199-224, 155-190)
template <class T>
__global__ __launch_bounds__(256) void k_orth_update(long n, int k, const State<T> *st, T *__restrict__ G,
                                                     T *__restrict__ U, T *__restrict__ r, T *__restrict__ x,
                                                     const T *__restrict__ m1, T *partials) {
    if (st->info >= 0) return;
    __shared__ T sa[SMAX], sg[SMAX], lds[4];
    const int s = st->s, jj = st->jj, ini = st->inispace;
    if (threadIdx.x < SMAX) {
        sa[threadIdx.x] = st->alpha[threadIdx.x];
        sg[threadIdx.x] = st->gamma[threadIdx.x];
    }
    __syncthreads();
    const T beta = st->beta[k], om = st->om;
    const bool next = (k + 1 < s);
    T e[1] = {T(0)};
    GRID_LOOP(q, n) {
        T gk = G[(long)k * n + q], uk = U[(long)k * n + q];
        for (int i = 0; i < k; i++) {
            gk = fmaT(-sa[i], ldnt(&G[(long)i * n + q]), gk);
            uk = fmaT(-sa[i], ldnt(&U[(long)i * n + q]), uk);
        }
        stnt(&G[(long)k * n + q], gk);   // read again only after other streams have passed
        stnt(&U[(long)k * n + q], uk);
        const T rq = fmaT(-beta, gk, r[q]);
        stnt(&x[q], fmaT(beta, uk, ldnt(&x[q])));   // x is read again only by the next update
        r[q] = rq;
        e[0] = e[0] + T(Tr<T>::abs2(rq));
        if (next) {
            if (jj > 0) {
                T v = rq;
                for (int i = k + 1; i < s; i++) v = fmaT(-sg[i], ldnt(&G[(long)i * n + q]), v);
                T t = om * precon_div(v, m1, q);
                for (int i = k + 1; i < s; i++) t = fmaT(sg[i], ldnt(&U[(long)i * n + q]), t);
                U[(long)(k + 1) * n + q] = t;
            } else if (!ini) {
                U[(long)(k + 1) * n + q] = precon_div(rq, m1, q);
            }
        }
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

// v = r/M1 before the omega step's product (idrs.inc:266-270)
template <class T>
__global__ __launch_bounds__(256) void k_form_v(long n, const State<T> *st, const T *__restrict__ r,
                                                const T *__restrict__ m1, T *__restrict__ v) {
    if (st->info >= 0) return;
    GRID_LOOP(q, n) v[q] = sdiv(r[q], m1[q]);
}

// r = r - om t; x = x + om v; ||r||^2 and P^T r partials (idrs.inc:287-292, 147)
template <class T, int NV>
__global__ __launch_bounds__(256) void k_omega_update(long n, const State<T> *st, const T *v, const T *__restrict__ t,
                                                      T *r, T *__restrict__ x, const T *__restrict__ P, int prow,
                                                      int conjP, T *partials, int nparts) {
    if (st->info >= 0) return;
    __shared__ T lds[(NV + 1) * 4];
    const T om = st->om;
    T e[NV + 1];
#pragma unroll
    for (int i = 0; i <= NV; i++) e[i] = T(0);
    GRID_LOOP(q, n) {
        const T vq = v[q];
        const T rq = fmaT(-om, t[q], r[q]);
        stnt(&x[q], fmaT(om, vq, ldnt(&x[q])));
        r[q] = rq;
        e[0] = e[0] + T(Tr<T>::abs2(rq));
        if constexpr (std::is_same<T, double>::value && NV % 2 == 0) {
            if (prow == NV) {   // whole, even-length P row: 16-byte loads
#pragma unroll
                for (int i = 0; i < NV; i += 2) {
                    const dbl2_nt p2 = __builtin_nontemporal_load((const dbl2_nt *)(P + q * NV + i));
                    e[1 + i] = fmaT(p2.x, rq, e[1 + i]);
                    e[2 + i] = fmaT(p2.y, rq, e[2 + i]);
                }
                continue;
            }
        }
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < prow) {
                T p = ldnt(&P[q * prow + i]);
                if (conjP) p = Tr<T>::conj(p);
                e[1 + i] = fmaT(p, rq, e[1 + i]);
            }
    }
    block_sum<T, NV + 1>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i <= NV; i++) partials[(long)i * nparts + blockIdx.x] = e[i];
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
template <class T>
static void launch_pdot(long n, const T *w, const T *P, int prow, int conjP, T *partials, hipStream_t st) {
    int g = vec_grid(n);
    if (prow > PDOT_MAX) return launch_pdot_wide<T>(n, w, P, prow, conjP, partials, 0, nullptr, st);
    if (prow <= 1) hipLaunchKernelGGL((k_pdot<T, 1>), dim3(g), dim3(256), 0, st, n, w, P, prow, conjP, partials, g);
    else if (prow <= 2) hipLaunchKernelGGL((k_pdot<T, 2>), dim3(g), dim3(256), 0, st, n, w, P, prow, conjP, partials, g);
    else if (prow <= 4) hipLaunchKernelGGL((k_pdot<T, 4>), dim3(g), dim3(256), 0, st, n, w, P, prow, conjP, partials, g);
    else if (prow <= 8) hipLaunchKernelGGL((k_pdot<T, 8>), dim3(g), dim3(256), 0, st, n, w, P, prow, conjP, partials, g);
    else hipLaunchKernelGGL((k_pdot<T, 16>), dim3(g), dim3(256), 0, st, n, w, P, prow, conjP, partials, g);
    LAUNCH_CHECK();
}

template <class T>
static void launch_omega_update(long n, const State<T> *st, const T *v, const T *t, T *r, T *x, const T *P, int prow,
                                int conjP, T *partials, hipStream_t stream) {
    int g = vec_grid(n);
    if (prow > PDOT_MAX) {
        // r, x and ||r||^2 here, P^T r from the chunked kernel (values 1..prow)
        hipLaunchKernelGGL((k_omega_update<T, 1>), dim3(g), dim3(256), 0, stream, n, st, v, t, r, x, P, 0, conjP,
                           partials, g);
        LAUNCH_CHECK();
        return launch_pdot_wide<T>(n, r, P, prow, conjP, partials, 1, &st->info, stream);
    }
#define OU(NV) hipLaunchKernelGGL((k_omega_update<T, NV>), dim3(g), dim3(256), 0, stream, n, st, v, t, r, x, P, prow, conjP, partials, g)
    if (prow <= 1) OU(1);
    else if (prow <= 2) OU(2);
    else if (prow <= 4) OU(4);
    else if (prow <= 8) OU(8);
    else OU(16);
#undef OU
    LAUNCH_CHECK();
}

// the dots of row j with rows 0..j-1; with scale_prev row j-1 is
// normalised in the same pass (returns false where that is not offered:
// the caller scales it first)
template <class T>
static bool launch_gs_dots(long n, int s, int j, const State<T> *sta, T *PT, T *partials, bool scale_prev,
                           hipStream_t st) {
    int g = vec_grid(n);
    if (j > PDOT_MAX) {
        if (scale_prev) return false;
        hipLaunchKernelGGL((k_gs_dots_wide<T>), dim3(g, ceil_div(j, PDOT_MAX)), dim3(256), 0, st, n, s, j, PT,
                           partials, g);
        LAUNCH_CHECK();
        return true;
    }
#define GD(NV)                                                                                                  \
    do {                                                                                                        \
        if (scale_prev)                                                                                         \
            hipLaunchKernelGGL((k_gs_dots<T, NV, true>), dim3(g), dim3(256), 0, st, n, s, j, sta, PT, partials, g); \
        else                                                                                                    \
            hipLaunchKernelGGL((k_gs_dots<T, NV, false>), dim3(g), dim3(256), 0, st, n, s, j, sta, PT, partials, g); \
    } while (0)
    if (j <= 1) GD(1);
    else if (j <= 2) GD(2);
    else if (j <= 4) GD(4);
    else if (j <= 8) GD(8);
    else GD(16);
#undef GD
    LAUNCH_CHECK();
    return true;
}

// MAKE_P body on the device, using 'st' for the Gram-Schmidt coefficients.
template <class T>
static void make_p_impl(long n, int s, int method, const T *PT_raw, const T *r, T *PT, unsigned long long seed,
                        State<T> *st, T *partials, hipStream_t stream) {
    const int g = vec_grid(n);
    const int rows = method == 1 ? s : 1;
    if (method == 2) {
        hipLaunchKernelGGL((k_copy_conj<T>), dim3(g), dim3(256), 0, stream, n, r, PT);
        LAUNCH_CHECK();
        return;
    }
    const long cnt = (long)rows * n;
    if (!PT_raw) {
        hipLaunchKernelGGL((k_fill_uniform<T>), dim3(vec_grid(cnt)), dim3(256), 0, stream, cnt, PT, seed);
        LAUNCH_CHECK();
    }
    // classical Gram-Schmidt over the rows, then normalise (one row for
    // 3/4).  The copy of a given P rides on row 0's norm pass, and each
    // row's normalisation on the next row's dot pass (the last row's on a
    // pass of its own): the same operations in the same order as separate
    // passes.
    for (int j = 0; j < rows; j++) {
        if (j > 0) {
            if (!launch_gs_dots<T>(n, rows, j, st, PT, partials, true, stream)) {
                hipLaunchKernelGGL((k_gs_scale<T>), dim3(g), dim3(256), 0, stream, n, rows, j - 1, st, PT);
                LAUNCH_CHECK();
                launch_gs_dots<T>(n, rows, j, st, PT, partials, false, stream);
            }
            hipLaunchKernelGGL((k_fin_gs<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, partials, g, j);
            LAUNCH_CHECK();
        }
        if (j == 0 && PT_raw)
            hipLaunchKernelGGL((k_gs_first<T>), dim3(g), dim3(256), 0, stream, n, rows, PT_raw, PT, partials);
        else
            hipLaunchKernelGGL((k_gs_update_norm<T>), dim3(g), dim3(256), 0, stream, n, rows, j, st, PT, partials);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_fin_rho<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, partials, g);
        LAUNCH_CHECK();
    }
    hipLaunchKernelGGL((k_gs_scale<T>), dim3(g), dim3(256), 0, stream, n, rows, rows - 1, st, PT);
    LAUNCH_CHECK();
}

template <class T>
void make_p_device(long n, int s, int method, const T *PT_raw, const T *r, T *PT, unsigned long long seed,
                   hipStream_t stream) {
    Buf stb(sizeof(State<T>)), part(sizeof(T) * MAX_NV * MAX_PARTS);
    HIP_TRY(hipMemsetAsync(stb.p, 0, sizeof(State<T>), stream));
    make_p_impl<T>(n, s, method, PT_raw, r, PT, seed, stb.as<State<T>>(), part.as<T>(), stream);
    HIP_TRY(hipStreamSynchronize(stream));
}

template <class T>
static T *upload_diag_as(const Operator *M1, long n, Buf &buf, hipStream_t stream, Arena *ar);

template <class T> __global__ void k_diag_to_T(long n, const double *d, T *out) { GRID_LOOP(q, n) out[q] = T(d[q]); }
__global__ void k_cdiag_to_c(long n, const c128 *d, c128 *out) { GRID_LOOP(q, n) out[q] = d[q]; }

template <class T>
static T *upload_diag_as(const Operator *M1, long n, Buf &buf, hipStream_t stream, Arena *ar) {
    if (!M1 || M1->which_precon != 1 || !M1->pdiag) return nullptr;
    if (M1->K->nrows != n) throw std::invalid_argument("M1 size does not match the system");
    buf = Buf(sizeof(T) * n, ar);
    int g = vec_grid(n);
    if (M1->pdiag_kind == ValKind::Complex) {
        if constexpr (std::is_same<T, c128>::value)
            hipLaunchKernelGGL(k_cdiag_to_c, dim3(g), dim3(256), 0, stream, n, (const c128 *)M1->pdiag, buf.as<c128>());
        else
            throw std::invalid_argument("complex preconditioner for a real solve");
    } else {
        hipLaunchKernelGGL((k_diag_to_T<T>), dim3(g), dim3(256), 0, stream, n, (const double *)M1->pdiag, buf.as<T>());
    }
    LAUNCH_CHECK();
    return buf.as<T>();
}

// Pinned status words and events of a solve's host loop, kept for reuse:
// hipHostMalloc and hipHostFree cost tens of microseconds each and the free
// waits for the device.  A pool, so that nested or concurrent solves take
// sets of their own; a set stays with the device it was made on.
namespace {
struct LoopRes {
    int dev = -1;
    int *h_info = nullptr;
    hipEvent_t pe[2] = {nullptr, nullptr}, e0 = nullptr, e1 = nullptr;
};
std::mutex g_loop_mu;
std::vector<LoopRes> g_loop_res;

LoopRes loop_res_take() {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> g(g_loop_mu);
        for (size_t i = 0; i < g_loop_res.size(); i++)
            if (g_loop_res[i].dev == dev) {
                LoopRes r = g_loop_res[i];
                g_loop_res.erase(g_loop_res.begin() + (long)i);
                return r;
            }
    }
    LoopRes r;
    r.dev = dev;
    HIP_TRY(hipHostMalloc((void **)&r.h_info, 2 * sizeof(int), hipHostMallocDefault));
    HIP_TRY(hipEventCreateWithFlags(&r.pe[0], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&r.pe[1], hipEventDisableTiming));
    HIP_TRY(hipEventCreate(&r.e0));
    HIP_TRY(hipEventCreate(&r.e1));
    return r;
}
void loop_res_give(const LoopRes &r) {
    std::lock_guard<std::mutex> g(g_loop_mu);
    g_loop_res.push_back(r);
}
// gives its set back to the pool on every way out of the solve
struct LoopResHold {
    LoopRes r;
    bool held = false;
    void take() {
        r = loop_res_take();
        held = true;
    }
    ~LoopResHold() {
        if (held) loop_res_give(r);
    }
};
}  // namespace

template <class T>
int idrs_solve(const Operator &A, long n, int s, const T *b, T *x, const SolveOpts<T> &o, SolveStats &out,
               hipStream_t stream) {
    if (s < 1 || s > SMAX) throw std::invalid_argument("s must be in 1..64");
    if (n < 1) throw std::invalid_argument("empty system");
    if (A.rows() != n || A.K->nrows != A.K->ncols) throw std::invalid_argument("matrix is not n x n");
    if (o.arena && A.general()) throw std::invalid_argument("the allocation-free solve takes standard operators only");
    if (o.has_m1 && o.M1 && o.M1->general())
        throw std::invalid_argument("M1 with mass, damping, block or deflation is not offered (v/M1 divides by a diagonal)");
    if (o.tol < 0) throw std::invalid_argument("Illegal value for parameter tolerance");
    const int method = o.variant;
    if (method < 1 || method > 4) throw std::invalid_argument("variant must be 1..4");
    const int prow = method == 1 ? s : 1;
    const int conjP = (method >= 3 && Tr<T>::is_complex) ? 1 : 0;
    const double kappa = (method == 1 || method == 3) ? (double)0.7f : 0.0;   // idrs.inc:112-124
    const int maxit = o.has_maxit ? o.maxit : (int)std::min<long>(2 * n, 1000);   // idrs.inc:37-38
    // the reference never stops on iter == maxit when maxit <= 0; bound the
    // host loop anyway
    const long cap_it = maxit > 0 ? maxit : 10 * n + 1000;
    if (o.resvec && maxit + 1 > o.resvec_len)
        throw std::invalid_argument("Length of vector with residual norms too small, should be maxit+1");
    const int user_omega = (o.omega && o.n_omega >= 1) ? 1 : 0;
    // H(nritz+1, nritz) wanted only for nritz >= 1 (idrs.inc:82-84)
    const bool out_h = o.H != nullptr && o.nritz > 0;

    Arena *ar = o.arena;
    if (ar && out_h) throw std::invalid_argument("the Hessenberg output is not offered by the allocation-free solve");
    SpmvTimer &tm = spmv_timer();
    if (ar && tm.enabled) throw std::invalid_argument("SpMV timing needs events: not in the allocation-free solve");
    const int gv = vec_grid(n);
    const int gs = spmv_grid(*A.K);
    const size_t vb = sizeof(T) * (size_t)n;
    // block partials [value][block]: every value of the widest kernel (its
    // NV template, the chunked P^T y above PDOT_MAX rows) for the largest grid
    const size_t part_cnt = ar ? (size_t)(std::max(s, PDOT_MAX) + 4) * (size_t)std::max(gv, gs)
                               : (size_t)MAX_NV * MAX_PARTS;
    const bool has_m1_diag = o.has_m1 && o.M1 && o.M1->which_precon == 1 && o.M1->pdiag;
    if (ar) {
        // the same pieces in the same order on a copy of the arena: a region
        // too small is reported before anything runs
        Arena sim{nullptr, ar->cap, ar->used};
        for (size_t b : {vb * s, vb * s, vb, vb, vb, vb, vb, vb * prow, sizeof(T) * part_cnt, sizeof(State<T>),
                         sizeof(double) * (cap_it + 2), sizeof(T) * MAX_NV * PRE})
            sim.take(b);
        if (has_m1_diag) {
            sim.take(sizeof(T) * n);
            sim.take(vb);
        }
        if (user_omega) sim.take(sizeof(T) * o.n_omega);
    }
    Buf bG(vb * s, ar), bU(vb * s, ar), br(vb, ar), bt(vb, ar), bw1(vb, ar), bw2(vb, ar), btmp(vb, ar), bP(vb * prow, ar);
    Buf bpart(sizeof(T) * part_cnt, ar), bst(sizeof(State<T>), ar), bres(sizeof(double) * (cap_it + 2), ar);
    Buf bpart2(sizeof(T) * MAX_NV * PRE, ar);
    Buf bv, bm1, bom, bh;
    T *G = bG.as<T>(), *U = bU.as<T>(), *r = br.as<T>(), *t = bt.as<T>(), *w1 = bw1.as<T>(), *w2 = bw2.as<T>();
    T *tmp = btmp.as<T>(), *P = bP.as<T>(), *part = bpart.as<T>();
    State<T> *st = bst.as<State<T>>();
    double *resvec = bres.as<double>();
    const T *m1 = o.has_m1 ? upload_diag_as<T>(o.M1, n, bm1, stream, ar) : nullptr;
    T *v = r;
    if (m1) {
        bv = Buf(vb, ar);
        v = bv.as<T>();
    }
    T *omega_d = nullptr;
    if (user_omega) {
        bom = Buf(sizeof(T) * o.n_omega, ar);
        omega_d = bom.as<T>();
        HIP_TRY(hipMemcpyAsync(omega_d, o.omega, sizeof(T) * o.n_omega, hipMemcpyHostToDevice, stream));
    }

    // solve time: HIP events, or (allocation-free) the host clock around a
    // synchronised solve
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::chrono::steady_clock::time_point t0;
    LoopResHold lres;
    if (ar) {
        HIP_TRY(hipStreamSynchronize(stream));
        t0 = std::chrono::steady_clock::now();
    } else {
        lres.take();
        e0 = lres.r.e0;
        e1 = lres.r.e1;
        HIP_TRY(hipEventRecord(e0, stream));
    }
    tm.reset();

    const size_t h_bytes = out_h ? sizeof(T) * (size_t)(o.nritz + 1) * o.nritz : 0;
    T *hd = nullptr;
    if (out_h) {
        bh = Buf(h_bytes, ar);
        hd = bh.as<T>();
        HIP_TRY(hipMemsetAsync(hd, 0, h_bytes, stream));   // H = 0 (idrs.inc:89)
    }
    hipLaunchKernelGGL((k_init<T>), dim3(1), dim3(64), 0, stream, st, s, method, prow, conjP, maxit, kappa, o.tol,
                       user_omega, user_omega ? o.n_omega : 0, o.U0 ? 1 : 0, (int)(cap_it + 2), hd, o.nritz);
    LAUNCH_CHECK();
    // x = 0 or x0; U = 0 or U0 (idrs.inc:52-65)
    if (o.x0) HIP_TRY(hipMemcpyAsync(x, o.x0, vb, hipMemcpyDeviceToDevice, stream));
    else HIP_TRY(hipMemsetAsync(x, 0, vb, stream));
    if (o.U0) HIP_TRY(hipMemcpyAsync(U, o.U0, vb * s, hipMemcpyDeviceToDevice, stream));
    else HIP_TRY(hipMemsetAsync(U, 0, vb * s, stream));
    // normb, tolb; r = b - A x; normr (idrs.inc:93-106)
    hipLaunchKernelGGL((k_norm_partials<T>), dim3(gv), dim3(256), 0, stream, n, b, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_fin_normb<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, part, gv);
    LAUNCH_CHECK();
    EpiArgs<T> none{};
    if (o.x0) op_apply<T>(A, x, tmp, w1, w2, EPI_NONE, none, nullptr, stream);
    hipLaunchKernelGGL((k_residual<T>), dim3(gv), dim3(256), 0, stream, n, b, o.x0 ? tmp : nullptr, r, part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_fin_init<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, part, gv, resvec);
    LAUNCH_CHECK();
    // P (idrs.inc:109-126) and f = P^T r for the first cycle (idrs.inc:147)
    make_p_impl<T>(n, s, method, o.PT_raw, r, P, o.p_seed, st, part, stream);
    launch_pdot<T>(n, r, P, prow, conjP, part, stream);
    hipLaunchKernelGGL((k_fin_f<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, part, gv);
    LAUNCH_CHECK();

    EpiArgs<T> epd{};
    epd.P = P;
    epd.s = s;
    epd.prow = prow;
    epd.conjP = conjP;
    epd.partials = part;
    EpiArgs<T> eom{};
    eom.r = r;
    eom.partials = part;
    const int *info = &st->info;
    // the product epilogues' partials (one per row slab): above 4096 slabs a
    // PRE-block first level runs before the one-workgroup scalar kernel
    T *part2 = bpart2.as<T>();
    const bool pre = gs > 4096;
    const T *epart = pre ? part2 : part;
    const int enparts = pre ? PRE : gs;
    auto prereduce = [&](int nv) {
        if (!pre) return;
        hipLaunchKernelGGL((k_prereduce<T>), dim3(PRE), dim3(256), 0, stream, st, part, gs, nv, part2);
        LAUNCH_CHECK();
    };

    auto enqueue_cycle = [&]() {
        hipLaunchKernelGGL((k_form_u<T>), dim3(gv), dim3(256), 0, stream, n, 0, st, r, G, U, m1);
        LAUNCH_CHECK();
        for (int k = 0; k < s; k++) {
            if (prow > PDOT_MAX) {
                // wider than the product epilogue holds: P^T G(:,k) in chunks
                op_apply<T>(A, U + (long)k * n, G + (long)k * n, w1, w2, EPI_NONE, none, info, stream);
                launch_pdot_wide<T>(n, G + (long)k * n, P, prow, conjP, part, 0, info, stream);
                hipLaunchKernelGGL((k_fin_mu<T>), dim3(1), dim3(FIN_THREADS), 0, stream, k, st, part, gv);
                LAUNCH_CHECK();
            } else {
                op_apply<T>(A, U + (long)k * n, G + (long)k * n, w1, w2, EPI_PDOT, epd, info, stream);
                prereduce(prow);
                hipLaunchKernelGGL((k_fin_mu<T>), dim3(1), dim3(FIN_THREADS), 0, stream, k, st, epart, enparts);
                LAUNCH_CHECK();
            }
            hipLaunchKernelGGL((k_orth_update<T>), dim3(gv), dim3(256), 0, stream, n, k, st, G, U, r, x, m1, part);
            LAUNCH_CHECK();
            hipLaunchKernelGGL((k_fin_norm<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, part, gv, resvec);
            LAUNCH_CHECK();
        }
        if (m1) {
            hipLaunchKernelGGL((k_form_v<T>), dim3(gv), dim3(256), 0, stream, n, st, r, m1, v);
            LAUNCH_CHECK();
        }
        op_apply<T>(A, v, t, w1, w2, EPI_OMEGA, eom, info, stream);
        prereduce(3);
        hipLaunchKernelGGL((k_fin_om<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, epart, enparts, omega_d);
        LAUNCH_CHECK();
        launch_omega_update<T>(n, st, v, t, r, x, P, prow, conjP, part, stream);
        hipLaunchKernelGGL((k_fin_cycle<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, part, gv, resvec);
        LAUNCH_CHECK();
    };

    // enqueue cycles, polling the device flag one cycle behind (pinned
    // flag, events); the allocation-free solve instead reads the flag after
    // every cycle with a blocking copy into host memory it already has
    int h_local[2] = {-1, -1};
    int *h_info = ar ? h_local : lres.r.h_info;
    hipEvent_t pe[2] = {lres.r.pe[0], lres.r.pe[1]};
    long max_cycles = cap_it / (s + 1) + 2;
    // Small systems are bound by the host's launch rate (~55 launches per
    // cycle): capture one cycle in a HIP graph and replay it -- the same
    // kernels with the same arguments in the same order.  Not while timing
    // single SpMVs (events between launches), nor in the allocation-free
    // solve (instantiating a graph allocates).
    const bool use_graph = !ar && !tm.enabled && (o.graph_mode == 1 || (o.graph_mode == 0 && n <= (1L << 21)));
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    if (use_graph) {
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        try {
            enqueue_cycle();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIP_TRY(hipStreamEndCapture(stream, &graph));
        HIP_TRY(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }
    bool done = false;
    try {
        for (long c = 0; c < max_cycles && !done; c++) {
            if (use_graph) HIP_TRY(hipGraphLaunch(gexec, stream));
            else enqueue_cycle();
            if (ar) {
                HIP_TRY(hipMemcpyAsync(&h_info[0], info, sizeof(int), hipMemcpyDeviceToHost, stream));
                HIP_TRY(hipStreamSynchronize(stream));
                if (h_info[0] >= 0) done = true;
                continue;
            }
            HIP_TRY(hipMemcpyAsync(&h_info[c % 2], info, sizeof(int), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipEventRecord(pe[c % 2], stream));
            if (c >= 1) {
                HIP_TRY(hipEventSynchronize(pe[(c - 1) % 2]));
                if (h_info[(c - 1) % 2] >= 0) done = true;
            }
        }
        // final true residual (idrs.inc:304-310)
        op_apply<T>(A, x, tmp, w1, w2, EPI_NONE, none, nullptr, stream);
        hipLaunchKernelGGL((k_residual<T>), dim3(gv), dim3(256), 0, stream, n, b, tmp, r, part);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_fin_final<T>), dim3(1), dim3(FIN_THREADS), 0, stream, st, part, gv);
        LAUNCH_CHECK();
        if (!ar) HIP_TRY(hipEventRecord(e1, stream));
        State<T> hs;
        HIP_TRY(hipMemcpyAsync(&hs, st, sizeof(State<T>), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        float ms = 0.f;
        if (ar) ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        else HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        if (o.resvec) {
            long cnt = std::min<long>(std::min<long>(o.resvec_len, hs.iter + 1), cap_it + 2);
            HIP_TRY(hipMemcpy(o.resvec, resvec, sizeof(double) * cnt, hipMemcpyDeviceToHost));
        }
        if (out_h) HIP_TRY(hipMemcpy(o.H, hd, h_bytes, hipMemcpyDeviceToHost));
        out.flag = hs.info < 0 ? 1 : hs.info;
        out.iterations = hs.iter;
        out.relres = hs.normr / hs.normb;
        out.solve_ms = ms;
        out.spmv_ms = tm.enabled ? tm.total_ms() : 0.0;
        out.spmv_launches = tm.launches();
    } catch (...) {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    return 0;
}

template int idrs_solve<double>(const Operator &, long, int, const double *, double *, const SolveOpts<double> &,
                                SolveStats &, hipStream_t);
template int idrs_solve<c128>(const Operator &, long, int, const c128 *, c128 *, const SolveOpts<c128> &,
                              SolveStats &, hipStream_t);
template int idrs_solve<float>(const Operator &, long, int, const float *, float *, const SolveOpts<float> &,
                               SolveStats &, hipStream_t);
template int idrs_solve<c64>(const Operator &, long, int, const c64 *, c64 *, const SolveOpts<c64> &, SolveStats &,
                             hipStream_t);
template void make_p_device<double>(long, int, int, const double *, const double *, double *, unsigned long long,
                                    hipStream_t);
template void make_p_device<c128>(long, int, int, const c128 *, const c128 *, c128 *, unsigned long long, hipStream_t);
template void make_p_device<c64>(long, int, int, const c64 *, const c64 *, c64 *, unsigned long long, hipStream_t);

}  // namespace idrs
