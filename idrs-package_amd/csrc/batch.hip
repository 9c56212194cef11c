// Batched IDR(s) on one GPU: the independent systems of a block-diagonal
// operator advance together, one workgroup per tile (a row range inside one
// system) for every streaming kernel and SpMV, one workgroup per system for
// every reduction + scalar step.
//
// Each system follows MODULES/IDRS/idrs.inc:1-311 exactly as the
// single-system solver does (same kernels' per-row arithmetic, same scalar
// logic from idrs_state.hpp); only the reduction trees differ (per tile, then
// per system).  A system that has met its tolerance (or maxit, or broke down)
// is frozen: its tiles return at once in every later launch, so the batch's
// cost per cycle falls as systems finish.  The host enqueues whole cycles and
// polls a single "all finished" flag one cycle behind, like the
// single-system driver.
#include "batch.hpp"

#include "idrs_state.hpp"
#include "mixed.hpp"
#include "spmv_epi.hpp"

#include <algorithm>
#include <cfloat>
#include <stdexcept>

namespace idrs {
namespace {

// device view of the partition
struct BV {
    int nsys, ntiles;
    const int *tile_sys;   // ntiles
    const long *tile_r0;   // ntiles: first row
    const long *tile_r1;   // ntiles: end row
    const int *tile0;      // nsys+1: tiles [tile0[k], tile0[k+1]) belong to system k
    const long *row0;      // nsys+1
};

constexpr int BT = 256;    // threads of the tile and per-system kernels

#define TILE_ROWS(q)                                                                                  \
    for (long q = bv.tile_r0[blockIdx.x] + threadIdx.x, q##_end = bv.tile_r1[blockIdx.x]; q < q##_end; \
         q += BT)

// Sum of partials[q * ntiles + t] over the tiles t of system blockIdx.x, for
// q < nv (fixed order); result in out[q], valid for all threads.
template <class T>
__device__ inline void reduce_system(const BV &bv, const T *__restrict__ partials, int nv, T *out) {
    __shared__ T lds[MAX_NV * (BT / WAVE)];
    const int t0 = bv.tile0[blockIdx.x], t1 = bv.tile0[blockIdx.x + 1];
    const int lane = threadIdx.x % WAVE, wid = threadIdx.x / WAVE;
    for (int q = 0; q < nv; q++) {
        T a = T(0);
        for (int t = t0 + threadIdx.x; t < t1; t += BT) a = a + partials[(long)q * bv.ntiles + t];
        a = wave_sum(a);
        if (lane == 0) lds[q * (BT / WAVE) + wid] = a;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        const int q = threadIdx.x;
        T v = lds[q * (BT / WAVE)];
        for (int w = 1; w < BT / WAVE; w++) v = v + lds[q * (BT / WAVE) + w];
        out[q] = v;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
// per-system state; without maximum_iterations each system gets the
// reference default min(2n, 1000) of its own size (idrs.inc:37-38)
template <class T>
__global__ void k_binit(BV bv, State<T> *st, int s, int method, int prow, int conjP, int has_maxit, int maxit,
                        double kappa, double tol, int user_omega, int n_omega) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= bv.nsys) return;
    const long n = bv.row0[b + 1] - bv.row0[b];
    const int mx = has_maxit ? maxit : (int)min(2 * n, 1000L);
    scalar_init(st + b, s, method, prow, conjP, mx, kappa, tol, user_omega, n_omega, 0, 0);
}

// r = b - ax (ax null: r = b), or only |b|^2 when r is null; partials per tile
template <class T>
__global__ __launch_bounds__(BT) void k_bresidual(BV bv, const T *__restrict__ b, const T *__restrict__ ax,
                                                  T *__restrict__ r, T *partials) {
    __shared__ T lds[BT / WAVE];
    T e[1] = {T(0)};
    TILE_ROWS(q) {
        const T rq = ax ? b[q] - ax[q] : b[q];
        if (r) r[q] = rq;
        e[0] = e[0] + T(Tr<T>::abs2(rq));
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

// what: 0 normb/tolb, 1 initial normr (idrs.inc:93-106), 2 final true
// residual (idrs.inc:304-307)
template <class T>
__global__ __launch_bounds__(BT) void k_bfin_scalar(BV bv, State<T> *st, const T *partials, int what) {
    __shared__ T red[1];
    reduce_system(bv, partials, 1, red);
    if (threadIdx.x != 0) return;
    State<T> *S = st + blockIdx.x;
    const double nr = sqrt(sre(red[0]));
    if (what == 0) {
        S->normb = nr;
        S->tolb = S->tol * nr;
    } else if (what == 1) {
        S->normr = nr;
        if (nr <= S->tolb) S->info = 0;
    } else {
        S->normr = nr;
        if (S->info == 0 && nr > S->tolb) S->info = 2;
    }
}

// per-system 2-norms from tile partials of |v|^2 (one workgroup per system)
template <class T>
__global__ __launch_bounds__(BT) void k_bsys_norm(BV bv, const T *partials, double *out) {
    __shared__ T red[1];
    reduce_system(bv, partials, 1, red);
    if (threadIdx.x == 0) out[blockIdx.x] = sqrt(sre(red[0]));
}

// y = c[k] v on the rows of every system k; y += c[k] v with accumulate
template <class T>
__global__ __launch_bounds__(BT) void k_bscale(BV bv, const double *__restrict__ c, const T *__restrict__ v, T *y,
                                               int accumulate) {
    const double ck = c[bv.tile_sys[blockIdx.x]];
    TILE_ROWS(q) y[q] = accumulate ? fma(ck, v[q], y[q]) : ck * v[q];
}

// MAKE_P (idrs_module.f90:839-937) per system: each system's rows draw the
// stream a single solve of that system would draw (index local to the system)
template <class T>
__global__ __launch_bounds__(BT) void k_bfill_uniform(BV bv, int rows, T *PT, unsigned long long seed) {
    const int sys = bv.tile_sys[blockIdx.x];
    const long r0 = bv.row0[sys], n = bv.row0[sys + 1] - r0;
    const long cnt = (long)rows * n;
    TILE_ROWS(q) {
        for (int i = 0; i < rows; i++) {
            const long loc = (q - r0) * rows + i;
            if constexpr (Tr<T>::is_complex)
                PT[q * rows + i] = T(uniform01(seed, (uint64_t)loc), uniform01(seed, (uint64_t)(cnt + loc)));
            else
                PT[q * rows + i] = uniform01(seed, (uint64_t)loc);
        }
    }
}

template <class T>
__global__ void k_bconj(long n, const T *__restrict__ r, T *__restrict__ PT) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        PT[q] = Tr<T>::conj(r[q]);
}

template <class T, int NV>
__global__ __launch_bounds__(BT) void k_bgs_dots(BV bv, int rows, int j, const T *__restrict__ PT, T *partials) {
    __shared__ T lds[NV * (BT / WAVE)];
    T e[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) e[k] = T(0);
    TILE_ROWS(q) {
        const T pj = PT[j + (long)rows * q];
#pragma unroll
        for (int k = 0; k < NV; k++)
            if (k < j) e[k] = fmaT(Tr<T>::conj(PT[k + (long)rows * q]), pj, e[k]);
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < NV; k++) partials[(long)k * bv.ntiles + blockIdx.x] = e[k];
}

template <class T>
__global__ __launch_bounds__(BT) void k_bfin_gs(BV bv, State<T> *st, const T *partials, int nv) {
    __shared__ T red[SMAX];
    reduce_system(bv, partials, nv, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < nv; k++) st[blockIdx.x].gs[k] = red[k];
}

template <class T>
__global__ __launch_bounds__(BT) void k_bgs_update_norm(BV bv, int rows, int j, const State<T> *st, T *PT,
                                                        T *partials) {
    __shared__ T lds[BT / WAVE], sa[SMAX];
    const State<T> *S = st + bv.tile_sys[blockIdx.x];
    if (threadIdx.x < SMAX) sa[threadIdx.x] = S->gs[threadIdx.x];
    __syncthreads();
    T e[1] = {T(0)};
    TILE_ROWS(q) {
        T p = PT[j + (long)rows * q];
        for (int k = 0; k < j; k++) p = fmaT(-sa[k], PT[k + (long)rows * q], p);
        PT[j + (long)rows * q] = p;
        e[0] = e[0] + T(Tr<T>::abs2(p));
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

template <class T>
__global__ __launch_bounds__(BT) void k_bfin_rho(BV bv, State<T> *st, const T *partials) {
    __shared__ T red[1];
    reduce_system(bv, partials, 1, red);
    if (threadIdx.x == 0) st[blockIdx.x].rho = sqrt(sre(red[0]));
}

template <class T>
__global__ __launch_bounds__(BT) void k_bgs_scale(BV bv, int rows, int j, const State<T> *st, T *PT) {
    const double rho = st[bv.tile_sys[blockIdx.x]].rho;
    TILE_ROWS(q) {
        const T p = PT[j + (long)rows * q];
        if constexpr (Tr<T>::is_complex) PT[j + (long)rows * q] = T(p.re / rho, p.im / rho);
        else PT[j + (long)rows * q] = p / rho;
    }
}

// partials of P^T r per tile (setup of the first cycle)
template <class T, int NV>
__global__ __launch_bounds__(BT) void k_bpdot(BV bv, const T *__restrict__ w, const T *__restrict__ P, int prow,
                                              int conjP, T *partials) {
    __shared__ T lds[NV * (BT / WAVE)];
    T e[NV];
#pragma unroll
    for (int i = 0; i < NV; i++) e[i] = T(0);
    TILE_ROWS(q) {
        const T wq = w[q];
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < prow) {
                T p = P[q * prow + i];
                if (conjP) p = Tr<T>::conj(p);
                e[i] = fmaT(p, wq, e[i]);
            }
    }
    block_sum<T, NV>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; i++) partials[(long)i * bv.ntiles + blockIdx.x] = e[i];
}

template <class T>
__global__ __launch_bounds__(BT) void k_bfin_f(BV bv, State<T> *st, const T *partials) {
    __shared__ T red[SMAX];
    State<T> *S = st + blockIdx.x;
    const int prow = S->prow;
    reduce_system(bv, partials, prow, red);
    if (threadIdx.x == 0)
        for (int i = 0; i < S->s; i++) S->f[i] = i < prow ? red[i] : T(0);
}

// ---------------------------------------------------------------------------
// per-system reduction + scalar steps of the cycle
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(BT) void k_bfin_mu(BV bv, int k, State<T> *st, const T *partials) {
    State<T> *S = st + blockIdx.x;
    if (S->info >= 0) return;
    __shared__ T red[SMAX];
    reduce_system(bv, partials, S->prow, red);
    if (threadIdx.x == 0) scalar_mu(S, k, red);
}

template <class T>
__global__ __launch_bounds__(BT) void k_bfin_norm(BV bv, State<T> *st, const T *partials) {
    State<T> *S = st + blockIdx.x;
    if (S->info >= 0) return;
    __shared__ T red[1];
    reduce_system(bv, partials, 1, red);
    if (threadIdx.x == 0) scalar_norm(S, sqrt(sre(red[0])), (double *)nullptr);
}

template <class T>
__global__ __launch_bounds__(BT) void k_bfin_om(BV bv, State<T> *st, const T *partials, const T *omega) {
    State<T> *S = st + blockIdx.x;
    if (S->info >= 0) return;
    __shared__ T red[3];
    reduce_system(bv, partials, 3, red);
    if (threadIdx.x == 0) scalar_om(S, red, omega);
}

template <class T>
__global__ __launch_bounds__(BT) void k_bfin_cycle(BV bv, State<T> *st, const T *partials) {
    State<T> *S = st + blockIdx.x;
    if (S->info >= 0) return;
    __shared__ T red[SMAX + 1];
    reduce_system(bv, partials, 1 + S->prow, red);
    if (threadIdx.x == 0) scalar_cycle(S, red, (double *)nullptr);
}

// flag[0] = 1 when every system has finished; flag[1 + k] = 1 while system
// k runs
template <class T>
__global__ __launch_bounds__(BT) void k_bdone(int nsys, const State<T> *st, int *flag) {
    __shared__ int open;
    if (threadIdx.x == 0) open = 0;
    __syncthreads();
    int mine = 0;
    for (int b = threadIdx.x; b < nsys; b += BT) {
        const int live = (st[b].info < 0) ? 1 : 0;
        flag[1 + b] = live;
        mine |= live;
    }
    if (mine) atomicOr(&open, 1);
    __syncthreads();
    if (threadIdx.x == 0) flag[0] = open ? 0 : 1;
}

// ---------------------------------------------------------------------------
// streaming kernels of the cycle (idrs.inc:155-299), per tile
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(BT) void k_bform_u(BV bv, long ntot, int k, const State<T> *st, const T *__restrict__ r,
                                                const T *__restrict__ G, T *__restrict__ U) {
    const State<T> *S = st + bv.tile_sys[blockIdx.x];
    if (S->info >= 0) return;
    __shared__ T sg[SMAX];
    const int s = S->s, jj = S->jj;
    if (threadIdx.x < SMAX) sg[threadIdx.x] = S->gamma[threadIdx.x];
    __syncthreads();
    const T om = S->om;
    if (jj > 0) {
        TILE_ROWS(q) {
            T v = r[q];
            for (int i = k; i < s; i++) v = fmaT(-sg[i], G[(long)i * ntot + q], v);
            T t = om * v;
            for (int i = k; i < s; i++) t = fmaT(sg[i], U[(long)i * ntot + q], t);
            U[(long)k * ntot + q] = t;
        }
    } else {
        TILE_ROWS(q) U[(long)k * ntot + q] = r[q];
    }
}

template <class T>
__global__ __launch_bounds__(BT) void k_borth_update(BV bv, long ntot, int k, const State<T> *st, T *__restrict__ G,
                                                     T *__restrict__ U, T *__restrict__ r, T *__restrict__ x,
                                                     T *partials) {
    const State<T> *S = st + bv.tile_sys[blockIdx.x];
    if (S->info >= 0) return;
    __shared__ T sa[SMAX], sg[SMAX], lds[BT / WAVE];
    const int s = S->s, jj = S->jj;
    if (threadIdx.x < SMAX) {
        sa[threadIdx.x] = S->alpha[threadIdx.x];
        sg[threadIdx.x] = S->gamma[threadIdx.x];
    }
    __syncthreads();
    const T beta = S->beta[k], om = S->om;
    const bool next = (k + 1 < s);
    T e[1] = {T(0)};
    TILE_ROWS(q) {
        T gk = G[(long)k * ntot + q], uk = U[(long)k * ntot + q];
        for (int i = 0; i < k; i++) {
            gk = fmaT(-sa[i], G[(long)i * ntot + q], gk);
            uk = fmaT(-sa[i], U[(long)i * ntot + q], uk);
        }
        G[(long)k * ntot + q] = gk;
        U[(long)k * ntot + q] = uk;
        const T rq = fmaT(-beta, gk, r[q]);
        x[q] = fmaT(beta, uk, x[q]);
        r[q] = rq;
        e[0] = e[0] + T(Tr<T>::abs2(rq));
        if (next) {
            if (jj > 0) {
                T v = rq;
                for (int i = k + 1; i < s; i++) v = fmaT(-sg[i], G[(long)i * ntot + q], v);
                T t = om * v;
                for (int i = k + 1; i < s; i++) t = fmaT(sg[i], U[(long)i * ntot + q], t);
                U[(long)(k + 1) * ntot + q] = t;
            } else {
                U[(long)(k + 1) * ntot + q] = rq;
            }
        }
    }
    block_sum<T, 1>(e, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = e[0];
}

template <class T, int NV>
__global__ __launch_bounds__(BT) void k_bomega_update(BV bv, const State<T> *st, const T *__restrict__ t, T *r,
                                                      T *__restrict__ x, const T *__restrict__ P, int prow, int conjP,
                                                      T *partials) {
    const State<T> *S = st + bv.tile_sys[blockIdx.x];
    if (S->info >= 0) return;
    __shared__ T lds[(NV + 1) * (BT / WAVE)];
    const T om = S->om;
    T e[NV + 1];
#pragma unroll
    for (int i = 0; i <= NV; i++) e[i] = T(0);
    TILE_ROWS(q) {
        const T vq = r[q];
        const T rq = fmaT(-om, t[q], vq);
        x[q] = fmaT(om, vq, x[q]);
        r[q] = rq;
        e[0] = e[0] + T(Tr<T>::abs2(rq));
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < prow) {
                T p = P[q * prow + i];
                if (conjP) p = Tr<T>::conj(p);
                e[1 + i] = fmaT(p, rq, e[1 + i]);
            }
    }
    block_sum<T, NV + 1>(e, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i <= NV; i++) partials[(long)i * bv.ntiles + blockIdx.x] = e[i];
}

#undef TILE_ROWS

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct DBuf {
    void *p = nullptr;
    explicit DBuf(size_t bytes) { HIP_TRY(hipMalloc(&p, bytes ? bytes : 1)); }
    ~DBuf() { (void)hipFree(p); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    template <class V> V *as() const { return (V *)p; }
};

#define NV_SWITCH(nv, KER, ...)                                                   \
    do {                                                                          \
        if ((nv) <= 1) hipLaunchKernelGGL((KER<T, 1>), __VA_ARGS__);              \
        else if ((nv) <= 2) hipLaunchKernelGGL((KER<T, 2>), __VA_ARGS__);         \
        else if ((nv) <= 4) hipLaunchKernelGGL((KER<T, 4>), __VA_ARGS__);         \
        else if ((nv) <= 8) hipLaunchKernelGGL((KER<T, 8>), __VA_ARGS__);         \
        else hipLaunchKernelGGL((KER<T, 16>), __VA_ARGS__);                       \
        LAUNCH_CHECK();                                                           \
    } while (0)

}  // namespace

template <class T>
void batch_idrs_solve(const Operator &A, int s, const T *b, T *x, const SolveOpts<T> &o,
                      std::vector<SolveStats> &stats, double &solve_ms, hipStream_t stream) {
    const Partition *part = A.K->part.get();
    if (!part) throw std::invalid_argument("the operator's matrix is not a batch (block-diagonal with a partition)");
    if (s < 1 || s > PDOT_MAX) throw std::invalid_argument("batched solves take s in 1..16");
    if (o.tol < 0) throw std::invalid_argument("Illegal value for parameter tolerance");
    if (o.has_m1 || o.U0 || o.resvec) throw std::invalid_argument("M1, U0 and resvec are not offered for batches");
    const int method = o.variant;
    if (method < 1 || method > 4) throw std::invalid_argument("variant must be 1..4");
    const int nsys = part->nsys, ntiles = part->ntiles;
    const long ntot = part->h_row0[nsys];
    const int prow = method == 1 ? s : 1;
    const int conjP = (method >= 3 && Tr<T>::is_complex) ? 1 : 0;
    const double kappa = (method == 1 || method == 3) ? (double)0.7f : 0.0;
    long nmax = 0;
    for (int q = 0; q < nsys; q++) nmax = std::max(nmax, part->h_row0[q + 1] - part->h_row0[q]);
    // host bound on the cycles: the largest per-system maxit
    const int maxit = o.has_maxit ? o.maxit : (int)std::min<long>(2 * nmax, 1000);
    const long cap_it = maxit > 0 ? maxit : 10 * nmax + 1000;
    const int user_omega = (o.omega && o.n_omega >= 1) ? 1 : 0;

    const size_t vb = sizeof(T) * (size_t)ntot;
    DBuf bG(vb * s), bU(vb * s), br(vb), bt(vb), bw1(vb), bw2(vb), btmp(vb), bP(vb * prow);
    DBuf bpart(sizeof(T) * (size_t)MAX_NV * ntiles), bst(sizeof(State<T>) * nsys), bflag(sizeof(int) * (nsys + 1));
    // two sets of compacted tile lists (tiles of unfinished systems only),
    // alternated so a set is rewritten only when no queued launch reads it
    DBuf ctile_sys[2] = {DBuf(sizeof(int) * ntiles), DBuf(sizeof(int) * ntiles)};
    DBuf ctile_r0[2] = {DBuf(sizeof(long) * ntiles), DBuf(sizeof(long) * ntiles)};
    DBuf ctile_r1[2] = {DBuf(sizeof(long) * ntiles), DBuf(sizeof(long) * ntiles)};
    DBuf ctile0[2] = {DBuf(sizeof(int) * (nsys + 1)), DBuf(sizeof(int) * (nsys + 1))};
    T *G = bG.as<T>(), *U = bU.as<T>(), *r = br.as<T>(), *t = bt.as<T>(), *w1 = bw1.as<T>(), *w2 = bw2.as<T>();
    T *tmp = btmp.as<T>(), *P = bP.as<T>(), *part_d = bpart.as<T>();
    State<T> *st = bst.as<State<T>>();
    int *flag = bflag.as<int>();
    std::unique_ptr<DBuf> bom;
    T *omega_d = nullptr;
    if (user_omega) {
        bom.reset(new DBuf(sizeof(T) * o.n_omega));
        omega_d = bom->as<T>();
        HIP_TRY(hipMemcpyAsync(omega_d, o.omega, sizeof(T) * o.n_omega, hipMemcpyHostToDevice, stream));
    }
    const BV bv_all{nsys, ntiles, part->tile_sys, part->tile_r0, part->tile_r0 + 1, part->tile0, part->row0};
    BV bv = bv_all;
    dim3 gt(ntiles);
    const dim3 gsys(nsys);
    TileCtx all{part, nullptr, 0};
    TileCtx live{part, &st[0].info, (int)(sizeof(State<T>) / sizeof(int))};
    static_assert(sizeof(State<T>) % sizeof(int) == 0, "State stride in ints");

    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, stream));
    hipLaunchKernelGGL((k_binit<T>), dim3((nsys + 255) / 256), dim3(256), 0, stream, bv, st, s, method, prow, conjP,
                       o.has_maxit ? 1 : 0, o.maxit, kappa, o.tol, user_omega, user_omega ? o.n_omega : 0);
    LAUNCH_CHECK();
    if (o.x0) HIP_TRY(hipMemcpyAsync(x, o.x0, vb, hipMemcpyDeviceToDevice, stream));
    else HIP_TRY(hipMemsetAsync(x, 0, vb, stream));
    HIP_TRY(hipMemsetAsync(U, 0, vb * s, stream));
    // normb, tolb; r = b - A x0; normr (idrs.inc:93-106)
    hipLaunchKernelGGL((k_bresidual<T>), gt, dim3(BT), 0, stream, bv, b, (const T *)nullptr, (T *)nullptr, part_d);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_bfin_scalar<T>), gsys, dim3(BT), 0, stream, bv, st, part_d, 0);
    LAUNCH_CHECK();
    EpiArgs<T> none{};
    if (o.x0) op_apply<T>(A, x, tmp, w1, w2, EPI_NONE, none, nullptr, stream, &all);
    hipLaunchKernelGGL((k_bresidual<T>), gt, dim3(BT), 0, stream, bv, b, o.x0 ? (const T *)tmp : nullptr, r, part_d);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((k_bfin_scalar<T>), gsys, dim3(BT), 0, stream, bv, st, part_d, 1);
    LAUNCH_CHECK();
    // P per system (idrs.inc:109-126), f = P^T r (idrs.inc:147)
    if (method == 2) {
        hipLaunchKernelGGL((k_bconj<T>), dim3(vec_grid(ntot)), dim3(256), 0, stream, ntot, r, P);
        LAUNCH_CHECK();
    } else {
        if (o.PT_raw) HIP_TRY(hipMemcpyAsync(P, o.PT_raw, vb * prow, hipMemcpyDeviceToDevice, stream));
        else {
            hipLaunchKernelGGL((k_bfill_uniform<T>), gt, dim3(BT), 0, stream, bv, prow, P, o.p_seed);
            LAUNCH_CHECK();
        }
        for (int j = 0; j < prow; j++) {
            if (j > 0) {
                NV_SWITCH(j, k_bgs_dots, gt, dim3(BT), 0, stream, bv, prow, j, (const T *)P, part_d);
                hipLaunchKernelGGL((k_bfin_gs<T>), gsys, dim3(BT), 0, stream, bv, st, part_d, j);
                LAUNCH_CHECK();
            }
            hipLaunchKernelGGL((k_bgs_update_norm<T>), gt, dim3(BT), 0, stream, bv, prow, j, st, P, part_d);
            LAUNCH_CHECK();
            hipLaunchKernelGGL((k_bfin_rho<T>), gsys, dim3(BT), 0, stream, bv, st, part_d);
            LAUNCH_CHECK();
            hipLaunchKernelGGL((k_bgs_scale<T>), gt, dim3(BT), 0, stream, bv, prow, j, st, P);
            LAUNCH_CHECK();
        }
    }
    NV_SWITCH(prow, k_bpdot, gt, dim3(BT), 0, stream, bv, (const T *)r, (const T *)P, prow, conjP, part_d);
    hipLaunchKernelGGL((k_bfin_f<T>), gsys, dim3(BT), 0, stream, bv, st, part_d);
    LAUNCH_CHECK();

    EpiArgs<T> epd{};
    epd.P = P;
    epd.s = s;
    epd.prow = prow;
    epd.conjP = conjP;
    epd.partials = part_d;
    EpiArgs<T> eom{};
    eom.r = r;
    eom.partials = part_d;

    auto enqueue_cycle = [&]() {
        hipLaunchKernelGGL((k_bform_u<T>), gt, dim3(BT), 0, stream, bv, ntot, 0, (const State<T> *)st, (const T *)r,
                           (const T *)G, U);
        LAUNCH_CHECK();
        for (int k = 0; k < s; k++) {
            op_apply<T>(A, U + (long)k * ntot, G + (long)k * ntot, w1, w2, EPI_PDOT, epd, nullptr, stream, &live);
            hipLaunchKernelGGL((k_bfin_mu<T>), gsys, dim3(BT), 0, stream, bv, k, st, (const T *)part_d);
            LAUNCH_CHECK();
            hipLaunchKernelGGL((k_borth_update<T>), gt, dim3(BT), 0, stream, bv, ntot, k, (const State<T> *)st, G, U,
                               r, x, part_d);
            LAUNCH_CHECK();
            hipLaunchKernelGGL((k_bfin_norm<T>), gsys, dim3(BT), 0, stream, bv, st, (const T *)part_d);
            LAUNCH_CHECK();
        }
        op_apply<T>(A, r, t, w1, w2, EPI_OMEGA, eom, nullptr, stream, &live);
        hipLaunchKernelGGL((k_bfin_om<T>), gsys, dim3(BT), 0, stream, bv, st, (const T *)part_d, (const T *)omega_d);
        LAUNCH_CHECK();
        NV_SWITCH(prow, k_bomega_update, gt, dim3(BT), 0, stream, bv, (const State<T> *)st, (const T *)t, r, x,
                  (const T *)P, prow, conjP, part_d);
        hipLaunchKernelGGL((k_bfin_cycle<T>), gsys, dim3(BT), 0, stream, bv, st, (const T *)part_d);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_bdone<T>), dim3(1), dim3(BT), 0, stream, nsys, (const State<T> *)st, flag);
        LAUNCH_CHECK();
    };

    // h_flag: two slots of nsys+1 ints (done flag, then one live flag per system)
    int *h_flag = nullptr;
    HIP_TRY(hipHostMalloc((void **)&h_flag, 2 * sizeof(int) * (nsys + 1), hipHostMallocDefault));
    int cset = 0;            // compacted set the next rewrite uses
    // Restricts every later tile launch to the unfinished systems' tiles once
    // they are at most 3/4 of the tiles launched so far.  Finished systems
    // keep an empty tile range (their reduction kernels return at once).
    // Each cycle replays a HIP graph of enqueue_cycle() (the late cycles run
    // few live tiles and are launch-bound); a compaction changes the tile
    // lists and grids, so the next cycle captures a new graph.  Replaced
    // graphs may still be queued: all are destroyed after the final sync.
    const bool use_graph = o.graph_mode != 2;
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    hipGraphExec_t cur = nullptr;
    auto capture = [&]() {
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        try {
            enqueue_cycle();
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamEndCapture(stream, &g));
        graphs.push_back(g);
        hipGraphExec_t e = nullptr;
        HIP_TRY(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
        execs.push_back(e);
        cur = e;
    };
    auto compact = [&](const int *liveflag) {
        std::vector<int> tsys, t0(nsys + 1);
        std::vector<long> r0v, r1v;
        for (int k = 0; k < nsys; k++) {
            t0[k] = (int)tsys.size();
            if (!liveflag[k]) continue;
            for (int t = part->h_tile0[k]; t < part->h_tile0[k + 1]; t++) {
                tsys.push_back(k);
                const long a = part->h_row0[k] + (long)(t - part->h_tile0[k]) * BATCH_TILE;
                r0v.push_back(a);
                r1v.push_back(std::min(a + (long)BATCH_TILE, part->h_row0[k + 1]));
            }
        }
        t0[nsys] = (int)tsys.size();
        const int nt = (int)tsys.size();
        if (nt == 0 || 4L * nt > 3L * bv.ntiles) return;
        HIP_TRY(hipMemcpy(ctile_sys[cset].p, tsys.data(), sizeof(int) * nt, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctile_r0[cset].p, r0v.data(), sizeof(long) * nt, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctile_r1[cset].p, r1v.data(), sizeof(long) * nt, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(ctile0[cset].p, t0.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice));
        bv.ntiles = nt;
        bv.tile_sys = ctile_sys[cset].as<int>();
        bv.tile_r0 = ctile_r0[cset].as<long>();
        bv.tile_r1 = ctile_r1[cset].as<long>();
        bv.tile0 = ctile0[cset].as<int>();
        live.ntiles = nt;
        live.tile_sys = bv.tile_sys;
        live.tile_r0 = bv.tile_r0;
        live.tile_r1 = bv.tile_r1;
        gt = dim3(nt);
        cset ^= 1;
        cur = nullptr;
    };
    hipEvent_t pe[2];
    HIP_TRY(hipEventCreateWithFlags(&pe[0], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&pe[1], hipEventDisableTiming));
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(stream);
        for (hipGraphExec_t e : execs) (void)hipGraphExecDestroy(e);
        for (hipGraph_t g : graphs) (void)hipGraphDestroy(g);
        (void)hipHostFree(h_flag);
        (void)hipEventDestroy(pe[0]);
        (void)hipEventDestroy(pe[1]);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    };
    try {
        const long max_cycles = cap_it / (s + 1) + 2;
        bool done = false;
        for (long c = 0; c < max_cycles && !done; c++) {
            if (use_graph) {
                if (!cur) capture();
                HIP_TRY(hipGraphLaunch(cur, stream));
            } else {
                enqueue_cycle();
            }
            int *slot = h_flag + (c % 2) * (nsys + 1);
            HIP_TRY(hipMemcpyAsync(slot, flag, sizeof(int) * (nsys + 1), hipMemcpyDeviceToHost, stream));
            HIP_TRY(hipEventRecord(pe[c % 2], stream));
            if (c >= 1) {
                HIP_TRY(hipEventSynchronize(pe[(c - 1) % 2]));
                const int *prev = h_flag + ((c - 1) % 2) * (nsys + 1);
                if (prev[0]) done = true;
                else compact(prev + 1);
            }
        }
        // final true residuals over every system (idrs.inc:304-310)
        bv = bv_all;
        gt = dim3(ntiles);
        op_apply<T>(A, x, tmp, w1, w2, EPI_NONE, none, nullptr, stream, &all);
        hipLaunchKernelGGL((k_bresidual<T>), gt, dim3(BT), 0, stream, bv, b, (const T *)tmp, r, part_d);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_bfin_scalar<T>), gsys, dim3(BT), 0, stream, bv, st, part_d, 2);
        LAUNCH_CHECK();
        HIP_TRY(hipEventRecord(e1, stream));
        std::vector<State<T>> hs(nsys);
        HIP_TRY(hipMemcpyAsync(hs.data(), st, sizeof(State<T>) * nsys, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        solve_ms = ms;
        stats.assign(nsys, SolveStats{});
        for (int q = 0; q < nsys; q++) {
            stats[q].flag = hs[q].info < 0 ? 1 : hs[q].info;
            stats[q].iterations = hs[q].iter;
            stats[q].relres = hs[q].normr / hs[q].normb;
            stats[q].solve_ms = ms;
        }
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}
#undef NV_SWITCH

void batch_mixed_solve(const Operator &A, int s, const double *b, double *x, const SolveOpts<double> &o,
                       double inner_tol, int max_outer, std::vector<MixedStats> &stats, double &solve_ms,
                       hipStream_t stream) {
    const Partition *part = A.K->part.get();
    if (!part) throw std::invalid_argument("the operator's matrix is not a batch (block-diagonal with a partition)");
    if (s < 1 || s > PDOT_MAX) throw std::invalid_argument("batched solves take s in 1..16");
    if (o.tol < 0) throw std::invalid_argument("Illegal value for parameter tolerance");
    if (!(inner_tol > 0 && inner_tol < 1)) throw std::invalid_argument("inner tolerance must be in (0, 1)");
    if (max_outer < 1) throw std::invalid_argument("at least one outer step");
    if (o.has_m1 || o.U0 || o.resvec || o.omega || o.x0 || o.PT_raw)
        throw std::invalid_argument("M1, U0, omega, resvec, x0 and PT_raw are not offered here");
    if (A.values_complex()) throw std::invalid_argument("mixed precision needs real matrix values");
    if (A.seed != std::complex<double>(0.0, 0.0)) throw std::invalid_argument("mixed precision without a seed shift");
    Operator &op = const_cast<Operator &>(A);
    op_make_lowprec(op);
    const int nsys = part->nsys, ntiles = part->ntiles;
    const long ntot = part->h_row0[nsys];
    long nmax = 0;
    for (int q = 0; q < nsys; q++) nmax = std::max(nmax, part->h_row0[q + 1] - part->h_row0[q]);
    const int maxit = o.has_maxit ? o.maxit : (int)std::min<long>(2 * nmax, 1000);
    const size_t vb = sizeof(double) * (size_t)ntot;
    DBuf br(vb), bt(vb), bw1(vb), bw2(vb), brh(vb), bdh(vb), bpart(sizeof(double) * (size_t)ntiles);
    DBuf bnorm(sizeof(double) * nsys), bfac(sizeof(double) * nsys);
    double *r = br.as<double>(), *t = bt.as<double>(), *w1 = bw1.as<double>(), *w2 = bw2.as<double>();
    double *rh = brh.as<double>(), *dh = bdh.as<double>(), *part_d = bpart.as<double>();
    double *norm_d = bnorm.as<double>(), *fac_d = bfac.as<double>();
    const BV bv{nsys, ntiles, part->tile_sys, part->tile_r0, part->tile_r0 + 1, part->tile0, part->row0};
    const dim3 gt(ntiles), gsys(nsys);
    TileCtx all{part, nullptr, 0};
    EpiArgs<double> none{};
    std::vector<double> normb(nsys), normr(nsys), prev(nsys), fac(nsys);
    std::vector<int> its(nsys, 0);
    stats.assign(nsys, MixedStats{});
    // per-system |b - ax| (ax null: |b|), r = b - ax when r is given
    auto sys_norms = [&](const double *ax, double *rv, std::vector<double> &out) {
        hipLaunchKernelGGL((k_bresidual<double>), gt, dim3(BT), 0, stream, bv, b, ax, rv, part_d);
        LAUNCH_CHECK();
        hipLaunchKernelGGL((k_bsys_norm<double>), gsys, dim3(BT), 0, stream, bv, (const double *)part_d, norm_d);
        LAUNCH_CHECK();
        HIP_TRY(hipMemcpyAsync(out.data(), norm_d, sizeof(double) * nsys, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
    };
    auto per_system = [&](const std::vector<double> &c, const double *v, double *y, int accumulate) {
        HIP_TRY(hipMemcpyAsync(fac_d, c.data(), sizeof(double) * nsys, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL((k_bscale<double>), gt, dim3(BT), 0, stream, bv, (const double *)fac_d, v, y, accumulate);
        LAUNCH_CHECK();
    };
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, stream));
    HIP_TRY(hipMemsetAsync(x, 0, vb, stream));
    sys_norms(nullptr, nullptr, normb);
    sys_norms(nullptr, r, normr);       // r = b (x = 0)
    while (true) {
        // a system stops at its FP64 tolerance, at max_outer or at its share of maxit
        int active = 0;
        double itol = 1.0;
        for (int q = 0; q < nsys; q++) {
            MixedStats &S = stats[q];
            if (S.flag >= 0) continue;
            if (normr[q] <= o.tol * normb[q]) S.flag = 0;
            else if (S.outer == max_outer || its[q] >= maxit) S.flag = 1;
            else {
                active++;
                // no tighter than the FP64 target needs (half of it, for the FP32 error)
                itol = std::min(itol, std::max(inner_tol, 0.5 * o.tol * normb[q] / normr[q]));
            }
        }
        if (active == 0) break;
        // inner right-hand sides r_k/|r_k|; a finished system gets zero,
        // which the inner solve finishes at once
        for (int q = 0; q < nsys; q++) fac[q] = stats[q].flag < 0 ? 1.0 / normr[q] : 0.0;
        per_system(fac, r, rh, 0);
        SolveOpts<double> io;
        io.tol = itol;
        io.has_maxit = true;
        io.maxit = maxit;
        io.variant = o.variant;
        io.p_seed = o.p_seed;
        io.graph_mode = o.graph_mode;
        std::vector<SolveStats> is;
        double ims = 0.0;
        op.values32 = true;
        try {
            batch_idrs_solve<double>(A, s, rh, dh, io, is, ims, stream);
        } catch (...) {
            op.values32 = false;
            throw;
        }
        op.values32 = false;
        for (int q = 0; q < nsys; q++) fac[q] = stats[q].flag < 0 ? normr[q] : 0.0;
        per_system(fac, dh, x, 1);
        // true FP64 residuals
        op_apply<double>(A, x, t, w1, w2, EPI_NONE, none, nullptr, stream, &all);
        prev = normr;
        sys_norms(t, r, normr);
        for (int q = 0; q < nsys; q++) {
            MixedStats &S = stats[q];
            if (S.flag >= 0) continue;
            its[q] += is[q].iterations;
            S.outer++;
            if (is[q].flag == 3 || !(normr[q] < prev[q])) S.flag = normr[q] <= o.tol * normb[q] ? 0 : 3;
        }
    }
    HIP_TRY(hipEventRecord(e1, stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    solve_ms = ms;
    for (int q = 0; q < nsys; q++) {
        stats[q].iterations = its[q];
        stats[q].relres = normr[q] / normb[q];
        stats[q].solve_ms = ms;
    }
}

template void batch_idrs_solve<double>(const Operator &, int, const double *, double *, const SolveOpts<double> &,
                                       std::vector<SolveStats> &, double &, hipStream_t);
template void batch_idrs_solve<c128>(const Operator &, int, const c128 *, c128 *, const SolveOpts<c128> &,
                                     std::vector<SolveStats> &, double &, hipStream_t);

}  // namespace idrs
