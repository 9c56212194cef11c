"""Benchmark: time-to-solution of the IDR(s) hot path on one MI355X.

Workload (default): 3-D convection-diffusion-reaction, 7-point central
differences on a 200^3 grid (N = 8e6, nnz = 55.76e6), Pe = 50, reaction 1,
FP64, stored both as offset-list CDS (DIA, the headline value) and as CRS
(reported beside it); IDR(4) with a degree-8 Chebyshev polynomial
preconditioner (foci 0.15, 1.85) and diagonal scaling; b ~ U[-1,1] (seed 0);
Gaussian shadow space (seed 0); ||r||/||b|| <= 1e-8.  A "step" is one full
solve with b, x and the matrix already resident in HBM.

Prints one JSON line.  roofline: the SpMV kernel (the dominant kernel),
achieved = algorithmic bytes of the SpMV launches of one solve / their summed
HIP event time (events on the solver's stream around each operator
application, i.e. its d+1 back-to-back SpMV launches; an extra solve after
the timed steps, so the events do not slow the timed ones), against the
8 TB/s HBM peak.  cpu_baseline: the oracle (the project's C restatement of
the reference path) built with OpenMP loops, timed on a full solve of the
same system on the host threads this process may use, with the
single-thread rate beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "idrs-package_amd"))

HBM_PEAK_GBS = 8000.0


def spmv_bytes(n, nnz, with_z, epi_vectors):
    """Algorithmic bytes of one FP64 CRS SpMV launch with int32 indices:
    values + column indices (12 nnz), row pointers (4(N+1)), x read once and
    y written once (16 N), the Horner vector z (8 N) and epilogue vectors."""
    return 12 * nnz + 4 * (n + 1) + 16 * n + (8 * n if with_z else 0) + 8 * n * epi_vectors


def solve_spmv_bytes(n, nnz, s, d, iterations, prow):
    """Sum of algorithmic bytes of the SpMV launches of one solve with
    'iterations' iterations (inner steps + omega steps), plus the final
    true-residual product.  Each operator application: d launches with z,
    one final launch with the epilogue (P^T G reads s vectors; the omega
    epilogue reads r)."""
    cycles_full, rem = divmod(iterations, s + 1)
    inner = cycles_full * s + min(rem, s)
    omega = cycles_full + (1 if rem > s else 0)
    per_op_z = d * spmv_bytes(n, nnz, True, 0)
    b_inner = inner * (per_op_z + spmv_bytes(n, nnz, False, prow))
    b_omega = omega * (per_op_z + spmv_bytes(n, nnz, False, 1))
    b_final = per_op_z + spmv_bytes(n, nnz, False, 0)
    launches = (inner + omega + 1) * (d + 1)
    return b_inner + b_omega + b_final, launches


def iteration_bytes(n, nnz, s, d):
    """Algorithmic bytes per IDR(s) iteration: (d+1) SpMVs, d Horner vector
    reads and the fused vector traffic 8N(3s^2+11s+3)/(s+1)."""
    b_spmv = 12 * nnz + 4 * (n + 1) + 16 * n
    b_op = (d + 1) * b_spmv + d * 8 * n
    return b_op + 8 * n * (3 * s * s + 11 * s + 3) / (s + 1)


def cpu_baseline(dia, b, PT, s, degree, foci, sample_iters, gpu_iters):
    """The oracle's restatement of the reference path timed on the host:
    a full solve of the same system (DIA storage, as the headline) with the
    OpenMP build on all the threads this process may use, and beside it the
    single-thread rate from a sample of the same solve, scaled to the full
    iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import orc

    host_cpus = len(os.sched_getaffinity(0))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or host_cpus
    alpha = orc.chebyshev(degree, np.asarray(foci))
    Ao = orc.OracleMatrix(orc.DIA, dia.vals.shape[0], dia=(dia.offsets, dia.vals), which_precon=1, degree=degree,
                          alpha=alpha)
    lib, used = orc.load_timed(threads)
    t0 = time.perf_counter()
    _, flag, relres, its, _ = orc.idrs(Ao, b, s, tolerance=1e-8, maximum_iterations=20000, PT_raw=PT, lib=lib)
    full = time.perf_counter() - t0
    print(f"# cpu baseline: {used} threads, {its} iterations, flag {flag}, relres {relres:.2e}, {full:.1f} s",
          file=sys.stderr, flush=True)
    lib, _ = orc.load_timed(1)
    t0 = time.perf_counter()
    orc.idrs(Ao, b, s, tolerance=1e-8, maximum_iterations=sample_iters, PT_raw=PT, lib=lib)
    one = time.perf_counter() - t0
    # the sample holds setup + sample_iters iterations + the final residual;
    # one operator application costs about one iteration
    one_full_ms = one / (sample_iters + 2) * (its + 2) * 1e3
    lib.orc_threads(threads)
    # flag 2 (the reference's "true residual above the tolerance"): the
    # recursive residual met 1e-8 at the same iteration as the GPU's, and
    # the final true residual, summed by threads in another order, did not
    why = (f", flag 2: the final true residual {relres:.3g} is above 1e-8 by the rounding of the threaded sums"
           if flag == 2 else "")
    return {
        "value": round(full * 1e3, 1),
        "unit": "ms (time-to-solution)",
        "cores": used,
        "kind": "port",
        "sample": f"the oracle (the C restatement of the reference path) built with OpenMP row and vector loops: "
                  f"the full IDR({s})+Chebyshev({degree}) solve of the same {dia.vals.shape[0]}-row system in DIA "
                  f"storage to 1e-8 on {used} threads ({its} iterations, flag {flag}{why}; the GPU took {gpu_iters})",
        "host_cpus": host_cpus,
        "iterations": its,
        "one_thread_ms": round(one_full_ms, 1),
        "one_thread_sample": f"{sample_iters} iterations on 1 thread in {one:.1f} s, scaled to {its} iterations",
    }


def read_traffic(fmt):
    """HBM bytes per SpMV launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), from the committed profile summary, or None."""
    path = os.path.join(ROOT, "profiles", f"spmv_{fmt}_200cube_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1].strip().split()[0])
            except ValueError:
                return None
    return None


def calibrate_copy():
    """A plain 512 MiB device-to-device copy, 5 times: each dispatch reads and
    writes exactly 536870912 bytes, which calibrates rocprofv3's FETCH_SIZE
    and WRITE_SIZE counters (KiB per dispatch)."""
    import torch

    a = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device="cuda:0").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_dispatch": a.numel() * 8, "dispatches": 5}))


def bench_batched(args):
    """One batched IDR(2) solve of all systems vs a loop of single-system
    solves on the same device-resident data."""
    import numpy as np
    import torch

    import idrs_amd

    t0 = time.time()
    from idrs_amd.problems import batched_systems

    systems = batched_systems(args.nsys)
    ntot = sum(c.n for c, _ in systems)
    nnz = sum(c.nnz for c, _ in systems)
    print(f"# {args.nsys} systems built: rows={ntot} nnz={nnz} ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    Ks = [idrs_amd.crs_format(c.n, c.n, c.row_ptr, c.col_ind, rval=c.val) for c, _ in systems]
    KB = idrs_amd.batch_format(Ks)
    group = KB.autotune() if not args.no_autotune else None
    AB = idrs_amd.Matrix(KB)
    b_all = torch.from_numpy(np.concatenate([bb for _, bb in systems])).to(dev)
    x_all = torch.empty_like(b_all)
    s, tol, maxit = 2, 1e-8, args.batch_maxit

    def batch_step(graph=True):
        return idrs_amd.idrs_batched(AB, b_all, s, tolerance=tol, maximum_iterations=maxit, graph=graph, out=x_all)

    for _ in range(args.warmup):
        batch_step()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        rb = batch_step()
    torch.cuda.synchronize()
    batch_ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    print(f"# batched: {batch_ms:.1f} ms, flags {np.bincount(rb.flag)}, iterations min/median/max "
          f"{rb.iterations.min()}/{int(np.median(rb.iterations))}/{rb.iterations.max()}, lanes per row {group}",
          file=sys.stderr, flush=True)
    # the same batch with eager launches (no graph replay)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    re_ = batch_step(graph=False)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t_start) * 1e3
    same = bool(np.array_equal(re_.iterations, rb.iterations))
    x_graph = x_all.clone()
    print(f"# batched, eager launches: {eager_ms:.1f} ms (same iteration counts: {same})", file=sys.stderr, flush=True)
    # mixed precision per system: batched inner solves on the FP32 copy of the
    # matrix values, FP64 residual replacement until every system's true
    # residual meets tol
    x_mixed = torch.empty_like(b_all)

    def mixed_step():
        return idrs_amd.idrs_batched_mixed(AB, b_all, s, tolerance=tol, maximum_iterations=maxit,
                                           inner_tolerance=args.batch_inner_tol, out=x_mixed)

    for _ in range(max(args.warmup, 1)):
        mixed_step()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        rmx = mixed_step()
    torch.cuda.synchronize()
    mixed_ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    mixed_diff = float((x_mixed - x_graph).norm() / x_graph.norm())
    print(f"# batched mixed precision: {mixed_ms:.1f} ms, flags {np.bincount(rmx.flag)}, worst true relres "
          f"{rmx.relres.max():.2e}, outer steps max {rmx.outer.max()}, ||x_mixed - x_fp64|| / ||x_fp64|| = "
          f"{mixed_diff:.2e}", file=sys.stderr, flush=True)
    if args.no_loop:
        return None
    # per-system loop over the same device data
    As = [idrs_amd.Matrix(K) for K in Ks]
    offs = np.cumsum([0] + [c.n for c, _ in systems])
    xs = torch.empty_like(b_all)
    bs = [b_all[offs[k]:offs[k + 1]] for k in range(args.nsys)]
    outs = [xs[offs[k]:offs[k + 1]] for k in range(args.nsys)]
    loop_its = np.zeros(args.nsys, dtype=int)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.nsys):
        r = idrs_amd.idrs(As[k], bs[k], s, tolerance=tol, maximum_iterations=maxit, out=outs[k])
        loop_its[k] = r.iterations
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t_start) * 1e3
    diff = float((x_graph - xs).norm() / xs.norm())
    worst = np.argsort(-np.abs(rb.iterations - loop_its))[:3]
    for k in worst:
        c = systems[k][0]
        print(f"# system {k}: n={c.n}, batch it={rb.iterations[k]} flag={rb.flag[k]} relres={rb.relres[k]:.2e}; "
              f"loop it={loop_its[k]}", file=sys.stderr, flush=True)
    print(f"# per-system loop: {loop_ms:.1f} ms; max |it batch - it loop| = "
          f"{int(np.abs(rb.iterations - loop_its).max())}; ||x_batch - x_loop|| / ||x_loop|| = {diff:.2e}",
          file=sys.stderr, flush=True)
    # algorithmic bytes of the work done (each system's own iterations)
    it_bytes = 0.0
    for (c, _), it in zip(systems, rb.iterations):
        b_spmv = 12 * c.nnz + 4 * (c.n + 1) + 16 * c.n
        it_bytes += it * (b_spmv + 8 * c.n * (3 * s * s + 11 * s + 3) / (s + 1))
    gbs = it_bytes / (batch_ms * 1e-3) / 1e9
    out = {
        "metric": "batched IDR(2) time-to-solution (ms) for 512 2-D convection-diffusion systems, vs a per-system loop",
        "value": round(batch_ms, 2),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batch_ms, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.nsys} 2-D 5-point convection-diffusion systems, side floor(32*16^u), "
                               f"Pe U[1,100], IDR(2), tol 1e-8, FP64 CRS", "rows": int(ntot), "nnz": int(nnz),
                   "per_system_loop_ms": round(loop_ms, 2), "speedup_vs_loop": round(loop_ms / batch_ms, 2),
                   "eager_batch_ms": round(eager_ms, 2), "lanes_per_row": group,
                   "iterations_median": int(np.median(rb.iterations)), "iterations_max": int(rb.iterations.max()),
                   "flags": {str(k): int(v) for k, v in enumerate(np.bincount(rb.flag))},
                   "x_rel_diff_vs_loop": diff,
                   "mixed_precision": {"ms": round(mixed_ms, 2), "speedup_vs_loop": round(loop_ms / mixed_ms, 2),
                                       "flags": {str(k): int(v) for k, v in enumerate(np.bincount(rmx.flag))},
                                       "worst_relres": float(rmx.relres.max()), "outer_max": int(rmx.outer.max()),
                                       "inner_iterations_median": int(np.median(rmx.iterations)),
                                       "inner_tolerance": args.batch_inner_tol,
                                       "x_rel_diff_vs_fp64_batch": mixed_diff}},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batched solve (algorithmic bytes of each system's iterations)"},
        "cpu_baseline": None,
    }
    return out


def bench_mixed(args):
    """The 200^3 IDR(4)+Chebyshev(8) solve to 1e-8 with single-precision inner
    solves and FP64 residual replacement, against the FP64 solve of the same
    system (DIA storage unless --format crs)."""
    import numpy as np
    import torch

    import idrs_amd
    from idrs_amd import problems as P

    m, s, d = args.grid, args.s, args.degree
    crs, dia = P.convection_diffusion_3d(m, 50.0, reaction=1.0)
    n = crs.n
    b = torch.from_numpy(np.random.default_rng(0).uniform(-1.0, 1.0, n)).to("cuda:0")
    x = torch.empty_like(b)
    fmt = "crs" if args.format == "crs" else "dia"
    K = (idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val) if fmt == "crs"
         else idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals))
    if fmt == "crs" and not args.no_autotune:
        K.autotune()
    A = idrs_amd.Matrix(K)
    A.set_precon(1)
    idrs_amd.real_preconditioner(A, 2, d, foci=(0.15, 1.85))

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t) * 1e3 / args.steps

    rf, f_ms = timed(lambda: idrs_amd.idrs(A, b, s, tolerance=1e-8, maximum_iterations=20000, out=x))
    x_f64 = x.clone()
    rm, m_ms = timed(lambda: idrs_amd.idrs_mixed(A, b, s, tolerance=1e-8, maximum_iterations=20000,
                                                 inner_tolerance=args.inner_tol,
                                                 fp32_vectors=args.mixed_vectors == "fp32", out=x))
    # rm.relres is the operator's true FP64 residual of the mixed solution;
    # its distance from the FP64 solve's solution is reported beside it
    err = float(torch.linalg.vector_norm(x - x_f64) / torch.linalg.vector_norm(x_f64))
    print(f"# fp64: {f_ms:.1f} ms, {rf.iterations} its; mixed: {m_ms:.1f} ms, {rm.iterations} inner its, "
          f"{rm.outer} outer, relres {rm.relres:.3e}, |x - x_fp64|/|x_fp64| {err:.2e}", file=sys.stderr, flush=True)
    out = {
        "metric": "mixed-precision time-to-solution (ms) at rel. residual 1e-8 vs the FP64 solve",
        "value": round(m_ms, 2), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(m_ms, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f32 values+vectors" if args.mixed_vectors == "fp32" else "f32 values, f64 vectors")
                 + " inner / f64 residual", "data": "synthetic",
        "config": {"workload": f"{m}^3 convection-diffusion-reaction, IDR({s}) + Chebyshev({d}) with diagonal "
                               f"scaling, tol 1e-8 (true FP64 residual), storage {fmt.upper()}",
                   "fp64_ms": round(f_ms, 2), "fp64_iterations": rf.iterations, "speedup_vs_fp64": round(f_ms / m_ms, 2),
                   "flag": rm.flag, "inner_iterations": rm.iterations, "outer_steps": rm.outer, "relres": rm.relres,
                   "x_rel_diff_vs_fp64": err, "inner_tolerance": args.inner_tol},
        "roofline": None, "cpu_baseline": None,
    }
    return out


def bench_small(args):
    """The reference's own CPU-runnable case: plain IDR(4), no preconditioner,
    32^3 convection-diffusion (Pe 10), b = A*ones, Gaussian shadow space
    (seed 0), tol 1e-8, CRS.  At this size a cycle is launch-bound, so the
    solve is timed with each cycle replayed from a HIP graph and with eager
    launches; the oracle solves the same system on one host core."""
    import numpy as np
    import torch

    import idrs_amd
    from idrs_amd import problems as P

    m = args.small_grid
    crs, _ = P.convection_diffusion_3d(m, 10.0)
    n = crs.n
    bh = crs.matvec(np.ones(n))
    PT = P.gaussian_shadow(n, args.s, seed=0)
    A = idrs_amd.Matrix(idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val))
    b = torch.from_numpy(bh).to("cuda:0")
    PTd = torch.from_numpy(np.ascontiguousarray(PT)).to("cuda:0")
    x = torch.empty_like(b)

    def timed(graph):
        for _ in range(args.warmup):
            idrs_amd.idrs(A, b, args.s, tolerance=1e-8, PT_raw=PTd, graph=graph, out=x)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            r = idrs_amd.idrs(A, b, args.s, tolerance=1e-8, PT_raw=PTd, graph=graph, out=x)
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t) * 1e3 / args.steps

    rg, g_ms = timed(True)
    re_, e_ms = timed(False)
    xg = x.cpu().numpy()
    err = float(np.linalg.norm(xg - 1.0) / np.sqrt(n))
    print(f"# graph: {g_ms:.2f} ms ({rg.iterations} its); eager: {e_ms:.2f} ms ({re_.iterations} its); "
          f"rms error vs ones {err:.2e}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc

        Ao = orc.OracleMatrix(orc.CRS, n, crs=crs)
        t0 = time.perf_counter()
        _, fo, ro, io, _ = orc.idrs(Ao, bh, args.s, tolerance=1e-8, PT_raw=PT)
        dt = (time.perf_counter() - t0) * 1e3
        print(f"# oracle: {dt:.1f} ms, {io} its, flag {fo}, relres {ro:.3e}", file=sys.stderr, flush=True)
        cpu = {"value": round(dt, 2), "unit": "ms (time-to-solution)", "cores": 1, "kind": "port",
               "sample": f"oracle (scalar C restatement) full IDR({args.s}) solve of the same system: "
                         f"{io} iterations, flag {fo}"}
    out = {
        "metric": "time-to-solution (ms) at rel. residual 1e-8, the reference's CPU case",
        "value": round(g_ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(g_ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{m}^3 convection-diffusion (Pe 10), plain IDR({args.s}), b = A*ones, "
                               "Gaussian shadow space (seed 0), tol 1e-8, CRS",
                   "iterations": rg.iterations, "flag": rg.flag, "relres": rg.relres,
                   "eager_ms": round(e_ms, 3), "graph_speedup": round(e_ms / g_ms, 2)},
        "roofline": None, "cpu_baseline": cpu,
    }
    return out


def bench_adjoint(args):
    """Differentiable solve: random sparse nonsymmetric A (N = 1e6, 16
    off-diagonal entries per row at uniform random columns, N(0,1) values,
    diagonal 1.2 x row abs-sum), L = 0.5 ||x - x_t||^2, x = A^-1 b.  A step
    is one forward IDR(4) solve to 1e-10, the loss and its backward (adjoint
    IDR(4) on A^T to 1e-10 and the pattern gradient) through torch
    autograd.  The gradients wrt the nonzero values and b are checked
    against central differences along seeded random directions."""
    import numpy as np
    import torch

    import idrs_amd
    from idrs_amd import problems as P

    t0 = time.time()
    n = args.adj_n
    crs = P.random_dominant(n, k=16, seed=0)
    print(f"# system built: n={n} nnz={crs.nnz} ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    system = idrs_amd.SparseSystem(n, crs.row_ptr, crs.col_ind, crs.val)
    groups = None
    if not args.no_autotune:
        groups = (system.K.autotune(), system.KT.autotune())
    rng = np.random.default_rng(1)
    b = torch.tensor(rng.standard_normal(n), device=dev)
    xt = torch.tensor(rng.standard_normal(n), device=dev)
    vals0 = torch.tensor(crs.val, device=dev)
    tol = 1e-10

    def step():
        v = vals0.clone().requires_grad_(True)
        bb = b.clone().requires_grad_(True)
        x = idrs_amd.idrs_solve(system, v, bb, s=4, tolerance=tol)
        L = 0.5 * torch.sum((x - xt) ** 2)
        L.backward()
        return v.grad, bb.grad, float(L.detach())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        gv, gb, L0 = step()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    # forward alone, for the split
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    with torch.no_grad():
        rf = idrs_amd.idrs(system.A, b, 4, tolerance=tol, maximum_iterations=2 * n)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - t_start) * 1e3
    print(f"# forward+backward step: {step_ms:.1f} ms (forward alone {fwd_ms:.1f} ms, {rf.iterations} its); "
          f"lanes per row {groups}", file=sys.stderr, flush=True)

    # the finite differences need x far more accurately than 1e-10 (its
    # error enters divided by h): one step of FP64 iterative refinement,
    # x = x1 + A^-1 (b - A x1), with the residual from an FP64 product
    fd_flags = set()

    def solution(v, bb):
        with torch.no_grad():
            system.set_values(v)
            r1 = idrs_amd.idrs(system.A, bb.contiguous(), 4, tolerance=tol, maximum_iterations=4000)
            res = bb.cpu().numpy() - system.A * r1.x.cpu().numpy()
            r2 = idrs_amd.idrs(system.A, torch.from_numpy(res).to(dev), 4, tolerance=tol, maximum_iterations=4000)
            fd_flags.update((r1.flag, r2.flag))
            return r1.x + r2.x

    # central differences along seeded random directions (values scaled by
    # their magnitude): truncation O(h^2), solve error O(1e-20/h).  The loss
    # difference is formed as 0.5 (x+ - x-).(x+ + x- - 2 x_t), which does
    # not cancel two large losses against each other
    h, worst, worst_pair = 1e-5, 0.0, None
    vabs = vals0.abs()
    drng = np.random.default_rng(2)
    for k in range(args.fd_dirs):
        dv = torch.tensor(drng.standard_normal(crs.nnz), device=dev) * vabs
        db = torch.tensor(drng.standard_normal(n), device=dev)
        xp, xm = solution(vals0 + h * dv, b + h * db), solution(vals0 - h * dv, b - h * db)
        fd = float(0.5 * torch.dot(xp - xm, xp + xm - 2.0 * xt)) / (2 * h)
        an = float(torch.dot(gv, dv) + torch.dot(gb, db))
        err = abs(fd - an) / max(abs(an), 1e-30)
        if err >= worst:
            worst, worst_pair = err, (fd, an)
    system.set_values(vals0)
    print(f"# gradient vs central differences on {args.fd_dirs} directions: worst rel. err {worst:.2e} "
          f"(fd {worst_pair[0]:.9e}, autograd {worst_pair[1]:.9e}); flags of the difference solves {sorted(fd_flags)}",
          file=sys.stderr, flush=True)
    out = {
        "metric": "differentiable solve: forward + adjoint IDR(4) step time (ms) at 1e-10",
        "value": round(step_ms, 2), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_ms, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"random nonsymmetric N={n}, 16 off-diagonals per row, diagonal 1.2 row abs-sum, "
                               "L = 0.5||x - x_t||^2, forward IDR(4) and adjoint IDR(4) on A^T to 1e-10, "
                               "torch autograd", "nnz": int(crs.nnz), "forward_ms": round(fwd_ms, 2),
                   "forward_iterations": rf.iterations, "loss": L0, "fd_directions": args.fd_dirs,
                   "fd_worst_rel_err": worst, "fd_solve_flags": sorted(fd_flags), "lanes_per_row": groups},
        "roofline": None, "cpu_baseline": None,
    }
    return out


def bench_multishift(args):
    """Shifted Helmholtz systems (A - sigma_k I) x_k = b, k = 1..16, complex128:
    MSIDRS(4) and MSQMRIDR(8) (one Krylov space for all shifts) against a
    loop of single-shift IDR(4) solves (the operator's seed set to sigma_k)."""
    import numpy as np
    import torch

    import idrs_amd
    from idrs_amd import problems as P

    m = args.ms_grid
    if args.ms_problem == "helmholtz":
        crs, _ = P.helmholtz_3d(m)
        sig = P.helmholtz_shifts(m, nshift=args.nshift)
        bh = np.zeros(crs.n, dtype=np.complex128)
        bh[(m // 2) * (1 + m + m * m)] = 1.0
        workload = (f"{m}^3 h^2-scaled 7-point Laplacian, sigma_k = (2 pi f_k h)^2 (1 - 0.05 i), ppw 4..30, "
                    "b = centre unit vector, tol 1e-8, complex128")
    elif args.ms_problem == "cdr":
        crs, _ = P.convection_diffusion_3d(m, 10.0)
        h = 1.0 / (m + 1)
        sig = -np.arange(args.nshift) * 0.01 * 6.0 / h ** 2
        bh = np.random.default_rng(0).uniform(-1.0, 1.0, crs.n)
        workload = (f"{m}^3 convection-diffusion (Pe 10), sigma_k = -k * 0.01 * diag, k = 0..{args.nshift - 1}, "
                    "b ~ U[-1,1], tol 1e-8, FP64")
    else:
        crs, _ = P.helmholtz_3d(m)
        sig = -np.arange(args.nshift) * 0.01
        bh = np.random.default_rng(0).uniform(-1.0, 1.0, crs.n)
        workload = (f"{m}^3 h^2-scaled 7-point Laplacian, sigma_k = -0.01 k, k = 0..{args.nshift - 1}, "
                    f"Chebyshev({args.ms_degree}) polynomial preconditioner on [0.05, 11.95] (MS_PARAMETERS "
                    "shifts for the multishift solvers, a per-shift polynomial for the loop), b ~ U[-1,1], "
                    "tol 1e-8, FP64")
    n = crs.n
    b = torch.from_numpy(bh).to("cuda:0")
    K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
    group = K.autotune() if not args.no_autotune else None
    A = idrs_amd.Matrix(K)
    tol, maxit = 1e-8, args.ms_maxit
    poly = args.ms_problem == "laplacian"
    foci = (0.05, 11.95)
    if poly:
        idrs_amd.real_preconditioner(A, 2, args.ms_degree, foci=foci)
        sig_solve, alpha_s = idrs_amd.ms_parameters(A, sig)
    else:
        sig_solve = sig
    torch.cuda.synchronize()

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, (time.perf_counter() - t) * 1e3

    def ms(fn, s_):
        r = fn(A, b, sig_solve, s_, tolerance=tol, maximum_iterations=maxit)
        return (r, idrs_amd.ms_scaleback(A, alpha_s, r.x)) if poly else (r, r.x)

    (rm, xm), ms_idrs = timed(lambda: ms(idrs_amd.msidrs, 4))
    print(f"# msidrs(4): {ms_idrs:.1f} ms, {rm.iterations} its, flag {rm.flag}, max relres {rm.relres.max():.2e}",
          file=sys.stderr, flush=True)
    (rq, xq), ms_qmr = timed(lambda: ms(idrs_amd.msqmridr, 8))
    print(f"# msqmridr(8): {ms_qmr:.1f} ms, {rq.iterations} its, flag {rq.flag}, max relres {rq.relres.max():.2e}",
          file=sys.stderr, flush=True)
    # shifts still running at each MSQMRIDR iteration (an untimed rerun with
    # the residual history): the per-iteration update reads and writes 8N
    # (1 + active (s+4)) bytes
    rr = idrs_amd.msqmridr(A, b, sig_solve, 8, tolerance=tol, maximum_iterations=maxit, resvec=True)
    hist = np.asarray(rr.resvec).reshape(len(sig_solve), -1)
    tolb = tol * float(torch.linalg.vector_norm(b))
    active = [int((hist[:, it - 1] >= tolb).sum()) for it in range(1, rr.iterations + 1)]
    print(f"# msqmridr(8) active shifts per iteration: {active}", file=sys.stderr, flush=True)
    # per-shift loop: IDR(4) on K - sigma_k I via the operator's seed (with the
    # polynomial rebuilt for the shifted operator, and its scale-back)
    loop_its, x = [], torch.empty_like(b)
    xs_loop = []
    torch.cuda.synchronize()
    t = time.perf_counter()
    for sk in sig:
        A.set_seed(complex(sk) if np.iscomplexobj(sig) else float(sk))
        if poly:
            idrs_amd.real_preconditioner(A, 2, args.ms_degree, foci=foci)
        r = idrs_amd.idrs(A, b, 4, tolerance=tol, maximum_iterations=maxit, out=x)
        if poly:
            xs_loop.append(idrs_amd.ms_scaleback(A, A.alpha.reshape(-1, 1), x.reshape(-1, 1)))
        loop_its.append(r.iterations)
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t) * 1e3
    A.set_seed(0j if np.iscomplexobj(sig) else 0.0)
    # true residuals of (K - sigma_k I) x_k = b for the multishift solutions
    # and for the per-shift loop's (recomputed on the host in FP64)
    Kt = crs.to_scipy()
    worst, worst_loop = 0.0, 0.0
    for name, xx in (("msidrs", xm), ("msqmridr", xq)):
        xh = xx.cpu().numpy()
        for k, sk in enumerate(sig):
            res = bh - (Kt @ xh[:, k] - sk * xh[:, k])
            worst = max(worst, np.linalg.norm(res) / np.linalg.norm(bh))
    if poly:
        for k, sk in enumerate(sig):
            xl = xs_loop[k].cpu().numpy()[:, 0]
            res = bh - (Kt @ xl - sk * xl)
            worst_loop = max(worst_loop, np.linalg.norm(res) / np.linalg.norm(bh))
    print(f"# worst true residual: multishift solutions {worst:.2e}, per-shift loop {worst_loop:.2e}",
          file=sys.stderr, flush=True)
    print(f"# per-shift IDR(4) loop: {loop_ms:.1f} ms, iterations {loop_its}", file=sys.stderr, flush=True)
    ok = [(t_, name) for t_, name, f_ in ((ms_idrs, "msidrs4", rm.flag), (ms_qmr, "msqmridr8", rq.flag)) if f_ == 0]
    best = min(ok)[0] if ok else min(ms_idrs, ms_qmr)
    out = {
        "metric": f"multishift time-to-solution (ms), {args.nshift} shifted systems, vs a per-shift IDR(4) loop",
        "value": round(best, 2), "unit": "ms", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(best, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128" if np.iscomplexobj(sig) else "f64", "data": "synthetic",
        "config": {"workload": workload, "converged_solvers": [name for _, name in ok],
                   "msidrs4_ms": round(ms_idrs, 2), "msidrs4_iterations": rm.iterations, "msidrs4_flag": rm.flag,
                   "msqmridr8_ms": round(ms_qmr, 2), "msqmridr8_iterations": rq.iterations, "msqmridr8_flag": rq.flag,
                   "per_shift_loop_ms": round(loop_ms, 2), "per_shift_iterations": loop_its,
                   "speedup_vs_loop": round(loop_ms / best, 2), "lanes_per_row": group,
                   "worst_true_relres_multishift": worst, "worst_true_relres_loop": worst_loop},
        "roofline": None, "cpu_baseline": None,
    }
    return out


def run_forms(args):
    """The north-star forms beside the headline solve, one timed step each
    after one warm-up: the batched solve (and its mixed-precision variant)
    against the per-system loop, the multishift solvers against the
    per-shift loop, the mixed-precision 200^3 solve against the FP64 one, and
    the differentiable solve (step time and the worst finite-difference
    error of its gradients).  A form that fails reports its error."""
    fargs = argparse.Namespace(**vars(args))
    fargs.steps, fargs.warmup = 1, 1
    fargs.format = "dia"
    summary = {}

    def run(name, fn, pick):
        t0 = time.time()
        print(f"# form {name} ...", file=sys.stderr, flush=True)
        try:
            out = fn(fargs)
            summary[name] = pick(out)
        except Exception as e:  # report, keep the headline line
            summary[name] = {"error": f"{type(e).__name__}: {e}"}
        summary[name]["wall_s"] = round(time.time() - t0, 1)

    def batched(o):
        c = o["config"]
        return {"ms": o["value"], "per_system_loop_ms": c["per_system_loop_ms"], "speedup_vs_loop": c["speedup_vs_loop"],
                "flags": c["flags"], "mixed_ms": c["mixed_precision"]["ms"],
                "mixed_speedup_vs_loop": c["mixed_precision"]["speedup_vs_loop"],
                "mixed_worst_relres": c["mixed_precision"]["worst_relres"], "roofline_frac": o["roofline"]["frac"]}

    def multishift(o):
        c = o["config"]
        return {"ms": o["value"], "per_shift_loop_ms": c["per_shift_loop_ms"], "speedup_vs_loop": c["speedup_vs_loop"],
                "msidrs4_ms": c["msidrs4_ms"], "msqmridr8_ms": c["msqmridr8_ms"],
                "worst_true_relres_multishift": c["worst_true_relres_multishift"],
                "worst_true_relres_loop": c["worst_true_relres_loop"], "workload": c["workload"]}

    def mixed(o):
        c = o["config"]
        return {"ms": o["value"], "fp64_ms": c["fp64_ms"], "speedup_vs_fp64": c["speedup_vs_fp64"],
                "relres": c["relres"], "x_rel_diff_vs_fp64": c["x_rel_diff_vs_fp64"], "dtype": o["dtype"]}

    def adjoint(o):
        c = o["config"]
        return {"step_ms": o["value"], "forward_ms": c["forward_ms"], "fd_directions": c["fd_directions"],
                "fd_worst_rel_err": c["fd_worst_rel_err"], "fd_solve_flags": c["fd_solve_flags"]}

    run("batched", bench_batched, batched)
    run("multishift", bench_multishift, multishift)
    run("mixed", bench_mixed, mixed)
    run("adjoint", bench_adjoint, adjoint)
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=200)
    ap.add_argument("--s", type=int, default=4)
    ap.add_argument("--degree", type=int, default=8)
    ap.add_argument("--format", choices=["both", "crs", "dia"], default="both",
                    help="storage: both (headline DIA, then CRS), or one of them")
    ap.add_argument("--form", choices=["solve", "batched", "mixed", "multishift", "small", "adjoint"], default="solve",
                    help="solve: the 200^3 IDR(4)+Chebyshev solve (default); batched: many small systems")
    ap.add_argument("--nsys", type=int, default=512, help="systems of the batched form")
    ap.add_argument("--small-grid", type=int, default=32, help="grid side of the small form")
    ap.add_argument("--adj-n", type=int, default=1000000, help="rows of the adjoint form")
    ap.add_argument("--fd-dirs", type=int, default=32, help="finite-difference directions of the adjoint form")
    ap.add_argument("--batch-maxit", type=int, default=4000)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default product layout")
    ap.add_argument("--no-loop", action="store_true", help="batched form: skip the per-system loop (profiling)")
    ap.add_argument("--inner-tol", type=float, default=1e-9, help="inner tolerance of the mixed form")
    ap.add_argument("--batch-inner-tol", type=float, default=1e-5,
                    help="inner tolerance of the batched form's mixed-precision solve")
    ap.add_argument("--mixed-vectors", choices=["fp64", "fp32"], default="fp64",
                    help="inner vectors of the mixed form (matrix values are FP32 either way)")
    ap.add_argument("--ms-grid", type=int, default=128, help="grid side of the multishift form")
    ap.add_argument("--nshift", type=int, default=16)
    ap.add_argument("--ms-problem", choices=["helmholtz", "cdr", "laplacian"], default="laplacian")
    ap.add_argument("--ms-degree", type=int, default=8, help="Chebyshev degree of the laplacian multishift problem")
    ap.add_argument("--ms-maxit", type=int, default=20000)
    ap.add_argument("--cpu-sample-iters", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="headline solve only (no batched/multishift/... forms)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain device copy of known size (scales rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): sharding the rows over GPUs "
              "(halo exchange per product, one all-reduce per reduction point) is designed, not built or "
              "measured.", file=sys.stderr)
        sys.exit(2)
    forms = {"batched": bench_batched, "mixed": bench_mixed, "multishift": bench_multishift, "small": bench_small,
             "adjoint": bench_adjoint}
    if args.form in forms:
        out = forms[args.form](args)
        if out is not None:
            print(json.dumps(out))
        return

    import numpy as np
    import torch

    import idrs_amd
    from idrs_amd import problems as P

    t0 = time.time()
    m, s, d = args.grid, args.s, args.degree
    foci = (0.15, 1.85)
    crs, dia = P.convection_diffusion_3d(m, 50.0, reaction=1.0)
    n, nnz = crs.n, crs.nnz
    rng = np.random.default_rng(0)
    b_h = rng.uniform(-1.0, 1.0, n)
    PT_h = P.gaussian_shadow(n, s, seed=0)
    print(f"# system built: n={n} nnz={nnz} ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    b = torch.from_numpy(b_h).to(dev)
    PT = torch.from_numpy(np.ascontiguousarray(PT_h)).to(dev)
    x = torch.empty_like(b)
    formats = ["dia", "crs"] if args.format == "both" else [args.format]

    def run(fmt):
        if fmt == "crs":
            K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
        else:
            K = idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals)
        # product layout chosen by timing: lanes per row (1: sliced-ELL copy)
        # and row slabs of one-thread-per-row products
        if not args.no_autotune:
            K.autotune()
        group, row_slabs = K.layout()
        A = idrs_amd.Matrix(K)
        A.set_precon(1)
        idrs_amd.real_preconditioner(A, 2, d, foci=foci)
        # value entries one product reads (the scaled CDS operator does not
        # store its all-ones main diagonal)
        nvals, unit = A.stored_values()
        torch.cuda.synchronize()

        def step(timed):
            return idrs_amd.idrs(A, b, s, tolerance=1e-8, maximum_iterations=20000, PT_raw=PT, out=x,
                                 time_spmv=timed)

        for i in range(args.warmup):
            r = step(False)
            print(f"# {fmt} warmup {i}: iterations={r.iterations} flag={r.flag} relres={r.relres:.3e} "
                  f"solve={r.solve_ms:.1f} ms", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        results = []
        for i in range(args.steps):
            r = step(False)
            results.append(r)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        ms_per_step = wall * 1e3 / args.steps
        for i, r in enumerate(results):
            print(f"# {fmt} step {i}: iterations={r.iterations} flag={r.flag} relres={r.relres:.3e} "
                  f"solve={r.solve_ms:.1f} ms", file=sys.stderr, flush=True)
        last = results[-1]
        assert last.flag == 0, f"solve did not converge: flag {last.flag}"
        # the SpMV roofline from one more solve, outside the timed steps, with
        # HIP events on the solver's stream around each operator application
        results = [step(True)]
        print(f"# {fmt} event-timed solve: spmv={results[0].spmv_ms:.1f} ms / {results[0].spmv_launches} launches",
              file=sys.stderr, flush=True)
        # roofline of the SpMV kernel over the timed launches
        tot_bytes, launches = solve_spmv_bytes(n, nnz, s, d, last.iterations, s)
        if fmt == "dia":
            tot_bytes = tot_bytes - launches * (12 * nnz + 4 * (n + 1)) + launches * 8 * nvals
        spmv_ms = sum(r.spmv_ms for r in results) / len(results)
        achieved = tot_bytes / (spmv_ms * 1e-3) / 1e9
        it_bytes = iteration_bytes(n, nnz, s, d)
        if fmt == "dia":
            it_bytes = it_bytes - (d + 1) * (12 * nnz + 4 * (n + 1)) + (d + 1) * 8 * nvals
        solve_gbs = last.iterations * it_bytes / (ms_per_step * 1e-3) / 1e9
        del A, K
        return {
            "ms_per_step": ms_per_step, "iterations": last.iterations, "relres": last.relres,
            "roofline": {
                "bound": "hbm",
                "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS,
                "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(fmt),
                "kernel": f"k_spmv_{fmt}",
                "bytes_per_launch": round(tot_bytes / launches),
                "stored_values": nvals, "implicit_unit_diagonal": unit, "lanes_per_row": group,
                "row_slabs": row_slabs,
                "avg_launch_us": round(spmv_ms * 1e3 / launches, 2),
            },
            "solve_gbs_model": round(solve_gbs, 1),
            "solve_frac_hbm_model": round(solve_gbs / HBM_PEAK_GBS, 4),
        }

    res = {fmt: run(fmt) for fmt in formats}
    head = res[formats[0]]

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline sample ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(dia, b_h, PT_h, s, d, foci, args.cpu_sample_iters, head["iterations"])

    forms = None
    if not args.no_forms:
        forms = run_forms(args)

    config = {
        "workload": f"{m}^3 convection-diffusion-reaction (Pe 50, reaction 1), FP64, IDR({s}) + "
                    f"Chebyshev({d}) with diagonal scaling, tol 1e-8; storage {formats[0].upper()} "
                    "(offset-list CDS)" if formats[0] == "dia" else
                    f"{m}^3 convection-diffusion-reaction (Pe 50, reaction 1), FP64, IDR({s}) + "
                    f"Chebyshev({d}) with diagonal scaling, tol 1e-8; storage CRS",
        "n": n, "nnz": nnz, "s": s, "degree": d, "format": formats[0],
        "iterations": head["iterations"], "relres": head["relres"],
        "solve_gbs_model": head["solve_gbs_model"],
        "solve_frac_hbm_model": head["solve_frac_hbm_model"],
    }
    for fmt in formats[1:]:
        r = res[fmt]
        config[fmt] = {"ms_per_step": round(r["ms_per_step"], 2), "iterations": r["iterations"],
                       "relres": r["relres"], "roofline": r["roofline"],
                       "solve_frac_hbm_model": r["solve_frac_hbm_model"]}
    out = {
        "metric": "time-to-solution (ms) at rel. residual 1e-8 + SpMV/vector HBM GB/s vs peak",
        "value": round(head["ms_per_step"], 2),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": config,
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
