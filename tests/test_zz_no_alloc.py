"""The allocation-free IDRS entry point (idrs_idrs_f64_noalloc) against the
ordinary one (idrs_idrs_f64 with device pointers), through the C ABI.

Both run the same kernels in the same order on the same grids, so their
results must be bit-identical: x, the flag, the iteration count, relres and
the residual history.  Two calls in a row share the scratch region, the
second with a smaller system, so what the first left behind must not leak
into the second.  A system whose workspace exceeds the region is refused
with an argument error before anything runs.
"""
import ctypes

import numpy as np
import pytest

import idrs_amd
from idrs_amd import _lib
from idrs_amd import problems as P

pytestmark = pytest.mark.gpu


def _opts(tol, maxit, PT=None, x0=None, resvec=None):
    o = _lib.idrs_opts_t()
    o.has_tolerance, o.tolerance = 1, tol
    o.has_maximum_iterations, o.maximum_iterations = 1, maxit
    o.device_pointers = 1
    if PT is not None:
        o.PT_raw = PT.data_ptr()
    if x0 is not None:
        o.x0 = x0.data_ptr()
    if resvec is not None:
        o.resvec, o.resvec_len = resvec.ctypes.data, len(resvec)
    return o


def _both(A, b, s, tol, maxit, PT=None, x0=None):
    import torch

    lib = _lib.load()
    out = []
    for f in (lib.idrs_idrs_f64, lib.idrs_idrs_f64_noalloc):
        x = torch.full_like(b, np.nan)
        rv = np.zeros(maxit + 1)
        o = _opts(tol, maxit, PT, x0, rv)
        st = _lib.idrs_stats_t()
        torch.cuda.synchronize()
        rc = f(A.handle, int(b.shape[0]), ctypes.c_void_p(b.data_ptr()), s, ctypes.byref(o),
               ctypes.c_void_p(x.data_ptr()), ctypes.byref(st))
        assert rc == 0, _lib.load().idrs_last_error().decode()
        torch.cuda.synchronize()
        out.append((x.cpu().numpy(), st.flag, st.iterations, st.relres, rv[: st.iterations + 1]))
    return out


def _check_equal(a, b):
    xa, fa, ia, ra, rva = a
    xb, fb, ib, rb, rvb = b
    assert fa == fb and ia == ib and ra == rb
    assert np.array_equal(xa, xb)
    assert np.array_equal(rva, rvb)


def test_noalloc_matches_ordinary_entry(gpu):
    import torch

    dev = torch.device("cuda:0")
    # first call: 24^3 convection-diffusion, diagonal scaling + Chebyshev(4), IDR(4)
    crs, _ = P.convection_diffusion_3d(24, 30.0, reaction=1.0)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    A = idrs_amd.Matrix(K, which_precon=1)
    idrs_amd.real_preconditioner(A, 2, 4, foci=[0.15, 1.85])
    b = torch.tensor(np.random.default_rng(1).uniform(-1, 1, crs.n), device=dev)
    PT = torch.tensor(P.gaussian_shadow(crs.n, 4, seed=0), device=dev)
    print(f"noalloc call 1: n={crs.n} s=4", flush=True)
    first = _both(A, b, 4, 1e-9, 600, PT=PT)
    _check_equal(*first)
    assert first[0][1] == 0 and first[0][3] <= 1e-9
    # second call, smaller: 10^3 on the sliced CRS copy, IDR(8) from x0, the
    # library's own shadow-space stream
    crs2, _ = P.convection_diffusion_3d(10, 5.0, reaction=2.0)
    K2 = idrs_amd.crs_format(crs2.n, crs2.n, crs2.row_ptr, crs2.col_ind, rval=crs2.val)
    K2.set_group(1)
    A2 = idrs_amd.Matrix(K2, which_precon=1)
    b2 = torch.tensor(np.random.default_rng(2).standard_normal(crs2.n), device=dev)
    x02 = torch.tensor(np.random.default_rng(3).standard_normal(crs2.n) * 1e-3, device=dev)
    print(f"noalloc call 2: n={crs2.n} s=8", flush=True)
    second = _both(A2, b2, 8, 1e-10, 400, x0=x02)
    _check_equal(*second)
    assert second[0][1] == 0
    # the true residual of the returned x, recomputed on the host
    x2 = second[1][0]
    Ks = crs2.to_scipy()
    d = np.diag(Ks.toarray())
    # the operator is K D^-1, so y = x solves (K D^-1) y = b
    assert np.linalg.norm(b2.cpu().numpy() - Ks @ (x2 / d)) / np.linalg.norm(b2.cpu().numpy()) <= 1e-10


def test_noalloc_refuses_too_large(gpu):
    """A 2e6-row IDR(4) solve needs ~0.3 GB of workspace, above the 64 MiB
    region: an argument error with a message, and x untouched."""
    import torch

    n = 2_000_000
    crs, _, _ = P.toeplitz_1d(n)
    K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
    A = idrs_amd.Matrix(K)
    b = torch.ones(n, dtype=torch.float64, device="cuda:0")
    x = torch.full_like(b, 7.0)
    o = _opts(1e-8, 50)
    st = _lib.idrs_stats_t()
    lib = _lib.load()
    torch.cuda.synchronize()
    rc = lib.idrs_idrs_f64_noalloc(A.handle, n, ctypes.c_void_p(b.data_ptr()), 4, ctypes.byref(o),
                                   ctypes.c_void_p(x.data_ptr()), ctypes.byref(st))
    assert rc == -1
    assert b"workspace" in lib.idrs_last_error()
    assert bool((x == 7.0).all())
