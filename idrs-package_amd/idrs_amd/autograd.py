"""Differentiable IDR(s) solve for torch: x = A(values)^-1 b.

The sparsity pattern of A is fixed (CRS); its values and the right-hand
side are torch tensors on the GPU.  Forward: IDRS on A.  Backward, for an
incoming dL/dx: the adjoint system A^T lambda = dL/dx is solved by IDRS on
the transposed CRS matrix (kept on the device, values refreshed through a
permutation), then

    dL/db      = lambda
    dL/dA_ij   = -lambda_i x_j      for every stored entry (i, j)

computed by one kernel over the pattern.  All of it runs in libidrs_hip.so;
torch supplies device memory and the autograd graph.
"""
import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import check
from .api import Matrix, UserMatrix, crs_format, idrs


class SparseSystem:
    """A square CRS pattern (1-based row_ptr/col_ind, as crs_format takes
    them) whose values are set per solve, with its transpose for the
    adjoint solve."""

    def __init__(self, n, row_ptr, col_ind, values):
        values = np.ascontiguousarray(values, dtype=np.float64)
        self.n = int(n)
        self.nnz = int(len(values))
        self.K = crs_format(n, n, row_ptr, col_ind, rval=values)
        h = ctypes.c_void_p()
        check(_lib.load().idrs_crs_transpose(self.K.handle, ctypes.byref(h)))
        self.KT = UserMatrix(h.value, False)
        self.A = Matrix(self.K)
        self.AT = Matrix(self.KT)
        # the tensor whose values K holds (a weak reference, its version
        # counter and storage address), and whether KT holds their transpose
        self._held = None
        self._t_fresh = True

    def set_values(self, values):
        """values: contiguous float64 CUDA tensor of nnz entries (CRS order).
        Setting again the tensor object K already holds, unmodified since
        (same version counter, same storage), copies nothing: an address
        alone would not do, since a freed tensor's memory is handed to the
        next one.  The transpose is refreshed when the adjoint solve needs
        it (refresh_transpose)."""
        import torch

        if values.numel() != self.nnz:
            raise ValueError(f"expected {self.nnz} values, got {values.numel()}")
        if self._held is not None:
            ref, ver, ptr = self._held
            if ref() is values and values._version == ver and values.data_ptr() == ptr:
                return
        v = values.detach().contiguous()
        # the library reads v on its own stream: torch's pending work on v
        # (and on anything else the call reads) must have finished
        torch.cuda.current_stream().synchronize()
        self._held = None
        self._t_fresh = False
        check(_lib.load().idrs_user_matrix_set_values_f64(self.K.handle, ctypes.c_void_p(v.data_ptr()), 1))
        if v.data_ptr() == values.data_ptr():
            self._held = (weakref.ref(values), values._version, values.data_ptr())

    def refresh_transpose(self):
        """KT = K^T for the values K holds (one permutation gather)."""
        if not self._t_fresh:
            check(_lib.load().idrs_crs_transpose_refresh(self.K.handle, self.KT.handle))
            self._t_fresh = True

    def pattern_grad(self, lam, x, out):
        import torch

        torch.cuda.current_stream().synchronize()
        check(_lib.load().idrs_crs_pattern_grad_f64(self.K.handle, ctypes.c_void_p(lam.data_ptr()),
                                                    ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                                    1))


def _solve(A, rhs, s, tol, maxit, what):
    import torch

    out = torch.empty_like(rhs)
    r = idrs(A, rhs.contiguous(), s, tolerance=tol, maximum_iterations=maxit, out=out)
    if r.flag not in (0,):
        raise RuntimeError(f"{what} solve did not converge: flag {r.flag}, relres {r.relres:.3e}, "
                           f"{r.iterations} iterations")
    return out


def _function():
    import torch

    class IdrsSolveFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, values, b, system, s, tol, maxit):
            system.set_values(values)
            x = _solve(system.A, b.detach(), s, tol, maxit, "forward")
            ctx.system, ctx.s, ctx.tol, ctx.maxit = system, s, tol, maxit
            ctx.save_for_backward(values, x)
            return x

        @staticmethod
        def backward(ctx, gx):
            values, x = ctx.saved_tensors
            system = ctx.system
            # the pattern's values may have been reset by a later forward
            system.set_values(values)
            system.refresh_transpose()
            lam = _solve(system.AT, gx.detach().contiguous(), ctx.s, ctx.tol, ctx.maxit, "adjoint")
            gvals = None
            if ctx.needs_input_grad[0]:
                gvals = torch.empty_like(values)
                system.pattern_grad(lam, x, gvals)
            gb = lam if ctx.needs_input_grad[1] else None
            return gvals, gb, None, None, None, None

    return IdrsSolveFn


_FN = None


def idrs_solve(system, values, b, s=4, tolerance=1e-10, maximum_iterations=None):
    """x = A(values)^-1 b by IDR(s), differentiable in values and b.
    values: float64 CUDA tensor (nnz, CRS order of system's pattern);
    b: float64 CUDA tensor (n)."""
    global _FN
    if _FN is None:
        _FN = _function()
    maxit = maximum_iterations if maximum_iterations is not None else max(1000, 2 * system.n)
    return _FN.apply(values, b, system, int(s), float(tolerance), int(maxit))
