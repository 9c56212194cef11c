"""GPU checks of the differentiable solve: the transposed CRS matrix, value
updates, the pattern gradient, the adjoint IDR(s) solve against the oracle on
an explicitly transposed matrix, and torch autograd gradients against central
finite differences and the closed form dL/dA_ij = -lambda_i x_j.

The reference has no adjoint; these pins are the mathematics: A^T from the
definition, IDRS on A^T from the oracle run on the explicit transpose, and
gradients from finite differences (error O(eps^2) + solve error / eps).
"""
import ctypes

import numpy as np
import pytest

import idrs_amd
import orc
from idrs_amd import _lib
from idrs_amd import problems as P

from helpers import oracle_crs, rel

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _transpose(K):
    h = ctypes.c_void_p()
    _lib.check(_lib.load().idrs_crs_transpose(K.handle, ctypes.byref(h)))
    return idrs_amd.UserMatrix(h.value, False)


def test_transpose_spmv_equals_explicit_transpose(gpu):
    crs = P.random_dominant(3000, k=9, seed=4)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    KT = _transpose(K)
    x = np.random.default_rng(1).standard_normal(crs.n)
    w = KT * x
    # the transposed rows list entries by increasing original row, as the
    # explicit transpose stores them: the products agree to rounding
    ref = oracle_crs(crs.transpose()).user_mv(x)
    dense = crs.to_scipy().T @ x
    assert np.allclose(w, ref, rtol=0, atol=64 * EPS * np.abs(crs.to_scipy().T).dot(np.abs(x)).max())
    assert rel(w, dense) < 1e-14


def test_set_values_and_refresh(gpu):
    crs = P.random_dominant(2000, k=5, seed=2)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    KT = _transpose(K)
    new = crs.val * np.random.default_rng(3).uniform(0.5, 1.5, crs.nnz)
    lib = _lib.load()
    _lib.check(lib.idrs_user_matrix_set_values_f64(K.handle, new.ctypes.data, 0))
    _lib.check(lib.idrs_crs_transpose_refresh(K.handle, KT.handle))
    crs2 = P.Crs(crs.n, crs.row_ptr, crs.col_ind, new)
    x = np.random.default_rng(4).standard_normal(crs.n)
    assert rel(K * x, crs2.matvec(x)) < 1e-14
    assert rel(KT * x, crs2.to_scipy().T @ x) < 1e-14
    assert np.array_equal(idrs_amd.real_diagonal(K), crs2.to_scipy().diagonal())


def test_pattern_grad_closed_form(gpu):
    crs = P.random_dominant(1500, k=7, seed=5)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    rng = np.random.default_rng(6)
    lam, x = rng.standard_normal(crs.n), rng.standard_normal(crs.n)
    g = np.empty(crs.nnz)
    _lib.check(_lib.load().idrs_crs_pattern_grad_f64(K.handle, lam.ctypes.data, x.ctypes.data, g.ctypes.data, 0))
    rows = np.repeat(np.arange(crs.n), np.diff(crs.row_ptr))
    assert np.array_equal(g, -(lam[rows] * x[crs.col_ind - 1]))


def test_adjoint_solve_matches_oracle(gpu):
    crs = P.random_dominant(4000, k=16, seed=7)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    AT = idrs_amd.Matrix(_transpose(K))
    g = np.random.default_rng(8).standard_normal(crs.n)
    ATo = oracle_crs(crs.transpose())
    r = idrs_amd.idrs(AT, g, 4, tolerance=0.0, maximum_iterations=12)
    o = orc.idrs(ATo, g, 4, tolerance=0.0, maximum_iterations=12)
    assert rel(r.x, o[0]) < 1e-10
    r = idrs_amd.idrs(AT, g, 4, tolerance=1e-10, maximum_iterations=500)
    o = orc.idrs(ATo, g, 4, tolerance=1e-10, maximum_iterations=500)
    assert r.flag == o[1] == 0 and abs(r.iterations - o[3]) <= 2


def test_autograd_gradients(gpu):
    """L = w . x(values, b): autograd against the closed form (dense solves)
    and against central differences along random directions."""
    import torch

    dev = torch.device("cuda:0")
    crs = P.random_dominant(2000, k=16, seed=9)
    sysm = idrs_amd.SparseSystem(crs.n, crs.row_ptr, crs.col_ind, crs.val)
    rng = np.random.default_rng(10)
    b_h, w_h = rng.standard_normal(crs.n), rng.standard_normal(crs.n)
    vals = torch.tensor(crs.val, device=dev, requires_grad=True)
    b = torch.tensor(b_h, device=dev, requires_grad=True)
    w = torch.tensor(w_h, device=dev)
    x = idrs_amd.idrs_solve(sysm, vals, b, s=4, tolerance=1e-12)
    L = torch.dot(w, x)
    L.backward()
    gv, gb = vals.grad.cpu().numpy(), b.grad.cpu().numpy()
    # closed form
    Ad = crs.to_scipy().toarray()
    xd = np.linalg.solve(Ad, b_h)
    lam = np.linalg.solve(Ad.T, w_h)
    rows = np.repeat(np.arange(crs.n), np.diff(crs.row_ptr))
    assert rel(x.detach().cpu().numpy(), xd) < 1e-10
    assert rel(gb, lam) < 1e-10
    assert rel(gv, -(lam[rows] * xd[crs.col_ind - 1])) < 1e-10

    # central differences along random directions of (values, b)
    def loss(v, bb):
        with torch.no_grad():
            return float(torch.dot(w, idrs_amd.idrs_solve(sysm, v, bb, s=4, tolerance=1e-12)))

    h = 1e-4
    for k in range(4):
        dv = torch.tensor(rng.standard_normal(crs.nnz) * np.abs(crs.val), device=dev)
        db = torch.tensor(rng.standard_normal(crs.n), device=dev)
        fd = (loss(vals.detach() + h * dv, b.detach() + h * db) - loss(vals.detach() - h * dv, b.detach() - h * db)) / (2 * h)
        an = float(np.dot(gv, dv.cpu().numpy()) + np.dot(gb, db.cpu().numpy()))
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3), (k, fd, an)


def test_set_values_waits_for_torch(gpu):
    """Values computed by torch just before set_values (still queued on
    torch's stream) reach the matrix: set_values waits for them, and the
    product afterwards is the one of the new values (1e6 rows, so the torch
    kernels take long enough to matter)."""
    import torch

    dev = torch.device("cuda:0")
    crs = P.random_dominant(1000000, k=16, seed=3)
    sysm = idrs_amd.SparseSystem(crs.n, crs.row_ptr, crs.col_ind, crs.val)
    sysm.K.set_group(1)
    v0 = torch.tensor(crs.val, device=dev)
    x = np.random.default_rng(6).standard_normal(crs.n)
    for k, scale in enumerate((0.5, 2.0, -1.5, 0.75)):
        v = v0 * scale + float(k)      # queued on torch's stream
        sysm.set_values(v)
        del v
        c2 = P.Crs(crs.n, crs.row_ptr, crs.col_ind, crs.val * scale + float(k))
        assert rel(sysm.K * x, c2.matvec(x)) < 1e-14, k


def test_sparse_system_held_values(gpu, monkeypatch):
    """SparseSystem.set_values copies nothing for the tensor it already
    holds, and copies again after an in-place change of that tensor (its
    version counter moves) or for another tensor; the transpose follows on
    refresh_transpose.  Both matrices use the sliced copy, whose values are
    gathered again into the buffer they already have."""
    import torch

    import idrs_amd.autograd as AG

    real = _lib.load()
    calls = []

    class Counting:
        def __getattr__(self, name):
            f = getattr(real, name)
            if name != "idrs_user_matrix_set_values_f64":
                return f
            return lambda *a: (calls.append(name), f(*a))[1]

    monkeypatch.setattr(AG._lib, "load", lambda: Counting())
    dev = torch.device("cuda:0")
    crs = P.random_dominant(5000, k=7, seed=8)
    sysm = idrs_amd.SparseSystem(crs.n, crs.row_ptr, crs.col_ind, crs.val)
    sysm.K.set_group(1)
    sysm.KT.set_group(1)
    x = np.random.default_rng(9).standard_normal(crs.n)

    def holds(vals, ncalls):
        c = P.Crs(crs.n, crs.row_ptr, crs.col_ind, vals)
        assert len(calls) == ncalls
        assert rel(sysm.K * x, c.matvec(x)) < 1e-14
        sysm.refresh_transpose()
        assert rel(sysm.KT * x, c.to_scipy().T @ x) < 1e-14

    v = torch.tensor(crs.val * 2.0, device=dev)
    sysm.set_values(v)
    holds(crs.val * 2.0, 1)
    sysm.set_values(v)                  # held and unchanged: no copy
    holds(crs.val * 2.0, 1)
    v.mul_(-0.5)                        # in place: the version counter moves
    sysm.set_values(v)
    holds(crs.val * -1.0, 2)
    w = v + 3.0
    sysm.set_values(w)
    holds(crs.val * -1.0 + 3.0, 3)
    sysm.set_values(v)
    holds(crs.val * -1.0, 4)
