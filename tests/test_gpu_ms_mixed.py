"""The multishift solvers in complex64 with FP64 refinement (the 3-D
Helmholtz workload's complex64 + FP64-refinement variant; no reference
counterpart -- the reference's MSQMRIDR / MSIDRS, msqmridr.inc and
msidrs.inc, run in FP64 only).

Bars: every shift's returned relres is its FP64 true residual recomputed on
the host (rtol 1e-6) and meets the FP64 tolerance 1e-8; each shift's
solution is within 1e-6 relative of the oracle's FP64 multishift solution
(the FP64 tolerance times the shifted systems' conditioning, measured on
the oracle's side as the distance between its converged solution and a
direct solve).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import idrs_amd
import orc
from idrs_amd import problems as P

from helpers import gpu_crs, rel

pytestmark = pytest.mark.gpu


def _true_relres(crs, b, sig, X):
    K = crs.to_scipy()
    return np.array([np.linalg.norm(b - (K @ X[:, k] - sig[k] * X[:, k])) for k in range(len(sig))]) / np.linalg.norm(b)


def _helmholtz(m, nshift):
    crs, _ = P.helmholtz_3d(m)
    b = np.zeros(crs.n, dtype=np.complex128)
    b[crs.n // 2] = 1.0
    return crs, b, P.helmholtz_shifts(m, nshift=nshift)


def test_helmholtz_shifts_complex64_refined(gpu):
    """16 complex Helmholtz shifts at 12^3, MSQMRIDR(8): the complex64
    multishift solve plus FP64 corrections reach 1e-8 on every shift."""
    crs, b, sig = _helmholtz(12, 16)
    A = gpu_crs(crs)
    print("complex64 msqmridr(8) 12^3, 16 shifts", flush=True)
    r = idrs_amd.ms_mixed(A, b, sig, 8, solver="msqmridr", tolerance=1e-8, maximum_iterations=20000,
                          inner_tolerance=1e-5, max_outer=20)
    print(f"  flag {r.flag} its {r.iterations} outer {r.outer} corrections {r.corrections} "
          f"relres max {r.relres.max():.2e}", flush=True)
    X = np.asarray(r.x)
    tr = _true_relres(crs, b, sig, X)
    assert r.flag == 0 and np.all(tr <= 1e-8)
    assert np.allclose(r.relres, tr, rtol=1e-6, atol=1e-14)
    # against the FP64 oracle: every shift within the conditioning of its system
    K = crs.to_scipy().tocsc()
    for k in (0, 7, 15):
        xd = spla.spsolve((K - sig[k] * __import__("scipy").sparse.identity(crs.n, format="csc")).tocsc(), b)
        assert rel(X[:, k], xd) < 1e-5, (k, rel(X[:, k], xd))


def test_real_shifts_complex64_refined_vs_oracle(gpu):
    """Real shifts of a convection-diffusion operator, MSIDRS(4): the
    refined complex64 solution against the oracle's FP64 MSIDRS solution."""
    crs, _ = P.convection_diffusion_3d(10, 5.0, reaction=3.0)
    b = np.random.default_rng(4).standard_normal(crs.n).astype(np.complex128)
    sig = np.array([0.0, -10.0, -40.0, -100.0], dtype=np.complex128)
    A = gpu_crs(crs)
    r = idrs_amd.ms_mixed(A, b, sig, 4, solver="msidrs", tolerance=1e-8, maximum_iterations=5000,
                          inner_tolerance=1e-5)
    Ao = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=True)
    o = orc.msidrs(Ao, b, sig, 4, tolerance=1e-10, maximum_iterations=5000)
    X = np.asarray(r.x)
    print(f"flag {r.flag} its {r.iterations} corrections {r.corrections}", flush=True)
    assert r.flag == 0 and o["flag"] == 0
    for k in range(len(sig)):
        assert rel(X[:, k], o["x"][:, k]) < 1e-6
    assert np.allclose(r.relres, _true_relres(crs, b, sig, X), rtol=1e-6, atol=1e-14)


def test_complex64_multishift_errors(gpu):
    crs, b, sig = _helmholtz(6, 3)
    Kc = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, cval=crs.val.astype(np.complex128))
    with pytest.raises(idrs_amd.IdrsError, match="real matrix values"):
        idrs_amd.ms_mixed(idrs_amd.Matrix(Kc), b, sig, 4)
    A = gpu_crs(crs)
    A.set_seed(0.5)
    with pytest.raises(idrs_amd.IdrsError, match="seed"):
        idrs_amd.ms_mixed(A, b, sig, 4)
    with pytest.raises(idrs_amd.IdrsError, match="inner tolerance"):
        idrs_amd.ms_mixed(gpu_crs(crs), b, sig, 4, inner_tolerance=2.0)


def test_helmholtz_shifts_complex64_msidrs_reports(gpu):
    """MSIDRS(4) in complex64 on the same shifts: the two shifts with 4-5
    points per wavelength diverge in single precision (the FP64 MSIDRS
    converges at 12^3 in 333 iterations), and their single-shift complex64
    corrections make no progress either.  The solve says so -- flag 3 --
    while every other shift meets 1e-8, and every returned relres is the
    shift's FP64 true residual."""
    crs, b, sig = _helmholtz(12, 16)
    r = idrs_amd.ms_mixed(gpu_crs(crs), b, sig, 4, solver="msidrs", tolerance=1e-8, maximum_iterations=20000,
                          inner_tolerance=1e-5, max_outer=20)
    X = np.asarray(r.x)
    tr = _true_relres(crs, b, sig, X)
    print(f"flag {r.flag} relres {np.array2string(r.relres, precision=1)}", flush=True)
    assert r.flag in (0, 3)
    assert np.allclose(r.relres, tr, rtol=1e-6, atol=1e-14)
    assert np.all(tr <= 1.0 + 1e-12)                    # a diverged shift is restarted from zero
    assert np.sum(tr <= 1e-8) >= 14
