"""Pins of the oracle's QMRIDR, MSIDRS, MSQMRIDR and dense helpers (CPU).

Checks: SOLVE against numpy's LU solve; GROT's rotation identity; QMRIDR's
first s iterations against an independent GMRES (the reference README
states the equivalence); finite termination; multishift solutions against
dense solves of (A - sigma I) x = b; the collinear residuals of MSIDRS and
its colfac output; MSQMRIDR with and without the inner MSIDRS.
"""
import numpy as np
import pytest

import orc
from idrs_amd import problems as P


def test_solve_gepp():
    rng = np.random.default_rng(0)
    for cplx in (False, True):
        A = rng.standard_normal((7, 7)) + (1j * rng.standard_normal((7, 7)) if cplx else 0)
        A[0, 0] = 1e-12  # forces a pivot
        b = rng.standard_normal(7) + (1j if cplx else 0)
        assert np.allclose(orc.solve(A, b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("a,b", [(3.0, 4.0), (-2.0, 0.5), (0.0, 2.0), (1 + 2j, 3 - 1j), (0j, 2 + 1j)])
def test_grot_rotation(a, b):
    c, s, r = orc.grot(a, b)
    # [c s; -conj(s) c] [a; b] = [r; 0]
    assert abs(c * a + s * b - r) < 1e-14 * (abs(a) + abs(b))
    assert abs(-np.conj(s) * a + c * b) < 1e-14 * (abs(a) + abs(b))
    assert abs(abs(c) ** 2 + abs(s) ** 2 - 1) < 1e-14


def _gmres_residuals(A, b, kmax):
    """min ||b - A x|| over x in K_k(A, b), k = 1..kmax (orthonormal Krylov
    basis from numpy QR)."""
    n = len(b)
    K = np.zeros((n, kmax), dtype=np.result_type(A, b))
    v = b.copy()
    for k in range(kmax):
        K[:, k] = v
        v = A @ v
    out = []
    for k in range(1, kmax + 1):
        Q, _ = np.linalg.qr(K[:, :k])
        AQ = A @ Q
        y, *_ = np.linalg.lstsq(AQ, b, rcond=None)
        out.append(np.linalg.norm(b - AQ @ y))
    return np.array(out)


@pytest.mark.parametrize("s", [3, 6])
def test_qmridr_first_s_steps_are_gmres(s):
    crs, _ = P.convection_diffusion_3d(5, 20.0, reaction=3.0)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs)
    b = np.random.default_rng(1).standard_normal(crs.n)
    x, flag, relres, it, rv = orc.qmridr(A, b, s, tolerance=1e-10, maximum_iterations=500, resvec=True)
    assert flag == 0
    ref = _gmres_residuals(crs.to_scipy().toarray(), b, s)
    assert np.allclose(rv[1:s + 1], ref, rtol=1e-8)


@pytest.mark.parametrize("s", [1, 2, 4, 8])
def test_qmridr_finite_termination(s):
    crs, _, b = P.toeplitz_1d()
    A = orc.OracleMatrix(orc.CRS, 60, crs=crs)
    x, flag, relres, it, _ = orc.qmridr(A, b, s, M1=A, tolerance=1e-8, maximum_iterations=4000)
    assert flag == 0 and it <= 60 + 60 // s
    assert np.allclose(x, 1.0, atol=1e-6)


def _helmholtz(m, nshift):
    crs, _ = P.helmholtz_3d(m)
    n = crs.n
    b = np.zeros(n, dtype=np.complex128)
    b[n // 2] = 1.0
    sig = P.helmholtz_shifts(m, nshift=nshift)
    return crs, b, sig


def test_msidrs_shift_solutions():
    crs, b, sig = _helmholtz(7, 5)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=True)
    out = orc.msidrs(A, b, sig, 4, tolerance=1e-9, maximum_iterations=2000)
    assert out["flag"] == 0 and np.all(out["relres"] <= 1e-9)
    K = crs.to_scipy().toarray()
    for k, s_ in enumerate(sig):
        xs = np.linalg.solve(K - s_ * np.eye(crs.n), b)
        r = b - (K - s_ * np.eye(crs.n)) @ out["x"][:, k]
        assert np.linalg.norm(r) / np.linalg.norm(b) == pytest.approx(out["relres"][k], rel=1e-6, abs=1e-15)
        assert np.linalg.norm(out["x"][:, k] - xs) / np.linalg.norm(xs) < 1e-6


def test_msidrs_collinear_residuals_and_colfac():
    """All shifted residuals are multiples of one vector; colfac holds the
    factors (msidrs.inc:207-226, 329)."""
    crs, b, sig = _helmholtz(6, 4)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=True)
    out = orc.msidrs(A, b, sig, 3, tolerance=0.0, maximum_iterations=9, colfac=True)
    K = crs.to_scipy().toarray()
    R = np.stack([b - (K - s_ * np.eye(crs.n)) @ out["x"][:, k] for k, s_ in enumerate(sig)], axis=1)
    phi = out["colfac"]
    for k in range(1, len(sig)):
        lhs, rhs = R[:, k] * phi[0], R[:, 0] * phi[k]
        assert np.linalg.norm(lhs - rhs) <= 1e-9 * np.linalg.norm(lhs)


def test_msidrs_real_and_variant2():
    crs, _ = P.convection_diffusion_3d(6, 5.0)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs)
    b = np.random.default_rng(2).standard_normal(crs.n)
    sig = np.array([0.0, -50.0, -200.0])
    K = crs.to_scipy().toarray()
    for variant in (1, 2):
        out = orc.msidrs(A, b, sig, 4 if variant == 1 else 1, variant=variant, tolerance=1e-9, maximum_iterations=3000)
        assert out["flag"] == 0
        for k, s_ in enumerate(sig):
            r = b - (K - s_ * np.eye(crs.n)) @ out["x"][:, k]
            assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-9


@pytest.mark.parametrize("inner", [0, 15])
def test_msqmridr_shift_solutions(inner):
    crs, b, sig = _helmholtz(7, 4)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=True)
    kw = dict(inner_s=4, inner_iterations=inner, inner_tolerance=1e-1) if inner else {}
    out = orc.msqmridr(A, b, sig, 6, tolerance=1e-9, maximum_iterations=3000, **kw)
    assert out["flag"] == 0 and np.all(out["relres"] <= 1e-9)
    K = crs.to_scipy().toarray()
    for k, s_ in enumerate(sig):
        xs = np.linalg.solve(K - s_ * np.eye(crs.n), b)
        assert np.linalg.norm(out["x"][:, k] - xs) / np.linalg.norm(xs) < 1e-6


# -------------------------------------------- the 3-D Helmholtz shifts ---

def _helmholtz_c3(m):
    """The multishift workload's recipe at a smaller grid: h^2-scaled 7-point
    -Laplacian on m^3, 16 shifts sigma_k = (2 pi f_k h)^2 (1 - 0.05 i) with
    4..30 points per wavelength, point source at the centre."""
    crs, _ = P.helmholtz_3d(m)
    b = np.zeros(crs.n, dtype=np.complex128)
    b[crs.n // 2] = 1.0
    return crs, b, P.helmholtz_shifts(m, 16)


def _true_relres(crs, b, sig, X):
    K = crs.to_scipy()
    return np.array([np.linalg.norm(b - (K @ X[:, k] - sig[k] * X[:, k])) for k in range(len(sig))]) / np.linalg.norm(b)


@pytest.mark.parametrize("m,its", [(16, 574), (24, 922)])
def test_helmholtz_shifts_msqmridr_converges(m, its):
    """MSQMRIDR(8) without inner iterations (msqmridr.inc: in_it = 0) reaches
    1e-8 on every shift: 574 iterations at 16^3 and 922 at 24^3; each
    shift's returned relres is its true residual."""
    crs, b, sig = _helmholtz_c3(m)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=True)
    o = orc.msqmridr(A, b, sig, 8, tolerance=1e-8, maximum_iterations=3000)
    assert o["flag"] == 0 and abs(o["iterations"] - its) <= 2
    tr = _true_relres(crs, b, sig, o["x"])
    assert tr.max() <= 1e-8
    assert np.allclose(o["relres"], tr, rtol=1e-10)


def test_helmholtz_shifts_inner_msidrs_and_msidrs_alone():
    """What the rest of the recipe reaches, recorded: with the inner
    MSIDRS(4) over [sigma, 0] (4 inner iterations per outer step) the outer
    MSQMRIDR(8) falls behind -- after 150 outer steps its worst shift is at
    1.8e-1 against 1.9e-2 for the plain MSQMRIDR(8) -- and MSIDRS(4)
    alone ends at 24^3 with flag 2 (its recursive residuals meet 1e-8 while
    the true residuals have grown above 1)."""
    crs, b, sig = _helmholtz_c3(16)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=True)
    plain = orc.msqmridr(A, b, sig, 8, tolerance=1e-8, maximum_iterations=150)
    inner = orc.msqmridr(A, b, sig, 8, tolerance=1e-8, maximum_iterations=150, inner_s=4, inner_iterations=4)
    tp = _true_relres(crs, b, sig, plain["x"]).max()
    ti = _true_relres(crs, b, sig, inner["x"]).max()
    print(f"150 outer steps: plain {tp:.2e}, inner MSIDRS(4) {ti:.2e}")
    assert plain["flag"] == inner["flag"] == 1
    assert ti > 0.1 and ti > 5 * tp
    crs, b, sig = _helmholtz_c3(24)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=True)
    o = orc.msidrs(A, b, sig, 4, tolerance=1e-8, maximum_iterations=3000)
    assert o["flag"] == 2 and _true_relres(crs, b, sig, o["x"]).max() > 1.0
