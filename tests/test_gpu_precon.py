"""GPU parity of the seed shift and of every polynomial kind against the
oracle and against closed forms.

- Seed: real_mv / complex_mv compute w = K t - seed t with t = v/D
  (matrix_module.f90:241-262), so the operator is (K - seed I) D^-1
  p((K - seed I) D^-1); the seed term alone is A_seed v - A_0 v = -seed D^-1 v
  (degree 0).
- Polynomials (real_preconditioner / complex_preconditioner,
  matrix_module.f90:551-722): Neumann (kind 1), Chebyshev with complex foci
  (kind 2, cpolpre with complex coefficients), Hessenberg (kind 3,
  compute_alpha).
Bars: products within 1e-12 of the oracle (normwise); the seed term within
the rounding of one product; IDR(s) iteration counts within 2 of the
oracle's and, at a fixed iteration count, solutions within 1e-10 relative.
"""
import numpy as np
import pytest

import idrs_amd
import orc
from idrs_amd import problems as P

from helpers import rel, rounding_spread
from test_oracle import hessenberg_zero_colsum

pytestmark = pytest.mark.gpu


def _gpu_operator(crs, layout, which_precon=1):
    n = crs.n
    kw = {"cval": crs.val} if np.iscomplexobj(crs.val) else {"rval": crs.val}
    K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, **kw)
    if layout == "sliced":
        K.set_group(1)
    elif layout == "lanes":
        K.set_group(8)
    return idrs_amd.Matrix(K, which_precon=which_precon)


def _gpu_dia_operator(dia, which_precon=1):
    n = dia.vals.shape[0]
    kw = {"cvals": dia.vals} if np.iscomplexobj(dia.vals) else {"rvals": dia.vals}
    return idrs_amd.Matrix(idrs_amd.dia_format(n, n, dia.offsets, **kw), which_precon=which_precon)


def _dense_operator(K, d, seed, alpha):
    S = (K - seed * np.eye(K.shape[0])) / d[None, :]
    return S, alpha


def _apply_dense(S, alpha, v):
    if alpha is None or len(alpha) == 1:
        pv = v if alpha is None else alpha[0] * v
    else:
        pv = alpha[-1] * v
        for a in alpha[-2::-1]:
            pv = S @ pv + a * v
    return S @ pv


@pytest.mark.parametrize("layout", ["sliced", "lanes", "dia"])
@pytest.mark.parametrize("seed", [0.35, -1.25, 0.4 - 0.9j])
@pytest.mark.parametrize("degree", [0, 4])
def test_seed_products(gpu, layout, seed, degree):
    crs, dia = P.convection_diffusion_3d(9, 15.0, reaction=0.5)
    n = crs.n
    cplx = np.iscomplexobj(seed)
    A = _gpu_dia_operator(dia) if layout == "dia" else _gpu_operator(crs, layout)
    A0 = _gpu_dia_operator(dia) if layout == "dia" else _gpu_operator(crs, layout)
    A.set_seed(seed)
    foci = np.array([0.2, 1.8], dtype=np.complex128 if cplx else np.float64)
    alpha = None
    if degree:
        if cplx:
            idrs_amd.complex_preconditioner(A, 2, degree, foci=foci)
            alpha = orc.chebyshev(degree, foci - seed)
        else:
            idrs_amd.real_preconditioner(A, 2, degree, foci=foci)
            alpha = orc.chebyshev(degree, foci - seed)
        assert np.allclose(A.alpha[: degree + 1], alpha, rtol=0, atol=1e-15 * np.abs(alpha).max())
    rng = np.random.default_rng(7)
    v = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    Kd = crs.to_scipy().toarray()
    d = np.diag(Kd).copy()
    S, _ = _dense_operator(Kd, d, seed, alpha)
    ref = _apply_dense(S, alpha, v)
    w = A * v
    assert rel(w, ref) < 1e-12
    Ao = orc.OracleMatrix(orc.CRS, n, crs=crs, which_precon=1, degree=degree, alpha=alpha, seed=seed,
                          complex_vectors=cplx)
    assert rel(w, Ao.matvec(v)) < 1e-12
    if degree == 0:
        # the seed term alone: A_seed v - A_0 v = -seed v/D
        diff = w - A0 * v
        assert np.allclose(diff, -seed * v / d, rtol=1e-9, atol=1e-12 * np.abs(A0 * v).max())


@pytest.mark.parametrize("seed", [-0.4, -0.5 + 0.2j])
def test_idrs_with_seed_matches_oracle(gpu, seed):
    """IDR(4) on K - seed I with Chebyshev(4) against the oracle.  K is the
    h^2-scaled convection-diffusion operator (spectrum inside (0, 12)), so
    the seed is on the scale of the spectrum, and the foci move by the seed
    as real_preconditioner / complex_preconditioner move them
    (matrix_module.f90:597-611)."""
    m = 10
    h = 1.0 / (m + 1)
    crs, _ = P.convection_diffusion_3d(m, 20.0, reaction=2.0)
    crs = P.Crs(crs.n, crs.row_ptr, crs.col_ind, crs.val * h * h)
    n = crs.n
    cplx = np.iscomplexobj(seed)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, n) + (1j * rng.uniform(-1, 1, n) if cplx else 0)
    PT = P.gaussian_shadow(n, 4, seed=1, complex=cplx)
    foci = np.array([0.05, 11.95], dtype=np.complex128 if cplx else np.float64)
    A = _gpu_operator(crs, "sliced", which_precon=0)
    A.set_seed(seed)
    (idrs_amd.complex_preconditioner if cplx else idrs_amd.real_preconditioner)(A, 2, 4, foci=foci)
    alpha = orc.chebyshev(4, foci - seed)
    assert np.allclose(A.alpha[:5] if cplx else A.alpha[:5].real, alpha, rtol=1e-14, atol=0)
    Ao = orc.OracleMatrix(orc.CRS, n, crs=crs, which_precon=0, degree=4, alpha=alpha, seed=seed,
                          complex_vectors=cplx)
    g = idrs_amd.idrs(A, b, 4, tolerance=1e-9, maximum_iterations=2000, PT_raw=PT)
    xo, fo, ro, io, _ = orc.idrs(Ao, b, 4, tolerance=1e-9, maximum_iterations=2000, PT_raw=PT)
    print(f"seed={seed}: gpu it={g.iterations} flag={g.flag} relres={g.relres:.2e}; oracle it={io} flag={fo}")
    assert g.flag == 0 and fo == 0
    assert abs(g.iterations - io) <= 2
    # At a fixed count: ten iterations, where the oracle's own spread under
    # 1e-16 changes of b is ~1e-11 (this recurrence amplifies rounding by
    # 1e5-1e16 at some later steps: near-breakdowns), held to the 1e-10 bar
    # or 100x the measured spread, whichever is larger.
    m = 10
    g = idrs_amd.idrs(A, b, 4, tolerance=0.0, maximum_iterations=m, PT_raw=PT)
    run = lambda bb: (orc.idrs(Ao, bb, 4, tolerance=0.0, maximum_iterations=m, PT_raw=PT)[0], m)
    xo = run(b)[0]
    spread, _ = rounding_spread(run, b)
    assert rel(g.x, xo) < max(1e-10, 100 * spread), (rel(g.x, xo), spread)
    # the returned relres is the residual of the seeded operator
    res = b - Ao.matvec(g.x)
    assert abs(np.linalg.norm(res) / np.linalg.norm(b) - g.relres) <= 1e-6 * g.relres


@pytest.mark.parametrize("degree,center", [(1, 1.0), (3, 1.3), (6, 0.9)])
def test_neumann_kind1(gpu, degree, center):
    """real_preconditioner(A, 1, degree, center): the coefficients of
    1 - t p(t) = (1 - t/c)^{d+1} (rneumann) and the operator they define,
    then an IDR(4) solve against the oracle."""
    crs, dia = P.convection_diffusion_3d(10, 10.0, reaction=1.0)
    n = crs.n
    for A in (_gpu_operator(crs, "sliced"), _gpu_dia_operator(dia)):
        idrs_amd.real_preconditioner(A, 1, degree, center=center)
        a = orc.neumann(degree, center)
        assert np.allclose(A.alpha[: degree + 1].real, a, rtol=1e-15, atol=0)
        t = np.linspace(-0.5, 2.0, 9)
        assert np.allclose(1 - t * np.polynomial.polynomial.polyval(t, A.alpha[: degree + 1].real),
                           (1 - t / center) ** (degree + 1), atol=1e-10)
        Ao = orc.OracleMatrix(orc.CRS, n, crs=crs, which_precon=1, degree=degree, alpha=a)
        v = np.random.default_rng(degree).standard_normal(n)
        assert rel(A * v, Ao.matvec(v)) < 1e-12
        assert rel(idrs_amd.scaleback(A, None, v), Ao.scaleback(v)) < 1e-12
    b = np.random.default_rng(5).uniform(-1, 1, n)
    PT = P.gaussian_shadow(n, 4, seed=2)
    g = idrs_amd.idrs(A, b, 4, tolerance=1e-9, maximum_iterations=1000, PT_raw=PT)
    xo, fo, _, io, _ = orc.idrs(Ao, b, 4, tolerance=1e-9, maximum_iterations=1000, PT_raw=PT)
    assert g.flag == 0 and fo == 0 and abs(g.iterations - io) <= 2
    m = min(io, 30)
    g = idrs_amd.idrs(A, b, 4, tolerance=0.0, maximum_iterations=m, PT_raw=PT)
    assert rel(g.x, orc.idrs(Ao, b, 4, tolerance=0.0, maximum_iterations=m, PT_raw=PT)[0]) < 1e-10


def _complex_helmholtz(m, sigma):
    """h^2-scaled 7-point -Laplacian minus a complex shift (complex values)."""
    crs, _ = P.helmholtz_3d(m)
    rows = np.repeat(np.arange(crs.n), np.diff(crs.row_ptr))
    val = crs.val.astype(np.complex128)
    val[rows == crs.col_ind - 1] -= sigma
    return P.Crs(crs.n, crs.row_ptr, crs.col_ind, val)


@pytest.mark.parametrize("degree", [2, 5])
def test_complex_chebyshev_foci(gpu, degree):
    """complex_preconditioner(A, 2, degree, foci) with complex foci on a
    complex matrix: cchebyshev's coefficients (complex), cpolpre with them,
    and an IDR(4) solve, against the oracle."""
    crs = _complex_helmholtz(9, 0.3 + 0.05j)
    n = crs.n
    foci = np.array([0.05 - 0.01j, 1.95 - 0.01j])
    A = _gpu_operator(crs, "sliced")
    idrs_amd.complex_preconditioner(A, 2, degree, foci=foci)
    a = orc.chebyshev(degree, foci)
    assert np.abs(a.imag).max() > 1e-6          # genuinely complex coefficients
    assert np.allclose(A.alpha[: degree + 1], a, rtol=1e-14, atol=0)
    Ao = orc.OracleMatrix(orc.CRS, n, crs=crs, which_precon=1, degree=degree, alpha=a, complex_vectors=True)
    rng = np.random.default_rng(9)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(A * v, Ao.matvec(v)) < 1e-12
    assert rel(idrs_amd.scaleback(A, None, v), Ao.scaleback(v)) < 1e-12
    b = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    PT = P.gaussian_shadow(n, 4, seed=3, complex=True)
    g = idrs_amd.idrs(A, b, 4, tolerance=1e-9, maximum_iterations=2000, PT_raw=PT)
    xo, fo, _, io, _ = orc.idrs(Ao, b, 4, tolerance=1e-9, maximum_iterations=2000, PT_raw=PT)
    assert g.flag == 0 and fo == 0 and abs(g.iterations - io) <= 2
    m = min(io, 30)
    g = idrs_amd.idrs(A, b, 4, tolerance=0.0, maximum_iterations=m, PT_raw=PT)
    assert rel(g.x, orc.idrs(Ao, b, 4, tolerance=0.0, maximum_iterations=m, PT_raw=PT)[0]) < 1e-10


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("degree", [2, 5])
def test_hessenberg_kind3(gpu, degree, cplx):
    """real_preconditioner / complex_preconditioner(A, 3, degree, H=H):
    compute_alpha's coefficients equal the oracle's, and the roots of
    1 - t p(t) are the eigenvalues of H(1:d+1, 1:d+1) for a Hessenberg H with
    zero column sums; the operator against the oracle."""
    H = hessenberg_zero_colsum(degree, 20 + degree, cplx)
    crs, _ = P.convection_diffusion_3d(8, 5.0, reaction=1.0)
    n = crs.n
    A = _gpu_operator(crs, "sliced")
    (idrs_amd.complex_preconditioner if cplx else idrs_amd.real_preconditioner)(A, 3, degree, H=H)
    a = orc.compute_alpha(degree, H)
    ga = A.alpha[: degree + 1] if cplx else A.alpha[: degree + 1].real
    assert np.allclose(ga, a, rtol=1e-13, atol=1e-14 * np.abs(a).max())
    roots = np.sort_complex(np.polynomial.polynomial.polyroots(np.concatenate([[1.0], -ga])))
    assert np.allclose(roots, np.sort_complex(np.linalg.eigvals(H[: degree + 1, : degree + 1])), rtol=1e-7, atol=1e-8)
    Ao = orc.OracleMatrix(orc.CRS, n, crs=crs, which_precon=1, degree=degree, alpha=a, complex_vectors=cplx)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    assert rel(A * v, Ao.matvec(v)) < 1e-12
