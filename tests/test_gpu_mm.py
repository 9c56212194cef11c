"""Matrix Market files through the GPU path: the reference's own ocean and
wedge data (EXAMPLES/MATRIX_MARKET/DATA, kept as fixtures) read by
mm_matrix into device CRS, multiplied, and solved, against the oracle on
the same entries in the reference reader's order and against scipy's
independent reader.

Bars: products entry-wise within the rounding bound 2(k+1) eps sum|a_ij x_j|
of the oracle's sequential COO product, and within 1e-13 of scipy's;
IDR(4) + Chebyshev(8) with the default foci (0.25, 1.75) of
real_preconditioner (matrix_module.f90:613-621): iterations within 2 of the
oracle's at 1e-8, solutions within 1e-10 at a fixed count.
"""
import numpy as np
import pytest
import scipy.io

import idrs_amd
import orc

from helpers import mm_expanded, mm_fixture, rel, rounding_spread, write_mtx

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps
COORD = ["stommel4", "stommel5", "stommel6", "sag5", "sag6"] + [f"wedge{k}_{m}" for k in (1, 2, 3, 4) for m in "KMC"]


@pytest.mark.parametrize("fmt", ["coo", "crs"])
@pytest.mark.parametrize("name", COORD)
def test_mm_matrix_products(gpu, tmp_path, name, fmt):
    fx = mm_fixture(name)
    path = write_mtx(fx, str(tmp_path / f"{name}.mtx"))
    K = idrs_amd.mm_matrix(path, fmt)
    I, J, V = mm_expanded(fx)
    n = int(fx["size"][0])
    assert K.nrows == n and K.nnz == len(I) and K.matrix_format == (3 if fmt == "coo" else 2)
    x = np.random.default_rng(3).standard_normal(n)
    w = K * x
    Ko = orc.OracleMatrix(orc.COO, n, coo=(I, J, V))
    ref = Ko.user_mv(x)
    mag = np.zeros(n)
    np.add.at(mag, I - 1, np.abs(V) * np.abs(x[J - 1]))
    cnt = np.bincount(I - 1, minlength=n)
    assert np.all(np.abs(w - ref) <= 2.0 * (cnt + 1) * EPS * mag + 1e-300)
    S = scipy.io.mmread(path).tocsr()
    assert rel(w, S @ x) < 1e-13


def _stommel(name):
    fx, bx = mm_fixture(name), mm_fixture(name + "_b")
    n = int(fx["size"][0])
    return fx, n, bx["vals"][:n, 0].copy()


@pytest.mark.parametrize("name", ["stommel6", "stommel5", "stommel4"])
def test_stommel_pp_idrs_default(gpu, tmp_path, name):
    """The stommel example's chain (stommel.f90: mm_matrix 'coo', diagonal
    scaling) with IDR(4) and the Chebyshev(8) polynomial at the default foci,
    first wind field as the right-hand side."""
    fx, n, b = _stommel(name)
    K = idrs_amd.mm_matrix(write_mtx(fx, str(tmp_path / f"{name}.mtx")), "coo")
    A = idrs_amd.Matrix(K, which_precon=1)
    idrs_amd.real_preconditioner(A, 2, 8)
    a = orc.chebyshev(8, np.array([0.25, 1.75]))
    assert np.allclose(A.alpha[:9].real, a, rtol=1e-15, atol=0)
    I, J, V = mm_expanded(fx)
    Ao = orc.OracleMatrix(orc.COO, n, coo=(I, J, V), which_precon=1, degree=8, alpha=a)
    r = idrs_amd.idrs(A, b, 4, tolerance=1e-8, maximum_iterations=4000)
    xo, fo, ro, io, _ = orc.idrs(Ao, b, 4, tolerance=1e-8, maximum_iterations=4000)
    print(f"{name}: gpu it={r.iterations} flag={r.flag} relres={r.relres:.2e}; oracle it={io} flag={fo}")
    slack = 2
    if abs(r.iterations - io) > 2:
        run = lambda bb: (lambda o: (o[0], o[3]))(orc.idrs(Ao, bb, 4, tolerance=1e-8, maximum_iterations=4000))
        slack = max(2, rounding_spread(run, b)[1])
    assert r.flag == fo == 0 and abs(r.iterations - io) <= slack
    # the solution of the unscaled system: x = scaleback(y), K x = b
    x = idrs_amd.scaleback(A, None, r.x)
    S = scipy.io.mmread(str(tmp_path / f"{name}.mtx")).tocsr()
    assert np.linalg.norm(b - S @ x) / np.linalg.norm(b) < 1e-7
    m = 20
    g = idrs_amd.idrs(A, b, 4, tolerance=0.0, maximum_iterations=m)
    run = lambda bb: (orc.idrs(Ao, bb, 4, tolerance=0.0, maximum_iterations=m)[0], m)
    d = rel(g.x, run(b)[0])
    assert d < 1e-10 or d < 100 * rounding_spread(run, b)[0], d


@pytest.mark.parametrize("name", ["sag6", "sag5"])
def test_sag_fixed_count(gpu, tmp_path, name):
    """The SAG ocean problems (ill-conditioned: the Chebyshev(8) IDR(4) solve
    stagnates on both sides) at a fixed iteration count."""
    fx, n, b = _stommel(name)
    K = idrs_amd.mm_matrix(write_mtx(fx, str(tmp_path / f"{name}.mtx")), "crs")
    A = idrs_amd.Matrix(K, which_precon=1)
    I, J, V = mm_expanded(fx)
    Ao = orc.OracleMatrix(orc.COO, n, coo=(I, J, V), which_precon=1)
    m = 15
    g = idrs_amd.idrs(A, b, 4, tolerance=0.0, maximum_iterations=m)
    run = lambda bb: (orc.idrs(Ao, bb, 4, tolerance=0.0, maximum_iterations=m)[0], m)
    d = rel(g.x, run(b)[0])
    assert d < 1e-10 or d < 100 * rounding_spread(run, b)[0], d
