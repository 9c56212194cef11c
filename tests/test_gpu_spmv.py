"""GPU parity of the sparse formats, the operator and MAKE_P against the oracle.

Floating-point bar: the GPU sums each row in stored order with fused
multiply-adds, the oracle rounds every product and sum separately, so each
entry must agree within the rounding bound 2*(k+1)*eps*sum_j |a_ij x_j|
(k = entries in the row); normwise checks use 1e-13.
"""
import numpy as np
import pytest

import idrs_amd
import orc
from idrs_amd import problems as P

from helpers import gpu_crs, oracle_crs, rel

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _bound(crs, x):
    rows = np.repeat(np.arange(crs.n), np.diff(crs.row_ptr))
    mag = np.zeros(crs.n)
    np.add.at(mag, rows, np.abs(crs.val) * np.abs(x[crs.col_ind - 1]))
    k = np.diff(crs.row_ptr)
    return 2.0 * (k + 1) * EPS * mag + 1e-300


def _ragged(n, seed, long_rows=()):
    """Random CRS with empty rows, short rows and a few very long rows (more
    nonzeros than one LDS slab holds)."""
    rng = np.random.default_rng(seed)
    lens = np.minimum(rng.integers(0, 12, size=n), n)
    lens[rng.integers(0, n, size=max(1, n // 50))] = 0
    for r in long_rows:
        lens[r] = min(n, 3000)
    cols = [np.sort(rng.choice(n, size=l, replace=False)) for l in lens]
    rp = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    ci = (np.concatenate(cols) + 1).astype(np.int32) if lens.sum() else np.zeros(0, np.int32)
    val = rng.standard_normal(int(lens.sum()))
    return P.Crs(n, rp, ci, val)


@pytest.mark.parametrize("n,seed,long_rows", [(1, 0, ()), (257, 1, ()), (5000, 2, (7, 4000)), (70001, 3, ())])
def test_crs_spmv_matches_oracle(gpu, n, seed, long_rows):
    crs = _ragged(n, seed, long_rows)
    x = np.random.default_rng(seed + 10).standard_normal(n)
    K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
    w = K * x
    ref = oracle_crs(crs).user_mv(x)
    assert np.all(np.abs(w - ref) <= _bound(crs, x))


def test_crs_pattern_and_complex(gpu):
    crs = _ragged(3000, 5)
    x = np.random.default_rng(1).standard_normal(3000) + 1j * np.random.default_rng(2).standard_normal(3000)
    # pattern matrix (rval absent): values 1 (user_module.f90:408-414)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind)
    pat = P.Crs(crs.n, crs.row_ptr, crs.col_ind, np.ones(crs.nnz))
    xr = x.real.copy()
    assert np.all(np.abs(K * xr - pat.matvec(xr)) <= _bound(pat, xr))
    # real matrix times complex vector (ccrs_mv rval branch, user_module.f90:318-324)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    ref = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=True).user_mv(x)
    assert rel(K * x, ref) < 1e-13
    # complex matrix
    cv = crs.val * (1 + 0.5j)
    ccrs = P.Crs(crs.n, crs.row_ptr, crs.col_ind, cv)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, cval=cv)
    ref = orc.OracleMatrix(orc.CRS, crs.n, crs=ccrs, complex_vectors=True).user_mv(x)
    assert rel(K * x, ref) < 1e-13


def test_coo_order_and_duplicates(gpu):
    """COO entries in arbitrary order with duplicates: each row is summed in
    list order, like the sequential scatter loop (user_module.f90:542-558)."""
    rng = np.random.default_rng(7)
    n, nnz = 4000, 30000
    indx = rng.integers(1, n + 1, size=nnz).astype(np.int32)
    jndx = rng.integers(1, n + 1, size=nnz).astype(np.int32)
    jndx[:100] = jndx[100:200]
    indx[:100] = indx[100:200]                      # duplicate entries
    val = rng.standard_normal(nnz)
    x = rng.standard_normal(n)
    K = idrs_amd.coo_format(n, n, nnz, indx, jndx, rval=val)
    w = K * x
    ref = orc.OracleMatrix(orc.COO, n, coo=(indx, jndx, val)).user_mv(x)
    # exact order is kept, so the only difference is fused vs separate rounding
    order = np.argsort(indx, kind="stable")
    rp = np.concatenate([[1], 1 + np.cumsum(np.bincount(indx - 1, minlength=n))]).astype(np.int32)
    crs = P.Crs(n, rp, jndx[order], val[order])
    assert np.all(np.abs(w - ref) <= _bound(crs, x))


def test_cds_band_and_dia(gpu):
    rng = np.random.default_rng(11)
    n, hb = 3001, 17
    band = rng.standard_normal((n, 2 * hb + 1))
    x = rng.standard_normal(n)
    K = idrs_amd.cds_format(n, n, hb, rband=band)
    ref = orc.OracleMatrix(orc.CDS, n, cds=(hb, band)).user_mv(x)
    assert rel(K * x, ref) < 1e-14
    # offset-list storage of a 7-point stencil equals its CRS product
    crs, dia = P.convection_diffusion_3d(12, 10.0, reaction=1.0)
    x = rng.standard_normal(crs.n)
    Kd = idrs_amd.dia_format(crs.n, crs.n, dia.offsets, rvals=dia.vals)
    ref = orc.OracleMatrix(orc.DIA, crs.n, dia=(dia.offsets, dia.vals)).user_mv(x)
    assert rel(Kd * x, ref) < 1e-14
    assert rel(Kd * x, crs.matvec(x)) < 1e-14


def test_complex_toeplitz_dia(gpu):
    """The reference's complex Toeplitz operator (EXAMPLES/TOEPLITZ/toeplitz.f90:41-42)."""
    n = 200
    off = np.array([-1, 0, 2, 3], dtype=np.int32)
    cv = np.array([3.6j, 4.0, 1.0, 0.7])
    vals = np.tile(cv, (n, 1)).astype(np.complex128)
    for d, o in enumerate(off):
        i = np.arange(n)
        vals[(i + o < 0) | (i + o >= n), d] = 0
    K = idrs_amd.dia_format(n, n, off, cvals=vals)
    x = np.random.default_rng(0).standard_normal(n) + 1j
    dense = np.zeros((n, n), dtype=np.complex128)
    for d, o in enumerate(off):
        dense += np.diag(np.full(n - abs(o), cv[d]), o)
    assert rel(K * x, dense @ x) < 1e-14


def test_diagonal_zero_rule(gpu):
    crs = _ragged(500, 3)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    d = idrs_amd.real_diagonal(K)
    ref = orc.OracleMatrix(orc.CRS, crs.n, crs=crs).diagonal()
    assert np.array_equal(d, ref)
    assert np.all(d != 0)


@pytest.mark.parametrize("degree", [0, 1, 4, 8])
def test_operator_chebyshev_diag(gpu, degree):
    """A*v = K D^-1 p(K D^-1) v with diagonal scaling and the Chebyshev
    polynomial (matrix_module.f90:135-304)."""
    crs, _ = P.convection_diffusion_3d(16, 50.0, reaction=1.0)
    foci = [0.15, 1.85]
    poly = (2, degree, foci) if degree else None
    A = gpu_crs(crs, diag_precon=True, poly=poly)
    Ao = oracle_crs(crs, diag_precon=True, poly=poly)
    v = np.random.default_rng(degree).standard_normal(crs.n)
    assert rel(A * v, Ao.matvec(v)) < 1e-12
    # operator(/) and scaleback
    assert rel(v / A, v / idrs_amd.real_diagonal(A.K)) < 1e-15
    assert rel(idrs_amd.scaleback(A, None, v), Ao.scaleback(v)) < 1e-12
    if degree:
        assert np.allclose(A.alpha.real, orc.chebyshev(degree, np.array(foci)), rtol=0, atol=0)


def test_make_p_orthonormal(gpu):
    n, s = 10007, 6
    raw = P.gaussian_shadow(n, s, seed=0)
    PTg = idrs_amd.make_p(n, s, 1, PT_raw=raw)
    PTo = orc.make_p(n, s, 1, PT_raw=raw)
    assert rel(PTg, PTo) < 1e-13
    assert np.allclose(PTg.T @ PTg, np.eye(s), atol=1e-13)
    # default stream: the library and the oracle draw the same raw numbers
    PTg = idrs_amd.make_p(n, s, 1)
    PTo = orc.make_p(n, s, 1)
    assert rel(PTg, PTo) < 1e-13
    # complex, and the BiCGSTAB shadow vector conj(r)
    PTg = idrs_amd.make_p(n, s, 1, complex=True)
    PTo = orc.make_p(n, s, 1, complex=True)
    assert rel(PTg, PTo) < 1e-13
    r = raw[:, 0] + 1j * raw[:, 1]
    assert np.array_equal(idrs_amd.make_p(n, 1, 2, r=r, complex=True)[:, 0], np.conj(r))


def test_autotune_keeps_results(gpu):
    """The autotuned lanes-per-row choice changes only the in-row combination
    order: products stay within the rounding bound."""
    crs = P.random_dominant(20000, k=12, seed=3)
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    g = K.autotune(reps=3)
    assert g in (1, 2, 4, 8, 16, 32, 64)
    assert K.layout()[0] == g and K.layout()[1] in (4096, 8192, 16384, 32768)
    x = np.random.default_rng(4).standard_normal(crs.n)
    w = K * x
    ref = oracle_crs(crs).user_mv(x)
    assert np.all(np.abs(w - ref) <= _bound(crs, x))


def test_scaled_dia_drops_unit_diagonal(gpu):
    """Diagonal scaling of a CDS matrix: K D^-1 has an all-ones main diagonal,
    which is then not stored (one value stream less per product); the
    product still equals the reference operator's (real_mv: K (v / D))."""
    rng = np.random.default_rng(12)
    crs, dia = P.convection_diffusion_3d(12, 10.0, reaction=1.0)
    n, nd = crs.n, len(dia.offsets)
    x = rng.standard_normal(n)
    A = idrs_amd.Matrix(idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals), which_precon=1)
    assert A.stored_values() == ((nd - 1) * n, True)
    Ao = orc.OracleMatrix(orc.DIA, n, dia=(dia.offsets, dia.vals), which_precon=1)
    assert rel(A * x, Ao.matvec(x)) < 1e-14
    # the same operator in CRS storage keeps every entry
    Ac = idrs_amd.Matrix(idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val), which_precon=1)
    assert Ac.stored_values() == (crs.nnz, False)
    assert rel(A * x, Ac * x) < 1e-14
    # complex values: a_ii / a_ii in complex division need not round to
    # exactly (1, 0); the diagonal is dropped only where it does
    cv = dia.vals * (1.0 + 0.3j)
    Az = idrs_amd.Matrix(idrs_amd.dia_format(n, n, dia.offsets, cvals=cv), which_precon=1)
    cnt, unit = Az.stored_values()
    assert cnt == (nd - int(unit)) * n
    Aoz = orc.OracleMatrix(orc.DIA, n, dia=(dia.offsets, cv), which_precon=1, complex_vectors=True)
    xz = x + 1j * rng.standard_normal(n)
    assert rel(Az * xz, Aoz.matvec(xz)) < 1e-14
    # a zero on the diagonal is replaced by 1 in D, so the scaled entry is 0:
    # the diagonal stays stored
    z0 = list(dia.offsets).index(0)
    v2 = dia.vals.copy()
    v2[37, z0] = 0.0
    A2 = idrs_amd.Matrix(idrs_amd.dia_format(n, n, dia.offsets, rvals=v2), which_precon=1)
    assert A2.stored_values() == (nd * n, False)
    Ao2 = orc.OracleMatrix(orc.DIA, n, dia=(dia.offsets, v2), which_precon=1)
    assert rel(A2 * x, Ao2.matvec(x)) < 1e-14
    # a shift moves D off the diagonal: stored
    A3 = idrs_amd.Matrix(idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals), which_precon=1, shift=0.5)
    assert A3.stored_values() == (nd * n, False)
    # without scaling nothing is dropped
    A4 = idrs_amd.Matrix(idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals))
    assert A4.stored_values() == (nd * n, False)


def test_scaled_dia_solve_and_mixed(gpu):
    """IDR(4) + Chebyshev(8) on the unit-diagonal DIA operator against the
    oracle, and the mixed-precision solve on it (FP32 copy of the packed
    values)."""
    crs, dia = P.convection_diffusion_3d(24, 6.2, reaction=1.0)
    n = crs.n
    b = np.random.default_rng(0).uniform(-1, 1, n)
    poly = (2, 8, [0.15, 1.85])
    A = idrs_amd.Matrix(idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals), which_precon=1)
    idrs_amd.real_preconditioner(A, 2, 8, foci=poly[2])
    assert A.stored_values()[1]
    Ao = orc.OracleMatrix(orc.DIA, n, dia=(dia.offsets, dia.vals), which_precon=1, degree=8,
                          alpha=orc.chebyshev(8, np.asarray(poly[2])))
    PT = P.gaussian_shadow(n, 4, seed=0)
    r = idrs_amd.idrs(A, b, 4, tolerance=1e-8, maximum_iterations=2000, PT_raw=PT)
    xo, fo, ro, io, _ = orc.idrs(Ao, b, 4, tolerance=1e-8, maximum_iterations=2000, PT_raw=PT)
    assert r.flag == fo == 0 and abs(r.iterations - io) <= 2
    g = idrs_amd.idrs(A, b, 4, tolerance=0.0, maximum_iterations=30, PT_raw=PT)
    go = orc.idrs(Ao, b, 4, tolerance=0.0, maximum_iterations=30, PT_raw=PT)
    assert rel(g.x, go[0]) < 1e-10
    m = idrs_amd.idrs_mixed(A, b, 4, tolerance=1e-8, maximum_iterations=3000, inner_tolerance=1e-6)
    assert m.flag == 0 and np.linalg.norm(b - A * m.x) <= 1.01e-8 * np.linalg.norm(b)


@pytest.mark.parametrize("n,seed,long_rows", [(1, 0, ()), (257, 1, ()), (5000, 2, (7, 4000)), (70001, 3, ())])
def test_sliced_copy_matches_oracle(gpu, n, seed, long_rows):
    """Group 1: the sliced-ELL copy, one lane per row (rows of up to 8
    entries unrolled, longer rows and empty rows included), each row summed
    in CRS order; switching the layout back and forth keeps the products."""
    crs = _ragged(n, seed, long_rows)
    x = np.random.default_rng(seed + 10).standard_normal(n)
    K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
    K.set_group(1)
    w = K * x
    ref = oracle_crs(crs).user_mv(x)
    assert np.all(np.abs(w - ref) <= _bound(crs, x))
    K.set_group(4)
    assert np.all(np.abs(K * x - ref) <= _bound(crs, x))
    K.set_group(1)
    assert np.array_equal(K * x, w)
    with pytest.raises(idrs_amd.IdrsError):
        K.set_group(3)


def test_sliced_copy_complex_operator_and_refresh(gpu):
    rng = np.random.default_rng(21)
    crs = _ragged(3000, 5)
    xz = rng.standard_normal(crs.n) + 1j * rng.standard_normal(crs.n)
    # real matrix times complex vector
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, rval=crs.val)
    K.set_group(1)
    w = K * xz
    assert np.all(np.abs(w.real - crs.matvec(xz.real)) <= _bound(crs, xz.real))
    assert np.all(np.abs(w.imag - crs.matvec(xz.imag)) <= _bound(crs, xz.imag))
    # complex values
    cv = crs.val * (1.0 + 0.5j)
    Kc = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, cval=cv)
    Kc.set_group(1)
    ref = oracle_crs(P.Crs(crs.n, crs.row_ptr, crs.col_ind, cv), complex_vectors=True).user_mv(xz)
    assert rel(Kc * xz, ref) < 1e-14
    # operator with diagonal scaling: group 1 set after and before the
    # scaling (the scaled values get their sliced copy either way)
    c3, _ = P.convection_diffusion_3d(12, 10.0, reaction=1.0)
    x = rng.standard_normal(c3.n)
    Ao = oracle_crs(c3, diag_precon=True)
    K3 = idrs_amd.crs_format(c3.n, c3.n, c3.row_ptr, c3.col_ind, rval=c3.val)
    A = idrs_amd.Matrix(K3, which_precon=1)
    K3.set_group(1)
    assert rel(A * x, Ao.matvec(x)) < 1e-14
    K4 = idrs_amd.crs_format(c3.n, c3.n, c3.row_ptr, c3.col_ind, rval=c3.val)
    K4.set_group(1)
    A4 = idrs_amd.Matrix(K4, which_precon=1)
    assert rel(A4 * x, Ao.matvec(x)) < 1e-14
    # new values reach the sliced copy
    import ctypes
    from idrs_amd import _lib
    nv = np.ascontiguousarray(c3.val * 1.5 + 0.25)
    assert _lib.load().idrs_user_matrix_set_values_f64(K4.handle, ctypes.c_void_p(nv.ctypes.data), 0) == 0
    c5 = P.Crs(c3.n, c3.row_ptr, c3.col_ind, nv)
    assert np.all(np.abs(K4 * x - oracle_crs(c5).user_mv(x)) <= _bound(c5, x))


@pytest.mark.parametrize("case", ["2d_five", "band_hb5"])
def test_scaled_dia_unit_diagonal_other_counts(gpu, case):
    """The implicit unit diagonal in the 5-diagonal kernel and in the
    generic-count kernel (an 11-diagonal CDS band), against the oracle's
    scaled operator."""
    rng = np.random.default_rng(30)
    if case == "2d_five":
        _, dia = P.convection_diffusion_2d(37, 20.0, angle=0.7)
        n, off, vals = dia.vals.shape[0], np.asarray(dia.offsets), dia.vals
        K = idrs_amd.dia_format(n, n, off, rvals=vals)
        Ao = orc.OracleMatrix(orc.DIA, n, dia=(off, vals), which_precon=1)
    else:
        n, hb = 4099, 5
        band = rng.standard_normal((n, 2 * hb + 1))
        band[:, hb] = 4.0 + rng.uniform(0, 1, n)
        K = idrs_amd.cds_format(n, n, hb, rband=band)
        Ao = orc.OracleMatrix(orc.CDS, n, cds=(hb, band), which_precon=1)
        off = np.arange(-hb, hb + 1)
    A = idrs_amd.Matrix(K, which_precon=1)
    assert A.stored_values() == ((len(off) - 1) * n, True)
    x = rng.standard_normal(n)
    assert rel(A * x, Ao.matvec(x)) < 1e-14


def test_row_slabs_keep_results(gpu):
    """The row-slab count of one-thread-per-row products (autotuned or set)
    changes only the order of the epilogue's block sums: DIA products stay
    bit-identical, and a solve stays within the parity bar."""
    crs, dia = P.convection_diffusion_3d(24, 6.2, reaction=1.0)
    n = crs.n
    K = idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals)
    assert K.layout() == (0, 32768)
    x = np.random.default_rng(5).standard_normal(n)
    w0 = K * x
    K.autotune(reps=3)
    assert K.layout()[1] in (4096, 8192, 16384, 32768)
    for sl in (4096, 8192, 16384, 32768):
        K.set_row_slabs(sl)
        assert np.array_equal(K * x, w0)
    with pytest.raises(idrs_amd.IdrsError):
        K.set_row_slabs(1000)
    b = np.random.default_rng(0).uniform(-1, 1, n)
    PT = P.gaussian_shadow(n, 4, seed=0)
    poly = (2, 8, [0.15, 1.85])
    Ao = oracle_crs(crs, diag_precon=True, poly=poly)
    go = orc.idrs(Ao, b, 4, tolerance=0.0, maximum_iterations=30, PT_raw=PT)
    for sl in (4096, 16384):
        K.set_row_slabs(sl)
        A = idrs_amd.Matrix(K, which_precon=1)
        idrs_amd.real_preconditioner(A, *poly[:2], foci=poly[2])
        g = idrs_amd.idrs(A, b, 4, tolerance=0.0, maximum_iterations=30, PT_raw=PT)
        assert rel(g.x, go[0]) < 1e-10


def _block_operator(case, rng):
    if case == "dia7_scaled_cheb":
        crs, dia = P.convection_diffusion_3d(12, 10.0, reaction=1.0)
        n = crs.n
        A = idrs_amd.Matrix(idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals), which_precon=1)
        idrs_amd.real_preconditioner(A, 2, 8, foci=[0.15, 1.85])
    elif case == "dia5_seed":
        _, dia = P.convection_diffusion_2d(37, 20.0, angle=0.7)
        n = dia.vals.shape[0]
        A = idrs_amd.Matrix(idrs_amd.dia_format(n, n, np.asarray(dia.offsets), rvals=dia.vals))
        A.set_seed(0.3)
    elif case == "dia3_complex_vectors":
        n = 4099
        vals = rng.standard_normal((n, 3))
        A = idrs_amd.Matrix(idrs_amd.dia_format(n, n, np.array([-1, 0, 1]), rvals=vals))
    elif case == "dia_generic":
        n, hb = 4099, 5
        band = rng.standard_normal((n, 2 * hb + 1))
        A = idrs_amd.Matrix(idrs_amd.cds_format(n, n, hb, rband=band))
    elif case == "sliced_scaled_seed_cheb":
        crs, _ = P.convection_diffusion_3d(12, 10.0, reaction=1.0)
        n = crs.n
        K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
        K.set_group(1)
        A = idrs_amd.Matrix(K, which_precon=1)
        A.set_seed(0.2)
        idrs_amd.real_preconditioner(A, 2, 4, foci=[0.15, 1.85])
    elif case == "sliced_ragged":
        crs = _ragged(5000, 2, (7, 4000))
        n = crs.n
        K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
        K.set_group(1)
        A = idrs_amd.Matrix(K)
    elif case == "sliced_complex":
        crs = _ragged(3000, 5)
        n = crs.n
        K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, cval=crs.val * (1.0 + 0.5j))
        K.set_group(1)
        A = idrs_amd.Matrix(K)
    else:  # lanes4
        crs = _ragged(3000, 6)
        n = crs.n
        K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
        K.set_group(4)
        A = idrs_amd.Matrix(K)
    return A, n


@pytest.mark.parametrize("case", ["dia7_scaled_cheb", "dia5_seed", "dia3_complex_vectors", "dia_generic",
                                  "sliced_scaled_seed_cheb", "sliced_ragged", "sliced_complex", "lanes4"])
def test_block_products_match_columns(gpu, case):
    """A*V for a block of 11 vectors (launches of 8 + 3 columns) equals A*v
    one column at a time, bit for bit: the block kernels (DIA with 3, 5 or 7
    diagonals, the sliced CRS copy) repeat the one-column row sum, Horner
    term, seed shift and scaling; the other layouts (generic DIA, lanes per
    row) run the one-column product per column."""
    rng = np.random.default_rng(40)
    A, n = _block_operator(case, rng)
    V = rng.standard_normal((n, 11))
    if case in ("dia3_complex_vectors", "sliced_complex"):
        V = V + 1j * rng.standard_normal((n, 11))
    W = A * V
    assert W.shape == V.shape
    for j in range(11):
        assert np.array_equal(W[:, j], A * V[:, j]), j
    assert np.array_equal(A * V[:, 4:5], W[:, 4:5])
    with pytest.raises(idrs_amd.IdrsError):
        A * V[:, :0]


@pytest.mark.parametrize("fmt", ["sliced", "dia"])
def test_multishift_blocks_of_shifts(gpu, fmt):
    """Eleven shifts of the Chebyshev-preconditioned 3-D Laplacian (block
    products of 8 + 3 shifts in the final residuals and the scale-back):
    ms_scaleback of all shifts equals each shift's scale-back on its own bit
    for bit, every (K - sigma I) x = b holds, and the residual norms the
    solvers return agree with norms recomputed one shift at a time."""
    crs, dia = P.helmholtz_3d(12)
    n = crs.n
    if fmt == "sliced":
        K = idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val)
        K.set_group(1)
    else:
        K = idrs_amd.dia_format(n, n, dia.offsets, rvals=dia.vals)
    A = idrs_amd.Matrix(K)
    idrs_amd.real_preconditioner(A, 2, 4, foci=(0.05, 11.95))
    b = np.random.default_rng(22).uniform(-1, 1, n)
    sig = -0.25 * np.arange(11)
    sig_p, alpha_s = idrs_amd.ms_parameters(A, sig)
    Ks = crs.to_scipy()
    for solver in (idrs_amd.msidrs, idrs_amd.msqmridr):
        r = solver(A, b, sig_p, 4, tolerance=1e-10, maximum_iterations=2000)
        assert r.flag == 0
        for k in range(11):
            res = b - (A * r.x[:, k] - sig_p[k] * r.x[:, k])
            assert abs(np.linalg.norm(res) / np.linalg.norm(b) - r.relres[k]) <= 1e-12 * max(r.relres[k], 1e-300) + 1e-15
        X = idrs_amd.ms_scaleback(A, alpha_s, r.x)
        for k in range(11):
            Xk = idrs_amd.ms_scaleback(A, alpha_s[:, k:k + 1], r.x[:, k:k + 1])
            assert np.array_equal(X[:, k], Xk[:, 0]), k
            res = b - (Ks @ X[:, k] - sig[k] * X[:, k])
            assert np.linalg.norm(res) / np.linalg.norm(b) < 1e-8
