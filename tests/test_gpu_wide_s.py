"""Shadow spaces wider than the one-pass P^T y kernels hold (s > 16, up to
the reference's own s = 64 of EXAMPLES/TOEPLITZ/toeplitz.f90:60-61, 82-83):
the P^T y sums then run as chunks of 16 rows, one grid row per chunk.

Bars as in the other parity tests: the complex Toeplitz accuracy problem
(tol 1e-12) reaches the accuracy on both sides with iteration counts within
3 %; at an equal fixed iteration count the solutions agree with the
oracle's within 1e-10 relative (FP64), per shift for the multishift
solvers.
"""
import numpy as np
import pytest

import idrs_amd
import orc
from idrs_amd import problems as P

from helpers import gpu_crs, oracle_crs, rel, rounding_spread

pytestmark = pytest.mark.gpu


def _toeplitz_complex(n=200):
    """toeplitz.f90:41-45: A%cind = (-1, 0, 2, 3), A%cval = (3.6i, 4, 1, 0.7), b = i."""
    off = np.array([-1, 0, 2, 3], dtype=np.int32)
    cv = np.array([3.6j, 4.0, 1.0, 0.7])
    vals = np.tile(cv, (n, 1)).astype(np.complex128)
    i = np.arange(n)
    for d, o in enumerate(off):
        vals[(i + o < 0) | (i + o >= n), d] = 0
    K = idrs_amd.dia_format(n, n, off, cvals=vals)
    mk = lambda: orc.OracleMatrix(orc.DIA, n, dia=(off, vals), complex_vectors=True)
    return K, mk, np.full(n, 1j)


@pytest.mark.parametrize("s", [32, 64])
def test_complex_toeplitz_idrs_wide_s(gpu, s):
    K, mk, b = _toeplitz_complex()
    A = idrs_amd.Matrix(K)
    r = idrs_amd.idrs(A, b, s, M1=idrs_amd.Matrix(K), tolerance=1e-12, maximum_iterations=4000)
    xo, fo, ro, io, _ = orc.idrs(mk(), b, s, M1=mk(), tolerance=1e-12, maximum_iterations=4000)
    print(f"IDRS s={s} gpu it={r.iterations} relres={r.relres:.2e} oracle it={io} relres={ro:.2e}")
    assert r.flag == fo == 0 and abs(r.iterations - io) <= max(2, 0.03 * io)
    assert r.relres <= 1e-12 and ro <= 1e-12


@pytest.mark.parametrize("s", [32, 64])
def test_complex_toeplitz_qmridr_wide_s(gpu, s):
    K, mk, b = _toeplitz_complex()
    A = idrs_amd.Matrix(K)
    r = idrs_amd.qmridr(A, b, s, M1=idrs_amd.Matrix(K), tolerance=1e-12, maximum_iterations=4000)
    xo, fo, ro, io, _ = orc.qmridr(mk(), b, s, M1=mk(), tolerance=1e-12, maximum_iterations=4000)
    print(f"QMRIDR s={s} gpu it={r.iterations} relres={r.relres:.2e} oracle it={io} relres={ro:.2e}")
    assert r.flag == fo == 0 and abs(r.iterations - io) <= max(2, 0.03 * io)
    assert r.relres <= 1e-12 and ro <= 1e-12


@pytest.mark.parametrize("s", [17, 40, 64])
def test_idrs_wide_s_fixed_count(gpu, s):
    """IDR(s) with P^T y in chunks (product, omega update, MAKE_P's
    Gram-Schmidt above 16 rows) against the oracle at a fixed count."""
    crs, _ = P.convection_diffusion_3d(9, 20.0, reaction=2.0)
    b = np.random.default_rng(s).standard_normal(crs.n)
    A, Ao = gpu_crs(crs, diag_precon=True, poly=(2, 4, [0.15, 1.85])), oracle_crs(crs, diag_precon=True,
                                                                                  poly=(2, 4, [0.15, 1.85]))
    PT = P.gaussian_shadow(crs.n, s, seed=s)
    m = 2 * (s + 1) + s // 2
    g = idrs_amd.idrs(A, b, s, tolerance=0.0, maximum_iterations=m, PT_raw=PT, resvec=True)
    o = orc.idrs(Ao, b, s, tolerance=0.0, maximum_iterations=m, PT_raw=PT, resvec=True)
    assert g.iterations == o[3] == m
    # the residual norm of this convection-dominated system first grows to
    # ten times its start, so rounding alone (b perturbed at 1e-16) moves the
    # oracle's own history by up to 40 % at the end of the run: the bars are
    # ten times that spread, per entry, and at least 1e-9 relative
    dx, dr = 0.0, np.zeros_like(o[4])
    rng = np.random.default_rng(0)
    for _ in range(3):
        b1 = b * (1 + 1e-16 * rng.standard_normal(b.shape))
        o1 = orc.idrs(oracle_crs(crs, diag_precon=True, poly=(2, 4, [0.15, 1.85])), b1, s, tolerance=0.0,
                      maximum_iterations=m, PT_raw=PT, resvec=True)
        dx = max(dx, rel(o1[0], o[0]))
        dr = np.maximum(dr, np.abs(o1[4] - o[4]) / o[4])
    assert rel(g.x, o[0]) < max(1e-10, 10 * dx)
    keep = o[4] > 1e-5 * o[4][0]
    assert np.all(np.abs(g.resvec[keep] - o[4][keep]) <= np.maximum(1e-9, 10 * dr[keep]) * o[4][keep])
    r = idrs_amd.idrs(A, b, s, tolerance=1e-9, maximum_iterations=2000, PT_raw=PT)
    ro = orc.idrs(Ao, b, s, tolerance=1e-9, maximum_iterations=2000, PT_raw=PT)
    slack = 2
    if abs(r.iterations - ro[3]) > slack:
        run = lambda bb: (lambda q: (q[0], q[3]))(orc.idrs(oracle_crs(crs, diag_precon=True, poly=(2, 4, [0.15, 1.85])),
                                                           bb, s, tolerance=1e-9, maximum_iterations=2000, PT_raw=PT))
        slack = max(slack, 2 * rounding_spread(run, b)[1])
    assert r.flag == ro[1] == 0 and abs(r.iterations - ro[3]) <= slack, (r.iterations, ro[3], slack)
    # the library's own stream at a wide s, and the orthonormal shadow space
    PTg = idrs_amd.make_p(crs.n, s, 1)
    assert rel(PTg, orc.make_p(crs.n, s, 1)) < 1e-12
    assert np.allclose(PTg.T @ PTg, np.eye(s), atol=1e-12)


@pytest.mark.parametrize("s", [20, 48])
def test_qmridr_wide_s_fixed_count(gpu, s):
    crs, _ = P.convection_diffusion_3d(9, 20.0, reaction=3.0)
    b = np.random.default_rng(3).standard_normal(crs.n)
    A, Ao = gpu_crs(crs), oracle_crs(crs)
    PT = P.gaussian_shadow(crs.n, s, seed=5)
    g = idrs_amd.qmridr(A, b, s, tolerance=0.0, maximum_iterations=s + 12, PT_raw=PT, resvec=True)
    o = orc.qmridr(Ao, b, s, tolerance=0.0, maximum_iterations=s + 12, PT_raw=PT, resvec=True)
    assert g.iterations == o[3]
    assert rel(g.x, o[0]) < 1e-10
    keep = o[4] > 1e-5 * o[4][0]
    assert np.allclose(g.resvec[keep], o[4][keep], rtol=1e-9, atol=0)


@pytest.mark.parametrize("solver,s,nsig", [("msidrs", 24, 3), ("msqmridr", 20, 3), ("msidrs", 24, 40),
                                           ("msqmridr", 40, 40)])
def test_multishift_wide_s_fixed_count(gpu, solver, s, nsig):
    """MSIDRS / MSQMRIDR with a wide shadow space; 40 shifts at s = 40 read
    their per-shift update coefficients from device memory (more than the
    update kernels stage in LDS).  (MSIDRS itself breaks down at s = 40 on
    these systems, in the oracle too: NaN after the first cycle.)"""
    crs, _ = P.convection_diffusion_3d(12, 5.0, reaction=1.0)
    b = np.random.default_rng(11).standard_normal(crs.n)
    A, Ao = gpu_crs(crs), oracle_crs(crs)
    sig = -np.linspace(0.0, 300.0, nsig)
    PT = P.gaussian_shadow(crs.n, s, seed=2)
    kw = dict(tolerance=0.0, maximum_iterations=s + 20, PT_raw=PT)
    g = getattr(idrs_amd, solver)(A, b, sig, s, **kw)
    o = getattr(orc, solver)(Ao, b, sig, s, **kw)
    assert g.iterations == o["iterations"]
    assert np.isfinite(o["x"]).all()
    for k in range(nsig):
        assert rel(g.x[:, k], o["x"][:, k]) < 1e-10, k
