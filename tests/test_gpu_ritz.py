"""IDRS's Hessenberg matrix, IDRS_RITZ and the polynomial preconditioner
from Ritz values, on the GPU, against the oracle (MODULES/IDRS/idrs.inc:
81-90, 179-181, 211-213, 231-237; MODULES/IDRS/ritz_module.f90:225-556;
MODULES/PP_MATRIX/matrix_module.f90:551-722, 1079-1282;
MODULES/PP_IDRS/real_preconditioner.inc:12-34).

Bars: at an equal fixed iteration count the Hessenberg matrices agree
within 1e-9 of their largest entry, or within ten times the oracle's own
rounding spread where that is larger (the entries are ratios of the small
s x s system's values: at s = 20 rounding alone, b perturbed at 1e-16, moves
the oracle's H by 2e-3 of its largest entry); at finite termination the
GPU's H(1:n,1:n) has A's eigenvalues; omegas within 1e-8 relative as a set
(their order breaks near-ties of |p| by rounding); recycle spaces within
1e-6 in principal angles; solutions within 1e-10 relative.
"""
import numpy as np
import pytest

import idrs_amd
import orc
import ritz
from idrs_amd import problems as P

from helpers import gpu_crs, oracle_crs, rel

pytestmark = pytest.mark.gpu


def _cd3(m=8, cplx=False):
    crs, _ = P.convection_diffusion_3d(m, 10.0, reaction=1.0)
    if cplx:
        crs.val = crs.val * (1.0 + 0.2j)
    return crs


def _rhs(n, cplx, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)


def _cosines(U, V):
    Qu, Qv = np.linalg.qr(U)[0], np.linalg.qr(V)[0]
    return np.linalg.svd(np.conj(Qu).T @ Qv)[1]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("s,m", [(1, 12), (4, 23), (8, 30), (20, 50)])
def test_hessenberg_matches_oracle(gpu, cplx, s, m):
    crs = _cd3(cplx=cplx)
    b = _rhs(crs.n, cplx, seed=s)
    nritz = (m // (s + 1)) * s + m % (s + 1)
    print(f"H s={s} its={m} nritz={nritz} complex={cplx}", flush=True)
    g = idrs_amd.idrs(gpu_crs(crs), b, s, tolerance=0.0, maximum_iterations=m, H=nritz + 3)
    o = orc.idrs(oracle_crs(crs), b, s, tolerance=0.0, maximum_iterations=m, H_nritz=nritz + 3)
    assert g.iterations == o[3] == m
    assert rel(g.x, o[0]) < 1e-10
    Ho = o[5]
    assert g.H.shape == Ho.shape
    diff = np.abs(g.H - Ho).max() / np.abs(Ho).max()
    if diff > 1e-9:
        spread, rng = 0.0, np.random.default_rng(0)
        for _ in range(3):
            b1 = b * (1 + 1e-16 * rng.standard_normal(b.shape))
            H1 = orc.idrs(oracle_crs(crs), b1, s, tolerance=0.0, maximum_iterations=m, H_nritz=nritz + 3)[5]
            spread = max(spread, np.abs(H1 - Ho).max() / np.abs(Ho).max())
        assert diff <= 10 * spread, (diff, spread)
    # columns beyond the inner steps taken stay zero
    assert np.all(g.H[:, nritz:] == 0) and np.any(g.H[:, nritz - 1] != 0)


@pytest.mark.parametrize("cplx", [False, True])
def test_hessenberg_finite_termination(gpu, cplx):
    n, s = 12, 4
    rng = np.random.default_rng(5)
    ev = np.linspace(1, 3, n) + (1j * np.linspace(-0.5, 0.5, n) if cplx else 0)
    Q = np.linalg.qr(rng.standard_normal((n, n)))[0]
    A = Q @ np.diag(ev) @ np.linalg.inv(Q)
    I, J = np.nonzero(np.ones((n, n)))
    kw = {"cval": A[I, J]} if cplx else {"rval": A[I, J]}
    K = idrs_amd.coo_format(n, n, n * n, (I + 1).astype(np.int32), (J + 1).astype(np.int32), **kw)
    b = _rhs(n, cplx, seed=1)
    g = idrs_amd.idrs(idrs_amd.Matrix(K), b, s, tolerance=0.0, maximum_iterations=n + n // s, H=n)
    ritzv = np.linalg.eigvals(g.H[:n, :n])
    for v in ev:
        assert np.min(np.abs(ritzv - v)) < 1e-7


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("with_m", [False, True])
def test_idrs_ritz_matches_oracle(gpu, cplx, with_m):
    crs = _cd3(m=10, cplx=cplx)
    n, s = crs.n, 4
    b = _rhs(n, cplx, seed=3)
    A, Ao = gpu_crs(crs), oracle_crs(crs)
    M = Mo = None
    if with_m:
        M, Mo = gpu_crs(crs, diag_precon=True), oracle_crs(crs, diag_precon=True)
    print(f"IDRS_RITZ complex={cplx} M={with_m}", flush=True)
    kw = dict(tolerance=1e-8, maximum_iterations=400)
    g = idrs_amd.idrs_ritz(A, b, s, M=M, n_omega=8, dim_u0=s, H_shape=(12, 10), resvec=True, **kw)
    o = ritz.idrs_ritz(Ao, b, s, M=Mo, n_omega=8, dim_u0=s, H_shape=(12, 10), **kw)
    assert g.flag == o["flag"] == 0 and g.iterations == o["iterations"]
    assert g.nritz == o["nritz"] == (g.iterations // (s + 1)) * s + g.iterations % (s + 1)
    assert rel(g.x, o["x"]) < 1e-10
    assert np.abs(g.H - o["H"]).max() <= 1e-9 * np.abs(o["H"]).max()
    assert np.isclose(g.omega[0], o["omega"][0], rtol=1e-8, atol=0)
    assert np.allclose(np.sort_complex(g.omega.astype(complex)), np.sort_complex(o["omega"].astype(complex)),
                       rtol=1e-8, atol=0)
    assert g.converged == o["converged"]
    assert np.allclose(np.conj(g.U0).T @ g.U0, np.eye(s), atol=1e-12)
    assert _cosines(g.U0, o["U0"]).min() > 1 - 1e-6


def test_idrs_ritz_errors(gpu):
    crs = _cd3(m=6)
    b = _rhs(crs.n, False)
    A = gpu_crs(crs)
    with pytest.raises(idrs_amd.IdrsError, match="reduce omega"):
        idrs_amd.idrs_ritz(A, b, 4, tolerance=0.0, maximum_iterations=3, n_omega=4)
    with pytest.raises(idrs_amd.IdrsError, match="reduce dimension of U0"):
        idrs_amd.idrs_ritz(A, b, 4, tolerance=0.0, maximum_iterations=3, dim_u0=4)
    with pytest.raises(idrs_amd.IdrsError, match="Hessenberg matrix too large"):
        idrs_amd.idrs_ritz(A, b, 4, tolerance=0.0, maximum_iterations=10, H_shape=(12, 10))


@pytest.mark.parametrize("kind", [1, 2])
def test_preconditioner_from_ritz_values(gpu, kind):
    """real_preconditioner.inc's chain: initial IDRS_RITZ iterations
    (ini_iter = n_ritz + n_ritz/s), the preconditioner's foci (Chebyshev)
    or center (Neumann) from analyse_spectrum of H(1:n_ritz,1:n_ritz), then
    the solve with the suggested omegas."""
    crs = _cd3(m=12)
    n, s, n_ritz, degree = crs.n, 4, 32, 8
    b = _rhs(n, False, seed=9)
    D = gpu_crs(crs, diag_precon=True)
    ini = n_ritz + n_ritz // s
    M1 = idrs_amd.Matrix(idrs_amd.crs_format(n, n, crs.row_ptr, crs.col_ind, rval=crs.val))
    g = idrs_amd.idrs_ritz(D, b, s, M=M1, tolerance=1e-8, maximum_iterations=ini, n_omega=s, H_shape=(ini + 1, ini))
    Do, M1o = oracle_crs(crs, diag_precon=True), oracle_crs(crs)
    o = ritz.idrs_ritz(Do, b, s, M=M1o, tolerance=1e-8, maximum_iterations=ini, n_omega=s, H_shape=(ini + 1, ini))
    assert g.iterations == o["iterations"] and g.nritz == o["nritz"] >= 16
    nr = g.nritz
    Hg, Ho = g.H[:nr, :nr], o["H"][:nr, :nr]
    idrs_amd.real_preconditioner(D, kind, degree, H=Hg)
    alpha_g = D.alpha.real
    f, c, _ = ritz.analyse_spectrum(Ho)
    alpha_o = orc.chebyshev(degree, f) if kind == 2 else orc.neumann(degree, c)
    assert np.allclose(alpha_g, alpha_o, rtol=1e-7, atol=1e-9 * np.abs(alpha_o).max())
    # the same analysis through the library's own host routine
    fl, cl, _ = idrs_amd.analyse_spectrum(Hg)
    assert np.allclose(fl, f, rtol=1e-8) and np.isclose(cl, c, rtol=1e-8)
    # the preconditioned solve with the suggested omegas converges on both sides
    r = idrs_amd.idrs(D, b, s, M1=M1, tolerance=1e-8, maximum_iterations=2000, omega=g.omega)
    Dp = orc.OracleMatrix(orc.CRS, n, crs=crs, which_precon=1, degree=degree, alpha=alpha_o)
    ro = orc.idrs(Dp, b, s, M1=M1o, tolerance=1e-8, maximum_iterations=2000, omega=o["omega"])
    assert r.flag == ro[1] == 0
    assert abs(r.iterations - ro[3]) <= max(2, 0.05 * ro[3])


@pytest.mark.parametrize("cplx", [False, True])
def test_recycle_space_solve(gpu, cplx):
    """PP_IDRS/recycle.inc: IDRS_RITZ returns U0(n, s); the next solve starts
    its first cycle from that search space (idrs.inc:55-65, 183-190).  The
    GPU's U0 drives both the GPU and the oracle solve (equal at a fixed
    count), and the recycled solve converges."""
    crs = _cd3(m=10, cplx=cplx)
    n, s = crs.n, 4
    A, Ao = gpu_crs(crs), oracle_crs(crs)
    g = idrs_amd.idrs_ritz(A, _rhs(n, cplx, seed=4), s, tolerance=1e-8, maximum_iterations=400, dim_u0=s)
    U0 = g.U0
    b2 = _rhs(n, cplx, seed=5)
    r = idrs_amd.idrs(A, b2, s, tolerance=0.0, maximum_iterations=25, U0=U0)
    ro = orc.idrs(Ao, b2, s, tolerance=0.0, maximum_iterations=25, U0=U0)
    assert rel(r.x, ro[0]) < 1e-10
    r = idrs_amd.idrs(A, b2, s, tolerance=1e-8, maximum_iterations=1000, U0=U0)
    ro = orc.idrs(Ao, b2, s, tolerance=1e-8, maximum_iterations=1000, U0=U0)
    assert r.flag == ro[1] == 0 and abs(r.iterations - ro[3]) <= 2
