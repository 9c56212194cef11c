"""The oracle's generalized operator and deflation (MODULES/PP_MATRIX/
matrix_module.f90: real_mv/complex_mv :169-264 with mass, damping and block
systems, rprecon/cprecon :306-400, Pmul :1369-1431, make_deflation
:1284-1352, add_deflation :1433-1623), pinned against dense closed forms:

  K - s M,  K + s C + s^2 M,  K + s C + s^2 I,  K - s I  (times D^-1),
  the block system [-C -K; I 0] - s [M 0; 0 I] on (v1/d, v2),
  the projectors I - AV E^-1 V^H (complex), I - V E^-1 V^H (real, as the
  reference writes it) and I - V V^H (multishift), E = V^H A V,
  add_deflation's V E^-1 V^H b + (I - V E^-1 V^H A) y,
and the property that makes deflation work: with the complex projector, y
solving P A y = P b gives add_deflation(b, y) solving A x = b.
"""
import numpy as np
import pytest

import orc

RNG = np.random.default_rng(7)


def _coo(A):
    n = A.shape[0]
    I, J = np.nonzero(A)
    return (I + 1).astype(np.int32), (J + 1).astype(np.int32), A[I, J]


def _rand(n, cplx, scale=1.0, diag=4.0):
    A = scale * RNG.standard_normal((n, n))
    if cplx:
        A = A + 1j * scale * RNG.standard_normal((n, n))
    A[np.abs(A) < 0.8 * scale] = 0
    return A + diag * np.eye(n)


def _om(A, cplx, **kw):
    return orc.OracleMatrix(orc.COO, A.shape[0], coo=_coo(A), complex_vectors=cplx, **kw)


@pytest.mark.parametrize("cplx", [False, True])
def test_shifted_operators_dense(cplx):
    n = 10
    K, M, C = _rand(n, cplx), _rand(n, cplx, 0.3, 2.0), _rand(n, cplx, 0.2, 1.0)
    d = RNG.uniform(1.0, 2.0, n) + (0.3j * RNG.uniform(size=n) if cplx else 0)
    s = (0.7 - 0.4j) if cplx else 0.7
    v = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    t = v / d
    dt = np.complex128 if cplx else np.float64
    cases = [
        (dict(mass=_om(M, cplx), damping=_om(C, cplx)), (K + s * C + s * s * M) @ t),
        (dict(damping=_om(C, cplx)), (K + s * C + s * s * np.eye(n)) @ t),
        (dict(mass=_om(M, cplx)), (K - s * M) @ t),
        (dict(), (K - s * np.eye(n)) @ t),
    ]
    for kw, want in cases:
        A = _om(K, cplx, which_precon=1, precon_diag=d, seed=s, **kw)
        got = A.system_mv(v.astype(dt))
        assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max()), kw.keys()
    # seed zero: the plain product K t whatever matrices are present
    A = _om(K, cplx, which_precon=1, precon_diag=d, mass=_om(M, cplx))
    assert np.allclose(A.system_mv(v), K @ t, atol=1e-12 * np.abs(K @ t).max())


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("with_mass", [False, True])
def test_block_system_dense(cplx, with_mass):
    nn = 8
    K, M, C = _rand(nn, cplx), _rand(nn, cplx, 0.3, 2.0), _rand(nn, cplx, 0.2, 1.0)
    d = RNG.uniform(1.0, 2.0, nn)
    s = (0.3 + 0.9j) if cplx else -0.6
    v = RNG.standard_normal(2 * nn) + (1j * RNG.standard_normal(2 * nn) if cplx else 0)
    kw = dict(mass=_om(M, cplx)) if with_mass else {}
    A = _om(K, cplx, which_precon=1, precon_diag=d, seed=s, damping=_om(C, cplx), block=True, **kw)
    assert A.rows == 2 * nn
    Z, I = np.zeros((nn, nn)), np.eye(nn)
    B = np.block([[-C, -K], [I, Z]])
    S = np.block([[M if with_mass else I, Z], [Z, I]])
    t = np.concatenate([v[:nn] / d, v[nn:]])
    want = (B - s * S) @ t
    assert np.allclose(A.system_mv(v), want, rtol=0, atol=1e-12 * np.abs(want).max())
    # v/A divides the first half only
    assert np.allclose(A.precon(v), t, rtol=0, atol=4e-16 * np.abs(t).max())


def _defl_setup(cplx, multishift=False, nd=3, n=12):
    K, M = _rand(n, cplx), _rand(n, cplx, 0.3, 2.0)
    s = (0.4 + 0.2j) if cplx else 0.4
    A = _om(K, cplx, seed=s, mass=_om(M, cplx), multishift=multishift)
    V = RNG.standard_normal((n, nd)) + (1j * RNG.standard_normal((n, nd)) if cplx else 0)
    if multishift:
        V = np.linalg.qr(V)[0]
    Aop = K - s * M                     # real_mv without a preconditioner
    return A, Aop, V


@pytest.mark.parametrize("cplx", [False, True])
def test_deflation_projectors_dense(cplx):
    A, Aop, V = _defl_setup(cplx)
    AV, E = A.make_deflation(V)
    assert np.allclose(AV, Aop @ V, atol=1e-12 * np.abs(AV).max())
    assert np.allclose(E, np.conj(V).T @ Aop @ V, atol=1e-12 * np.abs(E).max())
    n = V.shape[0]
    w = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    Einv = np.linalg.inv(E)
    if cplx:
        P = np.eye(n) - AV @ Einv @ np.conj(V).T
    else:
        P = np.eye(n) - V @ Einv @ V.T          # the real version as written
    assert np.allclose(A.pmul(w), P @ w, atol=1e-11 * np.abs(w).max())
    # operator(*) = Pmul(real_mv(polpre(v))): without a polynomial, P A v
    assert np.allclose(A.matvec(w), P @ Aop @ w, atol=1e-11 * np.abs(Aop @ w).max())
    if cplx:
        assert np.allclose(A.pmul(AV[:, 0]), 0, atol=1e-11 * np.abs(AV).max())
    # multishift: I - V V^H with orthonormal V
    Am, _, Vm = _defl_setup(cplx, multishift=True)
    Am.make_deflation(Vm)
    assert np.allclose(Am.pmul(w), w - Vm @ (np.conj(Vm).T @ w), atol=1e-12 * np.abs(w).max())


@pytest.mark.parametrize("cplx", [False, True])
def test_add_deflation_dense(cplx):
    A, Aop, V = _defl_setup(cplx)
    AV, E = A.make_deflation(V)
    n = V.shape[0]
    b = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    y = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    Ei = np.linalg.inv(E)
    Vh = np.conj(V).T
    want = V @ Ei @ Vh @ b + y - V @ Ei @ Vh @ (Aop @ y)
    assert np.allclose(A.add_deflation(b, y), want, atol=1e-11 * np.abs(want).max())
    if cplx:
        # y with P A y = P b (least squares on the deflated system), then
        # add_deflation(b, y) solves A x = b
        P = np.eye(n) - AV @ Ei @ Vh
        y = np.linalg.lstsq(P @ Aop, P @ b, rcond=None)[0]
        x = A.add_deflation(b, y)
        assert np.linalg.norm(Aop @ x - b) < 1e-10 * np.linalg.norm(b)
    # multishift: per shift sigma, Hs = E - (sigma - seed) I
    sig = np.array([0.1, -0.5, 1.3]) + (0.2j if cplx else 0)
    Y = RNG.standard_normal((n, 3)) + (1j * RNG.standard_normal((n, 3)) if cplx else 0)
    X = A.ms_add_deflation(sig, b, Y)
    seed = A.m.cseed[0] + 1j * A.m.cseed[1] if cplx else A.m.rseed
    for k in range(3):
        ss = sig[k] - seed
        Hi = np.linalg.inv(E - ss * np.eye(3))
        want = V @ Hi @ Vh @ b + Y[:, k] - V @ Hi @ Vh @ (Aop @ Y[:, k] - ss * Y[:, k])
        assert np.allclose(X[:, k], want, atol=1e-11 * np.abs(want).max())


def test_deflated_solve_complex():
    """PP_IDRS with V (pp_idrs_module.f90:156-161, run_idrs.inc,
    matrix_module.f90:808-849): IDR(4) on Pmul(A, .) with rhs Pmul(A, b),
    then scaleback with add_deflation; the true residual of A x = b."""
    n = 60
    K = _rand(n, True, 0.5, 6.0)
    M = _rand(n, True, 0.1, 1.0)
    s = 0.5 + 0.5j
    A = _om(K, True, seed=s, mass=_om(M, True))
    V = RNG.standard_normal((n, 2)) + 1j * RNG.standard_normal((n, 2))
    A.make_deflation(V)
    b = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    y, flag, relres, it, _ = orc.idrs(A, A.pmul(b), 4, tolerance=1e-12, maximum_iterations=400)
    x = A.scaleback_b(b, y)
    Aop = K - s * M
    assert flag == 0 and np.linalg.norm(Aop @ x - b) < 1e-10 * np.linalg.norm(b)
