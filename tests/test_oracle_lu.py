"""The oracle's shift-and-invert preconditioners (MODULES/PP_MATRIX/
user_module.f90:1383-1459 REAL_PRECON(2|3), :1604-1831 LU_decomp /
LU_solve, :1833-2336 get_band, real_band, band_LU_decomp, band_LU_solve;
matrix_module.f90:306-400 the block system's shift-and-invert inverse,
:402-549 ms_parameters, :724-806 ms_scaleback), pinned by what the
mathematics fixes:

  the factors multiply back to K [+ s C + s^2 M | + s C + s^2 I | - s M | - s I],
  the band restatement's factors equal the dense restatement's on a band
  matrix, solves satisfy P x = b, a zero pivot stops,
  the block inverse satisfies [-C-sM -K; I -sI] w = v,
  and the whole multishift chain: with eta = (sigma - seed)/(sigma - shift),
  y solving (A P^-1 - eta I) y = b (dense solve of the oracle's own operator)
  scales back to the solution of the shifted system (K - sigma M) x = b, and
  of K + sigma C + sigma^2 M through the block system.
"""
import numpy as np
import pytest

import orc

RNG = np.random.default_rng(17)


def _coo(A):
    I, J = np.nonzero(A)
    return (I + 1).astype(np.int32), (J + 1).astype(np.int32), A[I, J]


def _band(n, hb, cplx, scale, diag):
    A = scale * RNG.standard_normal((n, n))
    if cplx:
        A = A + 1j * scale * RNG.standard_normal((n, n))
    A[np.abs(np.subtract.outer(np.arange(n), np.arange(n))) > hb] = 0
    return A + diag * np.eye(n)


def _om(A, cplx=False, **kw):
    return orc.OracleMatrix(orc.COO, A.shape[0], coo=_coo(A), complex_vectors=cplx, **kw)


def _unpack(lu, n):
    """Dense L (unit lower) and U from precon_lu's factors."""
    which, hb, F = lu
    D = np.zeros((n, n), dtype=F.dtype)
    if which == 2:
        D[:] = F
    else:
        for i in range(n):
            for c in range(max(0, i - hb), min(n, i + hb + 1)):
                D[i, c] = F[i, c - i + hb]
    return np.tril(D, -1) + np.eye(n), np.triu(D)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("which", [2, 3])
def test_factors_multiply_back(cplx, which):
    n = 24
    K, M, C = _band(n, 3, cplx, 1.0, 6.0), _band(n, 2, cplx, 0.3, 2.0), _band(n, 1, cplx, 0.2, 1.0)
    s = (0.6 + 0.3j) if cplx else 0.6
    I = np.eye(n)
    cases = [
        (dict(), None, K),
        (dict(M=_om(M), C=_om(C)), s, K + s * C + s * s * M),
        (dict(C=_om(C)), s, K + s * C + s * s * I),
        (dict(M=_om(M)), s, K - s * M),
        (dict(), s, K - s * I),
        (dict(M=_om(M), C=_om(C)), None, K),      # no shift: K alone whatever M and C
    ]
    for kw, sh, P in cases:
        lu = orc.precon_lu(which, _om(K), shift=sh, complex_values=cplx, **kw)
        if which == 3:
            assert lu[1] == 3
        L, U = _unpack(lu, n)
        assert np.allclose(L @ U, P, rtol=0, atol=1e-13 * np.abs(P).max()), (kw.keys(), sh)
        b = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
        x = orc.lu_solve(lu, b)
        assert np.linalg.norm(P @ x - b) <= 1e-13 * np.linalg.norm(b) * np.linalg.cond(P)


@pytest.mark.parametrize("cplx", [False, True])
def test_band_restatement_equals_dense(cplx):
    n, hb = 40, 5
    K = _band(n, hb, cplx, 1.0, 8.0)
    M = _band(n, 2, cplx, 0.5, 1.0)
    sh = (1.5 - 0.2j) if cplx else 1.5
    lud = orc.precon_lu(2, _om(K), M=_om(M), shift=sh, complex_values=cplx)
    lub = orc.precon_lu(3, _om(K), M=_om(M), shift=sh, complex_values=cplx)
    assert lub[1] == hb
    Ld, Ud = _unpack(lud, n)
    Lb, Ub = _unpack(lub, n)
    # the same eliminations on the same entries: equal to the last bit
    assert np.array_equal(Ld, Lb) and np.array_equal(Ud, Ub)
    b = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    assert np.array_equal(orc.lu_solve(lud, b), orc.lu_solve(lub, b))


def test_get_band_and_zero_pivot():
    n = 12
    K = _band(n, 2, False, 1.0, 5.0)
    C = _band(n, 4, False, 0.1, 0.0)
    assert orc.get_band(_om(K)) == 2 and orc.get_band(_om(K), C=_om(C)) == 4
    Z = K.copy()
    Z[0, 0] = 0.0
    Z[0, 1] = 1.0
    I, J = np.nonzero(Z)
    I, J = np.append(I, 0), np.append(J, 0)        # an explicit zero keeps (1,1) stored
    Zo = orc.OracleMatrix(orc.COO, n, coo=((I + 1).astype(np.int32), (J + 1).astype(np.int32), Z[I, J]))
    for which in (2, 3):
        with pytest.raises(ValueError, match="Pivot equal to zero"):
            orc.precon_lu(which, Zo)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("with_mass", [False, True])
def test_block_shift_invert_inverse(cplx, with_mass):
    nn = 10
    K, M, C = _band(nn, 2, cplx, 1.0, 5.0), _band(nn, 1, cplx, 0.3, 2.0), _band(nn, 1, cplx, 0.2, 1.0)
    sh = (0.4 + 0.8j) if cplx else -0.7
    kw = dict(M=_om(M, cplx)) if with_mass else {}
    lu = orc.precon_lu(3, _om(K, cplx), C=_om(C, cplx), shift=sh, complex_values=cplx, **kw)
    okw = dict(mass=_om(M, cplx)) if with_mass else {}
    A = _om(K, cplx, damping=_om(C, cplx), block=True, lu=lu, shift_invert=True, si_shift=sh, **okw)
    v = RNG.standard_normal(2 * nn) + (1j * RNG.standard_normal(2 * nn) if cplx else 0)
    w = A.precon(v)
    I, Z = np.eye(nn), np.zeros((nn, nn))
    Ms = M if with_mass else I
    B = np.block([[-C - sh * Ms, -K], [I, -sh * I]])
    assert np.allclose(B @ w, v, rtol=0, atol=1e-12 * np.abs(v).max())
    # without shift-and-invert the block preconditioner divides the first half only
    A2 = _om(K, cplx, damping=_om(C, cplx), block=True, lu=lu, **okw)
    w2 = A2.precon(v)
    P = K + sh * C + sh * sh * Ms
    assert np.allclose(P @ w2[:nn], v[:nn], atol=1e-12 * np.abs(v).max()) and np.array_equal(w2[nn:], v[nn:])


def test_ms_parameters_shift_invert():
    n = 8
    K = _band(n, 1, False, 1.0, 4.0)
    A = _om(K, seed=0.5, shift_invert=True, si_shift=-1.0)
    sig = np.array([0.0, 2.0, -3.0])
    sp, _ = A.ms_parameters(sig)
    assert np.allclose(sp, (sig - 0.5) / (sig + 1.0), rtol=1e-15)
    with pytest.raises(ValueError):
        A.ms_parameters(np.array([0.0, -1.0]))
    with pytest.raises(ValueError):
        _om(K, seed=2.0, shift_invert=True, si_shift=2.0).ms_parameters(sig)


def _dense_op(A, n, cplx):
    dt = np.complex128 if cplx else np.float64
    return np.stack([A.matvec(np.eye(n, dtype=dt)[:, j]) for j in range(n)], axis=1)


@pytest.mark.parametrize("cplx", [False, True])
def test_shift_invert_multishift_chain(cplx):
    """(K - sigma M) x = b for several sigma through the shift-and-invert
    operator: eta from ms_parameters, y from the dense shifted system of the
    oracle's operator, x from ms_scaleback."""
    n = 20
    K, M = _band(n, 2, cplx, 1.0, 5.0), _band(n, 1, cplx, 0.2, 2.0)
    seed, sh = ((1.0 + 0.5j, -2.0 + 1j) if cplx else (1.0, -2.0))
    lu = orc.precon_lu(3, _om(K, cplx), M=_om(M, cplx), shift=sh, complex_values=cplx)
    A = _om(K, cplx, seed=seed, mass=_om(M, cplx), lu=lu, shift_invert=True, si_shift=sh)
    sig = np.array([0.0, 0.5, 3.0]) + (0.25j if cplx else 0)
    eta, alpha_s = A.ms_parameters(sig)
    Ad = _dense_op(A, n, cplx)
    b = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    Y = np.stack([np.linalg.solve(Ad - e * np.eye(n), b) for e in eta], axis=1)
    X = A.ms_scaleback_b(sig, alpha_s, b, Y)
    for k, s in enumerate(sig):
        r = np.linalg.norm(b - (K - s * M) @ X[:, k]) / np.linalg.norm(b)
        assert r < 1e-12, (k, r)


def test_shift_invert_block_chain():
    """The damped system K + sigma C + sigma^2 M through the block system
    with the shift-and-invert block preconditioner: right-hand side
    [-b; 0], solution the second half (pp_idrs_module.f90:422-447)."""
    nn = 12
    K, M, C = _band(nn, 2, True, 1.0, 5.0), _band(nn, 1, True, 0.2, 2.0), _band(nn, 1, True, 0.1, 1.0)
    seed, sh = 0.8j, 0.3 + 1.2j
    cplx = True
    lu = orc.precon_lu(3, _om(K, cplx), M=_om(M, cplx), C=_om(C, cplx), shift=sh, complex_values=True)
    A = _om(K, cplx, seed=seed, mass=_om(M, cplx), damping=_om(C, cplx), block=True, multishift=True, lu=lu,
            shift_invert=True, si_shift=sh)
    sig = np.array([0.4j, 1.6j, 2.0j])
    eta, alpha_s = A.ms_parameters(sig)
    Ad = _dense_op(A, 2 * nn, cplx)
    rhs = RNG.standard_normal(nn) + 1j * RNG.standard_normal(nn)
    bb = np.concatenate([-rhs, np.zeros(nn)])
    Y = np.stack([np.linalg.solve(Ad - e * np.eye(2 * nn), bb) for e in eta], axis=1)
    X = A.ms_scaleback_b(sig, alpha_s, bb, Y)
    for k, s in enumerate(sig):
        x = X[nn:, k]
        r = np.linalg.norm(rhs - (K + s * C + s * s * M) @ x) / np.linalg.norm(rhs)
        assert r < 1e-12, (k, r)
