"""Shift-and-invert LU preconditioners on the GPU against the oracle
(MODULES/PP_MATRIX/user_module.f90:1383-1459, 1531-1602 REAL_PRECON /
COMPLEX_PRECON(2|3); :1604-2336 the LU routines; matrix_module.f90:306-400
the block system's shift-and-invert inverse, :402-549 / :724-806 the
multishift parameters and scale-back), then the reference's examples that
use them: EXAMPLES/RANDOM/band_real.f90 (band LU of the matrix itself) and
the wedge problems with -lu (EXAMPLES/MATRIX_MARKET/wedge_neumann.f90,
wedge_sommerfeld.f90 through PP_IDRS's RPP_MS_IDRS / CPP_MS_IDRS steps).

Bars: the factors are formed in the reference's order on both sides.  The
band solve of a half bandwidth below 64 sweeps columns in the reference's
order, so on real vectors its result equals the oracle's bit for bit; the
dense and wider band solves sum each row in another order, and complex
products and quotients round differently from C's, so there preconditioner
applications agree within 1e-12 of the largest entry (times the factor's
growth on the random band example).  Solves: the oracle's iteration count
within 2, or within twice the larger of the oracle's and the GPU's own
rounding spread (on wedge_neumann -lu -multi the GPU alone takes 49, 51 or
52 iterations when b moves by 1e-16), true residuals of every shifted
system within ten times the oracle's (or 1e-5).
"""
import numpy as np
import pytest

import idrs_amd
import orc

from helpers import mm_expanded, mm_fixture, rounding_spread

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(23)


def _band(n, hb, cplx, scale, diag):
    A = scale * RNG.standard_normal((n, n))
    if cplx:
        A = A + 1j * scale * RNG.standard_normal((n, n))
    A[np.abs(np.subtract.outer(np.arange(n), np.arange(n))) > hb] = 0
    return A + diag * np.eye(n)


def _coo(A):
    I, J = np.nonzero(A)
    return (I + 1).astype(np.int32), (J + 1).astype(np.int32), A[I, J]


def _um(A):
    I, J, V = _coo(A)
    kw = {"cval": V} if np.iscomplexobj(V) else {"rval": V}
    return idrs_amd.coo_format(A.shape[0], A.shape[0], len(I), I, J, **kw)


def _om(A, cplx=False, **kw):
    return orc.OracleMatrix(orc.COO, A.shape[0], coo=_coo(A), complex_vectors=cplx, **kw)


def _close(a, b, tol=1e-12):
    return np.abs(np.asarray(a) - np.asarray(b)).max() <= tol * np.abs(b).max()


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("which", [2, 3])
@pytest.mark.parametrize("case", ["noshift", "MC", "C", "M", "I"])
def test_lu_precon_matches_oracle(gpu, cplx, which, case):
    n = 60
    K, M, C = _band(n, 4, cplx, 1.0, 6.0), _band(n, 2, cplx, 0.3, 2.0), _band(n, 6, cplx, 0.2, 1.0)
    sh = None if case == "noshift" else ((0.5 + 0.25j) if cplx else 0.5)
    gm = {"PM": _um(M)} if "M" in case else {}
    gc = {"PC": _um(C)} if "C" in case else {}
    om = {"M": _om(M)} if "M" in case else {}
    oc = {"C": _om(C)} if "C" in case else {}
    Ku = _um(K)
    A = idrs_amd.Matrix(Ku)
    print(f"LU which={which} case={case} complex={cplx}", flush=True)
    A.set_precon_lu(which, Ku, shift=sh, **gm, **gc)
    lu = orc.precon_lu(which, _om(K), shift=sh, complex_values=cplx, **om, **oc)
    if which == 3:
        assert lu[1] == (6 if "C" in case and sh is not None else 4)
    v = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    assert _close(v / A, orc.lu_solve(lu, v))
    if which == 3 and not cplx:
        assert np.array_equal(v / A, orc.lu_solve(lu, v))
    # the operator's product divides by the factors first (real_mv's t = v/A)
    O = _om(K, cplx or np.iscomplexobj(sh), lu=lu)
    assert _close(A * v, O.matvec(v))


def test_zero_pivot(gpu):
    n = 10
    K = _band(n, 2, False, 1.0, 5.0)
    K[0, 0] = 0.0                   # the first pivot (later pivots are updated before they are tested)
    I, J, V = _coo(K)
    I, J, V = np.append(I, 1).astype(np.int32), np.append(J, 1).astype(np.int32), np.append(V, 0.0)
    Ku = idrs_amd.coo_format(n, n, len(I), I, J, rval=V)
    A = idrs_amd.Matrix(Ku)
    for which in (2, 3):
        with pytest.raises(idrs_amd.IdrsError, match="Pivot equal to zero"):
            A.set_precon_lu(which, Ku)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("with_mass", [False, True])
def test_block_shift_invert(gpu, cplx, with_mass):
    nn = 40
    K, M, C = _band(nn, 3, cplx, 1.0, 5.0), _band(nn, 1, cplx, 0.3, 2.0), _band(nn, 2, cplx, 0.2, 1.0)
    sh = (0.4 + 0.8j) if cplx else -0.7
    seed = (0.3 - 0.2j) if cplx else 0.25
    Ku, Mu, Cu = _um(K), _um(M), _um(C)
    A = idrs_amd.Matrix(Ku)
    if with_mass:
        A.set_mass(Mu)
    A.set_damping(Cu)
    A.set_flags(True, True)
    A.set_seed(seed)
    A.set_precon_lu(3, Ku, PM=Mu if with_mass else None, PC=Cu, shift=sh)
    A.set_shift_invert(sh)
    lu = orc.precon_lu(3, _om(K), M=_om(M) if with_mass else None, C=_om(C), shift=sh, complex_values=cplx)
    okw = dict(mass=_om(M, cplx)) if with_mass else {}
    O = _om(K, cplx, seed=seed, damping=_om(C, cplx), block=True, multishift=True, lu=lu, shift_invert=True,
            si_shift=sh, **okw)
    v = RNG.standard_normal(2 * nn) + (1j * RNG.standard_normal(2 * nn) if cplx else 0)
    assert _close(v / A, O.precon(v))
    assert _close(A * v, O.matvec(v))
    # ms scale-back with the shift-and-invert factor
    sig = np.array([0.1, 0.9]) + (0.5j if cplx else 0)
    sp, alpha_s = idrs_amd.ms_parameters(A, sig)
    spo, _ = O.ms_parameters(sig)
    assert np.allclose(sp, spo, rtol=1e-15, atol=0)
    b = RNG.standard_normal(2 * nn) + (1j * RNG.standard_normal(2 * nn) if cplx else 0)
    Y = RNG.standard_normal((2 * nn, 2)) + (1j * RNG.standard_normal((2 * nn, 2)) if cplx else 0)
    assert _close(idrs_amd.ms_scaleback_b(A, sig, alpha_s, b, Y), O.ms_scaleback_b(sig, None, b, Y))


def test_random_band_real(gpu):
    """band_real.f90: a random band matrix (values uniform on [0, 1) inside the
    band, CDS storage), b = A 1, P = real_precon(3, A), pp_idrs(A, b, P=P):
    the band LU of A itself, so IDR(4) converges at once."""
    n, hb = 2000, 5
    R = RNG.uniform(size=(n, 2 * hb + 1))
    for d in range(-hb, hb + 1):
        i = np.arange(n)
        R[(i + d < 0) | (i + d >= n), d + hb] = 0.0
    R[:, hb] += 2 * hb + 1           # keeps the unpivoted LU's growth moderate
    K = idrs_amd.cds_format(n, n, hb, rband=R)
    A = idrs_amd.Matrix(K)
    A.set_precon_lu(3, K)
    O = orc.OracleMatrix(orc.CDS, n, cds=(hb, R))
    lu = orc.precon_lu(3, O)
    assert lu[1] == hb
    b = O.user_mv(np.ones(n))
    assert np.array_equal(b / A, orc.lu_solve(lu, b))
    Oa = orc.OracleMatrix(orc.CDS, n, cds=(hb, R), lu=lu)
    g = idrs_amd.idrs(A, b, 4, tolerance=1e-6, maximum_iterations=1000)
    o = orc.idrs(Oa, b, 4, tolerance=1e-6, maximum_iterations=1000)
    assert g.flag == o[1] == 0 and g.iterations <= 3 and o[3] <= 3
    x = idrs_amd.scaleback(A, b, g.x)
    assert np.abs(x - 1).max() < 1e-8


def _wedge(k, cplx):
    mats = []
    for m in "KMC":
        fx = mm_fixture(f"wedge{k}_{m}")
        I, J, V = mm_expanded(fx)
        n = int(fx["size"][0])
        um = idrs_amd.coo_format(n, n, len(I), np.asarray(I, np.int32), np.asarray(J, np.int32), rval=V)
        om = orc.OracleMatrix(orc.COO, n, coo=(np.asarray(I, np.int32), np.asarray(J, np.int32), V),
                              complex_vectors=cplx)
        D = np.zeros((n, n))
        np.add.at(D, (np.asarray(I) - 1, np.asarray(J) - 1), V)
        mats.append((um, om, D))
    return mats, mm_fixture(f"wedge{k}_b")["vals"][:, 0].copy()


def _ms_check(g, o, run, run_g, b, res_g, res_o):
    print(f"  gpu {g.iterations} flag {g.flag}, oracle {o['iterations']} flag {o['flag']}; "
          f"residuals {np.max(res_g):.2e} / {np.max(res_o):.2e}", flush=True)
    assert g.flag == o["flag"] == 0
    if abs(g.iterations - o["iterations"]) > 2:
        spread = max(rounding_spread(run, b)[1], rounding_spread(run_g, b, trials=8)[1])
        assert abs(g.iterations - o["iterations"]) <= 2 * spread
    for rg, ro in zip(res_g, res_o):
        assert rg <= max(1e-5, 10 * ro), (res_g, res_o)


def test_wedge_neumann_lu_multishift(gpu):
    """wedge_neumann.f90 -lu -multi: P = real_precon(3, K, M, shift) with
    shift = -2 s_1 s_N/(s_1 + s_N), seed s_N, RPP_MS_IDRS with
    shift-and-invert: ms_parameters' eta = (sigma - seed)/(sigma - shift),
    MSIDRS(4) on Pmul(A, b), scaleback; residuals of K - sigma M."""
    ((Ku, Ko, Kd), (Mu, Mo, Md), _), b = _wedge(3, False)
    freq = np.array([1.0, 2.0, 4.0])
    sigma = (2 * np.pi * freq) ** 2
    rshift = -2 * sigma[0] * sigma[-1] / (sigma[0] + sigma[-1])
    seed = sigma[-1]
    A = idrs_amd.Matrix(Ku)
    A.set_mass(Mu)
    A.set_flags(False, True)
    A.set_seed(seed)
    A.set_precon_lu(3, Ku, PM=Mu, shift=rshift)
    A.set_shift_invert(rshift)
    lu = orc.precon_lu(3, Ko, M=Mo, shift=rshift)
    fx = mm_fixture("wedge3_K")
    I, J, V = mm_expanded(fx)
    O = orc.OracleMatrix(orc.COO, len(b), coo=(np.asarray(I, np.int32), np.asarray(J, np.int32), V), seed=seed,
                         mass=Mo, multishift=True, lu=lu, shift_invert=True, si_shift=rshift)
    eta, alpha_s = idrs_amd.ms_parameters(A, sigma)
    eta_o, _ = O.ms_parameters(sigma)
    assert np.allclose(eta, eta_o, rtol=1e-15)
    kw = dict(tolerance=1e-6, maximum_iterations=1000)
    print("neumann lu multishift", flush=True)
    g = idrs_amd.msidrs(A, A.pmul(b), eta, 4, **kw)
    o = orc.msidrs(O, O.pmul(b), eta_o, 4, **kw)
    X = idrs_amd.ms_scaleback_b(A, sigma, alpha_s, b, np.asarray(g.x))
    Xo = O.ms_scaleback_b(sigma, None, b, o["x"])
    res = lambda Z: [np.linalg.norm(b - Kd @ Z[:, k] + s * (Md @ Z[:, k])) / np.linalg.norm(b)
                     for k, s in enumerate(sigma)]
    run = lambda bb: (lambda q: (q["x"], q["iterations"]))(orc.msidrs(O, O.pmul(bb), eta_o, 4, **kw))
    run_g = lambda bb: (lambda q: (np.asarray(q.x), q.iterations))(idrs_amd.msidrs(A, A.pmul(bb), eta, 4, **kw))
    _ms_check(g, o, run, run_g, b, res(X), res(Xo))


def test_wedge_sommerfeld_lu_multishift_block(gpu):
    """wedge_sommerfeld.f90 -lu -multi: the block system of 2n rows with
    P = complex_precon(3, K, M, C, shift), shift = 2 pi/(f_N + f_1)
    ((f_N - f_1) sqrt(f_1 f_N) + 2 f_1 f_N i), seed at the last frequency,
    the block shift-and-invert preconditioner, MSIDRS(4), scaleback;
    residuals of K + sigma C + sigma^2 M for every frequency."""
    ((Ku, Ko, Kd), (Mu, Mo, Md), (Cu, Co, Cd)), rhs = _wedge(2, True)
    n = len(rhs)
    freq = np.array([1.0, 2.0, 4.0])
    sigma = 2 * np.pi * freq * 1j
    f1, fN = freq[0], freq[-1]
    cshift = 2 * np.pi / (fN + f1) * complex((fN - f1) * np.sqrt(f1 * fN), 2 * f1 * fN)
    seed = sigma[-1]
    A = idrs_amd.Matrix(Ku)
    A.set_mass(Mu)
    A.set_damping(Cu)
    A.set_flags(True, True)
    A.set_seed(seed)
    A.set_precon_lu(3, Ku, PM=Mu, PC=Cu, shift=cshift)
    A.set_shift_invert(cshift)
    lu = orc.precon_lu(3, Ko, M=Mo, C=Co, shift=cshift, complex_values=True)
    fx = mm_fixture("wedge2_K")
    I, J, V = mm_expanded(fx)
    O = orc.OracleMatrix(orc.COO, n, coo=(np.asarray(I, np.int32), np.asarray(J, np.int32), V), complex_vectors=True,
                         seed=seed, mass=Mo, damping=Co, block=True, multishift=True, lu=lu, shift_invert=True,
                         si_shift=cshift)
    bb = np.concatenate([-rhs, np.zeros(n)]).astype(complex)
    eta, alpha_s = idrs_amd.ms_parameters(A, sigma)
    eta_o, _ = O.ms_parameters(sigma)
    assert np.allclose(eta, eta_o, rtol=1e-15)
    kw = dict(tolerance=1e-6, maximum_iterations=1000)
    print("sommerfeld lu multishift block", flush=True)
    g = idrs_amd.msidrs(A, A.pmul(bb), eta, 4, **kw)
    o = orc.msidrs(O, O.pmul(bb), eta_o, 4, **kw)
    X = idrs_amd.ms_scaleback_b(A, sigma, alpha_s, bb, np.asarray(g.x))
    Xo = O.ms_scaleback_b(sigma, None, bb, o["x"])
    res = lambda Z: [np.linalg.norm(rhs - (Kd + s * Cd + s * s * Md) @ Z[n:, k]) / np.linalg.norm(rhs)
                     for k, s in enumerate(sigma)]
    run = lambda b1: (lambda q: (q["x"], q["iterations"]))(orc.msidrs(O, O.pmul(b1), eta_o, 4, **kw))
    run_g = lambda b1: (lambda q: (np.asarray(q.x), q.iterations))(idrs_amd.msidrs(A, A.pmul(b1), eta, 4, **kw))
    _ms_check(g, o, run, run_g, bb, res(X), res(Xo))


def test_wedge_neumann_lu_single_frequency(gpu):
    """wedge_neumann.f90 -lu without -multi: P = real_precon(3, K, M, shift)
    passed to pp_idrs without the shift, so v/P is the band LU solve of
    K - shift M, the system K - sigma M at sigma = (2 pi)^2; IDR(4)."""
    ((Ku, Ko, Kd), (Mu, Mo, Md), _), b = _wedge(3, False)
    sigma = (2 * np.pi) ** 2
    rshift = -2 * sigma * (2 * np.pi * 4) ** 2 / (sigma + (2 * np.pi * 4) ** 2)
    A = idrs_amd.Matrix(Ku)
    A.set_mass(Mu)
    A.set_seed(sigma)
    A.set_precon_lu(3, Ku, PM=Mu, shift=rshift)
    lu = orc.precon_lu(3, Ko, M=Mo, shift=rshift)
    fx = mm_fixture("wedge3_K")
    I, J, V = mm_expanded(fx)
    O = orc.OracleMatrix(orc.COO, len(b), coo=(np.asarray(I, np.int32), np.asarray(J, np.int32), V), seed=sigma,
                         mass=Mo, lu=lu)
    kw = dict(tolerance=1e-6, maximum_iterations=1000)
    g = idrs_amd.idrs(A, b, 4, **kw)
    o = orc.idrs(O, b, 4, **kw)
    print(f"neumann lu single: gpu {g.iterations} oracle {o[3]}", flush=True)
    assert g.flag == o[1] == 0
    if abs(g.iterations - o[3]) > 2:
        run = lambda bb: (lambda q: (q[0], q[3]))(orc.idrs(O, bb, 4, **kw))
        run_g = lambda bb: (lambda q: (np.asarray(q.x), q.iterations))(idrs_amd.idrs(A, bb, 4, **kw))
        spread = max(rounding_spread(run, b)[1], rounding_spread(run_g, b, trials=8)[1])
        assert abs(g.iterations - o[3]) <= 2 * spread
    x = idrs_amd.scaleback(A, b, g.x)
    assert np.linalg.norm(b - Kd @ x + sigma * (Md @ x)) / np.linalg.norm(b) < 1e-5
