"""Mass and damping matrices, block systems and deflation on the GPU against
the oracle (MODULES/PP_MATRIX/matrix_module.f90:169-264, 306-400,
724-849, 1284-1623), first operator by operator on small systems, then the
reference's wedge examples (EXAMPLES/MATRIX_MARKET/wedge_neumann.f90,
wedge_sommerfeld.f90, over their data kept as fixtures) through the
PP_IDRS call sequences (pp_idrs_module.f90, real_matrix.inc,
run_idrs.inc, run_ms_idrs.inc).

Bars: operator applications within 1e-13 of the largest entry (the same
products and elementwise steps in the same order; the sums of the
deflation's small systems differ by their reduction order); solves: the
oracle's iteration count within 2, or within twice the oracle's own
rounding spread where that is larger (b perturbed at 1e-16 moves the
multishift block solves on wedge2 by 40-60 of about 450 iterations), and
the scaled-back solutions' true residuals of the original systems within
ten times the oracle's (or 1e-5).
"""
import numpy as np
import pytest

import idrs_amd
import orc

from helpers import mm_expanded, mm_fixture, rel, rounding_spread

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(11)


def _dense(n, cplx, scale, diag):
    A = scale * RNG.standard_normal((n, n)) + (1j * scale * RNG.standard_normal((n, n)) if cplx else 0)
    A[np.abs(A) < 0.8 * scale] = 0
    return A + diag * np.eye(n)


def _coo(A):
    I, J = np.nonzero(A)
    return (I + 1).astype(np.int32), (J + 1).astype(np.int32), A[I, J]


def _gpu_um(I, J, V, n):
    kw = {"cval": V} if np.iscomplexobj(V) else {"rval": V}
    return idrs_amd.coo_format(n, n, len(I), I, J, **kw)


def _orc_om(I, J, V, n, cplx, **kw):
    return orc.OracleMatrix(orc.COO, n, coo=(I, J, V), complex_vectors=cplx, **kw)


class Pair:
    """The same type(matrix) on the GPU and in the oracle."""

    def __init__(self, K, cplx, seed=0.0, M=None, C=None, block=False, multishift=False, pdiag_of=None):
        n = K[3]
        self.cplx = cplx
        self.KU = _gpu_um(*K)
        self.A = idrs_amd.Matrix(self.KU)
        okw = {}
        if M is not None:
            self.MU = _gpu_um(*M)
            self.A.set_mass(self.MU)
            okw["mass"] = _orc_om(*M, cplx)
        if C is not None:
            self.CU = _gpu_um(*C)
            self.A.set_damping(self.CU)
            okw["damping"] = _orc_om(*C, cplx)
        if seed != 0:
            self.A.set_seed(seed)
        self.A.set_flags(block, multishift)
        if pdiag_of is not None:
            self.PU = _gpu_um(*pdiag_of)
            self.A.set_precon_diag(self.PU)
            d = _orc_om(*pdiag_of, cplx).diagonal()
            okw.update(which_precon=1, precon_diag=d)
        self.O = _orc_om(K[0], K[1], K[2], n, cplx, seed=seed, block=block, multishift=multishift, **okw)


def _mats(n, cplx):
    K, M, C = _dense(n, cplx, 1.0, 4.0), _dense(n, cplx, 0.3, 2.0), _dense(n, cplx, 0.2, 1.0)
    P = np.diag(RNG.uniform(1.0, 2.0, n))
    return [(*_coo(X), n) for X in (K, M, C, P)]


def _close(a, b, tol=1e-13):
    return np.abs(np.asarray(a) - np.asarray(b)).max() <= tol * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("kind", ["mass", "damping", "mass+damping", "seed", "block", "block+mass"])
def test_system_mv_and_products(gpu, cplx, kind):
    n = 40
    K, M, C, P = _mats(n, cplx)
    s = (0.6 - 0.3j) if cplx else 0.6
    kw = dict(seed=s, pdiag_of=P)
    if "mass" in kind:
        kw["M"] = M
    if "damping" in kind or "block" in kind:
        kw["C"] = C
    if "block" in kind:
        kw["block"] = True
    pr = Pair(K, cplx, **kw)
    rows = 2 * n if "block" in kind else n
    assert pr.A.rows == pr.O.rows == rows
    v = RNG.standard_normal(rows) + (1j * RNG.standard_normal(rows) if cplx else 0)
    assert _close(pr.A.system_mv(v), pr.O.system_mv(v))
    assert _close(pr.A * v, pr.O.matvec(v))
    assert _close(v / pr.A, pr.O.precon(v), 4e-16)
    # with a Neumann polynomial: operator(*) = real_mv(polpre(v))
    idrs_amd.complex_preconditioner(pr.A, 1, 3, center=2.5) if cplx else idrs_amd.real_preconditioner(pr.A, 1, 3,
                                                                                                      center=2.5)
    alpha = pr.A.alpha if cplx else pr.A.alpha.real
    O2 = orc.OracleMatrix(orc.COO, n, coo=K[:3], complex_vectors=cplx, seed=s, degree=3, alpha=alpha,
                          block="block" in kind, which_precon=1, precon_diag=_orc_om(*P, cplx).diagonal(),
                          mass=_orc_om(*M, cplx) if "mass" in kind else None,
                          damping=_orc_om(*C, cplx) if ("damping" in kind or "block" in kind) else None)
    assert _close(pr.A * v, O2.matvec(v), 1e-12)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("multishift", [False, True])
def test_deflation_parity(gpu, cplx, multishift):
    n, nd = 48, 3
    K, M, C, P = _mats(n, cplx)
    s = (0.4 + 0.2j) if cplx else 0.4
    pr = Pair(K, cplx, seed=s, M=M, multishift=multishift)
    V = RNG.standard_normal((n, nd)) + (1j * RNG.standard_normal((n, nd)) if cplx else 0)
    if multishift:
        V = np.linalg.qr(V)[0]
    pr.A.make_deflation(V)
    pr.O.make_deflation(V)
    w = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    b = RNG.standard_normal(n) + (1j * RNG.standard_normal(n) if cplx else 0)
    assert _close(pr.A.pmul(w), pr.O.pmul(w), 1e-12)
    assert _close(pr.A * w, pr.O.matvec(w), 1e-12)
    assert _close(pr.A.add_deflation(b, w), pr.O.add_deflation(b, w), 1e-12)
    assert _close(idrs_amd.scaleback(pr.A, b, w), pr.O.scaleback_b(b, w), 1e-12)
    sig = np.array([0.1, -0.7]) + (0.3j if cplx else 0)
    Y = RNG.standard_normal((n, 2)) + (1j * RNG.standard_normal((n, 2)) if cplx else 0)
    X = idrs_amd.ms_scaleback_b(pr.A, sig, np.ones((1, 2)), b, Y)
    Xo = pr.O.ms_add_deflation(sig, b, Y)
    assert _close(X, Xo, 1e-12)


def _wedge(k):
    out = []
    for m in "KMC":
        fx = mm_fixture(f"wedge{k}_{m}")
        I, J, V = mm_expanded(fx)
        out.append((I, J, V, int(fx["size"][0])))
    b = mm_fixture(f"wedge{k}_b")["vals"][:, 0].copy()
    return out, b


def _dense_of(T, n):
    D = np.zeros((n, n))
    np.add.at(D, (T[0] - 1, T[1] - 1), T[2])
    return D


def _count_ok(g_its, o_its, run, b):
    """Iteration counts equal within 2, or within twice the oracle's own
    rounding spread on this problem."""
    return abs(g_its - o_its) <= 2 or abs(g_its - o_its) <= 2 * rounding_spread(run, b)[1]


def test_wedge_neumann_single_frequency(gpu):
    """wedge_neumann.f90 without -multi, with deflation: (K - sigma M) x = b,
    sigma = (2 pi f)^2, V = M 1 (K's constant null space),
    pp_idrs(K, b, M, P, seed = sigma, V) with P left unset: IDR(4) on
    Pmul(A, b), tol 1e-6, maxit 1000, then scaleback with add_deflation."""
    (K, M, C), b = _wedge(3)
    n = K[3]
    Kd, Md = _dense_of(K, n), _dense_of(M, n)
    V = _orc_om(*M, False).user_mv(np.ones(n)).reshape(n, 1)
    kw = dict(tolerance=1e-6, maximum_iterations=1000)
    for f in (1.0, 2.0):
        sigma = (2 * np.pi * f) ** 2
        pr = Pair(K, False, seed=sigma, M=M)
        pr.A.make_deflation(V)
        pr.O.make_deflation(V)
        rhs = pr.A.pmul(b)
        assert _close(rhs, pr.O.pmul(b), 1e-12)
        print(f"neumann f={f}", flush=True)
        g = idrs_amd.idrs(pr.A, rhs, 4, **kw)
        o = orc.idrs(pr.O, pr.O.pmul(b), 4, **kw)
        print(f"  gpu {g.iterations} flag {g.flag}, oracle {o[3]} flag {o[1]}", flush=True)
        assert g.flag == o[1] == 0
        run = lambda bb: (lambda q: (q[0], q[3]))(orc.idrs(pr.O, pr.O.pmul(bb), 4, **kw))
        assert _count_ok(g.iterations, o[3], run, b)
        x, xo = idrs_amd.scaleback(pr.A, b, g.x), pr.O.scaleback_b(b, o[0])
        res = [np.linalg.norm(b - Kd @ y + sigma * (Md @ y)) / np.linalg.norm(b) for y in (x, xo)]
        assert res[0] <= max(1e-5, 10 * res[1]), res


def test_wedge_sommerfeld_single_frequency(gpu):
    """wedge_sommerfeld.f90 without -multi, with deflation:
    (K + sigma C + sigma^2 M) x = b, sigma = 2 pi f i, V = 1/sqrt(n),
    pp_idrs(K, b, M, C, P, seed = sigma, V) with P left unset."""
    (K, M, C), b = _wedge(3)
    n = K[3]
    Kd, Md, Cd = _dense_of(K, n), _dense_of(M, n), _dense_of(C, n)
    V = np.full((n, 1), 1 / np.sqrt(n), dtype=np.complex128)
    bc = b.astype(np.complex128)
    kw = dict(tolerance=1e-6, maximum_iterations=1000)
    for f in (1.0, 4.0):
        sigma = 2 * np.pi * f * 1j
        pr = Pair(K, True, seed=sigma, M=M, C=C)
        pr.A.make_deflation(V)
        pr.O.make_deflation(V)
        print(f"sommerfeld f={f}", flush=True)
        g = idrs_amd.idrs(pr.A, pr.A.pmul(bc), 4, **kw)
        o = orc.idrs(pr.O, pr.O.pmul(bc), 4, **kw)
        print(f"  gpu {g.iterations} flag {g.flag}, oracle {o[3]} flag {o[1]}", flush=True)
        assert g.flag == o[1] == 0
        run = lambda bb: (lambda q: (q[0], q[3]))(orc.idrs(pr.O, pr.O.pmul(bb), 4, **kw))
        assert _count_ok(g.iterations, o[3], run, bc)
        x, xo = idrs_amd.scaleback(pr.A, bc, g.x), pr.O.scaleback_b(bc, o[0])
        S = Kd + sigma * Cd + sigma ** 2 * Md
        res = [np.linalg.norm(bc - S @ y) / np.linalg.norm(bc) for y in (x, xo)]
        assert res[0] <= max(1e-5, 10 * res[1]), res


@pytest.mark.parametrize("deflation", [False, True])
@pytest.mark.parametrize("solver", ["msidrs", "msqmridr"])
def test_wedge_sommerfeld_multishift_block(gpu, solver, deflation):
    """wedge_sommerfeld.f90 -multi: the block system of 2n rows
    ([-C -K; I 0] - sigma [M 0; 0 I]; right-hand side [-b; 0], solution
    the second half), lumped mass preconditioner complex_precon(1, M), seed
    at the last frequency, ms_parameters' shifts sigma - seed, the
    multishift solver on Pmul(A, b), the per-shift scale-back; with
    -deflation V = [0; 1/sqrt(n)] and the projector I - V V^H."""
    (K, M, C), b = _wedge(2)
    n = K[3]
    Kd, Md, Cd = _dense_of(K, n), _dense_of(M, n), _dense_of(C, n)
    sigma = 2 * np.pi * np.array([1.0, 2.0]) * 1j
    seed = sigma[-1]
    pr = Pair(K, True, seed=seed, M=M, C=C, block=True, multishift=True, pdiag_of=M)
    assert pr.A.rows == pr.O.rows == 2 * n
    if deflation:
        V = np.zeros((2 * n, 1), dtype=np.complex128)
        V[n:, 0] = 1 / np.sqrt(n)
        pr.A.make_deflation(V)
        pr.O.make_deflation(V)
    bb = np.concatenate([-b, np.zeros(n)]).astype(np.complex128)
    assert _close(pr.A.pmul(bb), pr.O.pmul(bb), 1e-12)
    sigma_p = sigma - seed
    kw = dict(tolerance=1e-6, maximum_iterations=1000)
    print(f"{solver} block deflation={deflation}", flush=True)
    g = getattr(idrs_amd, solver)(pr.A, pr.A.pmul(bb), sigma_p, 4, **kw)
    o = getattr(orc, solver)(pr.O, pr.O.pmul(bb), sigma_p, 4, **kw)
    print(f"  gpu {g.iterations} flag {g.flag}, oracle {o['iterations']} flag {o['flag']}", flush=True)
    assert g.flag == o["flag"] == 0
    run = lambda b1: (lambda q: (q["x"], q["iterations"]))(getattr(orc, solver)(pr.O, pr.O.pmul(b1), sigma_p, 4,
                                                                                   **kw))
    assert _count_ok(g.iterations, o["iterations"], run, bb)
    ones = np.ones((1, len(sigma)))
    X = idrs_amd.ms_scaleback_b(pr.A, sigma, ones, bb, np.asarray(g.x))
    Xo = pr.O.ms_add_deflation(sigma, bb, np.asarray(o["x"]))
    Xo = np.stack([pr.O.precon(Xo[:, k]) for k in range(len(sigma))], axis=1)
    for k, sg in enumerate(sigma):
        S = Kd + sg * Cd + sg ** 2 * Md
        res = [np.linalg.norm(b - S @ Y[n:, k]) / np.linalg.norm(b) for Y in (X, Xo)]
        assert res[0] <= max(1e-5, 10 * res[1]), (k, res)


def _room(grid):
    """EXAMPLES/STENCIL/user_module.f90 laplace_matrix / damping_matrix /
    point_source on one image, assembled as COO: 16 2^(grid-1) points per
    direction on a 4 m cube, the 7-point stencil with Sommerfeld walls
    (the ghost value mirrors the second node, v(0) = v(2), so a wall node's
    outer coefficient lands on its inner neighbour -- kept as a separate
    entry, in the stencil's term order c, w, e, s, n, f, b), the damping
    matrix (2/h) rp on wall nodes (rp = 1; the faces assign, not add), and
    the point source 1/(hx hy hz) at node Nn/2."""
    N = 16 * 2 ** (grid - 1)
    h = 4.0 / (N - 1)
    c, o = 2 / h ** 2 * 3, -1 / h ** 2
    idx = lambda i, j, k: i + N * (j + N * k)
    I, J, V = [], [], []
    for k in range(N):
        for j in range(N):
            for i in range(N):
                r = idx(i, j, k)
                terms = [(idx(i, j, k), c)]
                for (a, lim, mk) in ((i, N, lambda t: idx(t, j, k)), (j, N, lambda t: idx(i, t, k)),
                                     (k, N, lambda t: idx(i, j, t))):
                    lo = a - 1 if a > 0 else a + 1          # ghost mirrors the second node
                    hi = a + 1 if a < lim - 1 else a - 1
                    terms += [(mk(lo), o), (mk(hi), o)]
                for col, v in terms:
                    I.append(r + 1)
                    J.append(col + 1)
                    V.append(v)
    n = N ** 3
    K = (np.array(I, np.int32), np.array(J, np.int32), np.array(V), n)
    wall = np.zeros(n, bool)
    g = np.arange(N)
    ii, jj, kk = np.meshgrid(g, g, g, indexing="ij")
    wall[idx(ii, jj, kk).ravel()] = ((ii == 0) | (ii == N - 1) | (jj == 0) | (jj == N - 1) | (kk == 0)
                                     | (kk == N - 1)).ravel()
    w = np.nonzero(wall)[0]
    C = ((w + 1).astype(np.int32), (w + 1).astype(np.int32), np.full(len(w), 2.0 / h), n)
    b = np.zeros(n, complex)
    b[n // 2 - 1] = 1.0 / h ** 3
    return K, C, b


@pytest.mark.parametrize("deflation", [False, True])
def test_room_single_frequency(gpu, deflation):
    """room.f90 without -multi: (K + sigma C + sigma^2 I) x = b at
    sigma = 2 pi 100 i / 340 (grid 1, 16^3), pp_idrs(K, b, C=C, seed=sigma[, V])
    with no preconditioner: IDR(4) on Pmul(A, b), tol 1e-6; with
    -deflation V = 1/sqrt(n)."""
    K, C, b = _room(1)
    n = K[3]
    sigma = 1j * 2 * np.pi * 100.0 / 340.0
    pr = Pair(K, True, seed=sigma, C=C)
    if deflation:
        V = np.full((n, 1), 1 / np.sqrt(n), complex)
        pr.A.make_deflation(V)
        pr.O.make_deflation(V)
    kw = dict(tolerance=1e-6, maximum_iterations=1000)
    g = idrs_amd.idrs(pr.A, pr.A.pmul(b), 4, **kw)
    o = orc.idrs(pr.O, pr.O.pmul(b), 4, **kw)
    print(f"room deflation={deflation}: gpu {g.iterations} oracle {o[3]}", flush=True)
    assert g.flag == o[1] == 0
    run = lambda bb: (lambda q: (q[0], q[3]))(orc.idrs(pr.O, pr.O.pmul(bb), 4, **kw))
    assert _count_ok(g.iterations, o[3], run, b)
    x, xo = idrs_amd.scaleback(pr.A, b, g.x), pr.O.scaleback_b(b, o[0])
    Kd = np.zeros((n, n))
    np.add.at(Kd, (K[0] - 1, K[1] - 1), K[2])
    Cd = np.zeros(n)
    Cd[C[0] - 1] = C[2]
    res = [np.linalg.norm(b - (Kd @ y + sigma * (Cd * y) + sigma ** 2 * y)) / np.linalg.norm(b) for y in (x, xo)]
    assert res[0] <= max(1e-5, 10 * res[1]), res


def test_room_multishift_block(gpu):
    """room.f90 -multi -deflation at grid 2 (32^3, block system of 65536
    rows): [-C -K; I 0] - sigma I (no mass matrix), shifts
    2 pi f i / 340 for f = 100, 200, seed at the last, V = [0; 1/sqrt(n)],
    MSIDRS(4) on Pmul(A, b) with b = [-rhs; 0], scaleback; each frequency's
    true residual of K + sigma C + sigma^2 I."""
    K, C, rhs = _room(2)
    n = K[3]
    sigma = 1j * 2 * np.pi * np.array([100.0, 200.0]) / 340.0
    seed = sigma[-1]
    pr = Pair(K, True, seed=seed, C=C, block=True, multishift=True)
    V = np.zeros((2 * n, 1), complex)
    V[n:, 0] = 1 / np.sqrt(n)
    pr.A.make_deflation(V)
    pr.O.make_deflation(V)
    bb = np.concatenate([-rhs, np.zeros(n)])
    kw = dict(tolerance=1e-6, maximum_iterations=1000)
    g = idrs_amd.msidrs(pr.A, pr.A.pmul(bb), sigma - seed, 4, **kw)
    o = orc.msidrs(pr.O, pr.O.pmul(bb), sigma - seed, 4, **kw)
    print(f"room multishift: gpu {g.iterations} oracle {o['iterations']}", flush=True)
    assert g.flag == o["flag"] == 0
    run = lambda b1: (lambda q: (q["x"], q["iterations"]))(orc.msidrs(pr.O, pr.O.pmul(b1), sigma - seed, 4, **kw))
    assert _count_ok(g.iterations, o["iterations"], run, bb)
    X = idrs_amd.ms_scaleback_b(pr.A, sigma, np.ones((1, 2)), bb, np.asarray(g.x))
    import scipy.sparse as sps
    Ks = sps.csr_matrix((K[2], (K[0] - 1, K[1] - 1)), shape=(n, n))
    Cd = np.zeros(n)
    Cd[C[0] - 1] = C[2]
    for k, s in enumerate(sigma):
        x = X[n:, k]
        r = np.linalg.norm(rhs - (Ks @ x + s * (Cd * x) + s * s * x)) / np.linalg.norm(rhs)
        assert r < 1e-5, (k, r)
