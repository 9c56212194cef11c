"""Pins of the oracle (CPU).  The reference publishes no outputs and its
Fortran cannot be built here, so each oracle function is checked against
what the mathematics fixes: closed forms, brute force on small inputs,
library routines (numpy/scipy), the reference examples' known answers and
invariants of the IDR(s) method.  Each check is chosen so that a dropped
term, a wrong sign or index, or a transposed operand fails it.
"""
import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

import orc
from idrs_amd import problems as P

EPS = np.finfo(np.float64).eps


def dense(crs):
    return crs.to_scipy().toarray()


# ---------------------------------------------------------------- SpMV ---

def _rand_crs(n, seed, cplx=False):
    rng = np.random.default_rng(seed)
    lens = np.minimum(rng.integers(0, 9, size=n), n)
    cols = [np.sort(rng.choice(n, size=l, replace=False)) for l in lens]
    rp = np.concatenate([[1], 1 + np.cumsum(lens)]).astype(np.int32)
    ci = (np.concatenate(cols) + 1).astype(np.int32)
    v = rng.standard_normal(len(ci))
    if cplx:
        v = v + 1j * rng.standard_normal(len(ci))
    return P.Crs(n, rp, ci, v)


@pytest.mark.parametrize("cplx", [False, True])
def test_crs_spmv_brute_force(cplx):
    crs = _rand_crs(300, 1, cplx)
    x = np.random.default_rng(2).standard_normal(300) + (1j if cplx else 0) * np.random.default_rng(3).standard_normal(300)
    w = orc.OracleMatrix(orc.CRS, 300, crs=crs, complex_vectors=cplx).user_mv(x)
    ref = dense(crs) @ x
    assert np.allclose(w, ref, rtol=1e-13, atol=1e-13)
    # summation order is the stored order: equal to an explicit sequential loop
    seq = np.zeros(300, dtype=w.dtype)
    for i in range(300):
        acc = 0.0 * seq[0]
        for j in range(crs.row_ptr[i] - 1, crs.row_ptr[i + 1] - 1):
            acc = acc + crs.val[j] * x[crs.col_ind[j] - 1]
        seq[i] = acc
    assert np.array_equal(w, seq)


def test_coo_scatter_order_and_duplicates():
    rng = np.random.default_rng(4)
    n, nnz = 200, 900
    indx = rng.integers(1, n + 1, nnz).astype(np.int32)
    jndx = rng.integers(1, n + 1, nnz).astype(np.int32)
    indx[:50], jndx[:50] = indx[50:100], jndx[50:100]
    val = rng.standard_normal(nnz)
    x = rng.standard_normal(n)
    w = orc.OracleMatrix(orc.COO, n, coo=(indx, jndx, val)).user_mv(x)
    seq = np.zeros(n)
    for k in range(nnz):
        seq[indx[k] - 1] = seq[indx[k] - 1] + val[k] * x[jndx[k] - 1]
    assert np.array_equal(w, seq)
    D = np.zeros((n, n))
    np.add.at(D, (indx - 1, jndx - 1), val)
    assert np.allclose(w, D @ x, rtol=1e-12, atol=1e-12)


def test_cds_band_layout():
    """rband(i,dia) multiplies v(i+dia) (user_module.f90:727-733)."""
    rng = np.random.default_rng(5)
    n, hb = 50, 4
    band = rng.standard_normal((n, 2 * hb + 1))
    D = np.zeros((n, n))
    for i in range(n):
        for d in range(-hb, hb + 1):
            if 0 <= i + d < n:
                D[i, i + d] = band[i, d + hb]
    x = rng.standard_normal(n)
    assert np.allclose(orc.OracleMatrix(orc.CDS, n, cds=(hb, band)).user_mv(x), D @ x, rtol=1e-13, atol=1e-13)


def test_dia_equals_crs_for_stencil():
    crs, dia = P.convection_diffusion_3d(6, 7.0, reaction=0.5)
    x = np.random.default_rng(6).standard_normal(crs.n)
    wd = orc.OracleMatrix(orc.DIA, crs.n, dia=(dia.offsets, dia.vals)).user_mv(x)
    assert np.allclose(wd, dense(crs) @ x, rtol=1e-13, atol=1e-11)


def test_stencil_coefficients_closed_form():
    """32^3, Pe=10: centre 6/h^2 = 6534, east -1/h^2 + Pe/(2h) = -924,
    west -1089 - 165 = -1254 (h = 1/33)."""
    crs, _ = P.convection_diffusion_3d(32, 10.0)
    A = crs.to_scipy()
    assert crs.nnz == 223232
    i = 32 * 32 * 16 + 32 * 16 + 16
    assert A[i, i] == pytest.approx(6534.0)
    assert A[i, i + 1] == pytest.approx(-924.0)
    assert A[i, i - 1] == pytest.approx(-1254.0)
    assert A[i, i + 1024] == pytest.approx(-924.0) and A[i, i - 32] == pytest.approx(-1254.0)


def test_diagonal_and_zero_rule():
    crs = _rand_crs(100, 7)
    d = orc.OracleMatrix(orc.CRS, 100, crs=crs).diagonal()
    ref = np.diag(dense(crs)).copy()
    ref[ref == 0] = 1.0
    assert np.array_equal(d, ref)


# ------------------------------------------------ polynomial coefficients ---

def _p_of_t(alpha, t):
    return np.polynomial.polynomial.polyval(t, alpha)


@pytest.mark.parametrize("degree", [1, 2, 5, 8, 12])
def test_chebyshev_closed_form(degree):
    """1 - t p(t) = T_{d+1}((theta - t)/hw) / T_{d+1}(theta/hw), theta the
    centre and hw the half width of [f1, f2]."""
    f1, f2 = 0.15, 1.85
    a = orc.chebyshev(degree, np.array([f1, f2]))
    th, hw = (f1 + f2) / 2, (f2 - f1) / 2
    e = np.zeros(degree + 2)
    e[-1] = 1
    t = np.linspace(-0.5, 2.5, 41)
    lhs = 1 - t * _p_of_t(a, t)
    rhs = C.chebval((th - t) / hw, e) / C.chebval(th / hw, e)
    # monomial-basis evaluation loses digits as the degree grows
    assert np.allclose(lhs, rhs, rtol=1e-10, atol=1e-11 * np.abs(rhs).max())


def test_chebyshev_complex_foci():
    f = np.array([0.3 + 0.2j, 1.7 - 0.2j])
    d = 6
    a = orc.chebyshev(d, f)
    th, dl = (f[0] + f[1]) / 2, (f[0] - f[1]) / 2
    t = np.array([0.1, 0.5 + 0.1j, 1.0, 1.9j])

    def T(k, z):
        t0, t1 = np.ones_like(z), z
        for _ in range(k - 1):
            t0, t1 = t1, 2 * z * t1 - t0
        return t1 if k else t0

    lhs = 1 - t * _p_of_t(a, t)
    sig = th / dl
    rhs = T(d + 1, sig - t / dl) / T(d + 1, sig)
    assert np.allclose(lhs, rhs, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("degree,center", [(1, 1.0), (4, 2.0), (7, 0.5)])
def test_neumann_closed_form(degree, center):
    a = orc.neumann(degree, center)
    t = np.linspace(-1, 3, 17)
    assert np.allclose(1 - t * _p_of_t(a, t), (1 - t / center) ** (degree + 1), rtol=1e-10, atol=1e-10)


def test_operator_definition():
    """A*v = K D^-1 p(K D^-1) v (matrix_module.f90:135-304), dense check."""
    crs, _ = P.convection_diffusion_3d(5, 30.0, reaction=1.0)
    a = orc.chebyshev(8, np.array([0.15, 1.85]))
    Ao = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, which_precon=1, degree=8, alpha=a)
    K = dense(crs)
    Kd = K / np.diag(K)[None, :]
    v = np.random.default_rng(0).standard_normal(crs.n)
    pv = sum(a[i] * np.linalg.matrix_power(Kd, i) @ v for i in range(9))
    assert np.allclose(Ao.matvec(v), Kd @ pv, rtol=1e-10, atol=1e-10 * np.abs(Kd @ pv).max())
    # scaleback: x = D^-1 p(Kd) y
    assert np.allclose(Ao.scaleback(v), pv / np.diag(K), rtol=1e-10, atol=1e-14)


# -------------------------------------------------------------- MAKE_P ---

def _splitmix64_first(seed_state):
    m = (1 << 64) - 1
    z = (seed_state + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def test_uniform_stream_is_splitmix64():
    assert _splitmix64_first(0) == 0xE220A8397B1DCDAF  # published first output from state 0
    for seed, idx in [(0, 0), (1, 0), (1, 12345), (7, 2**40)]:
        state = (idx + seed * 0xD1B54A32D192ED03) & ((1 << 64) - 1)
        assert orc.uniform01(seed, idx) == (_splitmix64_first(state) >> 11) * 2.0 ** -53


@pytest.mark.parametrize("cplx", [False, True])
def test_make_p_is_gram_schmidt(cplx):
    """Classical Gram-Schmidt of the raw rows = Q of a QR factorisation with a
    positive diagonal of R."""
    n, s = 500, 5
    raw = P.gaussian_shadow(n, s, seed=3, complex=cplx)
    PT = orc.make_p(n, s, 1, PT_raw=raw, complex=cplx)
    Q, R = np.linalg.qr(raw)
    Q = Q * (np.sign(np.diag(R).real) if not cplx else np.exp(-1j * np.angle(np.diag(R))))[None, :] ** (1 if not cplx else -1)
    assert np.allclose(PT, Q, atol=1e-12)
    assert np.allclose(PT.conj().T @ PT, np.eye(s), atol=1e-13)


# ------------------------------------------------------------ IDR(s) ---

def _toeplitz():
    crs, _, b = P.toeplitz_1d()
    return crs, orc.OracleMatrix(orc.CRS, crs.n, crs=crs), b


@pytest.mark.parametrize("s", [1, 2, 4, 8])
def test_idrs_finite_termination(s):
    """IDR(s) terminates within N+N/s iterations (toeplitz.f90:111-165);
    the solution is ones."""
    crs, A, b = _toeplitz()
    x, flag, relres, it, _ = orc.idrs(A, b, s, M1=orc.OracleMatrix(orc.CRS, 60, crs=crs), tolerance=1e-8,
                                      maximum_iterations=4000)
    assert flag == 0 and relres < 1e-8
    assert it <= 60 + 60 // s
    assert np.allclose(x, 1.0, atol=1e-6)


def _bicgstab(A, b, iters):
    """Textbook BiCGSTAB (van der Vorst 1992), shadow vector r0, x0 = 0;
    returns ||r|| after each full iteration."""
    x = np.zeros_like(b)
    r = b.copy()
    rt = r.copy()
    rho = alpha = om = 1.0
    v = p = np.zeros_like(b)
    out = []
    for _ in range(iters):
        rho1 = rt @ r
        beta = (rho1 / rho) * (alpha / om)
        p = r + beta * (p - om * v)
        v = A @ p
        alpha = rho1 / (rt @ v)
        sv = r - alpha * v
        t = A @ sv
        om = (t @ sv) / (t @ t)
        x = x + alpha * p + om * sv
        r = sv - om * t
        rho = rho1
        out.append(np.linalg.norm(r))
    return np.array(out)


def test_variant2_is_bicgstab():
    """variant 2 (s=1, kappa=0, P=r0): the residual after each omega step
    equals BiCGSTAB's residual (one BiCGSTAB step = two IDR steps)."""
    crs, _ = P.convection_diffusion_3d(6, 20.0, reaction=1.0)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs)
    b = np.random.default_rng(1).standard_normal(crs.n)
    x, flag, relres, it, rv = orc.idrs(A, b, 1, variant=2, tolerance=1e-10, maximum_iterations=400, resvec=True)
    assert flag == 0
    ref = _bicgstab(dense(crs), b, it // 2)
    k = min(len(ref), 12)
    assert np.allclose(rv[2:2 * k + 1:2], ref[:k], rtol=1e-7)


@pytest.mark.parametrize("s", [2, 4])
def test_residual_orthogonal_to_shadow_space(s):
    """After inner step k of a cycle the residual is orthogonal to p_1..p_k
    (idrs.inc:221-229)."""
    crs, _ = P.convection_diffusion_3d(7, 5.0, reaction=2.0)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs)
    K = dense(crs)
    b = np.random.default_rng(2).standard_normal(crs.n)
    raw = P.gaussian_shadow(crs.n, s, seed=5)
    PT = orc.make_p(crs.n, s, 1, PT_raw=raw)
    for cycle in (1, 2):
        for k in range(1, s + 1):
            it = cycle * (s + 1) + k
            x, flag, _, iters, _ = orc.idrs(A, b, s, tolerance=0.0, maximum_iterations=it, PT_raw=raw)
            assert iters == it and flag == 1
            r = b - K @ x
            proj = PT[:, :k].T @ r / np.linalg.norm(r)
            assert np.abs(proj).max() < 1e-9, (cycle, k, proj)
            # the remaining directions are not orthogonal (the test has teeth)
            if k < s:
                assert np.abs(PT[:, k:].T @ r / np.linalg.norm(r)).max() > 1e-6


def test_manufactured_solution_and_true_residual():
    crs, _ = P.convection_diffusion_3d(12, 10.0)
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs)
    b = crs.matvec(np.ones(crs.n))
    x, flag, relres, it, rv = orc.idrs(A, b, 4, tolerance=1e-8, resvec=True,
                                       PT_raw=P.gaussian_shadow(crs.n, 4, seed=0))
    assert flag == 0
    true = np.linalg.norm(b - crs.matvec(x)) / np.linalg.norm(b)
    assert true <= 1e-8 and abs(true - relres) < 1e-12
    assert rv[0] == pytest.approx(np.linalg.norm(b), rel=1e-14)
    assert np.allclose(x, 1.0, atol=1e-5)


def test_flags_and_errors():
    crs, A, b = _toeplitz()
    assert orc.idrs(A, b, 2, maximum_iterations=1)[1] == 1
    x, flag, relres, it, _ = orc.idrs(A, b, 2, x0=np.ones(60))
    assert flag == 0 and it == 0
    assert orc.idrs(A, b, 2, omega=[0.0])[1] == 3
    with pytest.raises(ValueError):
        orc.idrs(A, b, 2, tolerance=-1.0)


def test_complex_toeplitz_accuracy():
    """Complex Toeplitz (toeplitz.f90:41-45) reaches ||r||/||b|| < 1e-12."""
    n = 200
    off = np.array([-1, 0, 2, 3], dtype=np.int32)
    cv = np.array([3.6j, 4.0, 1.0, 0.7])
    vals = np.tile(cv, (n, 1)).astype(np.complex128)
    i = np.arange(n)
    D = np.zeros((n, n), dtype=np.complex128)
    for d, o in enumerate(off):
        vals[(i + o < 0) | (i + o >= n), d] = 0
        D += np.diag(np.full(n - abs(o), cv[d]), o)
    A = orc.OracleMatrix(orc.DIA, n, dia=(off, vals), complex_vectors=True)
    b = np.full(n, 1j)
    for s in (1, 2, 4, 8):
        x, flag, relres, it, _ = orc.idrs(A, b, s, tolerance=1e-12, maximum_iterations=4000)
        assert flag == 0
        assert np.linalg.norm(b - D @ x) / np.linalg.norm(b) <= 1e-12


# ------------------------------------------------------- COMP_OM kappa ---

def _omega_step_ratios(crs, b, s, cycles, cplx=False):
    """(rho, ||r_new||/||r||) of the first omega steps of IDR(s), measured
    from outside: x before the omega step from a run stopped there, r = b -
    K x, t = K r (no preconditioner), rho = |(t,r)|/(||t|| ||r||); the ratio
    from the recursive residual norms of a run one step longer."""
    A = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, complex_vectors=cplx)
    K = crs.to_scipy()
    out = []
    for c in range(cycles):
        itb = (s + 1) * c + s
        xb, _, _, it, _ = orc.idrs(A, b, s, tolerance=0.0, maximum_iterations=itb)
        assert it == itb
        r = b - K @ xb
        t = K @ r
        rho = abs(np.vdot(t, r)) / (np.linalg.norm(t) * np.linalg.norm(r))
        rv = orc.idrs(A, b, s, tolerance=0.0, maximum_iterations=itb + 1, resvec=True)[4]
        out.append((rho, rv[itb + 1] / rv[itb]))
    return out


@pytest.mark.parametrize("cplx", [False, True])
def test_comp_om_kappa_rule(cplx):
    """The omega step's residual ratio is fixed by rho and kappa = 0.7
    (COMP_OM, idrs_module.f90:740-748): with om = (t,r)/(t,t), boosted by
    kappa/rho when rho < kappa, ||r - om t||/||r|| = sqrt(1 - 2 kappa rho +
    kappa^2) if rho < kappa, else sqrt(1 - rho^2).  The problems give omega
    steps with rho below 0.5, in (0.5, 0.7) and above 0.7, so a dropped boost,
    a different threshold or a different boost factor fails."""
    kappa = float(np.float32(0.7))
    rng = np.random.default_rng(0)
    cases = [P.convection_diffusion_3d(6, 20.0, reaction=1.0)[0],   # rho in 0.63..0.79
             P.convection_diffusion_3d(6, 200.0)[0],                # rho ~ 0.1
             P.random_dominant(300, 6, seed=1, factor=1.05)]        # rho > 0.7
    seen = []
    for crs in cases:
        b = rng.standard_normal(crs.n)
        if cplx:
            b = b + 1j * rng.standard_normal(crs.n)
        for rho, ratio in _omega_step_ratios(crs, b, 2, 6, cplx):
            pred = np.sqrt(1 - 2 * kappa * rho + kappa ** 2) if rho < kappa else np.sqrt(1 - rho ** 2)
            assert ratio == pytest.approx(pred, rel=1e-9), (rho, ratio, pred)
            seen.append(rho)
    seen = np.array(seen)
    assert (seen < 0.5).any() and ((seen > 0.5) & (seen < kappa)).any() and (seen > kappa).any()


# ------------------------------------------------------------ seed term ---

@pytest.mark.parametrize("seed,degree", [(0.3, 0), (-1.7, 3), (0.4 - 0.9j, 0), (1.1 + 0.5j, 4)])
def test_seeded_operator_closed_form(seed, degree):
    """real_mv / complex_mv with a seed (matrix_module.f90:241-262):
    w = K t - seed t with t = v/D, so A = (K - seed I) D^-1 p((K - seed I) D^-1)
    against a dense evaluation; with a sign error in the seed term the
    products differ by 2 seed D^-1 v."""
    crs, _ = P.convection_diffusion_3d(4, 15.0, reaction=0.5)
    n = crs.n
    cplx = np.iscomplexobj(seed)
    K = dense(crs)
    d = np.diag(K).copy()
    alpha = orc.chebyshev(degree, np.array([0.2, 1.8], dtype=np.complex128 if cplx else np.float64)) if degree else None
    Ao = orc.OracleMatrix(orc.CRS, n, crs=crs, which_precon=1, degree=degree, alpha=alpha, seed=seed,
                          complex_vectors=cplx)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    S = (K - seed * np.eye(n)) / d[None, :]
    pv = v if not degree else sum(alpha[i] * np.linalg.matrix_power(S, i) @ v for i in range(degree + 1))
    ref = S @ pv
    w = Ao.matvec(v)
    assert np.allclose(w, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    wrong = ((K + seed * np.eye(n)) / d[None, :]) @ pv
    assert not np.allclose(w, wrong, rtol=1e-6, atol=1e-6 * np.abs(ref).max())


# -------------------------------------------------------- compute_alpha ---

def hessenberg_zero_colsum(d, seed, cplx=False):
    """(d+2) x (d+1) upper Hessenberg matrix whose columns sum to zero, the
    form of the residual-polynomial recurrence t r_{i-1} = sum_j H(j,i) r_{j-1}
    with r_i(0) = 1 that compute_alpha reads."""
    rng = np.random.default_rng(seed)
    H = np.triu(rng.standard_normal((d + 2, d + 1)) + 2 * np.eye(d + 2, d + 1), -1)
    if cplx:
        H = H + 1j * np.triu(rng.standard_normal((d + 2, d + 1)), -1)
    H[d + 1, d] = 0
    for j in range(d + 1):
        H[j + 1, j] = -H[: j + 1, j].sum()
    return H


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("d", [1, 3, 6])
def test_compute_alpha_roots_are_hessenberg_eigenvalues(d, cplx):
    """1 - t p(t) = r_{d+1}(t) is the characteristic polynomial of the square
    part H(1:d+1, 1:d+1), scaled to r(0) = 1: its roots are that block's
    eigenvalues (matrix_module.f90:983-1057)."""
    H = hessenberg_zero_colsum(d, 10 + d, cplx)
    a = orc.compute_alpha(d, H)
    q = np.concatenate([[1.0], -a])           # 1 - t p(t), ascending powers
    roots = np.sort_complex(np.polynomial.polynomial.polyroots(q))
    ev = np.sort_complex(np.linalg.eigvals(H[: d + 1, : d + 1]))
    assert np.allclose(roots, ev, rtol=1e-8, atol=1e-9)
    # and the monic form: the polynomial equals det(tI - H)/det(-H)
    t = np.array([0.3, -1.1, 2.5])
    lhs = np.polynomial.polynomial.polyval(t, q)
    rhs = [np.linalg.det(tt * np.eye(d + 1) - H[: d + 1, : d + 1]) / np.linalg.det(-H[: d + 1, : d + 1]) for tt in t]
    assert np.allclose(lhs, rhs, rtol=1e-9, atol=1e-10)
