"""Oracle restatement of the reference's spectral routines (numpy, small
dense matrices): QR_factorization, QR_solve, QR_eig and SI_power of
MODULES/IDRS/dense_la_module.f90, filter_ritzval and analyse_spectrum of
MODULES/PP_MATRIX/matrix_module.f90, and compute_U0, global_QR_factorization
and IDRS_RITZ of MODULES/IDRS/ritz_module.f90.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
library's IDRS_RITZ and of its host spectral routines.  Statement order
follows the reference; column operations are numpy vector operations.
"""
import numpy as np

import orc

TINY = np.finfo(np.float64).tiny
EPS = np.finfo(np.float64).eps


def qr_factorization(V):
    """CQR_factorization / RQR_factorization (dense_la_module.f90:42-124):
    classical Gram-Schmidt with one re-orthogonalisation sweep (refine = 2).
    Returns (Q, R, info); info = j (1-based) when column j vanished."""
    V = np.array(V)
    m, n = V.shape
    Q = V.copy()
    R = np.zeros((n, n), dtype=V.dtype)
    for j in range(n):
        for _ in range(2):
            for i in range(j):
                t = np.vdot(Q[:, i], Q[:, j])                   # dot_product(Q(:,i), Q(:,j))
                Q[:, j] = Q[:, j] - t * Q[:, i]
                R[i, j] = R[i, j] + t
        R[j, j] = R[j, j] + np.sqrt(np.sum(np.abs(Q[:, j]) ** 2))   # norm2(abs(Q(:,j)))
        if abs(R[j, j]) > TINY:
            Q[:, j] = Q[:, j] / R[j, j]
        else:
            return Q, R, j + 1
    return Q, R, 0


def qr_solve(Q, R, B):
    """CQR_solve / RQR_solve (dense_la_module.f90:126-190): R^-1 Q^H B."""
    X = np.conj(Q).T @ B
    n = Q.shape[1]
    for j in range(n - 1, -1, -1):
        X[j, :] = X[j, :] / R[j, j]
        for i in range(j):
            X[i, :] = X[i, :] - R[i, j] * X[j, :]
    return X


def _largest_diag(V, m):
    shift = 0j
    for i in range(m):
        if abs(V[i, i]) > abs(shift):
            shift = V[i, i]
    return shift


def qr_eig(H, want_U=False):
    """QR_eig (dense_la_module.f90:192-403) of H(1:n,1:n), n = size(H,2):
    RQR_eig for real H (largest-diagonal shift, times 1+i at the start,
    maxit 100 n), CQR_eig for complex H (shift V(m,m), maxit 10 n).
    Returns (eig, U or None, converged)."""
    H = np.asarray(H)
    real = not np.iscomplexobj(H)
    n = H.shape[1]
    V = np.array(H[:n, :n], dtype=np.complex128)
    maxit = 100 * n if real else 10 * n
    U = np.eye(n, dtype=np.complex128) if want_U else None
    m, it = n, 0
    shift = _largest_diag(V, m) * (1 + 1j) if real else V[n - 1, n - 1]
    while m > 1 and it < maxit:
        for i in range(m):
            V[i, i] = V[i, i] - shift
        Q, R, info = qr_factorization(V[:m, :m])
        if info > 0:
            break
        V[:m, :m] = R @ Q
        for i in range(m):
            V[i, i] = V[i, i] + shift
        if want_U:
            S = np.eye(n, dtype=np.complex128)
            S[:m, :m] = Q
            U = U @ S
        if abs(V[m - 1, m - 2]) < EPS * abs(V[m - 1, m - 1]):
            m = m - 1
        shift = _largest_diag(V, m) if real else V[m - 1, m - 1]
        it = it + 1
    return np.diag(V).copy(), U, n - m + 1


def uniform_block(n, s, seed, offset, cplx=False):
    """Start block of SI_power: the oracle's uniform [0,1) stream (seed) at
    counter offset, column-major; complex: real parts, then imaginary."""
    cnt = n * s
    re = np.array([orc.uniform01(seed, offset + q) for q in range(cnt)])
    if cplx:
        im = np.array([orc.uniform01(seed, offset + cnt + q) for q in range(cnt)])
        re = re + 1j * im
    return re.reshape((n, s), order="F")


def si_power(H, s, shift=0.0, start=None):
    """CSI_power / RSI_power (dense_la_module.f90:455-603): s Schur vectors
    of H(1:n,1:n) closest to shift by the block shift-and-invert power
    method from the start block; stops when the sum of the R(i,i) changes
    by at most 1e-3 of its first value (or after 1000 steps)."""
    H = np.asarray(H)
    n = H.shape[1]
    A = np.array(H[:n, :n])
    for i in range(n):
        A[i, i] = A[i, i] - shift
    Q, R, _ = qr_factorization(A)
    Qs, Rs, _ = qr_factorization(start)
    trace0 = sum(abs(Rs[i, i]) for i in range(s))
    trace_old, err, it = trace0, 1.0, 0
    while err > 1e-3 and it < 1000:
        it = it + 1
        S = qr_solve(Q, R, Qs)
        Qs, Rs, _ = qr_factorization(S)
        trace = sum(abs(Rs[i, i]) for i in range(s))
        err = abs(trace - trace_old) / trace0
        trace_old = trace
    return Qs


def filter_ritzval(ritzval):
    """filter_ritzval (matrix_module.f90:1059-1077): farther than 1.5 sigma
    from the mean."""
    rv = np.asarray(ritzval, dtype=np.complex128)
    n = len(rv)
    center = np.sum(rv) / n
    sigma = np.sqrt(np.sum(np.abs(rv - center) ** 2) / n)
    return np.abs(rv - center) > 1.5 * sigma


def analyse_spectrum(H):
    """ranalyse_spectrum / canalyse_spectrum (matrix_module.f90:1079-1282) of
    H (Ritz values of H(1:n,1:n), n = size(H,2)); returns (foci, center,
    shift), real for real H."""
    H = np.asarray(H)
    real = not np.iscomplexobj(H)
    rv, _, _ = qr_eig(H)
    keep = ~filter_ritzval(rv)
    left, right = rv.real[keep].min(), rv.real[keep].max()
    bottom, top = rv.imag[keep].min(), rv.imag[keep].max()
    center_box = (right + left) / 2.0 if real else complex(right + left, top + bottom) / 2.0
    radius = np.abs(rv - center_box)[keep].max()
    center_circle = center_box
    if abs(center_box) / radius < 1:
        correction = abs(center_box) / radius
        center_circle = center_box / correction
    shift_circle = center_box - center_circle
    x0, y0 = (right - left) / 2.0, (top - bottom) / 2.0
    a, b = np.sqrt(x0 * (x0 + y0)), np.sqrt(y0 * (x0 + y0))
    center_ellipse = center_box
    if real:
        with np.errstate(divide="ignore"):
            inside = abs(center_box) / a < 1.0
        if inside:
            correction = abs(center_box) / a
            center_ellipse = center_box / correction
        if a > b:
            c = np.sqrt(a ** 2 - b ** 2)
        else:
            c = 0.15 * center_ellipse
        foci = np.array([center_ellipse - c, center_ellipse + c])
    else:
        with np.errstate(divide="ignore", invalid="ignore"):
            q = (center_box.real / a) ** 2 + (center_box.imag / b) ** 2
        if q < 1.0:
            center_ellipse = center_box / np.sqrt(q)
        if a > b:
            c = np.sqrt(a ** 2 - b ** 2)
            foci = np.array([center_ellipse - c, center_ellipse + c])
        else:
            c = np.sqrt(b ** 2 - a ** 2)
            foci = np.array([center_ellipse - 1j * c, center_ellipse + 1j * c])
    return foci, center_circle, shift_circle


def ritz_omegas(ritzval, real):
    """The omegas of IDRS_RITZ (ritz_module.f90:404-425 real, 266-286
    complex): 1/Ritz value (real part in the real version; 1 below
    epsilon), reordered greedily by the residual polynomial built so far."""
    rv = np.asarray(ritzval, dtype=np.complex128)
    n = len(rv)
    om = rv.real.copy() if real else rv.copy()
    om[np.abs(om) < EPS] = 1.0
    om = 1.0 / om
    p = np.ones(n, dtype=om.dtype)
    for i in range(n):
        mx = i + int(np.argmax(np.abs(p[i:])))
        m = om[mx]
        om[mx] = om[i]
        om[i] = m
        p = p * (1.0 - (rv.real if real else rv) * om[i])
    return om


def compute_u0(matvec, H, eigvec, b, precon=None):
    """rcompute_U0 / ccompute_U0 (ritz_module.f90:100-223): the Krylov basis
    of b regenerated from H(1:nritz+1, 1:nritz), combined with eigvec
    (nritz x d); columns normalised."""
    H = np.asarray(H)
    nritz = H.shape[1]
    d = eigvec.shape[1]
    dt = np.result_type(H.dtype, eigvec.dtype, np.asarray(b).dtype)
    neq = len(b)
    U0 = np.zeros((neq, d), dtype=dt)
    W = np.zeros((neq, d + 1), dtype=dt)
    v = np.array(b, dtype=dt)
    ind = np.zeros(nritz, dtype=int)
    k = 0
    for i in range(1, nritz + 1):
        k = k + 1
        if k > d + 1:
            k = 1
        ind[i - 1] = k
        W[:, k - 1] = v
        for j in range(d):
            U0[:, j] = U0[:, j] + eigvec[i - 1, j] * v
        low = max(1, i - d)
        av = matvec(precon(v) if precon else v)
        acc = W[:, ind[low - 1] - 1] * H[low - 1, i - 1]
        for j in range(low + 1, i + 1):
            acc = acc + W[:, ind[j - 1] - 1] * H[j - 1, i - 1]
        v = (av - acc) / H[i, i - 1]
    for j in range(d):
        U0[:, j] = U0[:, j] / np.sqrt(np.sum(np.abs(U0[:, j]) ** 2))
    return U0


def global_qr(V):
    """global_CQR / global_RQR_factorization (ritz_module.f90:36-98):
    Gram-Schmidt with the updated column (modified Gram-Schmidt)."""
    Q = np.array(V)
    n = Q.shape[1]
    R = np.zeros((n, n), dtype=Q.dtype)
    for j in range(n):
        for i in range(j):
            R[i, j] = np.vdot(Q[:, i], Q[:, j])
            Q[:, j] = Q[:, j] - R[i, j] * Q[:, i]
        R[j, j] = np.sqrt(np.sum(np.abs(Q[:, j]) ** 2))
        if not abs(R[j, j]) > TINY:
            raise ValueError(" V has dependent columns ")
        Q[:, j] = Q[:, j] / R[j, j]
    return Q, R


def idrs_ritz(A, b, s, M=None, tolerance=None, maximum_iterations=None, variant=None, x0=None, PT_raw=None,
              n_omega=0, dim_u0=0, H_shape=None, seed=1):
    """RIDRS_RITZ / CIDRS_RITZ (ritz_module.f90:225-556) over the oracle's
    IDRS.  The shift-and-invert start block comes from the uniform stream
    (seed) after the shadow space's s n values.  Returns a dict with x,
    flag, relres, iterations, nritz, omega, U0, H, converged."""
    n = len(b)
    cplx = A.complex
    maxit = maximum_iterations if maximum_iterations is not None else min(2 * n, 1000)
    x, flag, relres, it, _, Hf = orc.idrs(A, b, s, M1=M, tolerance=tolerance, maximum_iterations=maxit,
                                          variant=variant, x0=x0, PT_raw=PT_raw, H_nritz=maxit)
    nritz = (it // (s + 1)) * s + it % (s + 1)
    out = dict(x=x, flag=flag, relres=relres, iterations=it, nritz=nritz, omega=None, U0=None, H=None,
               converged=None)
    if n_omega > 0:
        if nritz < n_omega:
            raise ValueError("Too many parameters omega requested, reduce omega.")
        rv, _, conv = qr_eig(Hf[:n_omega, :n_omega])
        out["omega"] = ritz_omegas(rv, real=not cplx)
        out["converged"] = conv if not cplx else n_omega
    if dim_u0 > 0:
        if nritz < dim_u0:
            raise ValueError("Recycle space is too large, reduce dimension of U0.")
        if cplx:
            _, V, _ = qr_eig(Hf[:nritz, :nritz], want_U=True)
            vec = V[:, :dim_u0]
        else:
            prow = s if (variant in (None, 1)) else 1
            vec = si_power(Hf[:nritz, :nritz], dim_u0, 0.0, uniform_block(nritz, dim_u0, seed, prow * n))
        U = compute_u0(A.matvec, Hf[:nritz + 1, :nritz], vec, b, precon=M.precon if M is not None else None)
        if M is not None:
            for i in range(dim_u0):
                U[:, i] = M.precon(U[:, i])
        out["U0"], _ = global_qr(U)
    if H_shape is not None:
        r, c = H_shape
        if r * c > Hf.size or r > maxit + 1 or c > maxit:
            raise ValueError("Dimensions of Hessenberg matrix too large")
        out["H"] = Hf[:r, :c].copy()
    return out
