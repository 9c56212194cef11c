"""Python mirror of the reference IDRS package interface for the hot path.

Names, argument meaning and error behaviour follow the Fortran modules
(paths relative to the reference's IDRS/ directory):

  user_module   crs_format / coo_format / cds_format, real_diagonal,
                complex_diagonal, operator(*)       MODULES/PP_MATRIX/user_module.f90
  matrix_module type(matrix), real_preconditioner, complex_preconditioner,
                operator(*) / operator(/), scaleback  MODULES/PP_MATRIX/matrix_module.f90
  idrs_module   IDRS (RIDRS / CIDRS)                 MODULES/IDRS/idrs_module.f90

Every computation runs in libidrs_hip.so on the GPU; there is no CPU path.
Index arrays are 1-based like the reference's.  Vectors may be numpy arrays
(host) or torch CUDA tensors (device resident; no host round trip).
"""
from collections import namedtuple
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import IdrsError, check, idrs_krylov_opts_t, idrs_opts_t, idrs_stats_t

# H: the Hessenberg matrix H(nritz+1, nritz) when asked for, else None
IdrsResult = namedtuple("IdrsResult", "x flag relres iterations resvec solve_ms spmv_ms spmv_launches H",
                        defaults=(None,))
# IDRS_RITZ: nritz Ritz values available; omega, U0, H when asked for;
# converged: Ritz values the QR iteration converged for the omegas
RitzResult = namedtuple("RitzResult", "x flag relres iterations resvec nritz omega U0 H converged solve_ms")
# relres: per-shift relative residuals (array) for the multishift solvers;
# colfac: MSIDRS collinearity factors when asked for, else None
KrylovResult = namedtuple("KrylovResult", "x flag relres iterations resvec solve_ms colfac")
# per-system arrays flag, relres, iterations; solve_ms of the whole batch
BatchResult = namedtuple("BatchResult", "x flag relres iterations solve_ms")
# iterations: inner (single-precision) iterations; outer: FP64 residual replacements
MixedResult = namedtuple("MixedResult", "x flag relres iterations outer solve_ms")
BatchMixedResult = namedtuple("BatchMixedResult", "x flag relres iterations outer solve_ms")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _is_torch(a):
    return type(a).__module__.startswith("torch")


class UserMatrix:
    """type(user_matrix) (user_module.f90:134-167), stored on the device."""

    __array_ufunc__ = None  # numpy defers v*K, v/K to this class

    def __init__(self, handle, is_complex):
        self._h = ctypes.c_void_p(handle)
        self.is_complex = is_complex
        n_r, n_c, nnz, fmt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64(), ctypes.c_int()
        check(_lib.load().idrs_user_matrix_info(self._h, ctypes.byref(n_r), ctypes.byref(n_c), ctypes.byref(nnz), ctypes.byref(fmt)))
        self.nrows, self.ncols, self.nnz, self.matrix_format = n_r.value, n_c.value, nnz.value, fmt.value

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().idrs_user_matrix_destroy(h)
            except Exception:
                pass
            self._h = None

    def autotune(self, reps=10):
        """Pick the CRS kernel's lanes per row by timing (idrs_user_matrix_autotune);
        returns the chosen group size (1: the sliced copy, one lane per row)."""
        g = ctypes.c_int()
        check(_lib.load().idrs_user_matrix_autotune(self._h, int(reps), ctypes.byref(g)))
        return g.value

    def set_group(self, group):
        """Lanes per row of the CRS products (idrs_user_matrix_set_group);
        1 builds the sliced copy."""
        check(_lib.load().idrs_user_matrix_set_group(self._h, int(group)))

    def layout(self):
        """(lanes per row (1: sliced copy, 0: DIA), row slabs) of the products
        (idrs_user_matrix_layout)."""
        g, sl = ctypes.c_int(), ctypes.c_int()
        check(_lib.load().idrs_user_matrix_layout(self._h, ctypes.byref(g), ctypes.byref(sl)))
        return g.value, sl.value

    def set_row_slabs(self, row_slabs):
        """Row slabs of one-thread-per-row products: 4096, 8192, 16384 or 32768."""
        check(_lib.load().idrs_user_matrix_set_row_slabs(self._h, int(row_slabs)))

    def __mul__(self, v):
        """w = K*v (ruser_mv / cuser_mv, user_module.f90:197-241)."""
        v = np.asarray(v)
        cplx = self.is_complex or np.iscomplexobj(v)
        dt = np.complex128 if cplx else np.float64
        v = np.ascontiguousarray(v, dtype=dt)
        w = np.empty(self.nrows, dtype=dt)
        f = _lib.load().idrs_user_mv_c128 if cplx else _lib.load().idrs_user_mv_f64
        check(f(self._h, _ptr(v), _ptr(w)))
        return w


def _new_um(fn, *args, is_complex=False):
    h = ctypes.c_void_p()
    check(fn(*args, ctypes.byref(h)))
    return UserMatrix(h.value, is_complex)


def crs_format(nrows, ncols, row_ptr, col_ind, rval=None, cval=None):
    """A = crs_format(nrows, ncols, row_ptr, col_ind, rval|cval); 1-based
    (user_module.f90:829-863)."""
    lib = _lib.load()
    rp, ci = _i32(row_ptr), _i32(col_ind)
    if cval is not None:
        cv = np.ascontiguousarray(cval, dtype=np.complex128)
        return _new_um(lib.idrs_crs_format_c128, int(nrows), int(ncols), _ptr(rp), _ptr(ci), _ptr(cv), is_complex=True)
    rv = None if rval is None else np.ascontiguousarray(rval, dtype=np.float64)
    return _new_um(lib.idrs_crs_format_f64, int(nrows), int(ncols), _ptr(rp), _ptr(ci), _ptr(rv))


def coo_format(nrows, ncols, nnz, indx, jndx, rval=None, cval=None):
    """A = coo_format(nrows, ncols, nnz, indx, jndx, rval|cval); 1-based
    (user_module.f90:865-898)."""
    lib = _lib.load()
    ix, jx = _i32(indx), _i32(jndx)
    if cval is not None:
        cv = np.ascontiguousarray(cval, dtype=np.complex128)
        return _new_um(lib.idrs_coo_format_c128, int(nrows), int(ncols), int(nnz), _ptr(ix), _ptr(jx), _ptr(cv), is_complex=True)
    rv = None if rval is None else np.ascontiguousarray(rval, dtype=np.float64)
    return _new_um(lib.idrs_coo_format_f64, int(nrows), int(ncols), int(nnz), _ptr(ix), _ptr(jx), _ptr(rv))


def cds_format(nrows, ncols, half_band, rband=None, cband=None):
    """A = cds_format(nrows, ncols, half_band, rband|cband); band(nrows,
    -hb:hb) given as a numpy array of shape (nrows, 2*hb+1)
    (user_module.f90:932-958)."""
    lib = _lib.load()
    if cband is not None:
        b = np.asfortranarray(cband, dtype=np.complex128)
        return _new_um(lib.idrs_cds_format_c128, int(nrows), int(ncols), int(half_band), _ptr(b), is_complex=True)
    b = None if rband is None else np.asfortranarray(rband, dtype=np.float64)
    return _new_um(lib.idrs_cds_format_f64, int(nrows), int(ncols), int(half_band), _ptr(b))


def dia_format(nrows, ncols, offsets, rvals=None, cvals=None):
    """Offset-list diagonal storage, values of shape (nrows, ndiag) for
    strictly ascending offsets (extension of cds_format for wide bands)."""
    lib = _lib.load()
    off = _i32(offsets)
    if cvals is not None:
        v = np.asfortranarray(cvals, dtype=np.complex128)
        return _new_um(lib.idrs_dia_format_c128, int(nrows), int(ncols), len(off), _ptr(off), _ptr(v), is_complex=True)
    v = None if rvals is None else np.asfortranarray(rvals, dtype=np.float64)
    return _new_um(lib.idrs_dia_format_f64, int(nrows), int(ncols), len(off), _ptr(off), _ptr(v))


MM_REP = {1: "coordinate", 2: "array"}
MM_FIELD = {1: "real", 2: "complex", 3: "integer", 4: "pattern"}
MM_SYMM = {1: "general", 2: "symmetric", 3: "skew-symmetric", 4: "hermitian"}


def mm_info(path):
    """Header of a Matrix Market file: dict(rep, field, symm, nrows, ncols,
    entries) (mminfo, EXAMPLES/MATRIX_MARKET/mmio.f)."""
    r, f, sy, m, n = (ctypes.c_int() for _ in range(5))
    e = ctypes.c_int64()
    check(_lib.load().idrs_mm_info(os.fsencode(path), ctypes.byref(r), ctypes.byref(f), ctypes.byref(sy),
                                   ctypes.byref(m), ctypes.byref(n), ctypes.byref(e)))
    return dict(rep=MM_REP[r.value], field=MM_FIELD[f.value], symm=MM_SYMM[sy.value], nrows=m.value,
                ncols=n.value, entries=e.value)


def mm_matrix(path, which_format="coo"):
    """A = mm_matrix(mm_file, which_format) (EXAMPLES/MATRIX_MARKET/
    mm_module.f90:48-150): a coordinate file as a user matrix on the device
    ('coo' or 'crs'); an array file as a numpy array (the reference's dense
    user matrix, here host data such as a block of right-hand sides)."""
    info = mm_info(path)
    if info["rep"] == "array":
        cplx = info["field"] == "complex"
        a = np.empty((info["nrows"], info["ncols"]), dtype=np.complex128 if cplx else np.float64, order="F")
        check(_lib.load().idrs_mm_dense(os.fsencode(path), _ptr(a)))
        return a
    fmt = {"coo": 3, "crs": 2}[which_format]
    h = ctypes.c_void_p()
    check(_lib.load().idrs_mm_matrix(os.fsencode(path), fmt, ctypes.byref(h)))
    return UserMatrix(h.value, info["field"] == "complex")


def real_diagonal(K):
    """D = real_diagonal(K), zeros replaced by 1 (user_module.f90:960-1047)."""
    d = np.empty(K.nrows)
    check(_lib.load().idrs_real_diagonal(K.handle, _ptr(d)))
    return d


def complex_diagonal(K):
    """D = complex_diagonal(K) (user_module.f90:1049-1161)."""
    d = np.empty(K.nrows, dtype=np.complex128)
    check(_lib.load().idrs_complex_diagonal(K.handle, _ptr(d)))
    return d


class Matrix:
    """type(matrix) (matrix_module.f90:38-81): the operator the solvers use.

    A = Matrix(K) sets A%K = K; set_precon(1) sets A%P = REAL_PRECON(1, K)
    (diagonal scaling); real_preconditioner()/complex_preconditioner() add the
    Chebyshev/Neumann polynomial.  A * v is rmatvec/cmatvec, v / A is
    rprecon/cprecon.
    """

    __array_ufunc__ = None  # numpy defers v / A to __rtruediv__

    def __init__(self, K, which_precon=0, shift=None):
        self.K = K  # keeps the user matrix alive
        h = ctypes.c_void_p()
        check(_lib.load().idrs_matrix_create(K.handle, ctypes.byref(h)))
        self._h = h
        self.is_complex = K.is_complex
        self.n = K.nrows
        self.seed = 0.0
        self.shift_invert = False
        self.si_shift = 0.0
        if which_precon:
            self.set_precon(which_precon, shift)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().idrs_matrix_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_precon(self, which_precon, shift=None):
        """A%P = REAL_PRECON/COMPLEX_PRECON(which_precon, K[, shift])."""
        lib = _lib.load()
        if shift is not None and np.iscomplexobj(shift):
            s = np.array([np.real(shift), np.imag(shift)], dtype=np.float64)
            check(lib.idrs_matrix_set_precon_c128(self._h, int(which_precon), _ptr(s)))
        else:
            s = None if shift is None else np.array([float(shift)])
            check(lib.idrs_matrix_set_precon_f64(self._h, int(which_precon), _ptr(s)))

    def stored_values(self):
        """(value entries one product reads, unit diagonal implicit) --
        with diagonal scaling a CDS operator's all-ones main diagonal is not
        stored (idrs_matrix_stored_values)."""
        cnt, unit = ctypes.c_int64(), ctypes.c_int()
        check(_lib.load().idrs_matrix_stored_values(self._h, ctypes.byref(cnt), ctypes.byref(unit)))
        return cnt.value, bool(unit.value)

    def set_seed(self, seed):
        self.seed = seed
        lib = _lib.load()
        if np.iscomplexobj(seed):
            s = np.array([np.real(seed), np.imag(seed)], dtype=np.float64)
            check(lib.idrs_matrix_set_seed_c128(self._h, _ptr(s)))
        else:
            check(lib.idrs_matrix_set_seed_f64(self._h, float(seed)))

    @property
    def alpha(self):
        deg = ctypes.c_int()
        buf = np.zeros(2 * 64)
        check(_lib.load().idrs_matrix_get_alpha(self._h, _ptr(buf), 64, ctypes.byref(deg)))
        a = buf[: 2 * (deg.value + 1)].view(np.complex128)
        return a.copy()

    def _vec(self, v, cplx):
        dt = np.complex128 if cplx else np.float64
        return np.ascontiguousarray(v, dtype=dt)

    def __mul__(self, v):
        """w = A*v (rmatvec / cmatvec, matrix_module.f90:135-167).  A 2-D v
        (n x k) is a block of k vectors: every column is multiplied, up to 8
        columns sharing each read of the matrix, with the same result as
        one column at a time."""
        cplx = self.is_complex or np.iscomplexobj(v)
        dt = np.complex128 if cplx else np.float64
        if np.ndim(v) == 2:
            v = np.asfortranarray(v, dtype=dt)
            w = np.empty_like(v, order="F")
            f = _lib.load().idrs_matvec_block_c128 if cplx else _lib.load().idrs_matvec_block_f64
            check(f(self._h, v.shape[1], _ptr(v), _ptr(w)))
            return w
        v = self._vec(v, cplx)
        w = np.empty_like(v)
        f = _lib.load().idrs_matvec_c128 if cplx else _lib.load().idrs_matvec_f64
        check(f(self._h, _ptr(v), _ptr(w)))
        return w

    # --- generalized operators (PP_IDRS/real_matrix.inc, complex_matrix.inc) ---
    def set_mass(self, M):
        """A%M = M (A%mass_matrix): products K - seed M, or with C,
        K + seed C + seed^2 M (matrix_module.f90:203-214)."""
        self._M = M
        check(_lib.load().idrs_matrix_set_mass(self._h, M.handle if M is not None else None))

    def set_damping(self, C):
        """A%C = C (A%damping_matrix): K + seed C + seed^2 (M or I)."""
        self._C = C
        check(_lib.load().idrs_matrix_set_damping(self._h, C.handle if C is not None else None))

    def set_flags(self, block_system=False, multishift=False):
        """A%block_system (vectors of 2 K%nrows: [-C -K; I 0] - seed [M 0; 0 I])
        and A%multishift (deflation projector I - V V^H)."""
        check(_lib.load().idrs_matrix_set_flags(self._h, int(bool(block_system)), int(bool(multishift))))

    @property
    def rows(self):
        r = ctypes.c_int()
        check(_lib.load().idrs_matrix_rows(self._h, ctypes.byref(r)))
        return r.value

    def set_precon_diag(self, PK, PM=None, PC=None, shift=None):
        """A%P = REAL_PRECON / COMPLEX_PRECON(1, PK, PM, PC, shift)
        (user_module.f90:1327-1381): e.g. PK = M, a lumped mass matrix."""
        self._P = (PK, PM, PC)
        cplx = self.is_complex or np.iscomplexobj(shift)
        s = None
        if shift is not None:
            s = np.array([np.real(shift), np.imag(shift)]) if cplx else np.array([float(shift)])
        check(_lib.load().idrs_matrix_set_precon_diag(self._h, PK.handle, PM.handle if PM is not None else None,
                                                      PC.handle if PC is not None else None, _ptr(s), int(cplx)))

    def set_precon_lu(self, which, PK, PM=None, PC=None, shift=None):
        """A%P = REAL_PRECON / COMPLEX_PRECON(which, PK, PM, PC, shift)
        (user_module.f90:1383-1459, 1531-1602): which 2 dense LU, 3 band LU
        of PK [+ shift PC + shift^2 PM | ...], factored on the device.
        Raises IdrsError on a zero pivot."""
        self._P = (PK, PM, PC)
        cplx = self.is_complex or np.iscomplexobj(shift)
        s = None
        if shift is not None:
            s = np.array([np.real(shift), np.imag(shift)]) if cplx else np.array([float(shift)])
        check(_lib.load().idrs_matrix_set_precon_lu(self._h, int(which), PK.handle,
                                                    PM.handle if PM is not None else None,
                                                    PC.handle if PC is not None else None, _ptr(s), int(cplx)))

    def share_lu(self, other):
        """Use other's LU factors (no second factorization)."""
        self._P = getattr(other, "_P", None)
        check(_lib.load().idrs_matrix_share_lu(self._h, other.handle))

    def set_shift_invert(self, shift):
        """A%shift_invert = .true., A%rshift / A%cshift = shift (pp_idrs with
        present(P) and present(shift)); None switches it off."""
        self.shift_invert = shift is not None
        self.si_shift = 0.0 if shift is None else shift
        cplx = np.iscomplexobj(shift)
        s = None if shift is None else (np.array([np.real(shift), np.imag(shift)]) if cplx
                                        else np.array([float(shift)]))
        check(_lib.load().idrs_matrix_set_shift_invert(self._h, int(shift is not None), _ptr(s), int(cplx)))

    def _cpl(self, *arrs):
        return self.is_complex or any(np.iscomplexobj(a) for a in arrs)

    def make_deflation(self, V):
        """call make_deflation(A, V) (matrix_module.f90:1284-1352)."""
        cplx = self._cpl(V)
        V = np.asfortranarray(V, dtype=np.complex128 if cplx else np.float64)
        if V.ndim == 1:
            V = V.reshape(-1, 1, order="F")
        f = _lib.load().idrs_make_deflation_c128 if cplx else _lib.load().idrs_make_deflation_f64
        check(f(self._h, V.shape[1], _ptr(V)))

    def _gen(self, name, v, b=None):
        cplx = self._cpl(v, b if b is not None else 0.0)
        dt = np.complex128 if cplx else np.float64
        v = np.ascontiguousarray(v, dtype=dt)
        w = np.empty_like(v)
        f = getattr(_lib.load(), f"idrs_{name}_{'c128' if cplx else 'f64'}")
        if b is None:
            check(f(self._h, _ptr(v), _ptr(w)))
        else:
            bb = np.ascontiguousarray(b, dtype=dt)
            check(f(self._h, _ptr(bb), _ptr(v), _ptr(w)))
        return w

    def pmul(self, v):
        """Pmul(A, v) (matrix_module.f90:1369-1431)."""
        return self._gen("pmul", v)

    def system_mv(self, v):
        """real_mv(A, v) / complex_mv(A, v) alone (no polynomial, no Pmul)."""
        return self._gen("system_mv", v)

    def add_deflation(self, b, y):
        """add_deflation(A, b, y) (matrix_module.f90:1433-1506)."""
        return self._gen("add_deflation", y, b)

    def __rtruediv__(self, v):
        """w = v/A (rprecon / cprecon, matrix_module.f90:306-400)."""
        cplx = self.is_complex or np.iscomplexobj(v)
        v = self._vec(v, cplx)
        w = np.empty_like(v)
        f = _lib.load().idrs_precon_c128 if cplx else _lib.load().idrs_precon_f64
        check(f(self._h, _ptr(v), _ptr(w)))
        return w


def real_preconditioner(A, preconditioner=0, degree=0, foci=None, center=None, H=None):
    """call real_preconditioner(A, preconditioner, degree, foci, center, H)
    (matrix_module.f90:551-635): 0 none, 1 Neumann, 2 Chebyshev, 3 Hessenberg."""
    f = None if foci is None else np.ascontiguousarray(foci, dtype=np.float64)
    c = None if center is None else np.array([float(center)])
    h, ldh, nc = None, 0, 0
    if H is not None:
        h = np.asfortranarray(H, dtype=np.float64)
        ldh, nc = h.shape
    check(_lib.load().idrs_real_preconditioner(A.handle, int(preconditioner), int(degree), _ptr(f), _ptr(c), _ptr(h), ldh, nc))


def complex_preconditioner(A, preconditioner=0, degree=0, foci=None, center=None, H=None):
    """call complex_preconditioner(...) (matrix_module.f90:637-722)."""
    f = None if foci is None else np.ascontiguousarray(foci, dtype=np.complex128)
    c = None if center is None else np.array([complex(center)], dtype=np.complex128)
    h, ldh, nc = None, 0, 0
    if H is not None:
        h = np.asfortranarray(H, dtype=np.complex128)
        ldh, nc = h.shape
    check(_lib.load().idrs_complex_preconditioner(A.handle, int(preconditioner), int(degree), _ptr(f), _ptr(c), _ptr(h), ldh, nc))


def scaleback(A, b, y):
    """x = scaleback(A, b, y) = add_deflation(A, b, polpre(y)) / A
    (matrix_module.f90:808-849)."""
    cplx = A.is_complex or np.iscomplexobj(y) or np.iscomplexobj(b)
    dt = np.complex128 if cplx else np.float64
    y = np.ascontiguousarray(y, dtype=dt)
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.empty_like(y)
    f = _lib.load().idrs_scaleback_b_c128 if cplx else _lib.load().idrs_scaleback_b_f64
    check(f(A.handle, _ptr(b), _ptr(y), _ptr(x)))
    return x


def ms_add_deflation(A, sigma, b, Y):
    """X = add_deflation(A, sigma, b, Y) (rms_add_deflation /
    cms_add_deflation, matrix_module.f90:1508-1623)."""
    sig = np.atleast_1d(sigma)
    cplx = A.is_complex or np.iscomplexobj(sig) or np.iscomplexobj(Y) or np.iscomplexobj(b)
    dt = np.complex128 if cplx else np.float64
    Y = np.asfortranarray(Y, dtype=dt)
    X = np.empty_like(Y, order="F")
    b = np.ascontiguousarray(b, dtype=dt)
    if cplx:
        sp = np.empty(2 * len(sig))
        sp[0::2], sp[1::2] = np.real(sig), np.imag(sig)
    else:
        sp = np.ascontiguousarray(sig, dtype=np.float64)
    f = _lib.load().idrs_ms_add_deflation_c128 if cplx else _lib.load().idrs_ms_add_deflation_f64
    check(f(A.handle, len(sig), _ptr(sp), _ptr(b), _ptr(Y), _ptr(X)))
    return X


def ms_scaleback_b(A, sigma, alpha_s, b, Y):
    """X = scaleback(A, sigma, b, Y) (ms_rscaleback / ms_cscaleback,
    matrix_module.f90:724-806): per shift the polynomial alpha_s(:, k) (from
    ms_parameters), the deflation step for sigma_k, the shift-and-invert
    factor (seed - shift)/(sigma_k - shift), then /A."""
    sig = np.atleast_1d(sigma)
    cplx = A.is_complex or np.iscomplexobj(sig) or np.iscomplexobj(Y) or np.iscomplexobj(b)
    dt = np.complex128 if cplx else np.float64
    Y = np.asfortranarray(Y, dtype=dt)
    X = np.empty_like(Y, order="F")
    b = np.ascontiguousarray(b, dtype=dt)
    a = np.asfortranarray(alpha_s, dtype=np.complex128)      # (degree+1, n_sigma)
    deg = a.shape[0] - 1
    ap = np.empty(2 * a.size)
    ap[0::2], ap[1::2] = a.ravel(order="F").real, a.ravel(order="F").imag
    if cplx:
        sp = np.empty(2 * len(sig))
        sp[0::2], sp[1::2] = np.real(sig), np.imag(sig)
    else:
        sp = np.ascontiguousarray(sig, dtype=np.float64)
    f = _lib.load().idrs_ms_scaleback_b_c128 if cplx else _lib.load().idrs_ms_scaleback_b_f64
    check(f(A.handle, len(sig), _ptr(sp), deg, _ptr(ap), _ptr(b), _ptr(Y), _ptr(X)))
    return X


def make_p(n, s, method=1, PT_raw=None, r=None, seed=1, complex=False):
    """PT(s,n) from MAKE_P (idrs_module.f90:839-937), returned as an (n, rows)
    array (row q holds PT(:,q))."""
    rows = s if method == 1 else 1
    dt = np.complex128 if complex else np.float64
    raw = None if PT_raw is None else np.ascontiguousarray(PT_raw, dtype=dt)
    rr = None if r is None else np.ascontiguousarray(r, dtype=dt)
    out = np.empty((n, rows), dtype=dt)
    f = _lib.load().idrs_make_p_c128 if complex else _lib.load().idrs_make_p_f64
    check(f(int(n), int(s), int(method), _ptr(raw), _ptr(rr), int(seed), _ptr(out)))
    return out


def idrs(A, b, s, M1=None, tolerance=None, maximum_iterations=None, variant=None, x0=None, U0=None,
         omega=None, resvec=False, PT_raw=None, p_seed=None, time_spmv=False, graph=None, out=None, H=None):
    """x = IDRS(A, b, s, M1, tolerance, maximum_iterations, variant, flag,
    relres, iterations, x0, U0, omega, resvec, H)  (idrs_module.f90:107-251).

    H: nritz (an int) asks for the Hessenberg matrix H(nritz+1, nritz) of
    the inner steps (idrs.inc:81-90, 179-181, 211-213, 231-237).

    Returns IdrsResult(x, flag, relres, iterations, resvec, solve_ms, ...).
    b may be a numpy array (host) or a torch CUDA tensor (device-resident
    solve; x0, U0, PT_raw and out are then CUDA tensors too).
    PT_raw: raw shadow space, shape (n, s) (row q = PT(:,q)), orthonormalised
    by the library like MAKE_P does with its random numbers.
    graph: None replays each cycle from a HIP graph when n <= 2^21, True
    always, False never; the results are the same either way.
    """
    lib = _lib.load()
    device = _is_torch(b)
    n = int(b.shape[0])
    o = idrs_opts_t()
    keep = []
    if M1 is not None:
        o.has_M1 = 1
        o.M1 = M1.handle if isinstance(M1, Matrix) else None
    if tolerance is not None:
        if tolerance < 0:
            raise ValueError("Illegal value for parameter tolerance")
        o.has_tolerance, o.tolerance = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maximum_iterations, o.maximum_iterations = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    maxit = int(maximum_iterations) if maximum_iterations is not None else min(2 * n, 1000)
    if device:
        import torch

        cplx = b.is_complex()
        dt = torch.complex128 if cplx else torch.float64
        if b.dtype != dt or not b.is_contiguous():
            raise TypeError("device b must be a contiguous float64/complex128 CUDA tensor")
        x = out if out is not None else torch.empty_like(b)
        o.device_pointers = 1
        for name, arr in (("x0", x0), ("U0", U0), ("PT_raw", PT_raw)):
            if arr is not None:
                if not (_is_torch(arr) and arr.is_cuda and arr.dtype == dt and arr.is_contiguous()):
                    raise TypeError(f"{name} must be a contiguous CUDA tensor of dtype {dt}")
                setattr(o, name, arr.data_ptr())
        bp, xp = ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr())
        npdt = np.complex128 if cplx else np.float64
        # the library runs on its own stream: finish torch's producers first
        torch.cuda.current_stream().synchronize()
    else:
        cplx = A.is_complex or np.iscomplexobj(b)
        npdt = np.complex128 if cplx else np.float64
        b = np.ascontiguousarray(b, dtype=npdt)
        x = np.empty_like(b)
        if x0 is not None:
            a = np.ascontiguousarray(x0, dtype=npdt)
            keep.append(a)
            o.x0 = a.ctypes.data
        if U0 is not None:
            a = np.asfortranarray(U0, dtype=npdt)  # U0(n,s) column-major
            keep.append(a)
            o.U0 = a.ctypes.data
        if PT_raw is not None:
            a = np.ascontiguousarray(PT_raw, dtype=npdt)  # (n, rows) C order == PT(rows,n) column-major
            keep.append(a)
            o.PT_raw = a.ctypes.data
        bp, xp = _ptr(b), _ptr(x)
    if omega is not None:
        om = np.ascontiguousarray(np.atleast_1d(omega), dtype=npdt)
        keep.append(om)
        o.omega, o.n_omega = om.ctypes.data, len(om)
    rv = None
    if resvec:
        rv = np.zeros(maxit + 1)
        o.resvec, o.resvec_len = rv.ctypes.data, len(rv)
    o.p_seed = int(p_seed) if p_seed is not None else 0
    o.time_spmv = 1 if time_spmv else 0
    o.graph_mode = 0 if graph is None else (1 if graph else 2)
    Hm = None
    if H is not None:
        Hm = np.zeros((int(H) + 1, int(H)), dtype=npdt, order="F")
        o.H, o.H_nritz = Hm.ctypes.data, int(H)
    st = idrs_stats_t()
    f = lib.idrs_idrs_c128 if cplx else lib.idrs_idrs_f64
    check(f(A.handle, n, bp, int(s), ctypes.byref(o), xp, ctypes.byref(st)))
    if rv is not None:
        rv = rv[: st.iterations + 1]
    return IdrsResult(x, st.flag, st.relres, st.iterations, rv, st.solve_ms, st.spmv_ms, st.spmv_launches, Hm)


def idrs_ritz(A, b, s, M=None, tolerance=None, maximum_iterations=None, variant=None, x0=None, n_omega=0,
              dim_u0=0, H_shape=None, resvec=False, PT_raw=None, p_seed=None):
    """x = IDRS_RITZ(A, b, s, M, tolerance, maximum_iterations, variant,
    flag, relres, iterations, x0, U0, omega, resvec, H)
    (ritz_module.f90:225-556): the IDRS solve, then n_omega suggested
    omegas, the recycle space U0(n, dim_u0) and H = Hf(H_shape) of its
    Hessenberg matrix.  Host arrays (numpy)."""
    lib = _lib.load()
    n = int(len(b))
    cplx = A.is_complex or np.iscomplexobj(b)
    npdt = np.complex128 if cplx else np.float64
    b = np.ascontiguousarray(b, dtype=npdt)
    x = np.empty_like(b)
    o = idrs_opts_t()
    keep = []
    if M is not None:
        o.has_M1, o.M1 = 1, M.handle
    if tolerance is not None:
        if tolerance < 0:
            raise ValueError("Illegal value for parameter tolerance")
        o.has_tolerance, o.tolerance = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maximum_iterations, o.maximum_iterations = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    maxit = int(maximum_iterations) if maximum_iterations is not None else min(2 * n, 1000)
    if x0 is not None:
        a = np.ascontiguousarray(x0, dtype=npdt)
        keep.append(a)
        o.x0 = a.ctypes.data
    if PT_raw is not None:
        a = np.ascontiguousarray(PT_raw, dtype=npdt)
        keep.append(a)
        o.PT_raw = a.ctypes.data
    rv = None
    if resvec:
        rv = np.zeros(maxit + 1)
        o.resvec, o.resvec_len = rv.ctypes.data, len(rv)
    o.p_seed = int(p_seed) if p_seed is not None else 0
    om = np.zeros(max(n_omega, 1), dtype=npdt)
    U0 = np.zeros((n, max(dim_u0, 1)), dtype=npdt, order="F")
    hr, hc = H_shape if H_shape is not None else (0, 0)
    Hm = np.zeros((max(hr, 1), max(hc, 1)), dtype=npdt, order="F")
    nritz, conv = ctypes.c_int(0), ctypes.c_int(0)
    st = idrs_stats_t()
    f = lib.idrs_idrs_ritz_c128 if cplx else lib.idrs_idrs_ritz_f64
    check(f(A.handle, n, _ptr(b), int(s), ctypes.byref(o), _ptr(x), ctypes.byref(st), _ptr(om), int(n_omega),
            _ptr(U0), int(dim_u0), _ptr(Hm), int(hr), int(hc), ctypes.byref(nritz), ctypes.byref(conv)))
    if rv is not None:
        rv = rv[: st.iterations + 1]
    return RitzResult(x, st.flag, st.relres, st.iterations, rv, nritz.value, om[:n_omega] if n_omega else None,
                      U0 if dim_u0 else None, Hm if H_shape is not None else None,
                      conv.value if n_omega else None, st.solve_ms)


def _pairs(a, shape_cnt):
    a = np.asarray(a)
    p = np.empty(2 * shape_cnt)
    v = np.asarray(a, dtype=np.complex128).ravel(order="F")
    p[0::2], p[1::2] = v.real, v.imag
    return p


def qr_eig(H, want_U=False):
    """eig = QR_eig(H, U, converged) (dense_la_module.f90:192-403) of
    H(1:n,1:n), n = number of columns; RQR_eig for real H, CQR_eig for
    complex H.  Returns (eig, U or None, converged).  Host routine."""
    H = np.asarray(H)
    n, ldh = H.shape[1], H.shape[0]
    hp = _pairs(H, H.size)
    eig = np.empty(2 * n)
    U = np.empty(2 * n * n) if want_U else None
    conv = ctypes.c_int(0)
    check(_lib.load().idrs_qr_eig(n, _ptr(hp), ldh, int(np.iscomplexobj(H)), _ptr(eig),
                                  _ptr(U) if want_U else None, ctypes.byref(conv)))
    e = eig[0::2] + 1j * eig[1::2]
    Um = (U[0::2] + 1j * U[1::2]).reshape((n, n), order="F") if want_U else None
    return e, Um, conv.value


def filter_ritzval(ritzval):
    """outlier = filter_ritzval(ritzval) (matrix_module.f90:1059-1077)."""
    rv = np.asarray(ritzval, dtype=np.complex128)
    out = np.zeros(len(rv), dtype=np.int32)
    check(_lib.load().idrs_filter_ritzval(len(rv), _ptr(_pairs(rv, len(rv))), _ptr(out)))
    return out.astype(bool)


def analyse_spectrum(H):
    """(foci, center, shift) = analyse_spectrum(H) (matrix_module.f90:
    1079-1282); real values for real H."""
    H = np.asarray(H)
    cplx = np.iscomplexobj(H)
    f, c, sh = np.empty(4), np.empty(2), np.empty(2)
    check(_lib.load().idrs_analyse_spectrum(H.shape[1], _ptr(_pairs(H, H.size)), H.shape[0], int(cplx), _ptr(f),
                                            _ptr(c), _ptr(sh)))
    foci = f[0::2] + 1j * f[1::2]
    center, shift = complex(c[0], c[1]), complex(sh[0], sh[1])
    if not cplx:
        return foci.real, center.real, shift.real
    return foci, center, shift


def si_power(H, s, shift=0.0, seed=1, offset=0):
    """schurvec = SI_power(H, s, shift) (dense_la_module.f90:455-603),
    started from the library's uniform stream (seed) at counter offset."""
    H = np.asarray(H)
    cplx = np.iscomplexobj(H)
    n = H.shape[1]
    out = np.empty(2 * n * s)
    sh = np.array([np.real(shift), np.imag(shift)])
    check(_lib.load().idrs_si_power(n, _ptr(_pairs(H, H.size)), H.shape[0], int(cplx), int(s), _ptr(sh),
                                    int(seed), int(offset), _ptr(out)))
    v = (out[0::2] + 1j * out[1::2]).reshape((n, s), order="F")
    return v if cplx else v.real.copy()


def _krylov(kind, A, b, s, sigma=None, M1=None, tolerance=None, maximum_iterations=None, variant=None,
            inner_s=None, inner_tolerance=None, inner_iterations=None, x0=None, omega=None, resvec=False,
            colfac=False, PT_raw=None, p_seed=None, out=None):
    lib = _lib.load()
    device = _is_torch(b)
    n = int(b.shape[0])
    o = idrs_krylov_opts_t()
    keep = []
    if M1 is not None:
        o.has_M1 = 1
        o.M1 = M1.handle if isinstance(M1, Matrix) else None
    if tolerance is not None:
        if tolerance < 0:
            raise ValueError("Illegal value for parameter tolerance")
        o.has_tolerance, o.tolerance = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maximum_iterations, o.maximum_iterations = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    if inner_s is not None:
        o.has_inner_s, o.inner_s = 1, int(inner_s)
    if inner_tolerance is not None:
        o.has_inner_tolerance, o.inner_tolerance = 1, float(inner_tolerance)
    if inner_iterations is not None:
        o.has_inner_iterations, o.inner_iterations = 1, int(inner_iterations)
    maxit = int(maximum_iterations) if maximum_iterations is not None else min(2 * n, 1000)
    ms = kind != "qmridr"
    if device:
        import torch

        cplx = b.is_complex() or bool(sigma is not None and np.iscomplexobj(sigma)) or A.is_complex
        dt = torch.complex128 if cplx else torch.float64
        if b.dtype != dt or not b.is_contiguous():
            raise TypeError(f"device b must be a contiguous {dt} CUDA tensor")
        npdt = np.complex128 if cplx else np.float64
    else:
        cplx = A.is_complex or np.iscomplexobj(b) or bool(sigma is not None and np.iscomplexobj(sigma))
        npdt = np.complex128 if cplx else np.float64
        b = np.ascontiguousarray(b, dtype=npdt)
    ns = 1
    sp = None
    if ms:
        sig = np.ascontiguousarray(np.atleast_1d(sigma), dtype=npdt)
        keep.append(sig)
        ns, sp = len(sig), sig.ctypes.data
    if device:
        X = out if out is not None else torch.empty((ns, n) if ms else (n,), dtype=dt, device=b.device)
        o.device_pointers = 1
        for name, arr in (("x0", x0), ("PT_raw", PT_raw)):
            if arr is not None:
                if not (_is_torch(arr) and arr.is_cuda and arr.dtype == dt and arr.is_contiguous()):
                    raise TypeError(f"{name} must be a contiguous CUDA tensor of dtype {dt}")
                setattr(o, name, arr.data_ptr())
        bp, xp = ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(X.data_ptr())
        torch.cuda.current_stream().synchronize()
    else:
        X = np.empty((ns, n) if ms else n, dtype=npdt)  # row k = solution of shift k (column-major X(n,ns))
        if x0 is not None:
            a = np.ascontiguousarray(x0, dtype=npdt)
            keep.append(a)
            o.x0 = a.ctypes.data
        if PT_raw is not None:
            a = np.ascontiguousarray(PT_raw, dtype=npdt)
            keep.append(a)
            o.PT_raw = a.ctypes.data
        bp, xp = _ptr(b), _ptr(X)
    if omega is not None:
        om = np.ascontiguousarray(np.atleast_1d(omega), dtype=npdt)
        keep.append(om)
        o.omega, o.n_omega = om.ctypes.data, len(om)
    rv = None
    if resvec:
        rv = np.zeros((ns, maxit + 1))  # resvec(maxit+1, ns) column-major
        o.resvec, o.resvec_len = rv.ctypes.data, maxit + 1
    cf = None
    if colfac:
        cf = np.zeros(ns, dtype=npdt)
        o.colfac = cf.ctypes.data
    o.p_seed = int(p_seed) if p_seed is not None else 0
    st = idrs_stats_t()
    sfx = "c128" if cplx else "f64"
    if ms:
        rr = np.zeros(ns)
        f = getattr(lib, f"idrs_{kind}_{sfx}")
        check(f(A.handle, n, bp, ns, sp, int(s), ctypes.byref(o), xp, rr.ctypes.data, ctypes.byref(st)))
        relres = rr
    else:
        f = getattr(lib, f"idrs_qmridr_{sfx}")
        check(f(A.handle, n, bp, int(s), ctypes.byref(o), xp, ctypes.byref(st)))
        relres = st.relres
    if rv is not None:
        rv = rv[:, : st.iterations + 1]
        if not ms:
            rv = rv[0]
    if ms and not device:
        X = X.T  # (n, ns) view, like the reference's x(n, n_sigma)
    elif ms:
        X = X.t()
    return KrylovResult(X, st.flag, relres, st.iterations, rv, st.solve_ms, cf)


def qmridr(A, b, s, M1=None, tolerance=None, maximum_iterations=None, inner_s=None, inner_tolerance=None,
           inner_iterations=None, x0=None, omega=None, resvec=False, PT_raw=None, p_seed=None, out=None):
    """x = QMRIDR(A, b, s, M1, tolerance, maximum_iterations, flag, relres,
    iterations, inner_s, inner_tolerance, inner_iterations, x0, omega,
    resvec)  (idrs_module.f90:253-408).  Returns KrylovResult."""
    return _krylov("qmridr", A, b, s, M1=M1, tolerance=tolerance, maximum_iterations=maximum_iterations,
                   inner_s=inner_s, inner_tolerance=inner_tolerance, inner_iterations=inner_iterations, x0=x0,
                   omega=omega, resvec=resvec, PT_raw=PT_raw, p_seed=p_seed, out=out)


def msidrs(A, b, sigma, s, M1=None, tolerance=None, maximum_iterations=None, variant=None, omega=None,
           resvec=False, colfac=False, PT_raw=None, p_seed=None, out=None):
    """X = MSIDRS(A, b, sigma, s, M1, tolerance, maximum_iterations,
    variant, flag, relres, iterations, omega, resvec, colfac)
    (idrs_module.f90:572-713).  X has shape (n, n_sigma); relres and colfac
    hold one entry per shift; resvec has shape (n_sigma, iterations+1)."""
    return _krylov("msidrs", A, b, s, sigma=sigma, M1=M1, tolerance=tolerance,
                   maximum_iterations=maximum_iterations, variant=variant, omega=omega, resvec=resvec,
                   colfac=colfac, PT_raw=PT_raw, p_seed=p_seed, out=out)


def msqmridr(A, b, sigma, s, M1=None, tolerance=None, maximum_iterations=None, inner_s=None,
             inner_tolerance=None, inner_iterations=None, omega=None, resvec=False, PT_raw=None, p_seed=None,
             out=None):
    """X = MSQMRIDR(A, b, sigma, s, M1, tolerance, maximum_iterations, flag,
    relres, iterations, inner_s, inner_tolerance, inner_iterations, omega,
    resvec)  (idrs_module.f90:410-570)."""
    return _krylov("msqmridr", A, b, s, sigma=sigma, M1=M1, tolerance=tolerance,
                   maximum_iterations=maximum_iterations, inner_s=inner_s, inner_tolerance=inner_tolerance,
                   inner_iterations=inner_iterations, omega=omega, resvec=resvec, PT_raw=PT_raw, p_seed=p_seed,
                   out=out)


MsMixedResult = namedtuple("MsMixedResult", "x flag relres iterations outer corrections solve_ms")


def ms_mixed(A, b, sigma, s, solver="msqmridr", tolerance=None, maximum_iterations=None, variant=None,
             inner_s=None, inner_iterations=None, inner_tolerance_ms=None, inner_tolerance=1e-5, max_outer=10,
             p_seed=None):
    """Multishift solve in complex64 with FP64 refinement (no reference
    counterpart; idrs_ms_mixed_c128): MSQMRIDR or MSIDRS with complex64
    vectors and FP32 values for every shift to inner_tolerance, then FP64
    residuals and single-shift complex64 corrections until every shift meets
    the FP64 tolerance.  inner_s / inner_iterations / inner_tolerance_ms are
    MSQMRIDR's own inner-solver options.  Returns MsMixedResult with X
    (n, n_sigma) and each shift's FP64 true residual."""
    lib = _lib.load()
    n = int(len(b))
    sig = np.atleast_1d(np.asarray(sigma, dtype=np.complex128))
    o = idrs_krylov_opts_t()
    if tolerance is not None:
        o.has_tolerance, o.tolerance = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maximum_iterations, o.maximum_iterations = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    if inner_s is not None:
        o.has_inner_s, o.inner_s = 1, int(inner_s)
    if inner_iterations is not None:
        o.has_inner_iterations, o.inner_iterations = 1, int(inner_iterations)
    if inner_tolerance_ms is not None:
        o.has_inner_tolerance, o.inner_tolerance = 1, float(inner_tolerance_ms)
    o.p_seed = int(p_seed) if p_seed is not None else 0
    bb = np.ascontiguousarray(b, dtype=np.complex128)
    sp = np.empty(2 * len(sig))
    sp[0::2], sp[1::2] = sig.real, sig.imag
    X = np.empty((len(sig), n), dtype=np.complex128)     # row k = column k of X(n, n_sigma)
    rr = np.zeros(len(sig))
    st = idrs_stats_t()
    outer, corr = ctypes.c_int(), ctypes.c_int()
    method = {"msqmridr": 0, "msidrs": 1}[solver]
    check(lib.idrs_ms_mixed_c128(A.handle, method, n, _ptr(bb), len(sig), _ptr(sp), int(s), ctypes.byref(o),
                                 float(inner_tolerance), int(max_outer), _ptr(X), _ptr(rr), ctypes.byref(st),
                                 ctypes.byref(outer), ctypes.byref(corr)))
    return MsMixedResult(X.T, st.flag, rr, st.iterations, outer.value, corr.value, st.solve_ms)


def batch_format(mats):
    """Block-diagonal batch of square CRS/COO user matrices (copies).  Vectors
    of the batch are the systems' vectors concatenated in order."""
    mats = list(mats)
    if not mats:
        raise ValueError("a batch needs at least one matrix")
    cplx = any(m.is_complex for m in mats)
    arr = (ctypes.c_void_p * len(mats))(*[m.handle.value for m in mats])
    h = ctypes.c_void_p()
    check(_lib.load().idrs_batch_format(len(mats), arr, ctypes.byref(h)))
    um = UserMatrix(h.value, cplx)
    um._keep = mats
    return um


def batch_info(K):
    """(nsys, row0) of a batch user matrix; row0[k] = first row of system k."""
    lib = _lib.load()
    ns = ctypes.c_int()
    check(lib.idrs_batch_info(K.handle, ctypes.byref(ns), None))
    row0 = np.zeros(ns.value + 1, dtype=np.int64)
    if ns.value:
        check(lib.idrs_batch_info(K.handle, ctypes.byref(ns), row0.ctypes.data))
    return ns.value, row0


def idrs_batched(A, b, s, tolerance=None, maximum_iterations=None, variant=None, x0=None, omega=None,
                 PT_raw=None, p_seed=None, graph=True, out=None):
    """x_k = IDRS(A_k, b_k, s, ...) for every system of a batch (A built on a
    batch_format user matrix).  b (and x0, PT_raw, out) are the systems'
    vectors concatenated: numpy arrays, or CUDA tensors for a
    device-resident batch.  graph: replay each cycle from a HIP graph
    (recaptured when finished systems' tiles are dropped); the results are
    the same either way.  Returns BatchResult with per-system arrays."""
    lib = _lib.load()
    nsys, row0 = batch_info(A.K)
    if nsys == 0:
        raise ValueError("A is not built on a batch matrix")
    device = _is_torch(b)
    n = int(b.shape[0])
    if n != row0[-1]:
        raise ValueError(f"b has {n} entries, the batch {row0[-1]}")
    o = idrs_opts_t()
    keep = []
    if tolerance is not None:
        if tolerance < 0:
            raise ValueError("Illegal value for parameter tolerance")
        o.has_tolerance, o.tolerance = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maximum_iterations, o.maximum_iterations = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    o.graph_mode = 0 if graph else 2
    if device:
        import torch

        cplx = b.is_complex()
        dt = torch.complex128 if cplx else torch.float64
        if b.dtype != dt or not b.is_contiguous():
            raise TypeError("device b must be a contiguous float64/complex128 CUDA tensor")
        x = out if out is not None else torch.empty_like(b)
        o.device_pointers = 1
        for name, arr in (("x0", x0), ("PT_raw", PT_raw)):
            if arr is not None:
                if not (_is_torch(arr) and arr.is_cuda and arr.dtype == dt and arr.is_contiguous()):
                    raise TypeError(f"{name} must be a contiguous CUDA tensor of dtype {dt}")
                setattr(o, name, arr.data_ptr())
        bp, xp = ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr())
        npdt = np.complex128 if cplx else np.float64
        torch.cuda.current_stream().synchronize()
    else:
        cplx = A.is_complex or np.iscomplexobj(b)
        npdt = np.complex128 if cplx else np.float64
        b = np.ascontiguousarray(b, dtype=npdt)
        x = np.empty_like(b)
        for name, arr in (("x0", x0), ("PT_raw", PT_raw)):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=npdt)
                keep.append(a)
                setattr(o, name, a.ctypes.data)
        bp, xp = _ptr(b), _ptr(x)
    if omega is not None:
        om = np.ascontiguousarray(np.atleast_1d(omega), dtype=npdt)
        keep.append(om)
        o.omega, o.n_omega = om.ctypes.data, len(om)
    o.p_seed = int(p_seed) if p_seed is not None else 0
    st = (idrs_stats_t * nsys)()
    f = lib.idrs_idrs_batched_c128 if cplx else lib.idrs_idrs_batched_f64
    check(f(A.handle, bp, int(s), ctypes.byref(o), xp, st))
    flag = np.array([q.flag for q in st])
    relres = np.array([q.relres for q in st])
    its = np.array([q.iterations for q in st])
    return BatchResult(x, flag, relres, its, st[0].solve_ms if nsys else 0.0)


def idrs_mixed(A, b, s, tolerance=None, maximum_iterations=None, variant=None, x0=None, PT_raw=None, p_seed=None,
               inner_tolerance=1e-5, max_outer=10, fp32_vectors=None, out=None):
    """x = A^-1 b to the FP64 tolerance with inner IDR(s) solves on FP32
    matrix values and FP64 residual replacement (idrs_idrs_mixed_*).
    fp32_vectors: inner vectors in FP32 too (default: only for complex b,
    where it is required).  b: numpy array or CUDA tensor as for idrs()."""
    lib = _lib.load()
    device = _is_torch(b)
    n = int(b.shape[0])
    o = idrs_opts_t()
    keep = []
    if tolerance is not None:
        if tolerance < 0:
            raise ValueError("Illegal value for parameter tolerance")
        o.has_tolerance, o.tolerance = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maximum_iterations, o.maximum_iterations = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    if device:
        import torch

        cplx = b.is_complex()
        dt = torch.complex128 if cplx else torch.float64
        if b.dtype != dt or not b.is_contiguous():
            raise TypeError("device b must be a contiguous float64/complex128 CUDA tensor")
        x = out if out is not None else torch.empty_like(b)
        o.device_pointers = 1
        for name, arr in (("x0", x0), ("PT_raw", PT_raw)):
            if arr is not None:
                setattr(o, name, arr.data_ptr())
        bp, xp = ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr())
        torch.cuda.current_stream().synchronize()
    else:
        cplx = np.iscomplexobj(b)
        npdt = np.complex128 if cplx else np.float64
        b = np.ascontiguousarray(b, dtype=npdt)
        x = np.empty_like(b)
        for name, arr in (("x0", x0), ("PT_raw", PT_raw)):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=npdt)
                keep.append(a)
                setattr(o, name, a.ctypes.data)
        bp, xp = _ptr(b), _ptr(x)
    o.p_seed = int(p_seed) if p_seed is not None else 0
    st = idrs_stats_t()
    outer = ctypes.c_int()
    if fp32_vectors is None:
        fp32_vectors = bool(cplx)
    f = lib.idrs_idrs_mixed_c128 if cplx else lib.idrs_idrs_mixed_f64
    check(f(A.handle, n, bp, int(s), ctypes.byref(o), float(inner_tolerance), int(max_outer), int(bool(fp32_vectors)),
            xp, ctypes.byref(st), ctypes.byref(outer)))
    return MixedResult(x, st.flag, st.relres, st.iterations, outer.value, st.solve_ms)


def idrs_batched_mixed(A, b, s, tolerance=None, maximum_iterations=None, variant=None, p_seed=None,
                       inner_tolerance=1e-5, max_outer=10, graph=True, out=None):
    """x_k = A_k^-1 b_k for every system of a batch to the FP64 tolerance on
    each system's true residual, with batched inner IDR(s) solves on the FP32
    copy of the matrix values and per-system FP64 residual replacement
    (idrs_idrs_batched_mixed_f64).  b: the systems' right-hand sides
    concatenated, a float64 numpy array or CUDA tensor.  Returns
    BatchMixedResult with per-system arrays (iterations: inner iterations;
    outer: FP64 residual replacements)."""
    lib = _lib.load()
    nsys, row0 = batch_info(A.K)
    if nsys == 0:
        raise ValueError("A is not built on a batch matrix")
    device = _is_torch(b)
    n = int(b.shape[0])
    if n != row0[-1]:
        raise ValueError(f"b has {n} entries, the batch {row0[-1]}")
    o = idrs_opts_t()
    if tolerance is not None:
        if tolerance < 0:
            raise ValueError("Illegal value for parameter tolerance")
        o.has_tolerance, o.tolerance = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maximum_iterations, o.maximum_iterations = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    o.graph_mode = 0 if graph else 2
    o.p_seed = int(p_seed) if p_seed is not None else 0
    if device:
        import torch

        if b.dtype != torch.float64 or not b.is_contiguous():
            raise TypeError("device b must be a contiguous float64 CUDA tensor")
        x = out if out is not None else torch.empty_like(b)
        o.device_pointers = 1
        bp, xp = ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(x.data_ptr())
        torch.cuda.current_stream().synchronize()
    else:
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        bp, xp = _ptr(b), _ptr(x)
    st = (idrs_stats_t * nsys)()
    outer = (ctypes.c_int * nsys)()
    check(lib.idrs_idrs_batched_mixed_f64(A.handle, bp, int(s), ctypes.byref(o), float(inner_tolerance), int(max_outer),
                                          xp, st, outer))
    return BatchMixedResult(x, np.array([q.flag for q in st]), np.array([q.relres for q in st]),
                            np.array([q.iterations for q in st]), np.array(list(outer)),
                            st[0].solve_ms if nsys else 0.0)


def ms_parameters(A, sigma):
    """sigma_p, alpha_s = MS_PARAMETERS(A, sigma) (matrix_module.f90:402-549,
    without shift-and-invert).  With the operator's polynomial p (degree d)
    and q(t) = t p(t): q(t) - q(s) = (t - s) p_s(t), so the shifted systems
    (K - s I) x = b become K p(K) y - q(s) y = b with x = p_s(K) y; p_s is
    the quotient of synthetic division, sigma_p = q(s).  Shifts are taken
    relative to the operator's seed.  Returns sigma_p (n_sigma) and alpha_s
    ((d+1) x n_sigma, column k the coefficients of p_{s_k})."""
    sigma = np.atleast_1d(np.asarray(sigma))
    alpha = A.alpha
    d = len(alpha) - 1
    si = getattr(A, "shift_invert", False)
    cplx = (np.iscomplexobj(sigma) or np.iscomplexobj(A.seed) or np.any(np.imag(alpha) != 0)
            or (si and np.iscomplexobj(A.si_shift)))
    dt = np.complex128 if cplx else np.float64
    if si:
        # shift-and-invert (matrix_module.f90:420-446): eta = (sigma - seed)/(sigma - shift)
        if A.si_shift == A.seed or np.any(sigma == A.si_shift):
            raise IdrsError(-1, "Preconditioner shift should not be equal to seed or to one of the shifts")
        sig_s = ((sigma - A.seed) / (sigma - A.si_shift)).astype(dt)
    else:
        sig_s = (sigma - A.seed).astype(dt)
    alpha = alpha if cplx else alpha.real
    alpha_s = np.zeros((d + 1, len(sigma)), dtype=dt)
    sigma_p = np.zeros(len(sigma), dtype=dt)
    for k, sk in enumerate(sig_s):
        if d == 0:
            alpha_s[0, k] = alpha[0]
        else:
            alpha_s[d, k] = alpha[d]
            for i in range(d - 1, -1, -1):
                alpha_s[i, k] = alpha[i] + sk * alpha_s[i + 1, k]
        sigma_p[k] = sk * alpha_s[0, k]
    return sigma_p, alpha_s


def ms_scaleback(A, alpha_s, Y, out=None):
    """X(:,k) = p_k(K') Y(:,k) / D (ms_rscaleback / ms_cscaleback,
    matrix_module.f90:724-806) for the alpha_s of ms_parameters.  Y: (n,
    n_sigma) numpy array (any layout) or the (n, n_sigma) view of a CUDA
    tensor that msidrs/msqmridr returned (column k contiguous)."""
    lib = _lib.load()
    alpha_s = np.asarray(alpha_s)
    d, ns = alpha_s.shape[0] - 1, alpha_s.shape[1]
    a = np.ascontiguousarray(alpha_s.T.astype(np.complex128))   # (ns, d+1): column k contiguous
    device = _is_torch(Y)
    if device:
        import torch

        cplx = Y.is_complex()
        Yc = Y.t().contiguous()                     # (ns, n): row k = column k of Y
        Xc = torch.empty_like(Yc)
        torch.cuda.current_stream().synchronize()
        f = lib.idrs_ms_scaleback_c128 if cplx else lib.idrs_ms_scaleback_f64
        check(f(A.handle, ns, d, a.ctypes.data, ctypes.c_void_p(Yc.data_ptr()), ctypes.c_void_p(Xc.data_ptr()), 1))
        return Xc.t()
    cplx = np.iscomplexobj(Y) or A.is_complex
    dt = np.complex128 if cplx else np.float64
    Yc = np.ascontiguousarray(np.asarray(Y, dtype=dt).T)
    Xc = np.empty_like(Yc)
    f = lib.idrs_ms_scaleback_c128 if cplx else lib.idrs_ms_scaleback_f64
    check(f(A.handle, ns, d, a.ctypes.data, _ptr(Yc), _ptr(Xc), 0))
    return Xc.T
