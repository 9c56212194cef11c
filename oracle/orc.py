"""ctypes front end of the oracle (oracle/liborc.so, built from orc.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker of the HIP library.  The
functions mirror the reference routines restated in orc_generic.inc.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
# the same restatement with OpenMP loops: timed as the CPU baseline, never
# used as the checker (its parallel sums round in another order)
LIB_OMP = os.path.join(HERE, "liborc_omp.so")

CRS, COO, CDS, DIA = 2, 3, 4, 5


class UserMatrix(ctypes.Structure):
    _fields_ = [
        ("format", c_int), ("nrows", c_int), ("ncols", c_int),
        ("rval", c_void_p), ("cval", c_void_p),
        ("row_ptr", c_void_p), ("col_ind", c_void_p),
        ("nnz", c_int), ("indx", c_void_p), ("jndx", c_void_p),
        ("half_band", c_int),
        ("ndiag", c_int), ("offsets", c_void_p),
    ]


class Matrix(ctypes.Structure):
    _fields_ = [
        ("K", UserMatrix),
        ("which_precon", c_int),
        ("rdiag", c_void_p), ("cdiag", c_void_p),
        ("rseed", c_double), ("cseed", c_double * 2),
        ("degree", c_int),
        ("ralpha_p", c_void_p), ("calpha_p", c_void_p),
        ("M", c_void_p), ("C", c_void_p),
        ("block_system", c_int), ("multishift", c_int),
        ("n_deflation", c_int),
        ("V", c_void_p), ("AV", c_void_p), ("VtAV", c_void_p),
        ("lu", c_void_p), ("lu_hb", c_int),
        ("shift_invert", c_int), ("rshift", c_double), ("cshift", c_double * 2),
    ]


class Opts(ctypes.Structure):
    _fields_ = [
        ("has_M1", c_int), ("M1", c_void_p),
        ("has_tol", c_int), ("tol", c_double),
        ("has_maxit", c_int), ("maxit", c_int),
        ("has_variant", c_int), ("variant", c_int),
        ("x0", c_void_p), ("U0", c_void_p),
        ("omega", c_void_p), ("n_omega", c_int),
        ("resvec", c_void_p), ("resvec_len", c_int),
        ("PT_raw", c_void_p),
        ("has_inner_s", c_int), ("inner_s", c_int),
        ("has_inner_tol", c_int), ("inner_tol", c_double),
        ("has_inner_it", c_int), ("inner_it", c_int),
        ("relres_vec", c_void_p),
        ("colfac", c_void_p),
        ("H", c_void_p), ("H_nritz", c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [("flag", c_int), ("relres", c_double), ("iterations", c_int)]


_lib = None
_lib_omp = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so", "liborc_omp.so"], check=True)


def load_timed(threads=0):
    """liborc_omp.so with `threads` OpenMP threads (0: OMP_NUM_THREADS or the
    runtime's default); returns (library, threads in use)."""
    global _lib_omp
    if _lib_omp is None:
        if not os.path.exists(LIB_OMP):
            build()
        _lib_omp = _setup(ctypes.CDLL(LIB_OMP))
    return _lib_omp, _lib_omp.orc_threads(int(threads))


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _setup(ctypes.CDLL(LIB))
    return _lib


def _setup(lib):
    lib.orc_uniform01.restype = c_double
    lib.orc_uniform01.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    lib.orc_r_norm.restype = c_double
    lib.orc_c_norm.restype = c_double
    for p in ("r", "c"):
        getattr(lib, f"orc_{p}_idrs").argtypes = [POINTER(Matrix), c_int, c_void_p, c_int, POINTER(Opts), c_void_p, POINTER(Stats)]
        getattr(lib, f"orc_{p}_qmridr").argtypes = [POINTER(Matrix), c_int, c_void_p, c_int, POINTER(Opts), c_void_p, POINTER(Stats)]
        for name in ("msidrs", "msqmridr"):
            getattr(lib, f"orc_{p}_{name}").argtypes = [POINTER(Matrix), c_int, c_void_p, c_int, c_void_p, c_int,
                                                         POINTER(Opts), c_void_p, POINTER(Stats)]
        getattr(lib, f"orc_{p}_compute_alpha").restype = c_int
    lib.orc_threads.restype = c_int
    lib.orc_threads.argtypes = [c_int]
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


class OracleMatrix:
    """Holds numpy arrays alive behind an orc_matrix."""

    def __init__(self, fmt, n, ncols=None, *, crs=None, coo=None, cds=None, dia=None, which_precon=0,
                 degree=0, alpha=None, seed=0.0, complex_vectors=False, mass=None, damping=None, block=False,
                 multishift=False, precon_diag=None, lu=None, shift_invert=False, si_shift=0.0):
        """mass, damping: OracleMatrix objects whose K is A%M / A%C;
        precon_diag: P%rband(:,0) given directly (e.g. a lumped mass
        matrix's diagonal, real_precon(1, M)) instead of K's diagonal;
        lu: (which_precon 2 or 3, half bandwidth, factors) from precon_lu;
        shift_invert, si_shift: A%shift_invert and A%rshift / A%cshift."""
        self.keep = []
        self.complex = complex_vectors
        um = UserMatrix()
        um.format, um.nrows, um.ncols = fmt, n, ncols if ncols is not None else n

        def hold(a, dt):
            a = np.ascontiguousarray(a, dtype=dt) if dt != "F" else a
            self.keep.append(a)
            return a.ctypes.data

        def setvals(v):
            if np.iscomplexobj(v):
                um.cval = hold(v, np.complex128)
            else:
                um.rval = hold(v, np.float64)

        if fmt == CRS:
            um.row_ptr, um.col_ind = hold(crs.row_ptr, np.int32), hold(crs.col_ind, np.int32)
            setvals(crs.val)
        elif fmt == COO:
            indx, jndx, v = coo
            um.nnz = len(indx)
            um.indx, um.jndx = hold(indx, np.int32), hold(jndx, np.int32)
            setvals(v)
        elif fmt == CDS:
            hb, band = cds
            um.half_band = hb
            setvals(np.asfortranarray(band).ravel(order="F"))
        elif fmt == DIA:
            off, vals = dia
            um.ndiag = len(off)
            um.offsets = hold(off, np.int32)
            setvals(np.asfortranarray(vals).ravel(order="F"))
        self.um = um
        m = Matrix()
        m.K = um
        m.which_precon = which_precon
        lib = load()
        if which_precon == 1 and precon_diag is not None:
            if complex_vectors:
                m.cdiag = hold(np.asarray(precon_diag, dtype=np.complex128), np.complex128)
            else:
                m.rdiag = hold(np.asarray(precon_diag, dtype=np.float64), np.float64)
        elif which_precon == 1:
            if complex_vectors:
                d = np.empty(n, dtype=np.complex128)
                lib.orc_c_diagonal(ctypes.byref(um), c_void_p(d.ctypes.data))
                m.cdiag = hold(d, np.complex128)
            else:
                d = np.empty(n)
                lib.orc_r_diagonal(ctypes.byref(um), c_void_p(d.ctypes.data))
                m.rdiag = hold(d, np.float64)
        if complex_vectors:
            m.cseed[0], m.cseed[1] = np.real(seed), np.imag(seed)
        else:
            m.rseed = float(np.real(seed))
        m.degree = degree
        if degree > 0:
            if complex_vectors:
                m.calpha_p = hold(np.asarray(alpha, dtype=np.complex128), np.complex128)
            else:
                m.ralpha_p = hold(np.asarray(alpha, dtype=np.float64), np.float64)
        if mass is not None:
            self.keep.append(mass)
            m.M = ctypes.addressof(mass.um)
        if damping is not None:
            self.keep.append(damping)
            m.C = ctypes.addressof(damping.um)
        m.block_system = 1 if block else 0
        m.multishift = 1 if multishift else 0
        if lu is not None:
            which, hb, arr = lu
            m.which_precon = which
            m.lu_hb = hb
            m.lu = hold(np.asarray(arr).ravel(order="F"), np.complex128 if complex_vectors else np.float64)
        m.shift_invert = 1 if shift_invert else 0
        if complex_vectors:
            m.cshift[0], m.cshift[1] = np.real(si_shift), np.imag(si_shift)
        else:
            m.rshift = float(np.real(si_shift))
        self.seed = seed
        self.m = m
        self.n = n

    def matvec(self, v):
        dt = np.complex128 if self.complex else np.float64
        v = np.ascontiguousarray(v, dtype=dt)
        w = np.empty_like(v)
        f = load().orc_c_matvec if self.complex else load().orc_r_matvec
        f(ctypes.byref(self.m), c_void_p(v.ctypes.data), c_void_p(w.ctypes.data))
        return w

    def precon(self, v):
        """v / A (rprecon): divides by the preconditioner diagonal, or copies."""
        dt = np.complex128 if self.complex else np.float64
        v = np.ascontiguousarray(v, dtype=dt)
        w = np.empty_like(v)
        f = load().orc_c_precon if self.complex else load().orc_r_precon
        f(ctypes.byref(self.m), c_void_p(v.ctypes.data), c_void_p(w.ctypes.data))
        return w

    @property
    def rows(self):
        return (load().orc_c_rows if self.complex else load().orc_r_rows)(ctypes.byref(self.m))

    def _dt(self):
        return np.complex128 if self.complex else np.float64

    def _call(self, name, *arrays):
        f = getattr(load(), ("orc_c_" if self.complex else "orc_r_") + name)
        return f(ctypes.byref(self.m), *[c_void_p(a.ctypes.data) for a in arrays])

    def system_mv(self, v):
        """real_mv / complex_mv alone (no polynomial, no Pmul)."""
        v = np.ascontiguousarray(v, dtype=self._dt())
        w = np.empty(self.rows, dtype=self._dt())
        self._call("system_mv", v, w)
        return w

    def pmul(self, v):
        w = np.array(v, dtype=self._dt())
        self._call("pmul", w)
        return w

    def make_deflation(self, V):
        """make_deflation(A, V): A%V = V, A%AV = real_mv(A, V), A%VtAV."""
        V = np.asfortranarray(V, dtype=self._dt())
        nd = V.shape[1]
        self._V, self._AV = V, np.zeros_like(V, order="F")
        self._VtAV = np.zeros((nd, nd), dtype=self._dt(), order="F")
        self.m.n_deflation = nd
        self.m.V, self.m.AV, self.m.VtAV = self._V.ctypes.data, self._AV.ctypes.data, self._VtAV.ctypes.data
        f = getattr(load(), ("orc_c_" if self.complex else "orc_r_") + "make_deflation")
        f(ctypes.byref(self.m))
        return self._AV, self._VtAV

    def add_deflation(self, b, y):
        b = np.ascontiguousarray(b, dtype=self._dt())
        y = np.ascontiguousarray(y, dtype=self._dt())
        x = np.empty_like(y)
        self._call("add_deflation", b, y, x)
        return x

    def ms_add_deflation(self, sigma, b, Y):
        sig = np.ascontiguousarray(np.atleast_1d(sigma), dtype=self._dt())
        b = np.ascontiguousarray(b, dtype=self._dt())
        Y = np.asfortranarray(Y, dtype=self._dt())
        X = np.empty_like(Y, order="F")
        f = getattr(load(), ("orc_c_" if self.complex else "orc_r_") + "ms_add_deflation")
        f(ctypes.byref(self.m), c_int(len(sig)), c_void_p(sig.ctypes.data), c_void_p(b.ctypes.data),
          c_void_p(Y.ctypes.data), c_void_p(X.ctypes.data))
        return X

    def scaleback_b(self, b, y):
        """rscaleback / cscaleback with the deflation step."""
        b = np.ascontiguousarray(b, dtype=self._dt())
        y = np.ascontiguousarray(y, dtype=self._dt())
        x = np.empty_like(y)
        self._call("scaleback_b", b, y, x)
        return x

    def ms_parameters(self, sigma, alpha_p=None):
        """(sigma_p, alpha_s) of real_ms_parameters / complex_ms_parameters;
        raises where the reference stops (shift equal to the seed or to a
        shift)."""
        sig = np.ascontiguousarray(np.atleast_1d(sigma), dtype=self._dt())
        sp = np.empty_like(sig)
        a = np.zeros((max(self.m.degree, 0) + 1, len(sig)), dtype=self._dt(), order="F")
        f = getattr(load(), ("orc_c_" if self.complex else "orc_r_") + "ms_parameters")
        f.restype = c_int
        if f(ctypes.byref(self.m), c_int(len(sig)), c_void_p(sig.ctypes.data), c_void_p(sp.ctypes.data),
             c_void_p(a.ctypes.data)):
            raise ValueError("Preconditioner shift should not be equal to seed or to one of the shifts")
        return sp, a

    def ms_scaleback_b(self, sigma, alpha_s, b, Y):
        """ms_rscaleback / ms_cscaleback: X(:,k) from Y(:,k) for shift sigma(k)."""
        sig = np.ascontiguousarray(np.atleast_1d(sigma), dtype=self._dt())
        b = np.ascontiguousarray(b, dtype=self._dt())
        Y = np.asfortranarray(Y, dtype=self._dt())
        X = np.empty_like(Y, order="F")
        a = None if alpha_s is None else np.asfortranarray(alpha_s, dtype=self._dt())
        f = getattr(load(), ("orc_c_" if self.complex else "orc_r_") + "ms_scaleback_b")
        f(ctypes.byref(self.m), c_int(len(sig)), c_void_p(sig.ctypes.data),
          c_void_p(a.ctypes.data if a is not None else None), c_void_p(b.ctypes.data), c_void_p(Y.ctypes.data),
          c_void_p(X.ctypes.data))
        return X

    def user_mv(self, v):
        dt = np.complex128 if self.complex else np.float64
        v = np.ascontiguousarray(v, dtype=dt)
        w = np.empty(self.um.nrows, dtype=dt)
        f = load().orc_c_user_mv if self.complex else load().orc_r_user_mv
        f(ctypes.byref(self.um), c_void_p(v.ctypes.data), c_void_p(w.ctypes.data))
        return w

    def diagonal(self):
        dt = np.complex128 if self.complex else np.float64
        d = np.empty(self.um.nrows, dtype=dt)
        f = load().orc_c_diagonal if self.complex else load().orc_r_diagonal
        f(ctypes.byref(self.um), c_void_p(d.ctypes.data))
        return d

    def scaleback(self, y):
        dt = np.complex128 if self.complex else np.float64
        y = np.ascontiguousarray(y, dtype=dt)
        x = np.empty_like(y)
        f = load().orc_c_scaleback if self.complex else load().orc_r_scaleback
        f(ctypes.byref(self.m), c_void_p(y.ctypes.data), c_void_p(x.ctypes.data))
        return x


def get_band(K, M=None, C=None):
    """get_band over K (and C, M): the half bandwidth REAL_PRECON(3, ...) uses."""
    lib = load()
    lib.orc_get_band.restype = c_int
    return lib.orc_get_band(ctypes.byref(K.um), ctypes.byref(M.um) if M is not None else None,
                            ctypes.byref(C.um) if C is not None else None)


def precon_lu(which, K, M=None, C=None, shift=None, complex_values=False):
    """REAL_PRECON / COMPLEX_PRECON(which = 2 dense LU | 3 band LU, K, M, C,
    shift) of OracleMatrix objects' user matrices: returns (which, hb,
    factors) for OracleMatrix(lu=...); factors (n, n) or (n, 2 hb + 1)
    column-major, the band's column d + hb holding diagonal d.  Raises on a
    zero pivot."""
    lib = load()
    n = K.um.nrows
    hb = n - 1 if which == 2 else get_band(K, M, C)
    cplx = complex_values or np.iscomplexobj(shift)
    dt = np.complex128 if cplx else np.float64
    lu = np.zeros((n, n) if which == 2 else (n, 2 * hb + 1), dtype=dt, order="F")
    f = getattr(lib, f"orc_{'c' if cplx else 'r'}_precon_lu")
    f.restype = c_int
    sh = None if shift is None else np.array([shift], dtype=dt)
    rc = f(c_int(which), c_void_p(ctypes.addressof(K.um)),
           c_void_p(ctypes.addressof(M.um)) if M is not None else None,
           c_void_p(ctypes.addressof(C.um)) if C is not None else None,
           c_void_p(sh.ctypes.data) if sh is not None else None, c_int(hb), c_void_p(lu.ctypes.data))
    if rc:
        raise ValueError("Pivot equal to zero in LU decomposition")
    return which, hb, lu


def lu_solve(lu, b):
    """RLU_solve / band_LU_solve with the factors of precon_lu."""
    which, hb, arr = lu
    cplx = np.iscomplexobj(arr)
    dt = np.complex128 if cplx else np.float64
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.empty_like(b)
    f = getattr(load(), f"orc_{'c' if cplx else 'r'}_lu_solve")
    f(c_int(which), c_int(len(b)), c_int(hb), c_void_p(arr.ctypes.data), c_void_p(b.ctypes.data),
      c_void_p(x.ctypes.data))
    return x


def chebyshev(degree, foci):
    lib = load()
    if np.iscomplexobj(foci):
        f = np.ascontiguousarray(foci, dtype=np.complex128)
        a = np.empty(degree + 1, dtype=np.complex128)
        lib.orc_c_chebyshev(c_int(degree), c_void_p(f.ctypes.data), c_void_p(a.ctypes.data))
    else:
        f = np.ascontiguousarray(foci, dtype=np.float64)
        a = np.empty(degree + 1)
        lib.orc_r_chebyshev(c_int(degree), c_void_p(f.ctypes.data), c_void_p(a.ctypes.data))
    return a


def neumann(degree, center):
    lib = load()
    a = np.empty(degree + 1)
    lib.orc_r_neumann.argtypes = [c_int, c_double, c_void_p]
    lib.orc_r_neumann(degree, float(center), c_void_p(a.ctypes.data))
    return a


def make_p(n, s, method, PT_raw=None, r=None, complex=False):
    lib = load()
    rows = s if method == 1 else 1
    dt = np.complex128 if complex else np.float64
    raw = None if PT_raw is None else np.ascontiguousarray(PT_raw, dtype=dt)
    rr = None if r is None else np.ascontiguousarray(r, dtype=dt)
    out = np.empty((n, rows), dtype=dt)
    f = lib.orc_c_make_p if complex else lib.orc_r_make_p
    f(c_int(n), c_int(s), c_int(method), c_void_p(_p(raw)), c_void_p(_p(rr)), c_void_p(out.ctypes.data))
    return out


def idrs(A, b, s, M1=None, tolerance=None, maximum_iterations=None, variant=None, x0=None, U0=None, omega=None,
         resvec=False, PT_raw=None, lib=None, H_nritz=None):
    """Oracle IDRS with the reference's optional arguments; returns
    (x, flag, relres, iterations, resvec), and H(H_nritz+1, H_nritz) as a
    sixth item when H_nritz is given.  lib: load_timed()'s library to run
    the OpenMP build instead of the checker."""
    lib = lib or load()
    cplx = A.complex
    dt = np.complex128 if cplx else np.float64
    n = len(b)
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.empty_like(b)
    o = Opts()
    keep = []

    def hold(a, order="C"):
        a = np.asarray(a, dtype=dt)
        a = np.asfortranarray(a) if order == "F" else np.ascontiguousarray(a)
        keep.append(a)
        return a.ctypes.data

    if M1 is not None:
        o.has_M1, o.M1 = 1, ctypes.addressof(M1.m)
    if tolerance is not None:
        o.has_tol, o.tol = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maxit, o.maxit = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    if x0 is not None:
        o.x0 = hold(x0)
    if U0 is not None:
        o.U0 = hold(U0, "F")
    if omega is not None:
        o.omega, o.n_omega = hold(np.atleast_1d(omega)), len(np.atleast_1d(omega))
    maxit = maximum_iterations if maximum_iterations is not None else min(2 * n, 1000)
    rv = None
    if resvec:
        rv = np.zeros(maxit + 1)
        o.resvec, o.resvec_len = rv.ctypes.data, len(rv)
    if PT_raw is not None:
        o.PT_raw = hold(PT_raw)
    H = None
    if H_nritz is not None:
        H = np.zeros((H_nritz + 1, H_nritz), dtype=dt, order="F")
        o.H, o.H_nritz = H.ctypes.data, int(H_nritz)
    st = Stats()
    f = lib.orc_c_idrs if cplx else lib.orc_r_idrs
    rc = f(ctypes.byref(A.m), n, c_void_p(b.ctypes.data), s, ctypes.byref(o), c_void_p(x.ctypes.data), ctypes.byref(st))
    if rc != 0:
        raise ValueError("oracle idrs rejected its arguments")
    if rv is not None:
        rv = rv[: st.iterations + 1]
    if H is not None:
        return x, st.flag, st.relres, st.iterations, rv, H
    return x, st.flag, st.relres, st.iterations, rv


def _opts(A, n, dt, keep, M1=None, tolerance=None, maximum_iterations=None, variant=None, x0=None, omega=None,
          PT_raw=None, inner_s=None, inner_tolerance=None, inner_iterations=None):
    o = Opts()

    def hold(a):
        a = np.ascontiguousarray(np.asarray(a, dtype=dt))
        keep.append(a)
        return a.ctypes.data

    if M1 is not None:
        o.has_M1, o.M1 = 1, ctypes.addressof(M1.m)
    if tolerance is not None:
        o.has_tol, o.tol = 1, float(tolerance)
    if maximum_iterations is not None:
        o.has_maxit, o.maxit = 1, int(maximum_iterations)
    if variant is not None:
        o.has_variant, o.variant = 1, int(variant)
    if x0 is not None:
        o.x0 = hold(x0)
    if omega is not None:
        o.omega, o.n_omega = hold(np.atleast_1d(omega)), len(np.atleast_1d(omega))
    if PT_raw is not None:
        o.PT_raw = hold(PT_raw)
    if inner_s is not None:
        o.has_inner_s, o.inner_s = 1, int(inner_s)
    if inner_tolerance is not None:
        o.has_inner_tol, o.inner_tol = 1, float(inner_tolerance)
    if inner_iterations is not None:
        o.has_inner_it, o.inner_it = 1, int(inner_iterations)
    return o


def qmridr(A, b, s, resvec=False, **kw):
    """Oracle QMRIDR (qmridr.inc); returns (x, flag, relres, iterations, resvec)."""
    lib = load()
    dt = np.complex128 if A.complex else np.float64
    n = len(b)
    b = np.ascontiguousarray(b, dtype=dt)
    x = np.empty_like(b)
    keep = []
    o = _opts(A, n, dt, keep, **kw)
    maxit = kw.get("maximum_iterations") or min(2 * n, 1000)
    rv = None
    if resvec:
        rv = np.zeros(maxit + 1)
        o.resvec, o.resvec_len = rv.ctypes.data, len(rv)
    st = Stats()
    f = lib.orc_c_qmridr if A.complex else lib.orc_r_qmridr
    if f(ctypes.byref(A.m), n, c_void_p(b.ctypes.data), s, ctypes.byref(o), c_void_p(x.ctypes.data), ctypes.byref(st)):
        raise ValueError("oracle qmridr rejected its arguments")
    return x, st.flag, st.relres, st.iterations, (rv[: st.iterations + 1] if rv is not None else None)


def _ms(name, A, b, sigma, s, resvec=False, colfac=False, **kw):
    lib = load()
    dt = np.complex128 if A.complex else np.float64
    n = len(b)
    b = np.ascontiguousarray(b, dtype=dt)
    sig = np.ascontiguousarray(np.atleast_1d(sigma), dtype=dt)
    ns = len(sig)
    x = np.zeros((n, ns), dtype=dt, order="F")
    keep = []
    o = _opts(A, n, dt, keep, **kw)
    maxit = kw.get("maximum_iterations") or min(2 * n, 1000)
    rv = None
    if resvec:
        rv = np.zeros((maxit + 1, ns), order="F")
        o.resvec, o.resvec_len = rv.ctypes.data, maxit + 1
    rr = np.zeros(ns)
    o.relres_vec = rr.ctypes.data
    cf = None
    if colfac:
        cf = np.zeros(ns, dtype=dt)
        o.colfac = cf.ctypes.data
    st = Stats()
    f = getattr(lib, f"orc_{'c' if A.complex else 'r'}_{name}")
    if f(ctypes.byref(A.m), n, c_void_p(b.ctypes.data), ns, c_void_p(sig.ctypes.data), s, ctypes.byref(o),
         c_void_p(x.ctypes.data), ctypes.byref(st)):
        raise ValueError(f"oracle {name} rejected its arguments")
    out = dict(x=x, flag=st.flag, relres=rr, iterations=st.iterations)
    if rv is not None:
        out["resvec"] = rv[: st.iterations + 1]
    if cf is not None:
        out["colfac"] = cf
    return out


def msidrs(A, b, sigma, s, **kw):
    """Oracle MSIDRS (msidrs.inc); returns dict(x (n,n_sigma), flag, relres (n_sigma), iterations[, resvec, colfac])."""
    return _ms("msidrs", A, b, sigma, s, **kw)


def msqmridr(A, b, sigma, s, **kw):
    """Oracle MSQMRIDR (msqmridr.inc)."""
    return _ms("msqmridr", A, b, sigma, s, **kw)


def solve(Am, bv):
    lib = load()
    cplx = np.iscomplexobj(Am) or np.iscomplexobj(bv)
    dt = np.complex128 if cplx else np.float64
    Am = np.asfortranarray(Am, dtype=dt)
    bv = np.ascontiguousarray(bv, dtype=dt)
    x = np.empty_like(bv)
    f = lib.orc_c_solve if cplx else lib.orc_r_solve
    f(c_int(len(bv)), c_void_p(Am.ctypes.data), c_void_p(bv.ctypes.data), c_void_p(x.ctypes.data))
    return x


def grot(a, b):
    lib = load()
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        c, s_, r = (np.zeros(1, np.complex128) for _ in range(3))
        class Z(ctypes.Structure):  # double _Complex by value (two SSE words)
            _fields_ = [("re", c_double), ("im", c_double)]

        lib.orc_c_grot.argtypes = [Z, Z, c_void_p, c_void_p, c_void_p]
        a, b = complex(a), complex(b)
        lib.orc_c_grot(Z(a.real, a.imag), Z(b.real, b.imag),
                       c_void_p(c.ctypes.data), c_void_p(s_.ctypes.data), c_void_p(r.ctypes.data))
    else:
        c, s_, r = (np.zeros(1) for _ in range(3))
        lib.orc_r_grot.argtypes = [c_double, c_double, c_void_p, c_void_p, c_void_p]
        lib.orc_r_grot(float(a), float(b), c_void_p(c.ctypes.data), c_void_p(s_.ctypes.data), c_void_p(r.ctypes.data))
    return c[0], s_[0], r[0]


def uniform01(seed, index):
    return load().orc_uniform01(seed, index)


def compute_alpha(degree, H):
    """rcompute_alpha / ccompute_alpha (matrix_module.f90:983-1057) of the
    Hessenberg matrix H (ldh x ncol); returns alpha(0:degree)."""
    lib = load()
    cplx = np.iscomplexobj(H)
    dt = np.complex128 if cplx else np.float64
    Hf = np.asfortranarray(H, dtype=dt)
    a = np.empty(degree + 1, dtype=dt)
    f = lib.orc_c_compute_alpha if cplx else lib.orc_r_compute_alpha
    rc = f(c_int(degree), c_void_p(Hf.ctypes.data), c_int(Hf.shape[0]), c_int(Hf.shape[1]), c_void_p(a.ctypes.data))
    if rc != 0:
        raise ValueError("H too small for the degree")
    return a
