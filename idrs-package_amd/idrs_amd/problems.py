"""Synthetic test systems (seeded numpy generators, host side).

The operators restate the reference's examples:
  - 3-D convection-diffusion-reaction, 7-point central differences on the
    unit cube with Dirichlet boundaries, lexicographic x-fastest ordering
    (EXAMPLES/STENCIL/user_module.f90:520-567 cdr_matrix, :65-187 stencil);
  - 2-D 5-point convection-diffusion (same discretisation in two dimensions);
  - 3-D Helmholtz: h^2-scaled 7-point -Laplacian (laplace_matrix,
    EXAMPLES/STENCIL/user_module.f90:569-620) with shifts;
  - 1-D convection-diffusion Toeplitz matrix (EXAMPLES/TOEPLITZ/toeplitz.f90:111-128);
  - random diagonally dominant sparse matrix;
  - the batch of uneven 2-D systems of the batched workload.
CRS arrays are 1-based (the reference's crs_format convention) with columns
ascending within each row.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class Crs:
    n: int
    row_ptr: np.ndarray   # int32, 1-based, n+1
    col_ind: np.ndarray   # int32, 1-based
    val: np.ndarray       # float64 or complex128

    @property
    def nnz(self):
        return int(self.row_ptr[-1] - 1)

    def matvec(self, v):
        """w = A v on the host, in stored order (float64 accumulation)."""
        rows = np.repeat(np.arange(self.n), np.diff(self.row_ptr))
        w = np.zeros(self.n, dtype=np.result_type(self.val, v))
        np.add.at(w, rows, self.val * v[self.col_ind - 1])
        return w

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.val, self.col_ind - 1, self.row_ptr - 1), shape=(self.n, self.n))

    def transpose(self):
        """Explicit A^T in CRS (columns ascending per row)."""
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.row_ptr))
        cols = self.col_ind.astype(np.int64) - 1
        order = np.lexsort((rows, cols))
        t_rows, t_cols, t_val = cols[order], rows[order], self.val[order]
        counts = np.bincount(t_rows, minlength=self.n)
        rp = np.concatenate([[1], 1 + np.cumsum(counts)]).astype(np.int32)
        return Crs(self.n, rp, (t_cols + 1).astype(np.int32), np.ascontiguousarray(t_val))


@dataclass
class Dia:
    n: int
    offsets: np.ndarray   # int32 ascending
    vals: np.ndarray      # (n, ndiag) Fortran order, zero where the entry is outside the matrix


def _stencil(shape, coefs):
    """Generic Dirichlet stencil on a lexicographic (x fastest) grid.
    coefs: list of (offset_vector, value) in ascending column order."""
    dims = len(shape)
    n = int(np.prod(shape))
    idx = np.arange(n, dtype=np.int64)
    coords = []
    rem = idx
    for d in range(dims):
        coords.append(rem % shape[d])
        rem = rem // shape[d]
    strides = np.cumprod([1] + list(shape[:-1]))
    k = len(coefs)
    cols = np.empty((n, k), dtype=np.int64)
    vals = np.empty((n, k), dtype=np.result_type(*[c[1] for c in coefs]))
    mask = np.ones((n, k), dtype=bool)
    offsets = []
    for j, (vec, val) in enumerate(coefs):
        off = int(sum(v * s for v, s in zip(vec, strides)))
        offsets.append(off)
        ok = np.ones(n, dtype=bool)
        for d, v in enumerate(vec):
            if v:
                c = coords[d] + v
                ok &= (c >= 0) & (c < shape[d])
        mask[:, j] = ok
        cols[:, j] = idx + off
        vals[:, j] = val
    counts = mask.sum(axis=1)
    row_ptr = np.empty(n + 1, dtype=np.int64)
    row_ptr[0] = 1
    np.cumsum(counts, out=row_ptr[1:])
    row_ptr[1:] += 1
    crs = Crs(n, row_ptr.astype(np.int32), (cols[mask] + 1).astype(np.int32), np.ascontiguousarray(vals[mask]))
    dvals = np.where(mask, vals, 0).astype(vals.dtype)
    order = np.argsort(offsets)
    dia = Dia(n, np.asarray(offsets, dtype=np.int32)[order], np.asfortranarray(dvals[:, order]))
    return crs, dia


def convection_diffusion_3d(m, peclet, reaction=0.0, diffusion=1.0, beta=(1.0, 1.0, 1.0)):
    """-diffusion*Lap(u) + Pe*beta.grad(u) + reaction*u on an m^3 interior grid
    (h = 1/(m+1)); returns (Crs, Dia).  Coefficients as cdr_matrix:
    c = 6 eps/h^2 + r, east = -eps/h^2 + b_x/(2h), west = -eps/h^2 - b_x/(2h)."""
    h = 1.0 / (m + 1)
    e2 = diffusion / h ** 2
    bx, by, bz = (peclet * b for b in beta)
    c = 6.0 * e2 + reaction
    coefs = [
        ((0, 0, -1), -e2 - bz / (2 * h)),  # front
        ((0, -1, 0), -e2 - by / (2 * h)),  # south
        ((-1, 0, 0), -e2 - bx / (2 * h)),  # west
        ((0, 0, 0), c),
        ((1, 0, 0), -e2 + bx / (2 * h)),   # east
        ((0, 1, 0), -e2 + by / (2 * h)),   # north
        ((0, 0, 1), -e2 + bz / (2 * h)),   # back
    ]
    return _stencil((m, m, m), coefs)


def convection_diffusion_2d(m, peclet, angle=0.0, diffusion=1.0):
    """2-D 5-point convection-diffusion with convection Pe*(cos a, sin a)."""
    h = 1.0 / (m + 1)
    e2 = diffusion / h ** 2
    bx, by = peclet * np.cos(angle), peclet * np.sin(angle)
    coefs = [
        ((0, -1), -e2 - by / (2 * h)),
        ((-1, 0), -e2 - bx / (2 * h)),
        ((0, 0), 4.0 * e2),
        ((1, 0), -e2 + bx / (2 * h)),
        ((0, 1), -e2 + by / (2 * h)),
    ]
    return _stencil((m, m), coefs)


def helmholtz_3d(m):
    """h^2-scaled 7-point -Laplacian (6 on the diagonal, -1 off) on m^3."""
    coefs = [((0, 0, -1), -1.0), ((0, -1, 0), -1.0), ((-1, 0, 0), -1.0), ((0, 0, 0), 6.0),
             ((1, 0, 0), -1.0), ((0, 1, 0), -1.0), ((0, 0, 1), -1.0)]
    return _stencil((m, m, m), coefs)


def helmholtz_shifts(m, nshift=16, ppw_lo=4.0, ppw_hi=30.0, damping=0.05):
    """sigma_k = (2 pi f_k h)^2 (1 - damping i), kh = 2 pi / ppw_k with
    ppw_k evenly spaced in [ppw_lo, ppw_hi]."""
    ppw = np.linspace(ppw_lo, ppw_hi, nshift)
    kh = 2 * np.pi / ppw
    return (kh ** 2) * (1 - 1j * damping)


def toeplitz_1d(n=60, pe=0.5):
    """Tridiagonal (-1-Pe, 2, -1+Pe), b(1)=1+Pe, b(n)=1-Pe: solution ones
    (EXAMPLES/TOEPLITZ/toeplitz.f90:111-128)."""
    crs, dia = _stencil((n,), [((-1,), -1.0 - pe), ((0,), 2.0), ((1,), -1.0 + pe)])
    b = np.zeros(n)
    b[0] = 1.0 + pe
    b[-1] = 1.0 - pe
    return crs, dia, b


def random_dominant(n, k=16, seed=0, factor=1.2):
    """k off-diagonal entries per row at uniform random columns (diagonal
    excluded), N(0,1) values, diagonal = factor * row abs-sum."""
    rng = np.random.default_rng(seed)
    cols = rng.integers(0, n - 1, size=(n, k), dtype=np.int64)
    rows = np.arange(n, dtype=np.int64)[:, None]
    cols = cols + (cols >= rows)              # skip the diagonal
    vals = rng.standard_normal((n, k))
    diag = factor * np.abs(vals).sum(axis=1)
    allc = np.concatenate([cols, rows], axis=1)
    allv = np.concatenate([vals, diag[:, None]], axis=1)
    order = np.argsort(allc, axis=1, kind="stable")
    allc = np.take_along_axis(allc, order, axis=1)
    allv = np.take_along_axis(allv, order, axis=1)
    row_ptr = (1 + np.arange(n + 1, dtype=np.int64) * (k + 1)).astype(np.int32)
    return Crs(n, row_ptr, (allc.ravel() + 1).astype(np.int32), np.ascontiguousarray(allv.ravel()))


def batched_systems(nsys, seed=0):
    """The batched workload: nsys 2-D 5-point convection-diffusion systems,
    side floor(32 * 16^u) with u ~ U(0,1) (log-uniform in [32, 512]),
    Pe ~ U[1,100], convection angle ~ U[0, 2 pi), b ~ U[-1,1] (all seeded).
    Returns a list of (Crs, b)."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 1.0, nsys)
    pe = rng.uniform(1.0, 100.0, nsys)
    ang = rng.uniform(0.0, 2 * np.pi, nsys)
    out = []
    for k in range(nsys):
        m = int(np.floor(32 * 16 ** u[k]))
        crs, _ = convection_diffusion_2d(m, pe[k], angle=ang[k])
        out.append((crs, rng.uniform(-1.0, 1.0, crs.n)))
    return out


def gaussian_shadow(n, s, seed=0, complex=False):
    """Raw Gaussian shadow space, shape (n, s) (row q = PT(:,q))."""
    rng = np.random.default_rng(seed)
    if complex:
        return rng.standard_normal((n, s)) + 1j * rng.standard_normal((n, s))
    return rng.standard_normal((n, s))
