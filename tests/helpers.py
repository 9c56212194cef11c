"""Shared helpers of the parity tests: build the same system for the HIP
library (idrs_amd) and for the oracle."""
import numpy as np

import idrs_amd
import orc


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def gpu_crs(crs, diag_precon=False, poly=None, complex_poly=False):
    K = idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind,
                            **({"cval": crs.val} if np.iscomplexobj(crs.val) else {"rval": crs.val}))
    A = idrs_amd.Matrix(K)
    if diag_precon:
        A.set_precon(1)
    if poly is not None:
        kind, degree, foci = poly
        if complex_poly:
            idrs_amd.complex_preconditioner(A, kind, degree, foci=foci)
        else:
            idrs_amd.real_preconditioner(A, kind, degree, foci=foci)
    return A


def oracle_crs(crs, diag_precon=False, poly=None, complex_vectors=False):
    alpha, degree = None, 0
    if poly is not None:
        kind, degree, foci = poly
        assert kind == 2
        alpha = orc.chebyshev(degree, np.asarray(foci))
    return orc.OracleMatrix(orc.CRS, crs.n, crs=crs, which_precon=1 if diag_precon else 0, degree=degree,
                            alpha=alpha, complex_vectors=complex_vectors or np.iscomplexobj(crs.val))


def rounding_spread(run, b, trials=4, seed=0):
    """How far rounding alone moves a solver on this problem: run(b') for
    b' = b (1 + 1e-16 g), g ~ N(0,1), against run(b).  run returns (x,
    iterations); returns (max relative change of x, max |change of the
    iteration count|).  Used where a recurrence amplifies rounding so much
    that a fixed bound between two implementations would be meaningless."""
    rng = np.random.default_rng(seed)
    x0, it0 = run(b)
    dx, dit = 0.0, 0
    for _ in range(trials):
        g = rng.standard_normal(b.shape)
        if np.iscomplexobj(b):
            g = g + 1j * rng.standard_normal(b.shape)
        x1, it1 = run(b * (1 + 1e-16 * g))
        dx = max(dx, rel(x1, x0))
        dit = max(dit, abs(it1 - it0))
    return dx, dit


def mm_fixture(name):
    """A Matrix Market fixture of tests/golden/mm (header words, size line,
    entries in file order)."""
    import os

    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mm", name + ".npz"))
    return {k: d[k] for k in d.files}


def write_mtx(fx, path):
    """Writes a fixture back as Matrix Market text (17 significant digits:
    the same doubles when parsed)."""
    h, size = list(fx["header"]), list(fx["size"])
    with open(path, "w") as f:
        f.write("%%MatrixMarket " + " ".join(h) + "\n% fixture\n" + " ".join(str(int(v)) for v in size) + "\n")
        if h[1] == "coordinate":
            for (i, j), v in zip(fx["ij"], fx["vals"]):
                f.write(f"{i} {j} " + " ".join(f"{x:.17g}" for x in v) + "\n")
        else:
            for v in fx["vals"]:
                f.write(" ".join(f"{x:.17g}" for x in v) + "\n")
    return path


def mm_expanded(fx):
    """The fixture's entries as the reference's mm_coo lists them: file
    order, each off-diagonal entry of a symmetric / skew-symmetric /
    hermitian file followed by its mirror; 1-based (indx, jndx, values)."""
    h = list(fx["header"])
    field, symm = h[2], h[3]
    ij, vals = fx["ij"], fx["vals"]
    if field == "complex":
        v = (vals[:, 0].astype(np.float32).astype(np.float64)
             + 1j * vals[:, 1].astype(np.float32).astype(np.float64))
    elif field == "pattern":
        v = np.ones(len(ij))
    else:
        v = vals[:, 0].copy()
    I, J, V = [], [], []
    for (i, j), x in zip(ij, v):
        I.append(i)
        J.append(j)
        V.append(x)
        if symm != "general" and i != j:
            I.append(j)
            J.append(i)
            V.append(x if symm == "symmetric" else (-x if symm == "skew-symmetric" else np.conj(x)))
    return np.array(I, dtype=np.int32), np.array(J, dtype=np.int32), np.array(V)
