"""CPU checks of the Matrix Market reader (the header and array files are
read on the host, no device needed): headers and array values of the
reference's own data files against the fixtures and against scipy's
independent reader, the symmetric expansions of array files against their
closed forms, and errors for malformed files."""
import gzip
import os
import shutil

import numpy as np
import pytest
import scipy.io

import idrs_amd

from helpers import mm_fixture, write_mtx

HERE = os.path.dirname(os.path.abspath(__file__))
COORD = ["stommel4", "stommel5", "stommel6", "sag5", "sag6"] + [f"wedge{k}_{m}" for k in (1, 2, 3, 4) for m in "KMC"]
ARRAY = ["stommel6_b", "sag6_b", "wedge3_b"]


@pytest.mark.parametrize("name", COORD + ARRAY)
def test_header(name, tmp_path):
    fx = mm_fixture(name)
    info = idrs_amd.mm_info(write_mtx(fx, str(tmp_path / f"{name}.mtx")))
    h = list(fx["header"])
    assert (info["rep"], info["field"], info["symm"]) == (h[1], h[2], h[3])
    assert info["nrows"] == fx["size"][0] and info["ncols"] == fx["size"][1]
    assert info["entries"] == (fx["size"][2] if h[1] == "coordinate" else fx["size"][0] * fx["size"][1])


@pytest.mark.parametrize("name", ARRAY)
def test_array_values(name, tmp_path):
    fx = mm_fixture(name)
    path = write_mtx(fx, str(tmp_path / f"{name}.mtx"))
    a = idrs_amd.mm_matrix(path)
    m, n = (int(v) for v in fx["size"][:2])
    assert a.shape == (m, n)
    assert np.array_equal(a, fx["vals"][:, 0].reshape(n, m).T)
    assert np.array_equal(a, scipy.io.mmread(path))


@pytest.mark.parametrize("name", ["wedge1_K.mtx.gz", "wedge2_M.mtx.gz"])
def test_verbatim_reference_files(name, tmp_path):
    """Untouched reference text: the header and sizes as scipy reads them."""
    path = str(tmp_path / name[:-3])
    with gzip.open(os.path.join(HERE, "golden", "mm", name), "rb") as src, open(path, "wb") as dst:
        shutil.copyfileobj(src, dst)
    info = idrs_amd.mm_info(path)
    rows, cols, entries, rep, field, symm = scipy.io.mminfo(path)
    assert (info["nrows"], info["ncols"], info["entries"]) == (rows, cols, entries)
    assert (info["rep"], info["field"], info["symm"]) == (rep, field, symm)


def _array_file(path, field, symm, vals, m):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix array {field} {symm}\n{m} {m}\n")
        for v in vals:
            f.write((f"{v.real:.17g} {v.imag:.17g}" if field == "complex" else f"{v:.17g}") + "\n")
    return path


@pytest.mark.parametrize("symm", ["general", "symmetric", "skew-symmetric", "hermitian"])
def test_array_expansion(symm, tmp_path):
    """mm_dense (mm_module.f90:152-310): column-major, the lower triangle
    for the symmetric kinds (strictly lower for skew), mirrored as the value,
    its negative or its conjugate -- the hermitian diagonal ends up
    conjugated too, as the reference writes (j,i) after (i,j)."""
    m = 5
    rng = np.random.default_rng(1)
    field = "complex" if symm == "hermitian" else "real"
    full = rng.standard_normal((m, m)) + (1j * rng.standard_normal((m, m)) if field == "complex" else 0)
    if field == "complex":
        full = (full.real.astype(np.float32) + 1j * full.imag.astype(np.float32)).astype(np.complex128)
    vals = []
    for j in range(m):
        lo = 0 if symm == "general" else (j + 1 if symm == "skew-symmetric" else j)
        vals += [full[i, j] for i in range(lo, m)]
    a = idrs_amd.mm_matrix(_array_file(str(tmp_path / "a.mtx"), field, symm, vals, m))
    ref = np.zeros((m, m), dtype=full.dtype)
    for j in range(m):
        for i in range(m):
            if symm == "general" or i > j or (i == j and symm != "skew-symmetric"):
                ref[i, j] = full[i, j]
    if symm == "symmetric":
        ref = np.tril(ref) + np.tril(ref, -1).T
    elif symm == "skew-symmetric":
        ref = ref - ref.T
    elif symm == "hermitian":
        ref = np.tril(ref, -1) + np.tril(ref, -1).conj().T + np.diag(np.diag(ref).conj())
    assert np.array_equal(a, ref)


def test_malformed_files(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("not a banner\n1 1 1\n")
    with pytest.raises(idrs_amd.IdrsError):
        idrs_amd.mm_info(str(p))
    p.write_text("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 1.0\n")
    with pytest.raises(idrs_amd.IdrsError):
        idrs_amd.mm_matrix(str(p))
    with pytest.raises(idrs_amd.IdrsError):
        idrs_amd.mm_info(str(tmp_path / "missing.mtx"))
