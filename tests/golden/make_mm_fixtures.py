"""Writes the Matrix Market fixtures of the tests from the reference's own
data files (EXAMPLES/MATRIX_MARKET/DATA under the reference checkout, read
as text here; the files are data, not code).

Each fixture is a compressed .npz holding the file's header words, its size
line and its entries in file order, lossless (values as parsed FP64; the
tests write them back as Matrix Market text with 17 significant digits,
which parses to the same doubles).  Two small files are also kept verbatim
(gzipped) to test the reader on untouched reference text.

Run in a checkout that has the reference next to it:
    python tests/golden/make_mm_fixtures.py /path/to/reference
"""
import gzip
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "mm")

FILES = [
    "OCEAN_DATA/STOMMEL/stommel4.mtx", "OCEAN_DATA/STOMMEL/stommel4_b.mtx",
    "OCEAN_DATA/STOMMEL/stommel5.mtx", "OCEAN_DATA/STOMMEL/stommel5_b.mtx",
    "OCEAN_DATA/STOMMEL/stommel6.mtx", "OCEAN_DATA/STOMMEL/stommel6_b.mtx",
    "OCEAN_DATA/SAG/sag5.mtx", "OCEAN_DATA/SAG/sag5_b.mtx",
    "OCEAN_DATA/SAG/sag6.mtx", "OCEAN_DATA/SAG/sag6_b.mtx",
] + [f"WEDGE_DATA/wedge{k}_{m}.mtx" for k in (1, 2, 3, 4) for m in ("K", "M", "C", "b")]
VERBATIM = ["WEDGE_DATA/wedge1_K.mtx", "WEDGE_DATA/wedge2_M.mtx"]


def parse(path):
    with open(path) as f:
        header = f.readline().split()
        line = f.readline()
        while line.lstrip().startswith("%") or not line.strip():
            line = f.readline()
        size = [int(v) for v in line.split()]
        rest = f.read().split()
    rep, field = header[2].lower(), header[3].lower()
    w = {"real": 1, "integer": 1, "complex": 2, "pattern": 0}[field]
    if rep == "coordinate":
        k = 2 + w
        a = np.array(rest, dtype=object).reshape(size[2], k)
        ij = a[:, :2].astype(np.int64).astype(np.int32)
        vals = a[:, 2:].astype(np.float64)
    else:
        ij = np.zeros((0, 2), dtype=np.int32)
        vals = np.array(rest, dtype=np.float64).reshape(-1, w)
    return dict(header=np.array(header[1:]), size=np.array(size, dtype=np.int64), ij=ij, vals=vals)


def main(ref):
    data = os.path.join(ref, "IDRS", "EXAMPLES", "MATRIX_MARKET", "DATA")
    os.makedirs(OUT, exist_ok=True)
    for rel in FILES:
        d = parse(os.path.join(data, rel))
        name = os.path.basename(rel)[:-4]
        np.savez_compressed(os.path.join(OUT, name + ".npz"), **d)
        print(name, d["header"], d["size"])
    for rel in VERBATIM:
        with open(os.path.join(data, rel), "rb") as src, gzip.open(os.path.join(OUT, os.path.basename(rel) + ".gz"), "wb") as dst:
            shutil.copyfileobj(src, dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
