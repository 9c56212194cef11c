"""The Fortran drop-in (user_module / matrix_module / idrs_module over
libidrs_hip.so) runs reference-style programs (tests/fortran/drop_in.f90)
on the GPU and matches the oracle run on the same inputs."""
import os
import re
import subprocess

import numpy as np
import pytest

import orc
from idrs_amd import problems as P

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "idrs-package_amd", "fortran", "bin", "drop_in")


@pytest.fixture(scope="module")
def cases(gpu):
    assert os.path.exists(BIN), "Fortran drop-in program not built (amdflang)"
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stderr
    res = {}
    for line in out.stdout.splitlines():
        m = re.match(r"CASE (\S+) iter\s+(\d+) flag\s+(\d+) relres\s+(\S+) err\s+(\S+)", line)
        if m:
            res[m.group(1)] = dict(iter=int(m.group(2)), flag=int(m.group(3)), relres=float(m.group(4)),
                                   err=float(m.group(5)))
    return res


@pytest.mark.parametrize("s", [1, 2, 4, 8])
def test_toeplitz_finite_termination(cases, s):
    c = cases[f"toeplitz_s{s}"]
    crs, _, b = P.toeplitz_1d()
    Ao = orc.OracleMatrix(orc.CRS, 60, crs=crs)
    _, fo, _, io, _ = orc.idrs(Ao, b, s, M1=Ao, tolerance=1e-8, maximum_iterations=4000)
    assert c["flag"] == fo == 0 and abs(c["iter"] - io) <= 2
    assert c["iter"] <= 60 + 60 // s and c["err"] < 1e-6


def test_bicgstab(cases):
    assert cases["toeplitz_bicgstab"]["flag"] == 0


def test_cdr_pp_idrs_sequence(cases):
    """crs_format -> real_precon(1) -> real_preconditioner(Chebyshev 8) ->
    idrs(IDR(4)) -> scaleback, as PP_IDRS does."""
    c = cases["cdr_cheb8"]
    crs, _ = P.convection_diffusion_3d(16, 5.0, reaction=1.0)
    b = crs.matvec(np.ones(crs.n))
    a = orc.chebyshev(8, np.array([0.15, 1.85]))
    Ao = orc.OracleMatrix(orc.CRS, crs.n, crs=crs, which_precon=1, degree=8, alpha=a)
    _, fo, _, io, _ = orc.idrs(Ao, b, 4, M1=orc.OracleMatrix(orc.CRS, crs.n, crs=crs), tolerance=1e-8,
                               maximum_iterations=2000)
    assert c["flag"] == fo == 0 and abs(c["iter"] - io) <= 2
    assert c["relres"] < 1e-7 and c["err"] < 1e-6


def test_complex_toeplitz(cases):
    c = cases["toeplitz_complex_s4"]
    assert c["flag"] == 0 and c["err"] <= 1e-12


def test_qmridr(cases):
    c = cases["qmridr_s4"]
    crs, _, b = P.toeplitz_1d()
    Ao = orc.OracleMatrix(orc.CRS, 60, crs=crs)
    _, fo, _, io, _ = orc.qmridr(Ao, b, 4, M1=Ao, tolerance=1e-8, maximum_iterations=4000)
    assert c["flag"] == fo == 0 and abs(c["iter"] - io) <= 2 and c["err"] < 1e-6


@pytest.mark.parametrize("solver", ["msidrs", "msqmridr"])
def test_multishift(cases, solver):
    crs, _, b = P.toeplitz_1d()
    Ao = orc.OracleMatrix(orc.CRS, 60, crs=crs)
    sig = np.array([0.0, -0.5, -2.0])
    ref = getattr(orc, solver)(Ao, b, sig, 4, tolerance=1e-8, maximum_iterations=4000)
    for k in range(3):
        c = cases[f"{solver}_shift{k + 1}"]
        assert c["flag"] == ref["flag"] == 0 and abs(c["iter"] - ref["iterations"]) <= 2
        assert c["err"] < 1e-7 and c["relres"] < 1e-7


def test_multishift_pp_sequence(cases):
    """ms_parameters -> msidrs on the preconditioned shifted systems ->
    scaleback (PP_IDRS/run_ms_idrs.inc): each shift's true residual of
    (K - sigma I) x = b, recomputed from the scaled-back solution."""
    for k in range(3):
        c = cases[f"ms_pp_shift{k + 1}"]
        assert c["flag"] == 0 and c["err"] < 1e-7
