"""GPU checks of the mixed-precision solve (single-precision inner IDR(s),
FP64 residual replacement).  The reference is FP64 only; the pins are the
FP64 quantities the method promises: the true residual b - A x computed on
the host in FP64 meets the tolerance, and x agrees with the FP64 solve of
the same system to the accuracy the tolerance implies.
"""
import numpy as np
import pytest

import idrs_amd
from idrs_amd import problems as P

from helpers import gpu_crs, rel

pytestmark = pytest.mark.gpu


def _true_relres(crs, x, b):
    return np.linalg.norm(b - crs.matvec(x)) / np.linalg.norm(b)


@pytest.mark.parametrize("fp32_vectors", [False, True])
@pytest.mark.parametrize("s,variant", [(4, 1), (2, 1), (1, 2)])
def test_mixed_reaches_fp64_tolerance(gpu, s, variant, fp32_vectors):
    crs, _ = P.convection_diffusion_3d(20, 10.0, reaction=1.0)
    b = crs.matvec(np.ones(crs.n))
    A = gpu_crs(crs)
    r = idrs_amd.idrs_mixed(A, b, s, tolerance=1e-10, maximum_iterations=5000, variant=variant,
                            fp32_vectors=fp32_vectors)
    f = idrs_amd.idrs(A, b, s, tolerance=1e-10, maximum_iterations=5000, variant=variant)
    print(f"s={s} v={variant}: mixed it={r.iterations} outer={r.outer} relres={r.relres:.2e}; fp64 it={f.iterations}")
    assert r.flag == 0 and r.relres <= 1e-10
    assert _true_relres(crs, r.x, b) <= 1.01e-10
    assert 1 <= r.outer <= 5
    assert rel(r.x, f.x) < 1e-8


@pytest.mark.parametrize("group", [None, 1])
@pytest.mark.parametrize("fp32_vectors", [False, True])
def test_mixed_chebyshev_operator(gpu, fp32_vectors, group):
    """The pp_idrs operator (diagonal scaling + Chebyshev(8)) on FP32 values
    inside; group 1: the FP32 copies of the sliced values."""
    crs, _ = P.convection_diffusion_3d(24, 6.2, reaction=1.0)
    b = np.random.default_rng(0).uniform(-1, 1, crs.n)
    A = gpu_crs(crs, diag_precon=True, poly=(2, 8, (0.15, 1.85)))
    if group:
        A.K.set_group(group)
    r = idrs_amd.idrs_mixed(A, b, 4, tolerance=1e-8, maximum_iterations=3000, inner_tolerance=1e-4,
                            fp32_vectors=fp32_vectors)
    f = idrs_amd.idrs(A, b, 4, tolerance=1e-8, maximum_iterations=3000)
    print(f"mixed it={r.iterations} outer={r.outer}; fp64 it={f.iterations}")
    assert r.flag == 0 and r.relres <= 1e-8
    # the operator's residual, recomputed through the library's FP64 product
    assert np.linalg.norm(b - A * r.x) <= 1.01e-8 * np.linalg.norm(b)
    assert r.iterations <= 3 * f.iterations


def test_mixed_complex_vectors_and_device(gpu):
    import torch

    crs, _ = P.helmholtz_3d(12)
    sigma = -0.5
    shifted = P.Crs(crs.n, crs.row_ptr, crs.col_ind,
                    np.where(np.repeat(np.arange(crs.n), np.diff(crs.row_ptr)) == crs.col_ind - 1,
                             crs.val - sigma, crs.val))
    rng = np.random.default_rng(3)
    b = rng.standard_normal(crs.n) + 1j * rng.standard_normal(crs.n)
    A = gpu_crs(shifted)
    r = idrs_amd.idrs_mixed(A, b, 4, tolerance=1e-11, maximum_iterations=4000)
    assert r.flag == 0 and _true_relres(shifted, r.x, b) <= 1.01e-11
    d = idrs_amd.idrs_mixed(A, torch.from_numpy(b).cuda(), 4, tolerance=1e-11, maximum_iterations=4000)
    assert d.iterations == r.iterations and rel(d.x.cpu().numpy(), r.x) < 1e-14


def test_mixed_rejects(gpu):
    crs, _ = P.convection_diffusion_3d(4, 1.0)
    A = gpu_crs(crs)
    with pytest.raises(idrs_amd.IdrsError):
        idrs_amd.idrs_mixed(A, np.ones(crs.n), 4, inner_tolerance=2.0)
    Ac = idrs_amd.Matrix(idrs_amd.crs_format(crs.n, crs.n, crs.row_ptr, crs.col_ind, cval=crs.val + 0j))
    with pytest.raises(idrs_amd.IdrsError):
        idrs_amd.idrs_mixed(Ac, np.ones(crs.n) + 0j, 4)
    # FP32 values with FP64 vectors: real systems only
    with pytest.raises(idrs_amd.IdrsError):
        idrs_amd.idrs_mixed(A, np.ones(crs.n) + 0j, 4, fp32_vectors=False)
