"""MI355X-native IDR(s) solver hot path (HIP kernels behind a C ABI).

The public API mirrors the reference IDRS package (user_module converters,
type(matrix) with operator(*) and operator(/), polynomial preconditioners,
the IDRS solver); see api.py.
"""
from ._lib import IdrsError, load as load_library  # noqa: F401
from .api import (  # noqa: F401
    BatchMixedResult,
    BatchResult,
    IdrsResult,
    KrylovResult,
    MixedResult,
    Matrix,
    UserMatrix,
    batch_format,
    batch_info,
    cds_format,
    complex_diagonal,
    complex_preconditioner,
    coo_format,
    crs_format,
    dia_format,
    idrs,
    idrs_batched,
    idrs_batched_mixed,
    idrs_mixed,
    make_p,
    mm_info,
    mm_matrix,
    ms_parameters,
    ms_scaleback,
    msidrs,
    msqmridr,
    qmridr,
    real_diagonal,
    real_preconditioner,
    scaleback,
)
from .autograd import SparseSystem, idrs_solve  # noqa: F401,E402
