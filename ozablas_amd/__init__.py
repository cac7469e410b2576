"""ozablas_amd — FP64 GEMM emulated with INT8 matrix cores (Ozaki schemes I and II) on MI355X.

The compute path is the gfx950 library ``lib/libozablas_hip.so`` (hand-written HIP kernels behind
the C ABI of ``include/ozablas_hip.h``); the C++ drop-in API of the reference lives in
``lib/libozablas.so``.  This module binds the C ABI with ctypes for Python callers (tests, the
benchmark).  Importing it fails loudly when the library has not been built: there is no
fallback path.

Python names mirror the reference's C++ API (include/ozablas/ozablas.hpp:20,31,
include/ozablas/core/{executor,workspace}.hpp): ``HipExecutor``, ``WorkspaceScheme1``,
``WorkspaceScheme2``, ``OzaTimings``, ``ozaki_scheme1_gemm``, ``ozaki_scheme2_gemm``.
"""
from ._lib import (  # noqa: F401
    LIB_PATH,
    MAX_SLICES,
    DeviceBuffer,
    GroupScheme2,
    HipExecutor,
    OzaError,
    OzaTimings,
    ServerScheme2,
    WorkspaceScheme1,
    WorkspaceScheme2,
    device_count,
    COL_MAJOR,
    OP_N,
    OP_T,
    ROW_MAJOR,
    gemm,
    gemm_workspace_bytes,
    generate_matrix,
    group_autotune,
    group_plan,
    group_costs,
    lib,
    ozaki_scheme1_gemm,
    ozaki_scheme2_gemm,
    scheme1_beta,
    scheme2_kparam,
    scheme2_slices_for_accuracy,
    scheme2_slices_for_digits,
    workspace_bytes,
)
