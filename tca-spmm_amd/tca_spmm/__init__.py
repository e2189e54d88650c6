"""tca_spmm -- MI355X (gfx950) row-balanced CSR x dense SpMM.

Python mirror of the reference's host operator (mlsys-lab-sogang/TCA-SpMM,
src/our_spmm.cu our_spmm_balanced) over the C ABI of libtca_spmm.so.  Every
computation runs in the HIP library; a missing library raises ImportError.
"""
from ._lib import BF16, E4M3, F16, F32, TcaError, lib  # noqa: F401
from . import autotune, bell, gen, workloads  # noqa: F401
from .bell import bell_from_csr, bell_shape  # noqa: F401
from .sparse24 import hipsparselt_spmm_24  # noqa: F401

lib()  # fail loudly at import when the native library is absent

try:  # torch-facing API (torch is the device-memory layer, not the compute)
    from .core import (Plan, PlanCache, csr_transpose_device, hipblaslt_gemm, rocsparse_sddmm, rocsparse_spmm_t,  # noqa: F401
                       quantize_cols_e4m3, quantize_rows_e4m3, rocsparse_spmm, sddmm, sgd_update, spmm)
    from .autograd import SparseLinear, SpmmFunction, sparse_matmul  # noqa: F401
    from .batch import Batch, sddmm_item, spmm_item  # noqa: F401
    from .bell import BlockedEll, bell_spmm  # noqa: F401
    from .sparse24 import Sparse24, spmm_24  # noqa: F401
except ImportError:  # pragma: no cover - torch absent
    pass

__all__ = ["Plan", "PlanCache", "csr_transpose_device", "spmm", "sddmm", "sgd_update", "quantize_rows_e4m3", "quantize_cols_e4m3", "rocsparse_spmm",
           "hipblaslt_gemm", "rocsparse_sddmm", "rocsparse_spmm_t", "hipsparselt_spmm_24", "SparseLinear", "sparse_matmul", "Batch", "spmm_item", "sddmm_item", "gen", "workloads", "bell", "bell_from_csr", "bell_shape", "BlockedEll", "bell_spmm", "Sparse24", "spmm_24", "TcaError", "F16", "BF16", "E4M3", "F32"]
