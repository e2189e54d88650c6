"""ctypes binding of libtca_spmm.so (the C ABI declared in include/tca_spmm.h).

The library is loaded from this package directory only; if it is missing the
import fails loudly -- there is no fallback computation anywhere in the package.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtca_spmm.so")

F16, BF16, E4M3, F32 = 0, 1, 2, 3
DTYPE_NAMES = {F16: "f16", BF16: "bf16", E4M3: "e4m3", F32: "f32"}
PATTERN_BERNOULLI, PATTERN_LOGNORMAL, PATTERN_UNIFORM = 0, 1, 2
OP_SPMM, OP_SDDMM = 0, 1

STATUS = {0: "TCA_OK", 1: "TCA_ERR_INVALID_ARG", 2: "TCA_ERR_UNSUPPORTED", 3: "TCA_ERR_HIP",
          4: "TCA_ERR_ALLOC", 5: "TCA_ERR_IO", 6: "TCA_ERR_FORMAT", 7: "TCA_ERR_LIBRARY"}


class TcaError(RuntimeError):
    """A C ABI call returned a non-OK tca_status."""

    def __init__(self, status, where, detail):
        self.status = status
        self.name = STATUS.get(status, str(status))
        super().__init__(f"{where}: {self.name}: {detail}")


class PlanInfo(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_int), ("n_cols", ctypes.c_int), ("nnz", ctypes.c_int),
                ("rows_per_wave", ctypes.c_int), ("n_groups", ctypes.c_int), ("n_chunks", ctypes.c_int),
                ("max_group_chunks", ctypes.c_int), ("transposed", ctypes.c_int),
                ("device_bytes", ctypes.c_size_t), ("build_ms", ctypes.c_double), ("dense_elems", ctypes.c_size_t)]


class BatchItem(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int), ("plan", ctypes.c_void_p), ("tile_n", ctypes.c_int),
                ("a_dtype", ctypes.c_int), ("a", ctypes.c_void_p), ("lda", ctypes.c_int),
                ("b_dtype", ctypes.c_int), ("b", ctypes.c_void_p), ("ldb", ctypes.c_int),
                ("N", ctypes.c_int), ("out", ctypes.c_void_p), ("ldc", ctypes.c_int), ("a_dense", ctypes.c_int),
                ("values_order", ctypes.c_void_p), ("sgd_w", ctypes.c_void_p), ("sgd_w_dtype", ctypes.c_int),
                ("sgd_lr", ctypes.c_float)]


_vp = ctypes.c_void_p
_i = ctypes.c_int
_i32p = ctypes.POINTER(ctypes.c_int32)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_f32p = ctypes.POINTER(ctypes.c_float)
_u64 = ctypes.c_uint64
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); status-returning functions use c_int
SIGNATURES = {
    "tca_status_string": (ctypes.c_char_p, [_i]),
    "tca_last_error": (ctypes.c_char_p, []),
    "tca_version": (_i, []),
    "tca_batch_item_size": (ctypes.c_size_t, []),
    "tca_host_free": (None, [_vp]),
    "tca_smtx_read": (_i, [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                           ctypes.POINTER(_i32p), ctypes.POINTER(_i32p)]),
    "tca_smtx_write": (_i, [ctypes.c_char_p, _i, _i, _i, _i32p, _i32p]),
    "tca_mtx_read": (_i, [ctypes.c_char_p, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                          ctypes.POINTER(_i32p), ctypes.POINTER(_i32p), ctypes.POINTER(_u16p)]),
    "tca_dense_to_csr_f16": (_i, [_u16p, _i, _i, _i32p, _i32p, _u16p, _i, ctypes.POINTER(_i)]),
    "tca_csr_to_dense_f16": (_i, [_i, _i, _i32p, _i32p, _u16p, _u16p]),
    "tca_compare_mat_f32": (_i, [_f32p, _f32p, _i, ctypes.c_float, ctypes.POINTER(_i)]),
    "tca_rng_u64": (_u64, [_u64, _u64, _u64]),
    "tca_gen_pattern": (_i, [_i, _i, _i, ctypes.c_double, ctypes.c_double, _u64, _u64, _i32p,
                             ctypes.POINTER(_i32p), ctypes.POINTER(_i)]),
    "tca_gen_normal_f32": (_i, [_u64, _u64, ctypes.c_size_t, ctypes.c_double, ctypes.c_double, _f32p]),
    "tca_gen_uniform_f32": (_i, [_u64, _u64, ctypes.c_size_t, ctypes.c_double, ctypes.c_double, _f32p]),
    "tca_f32_to_f16_host": (_i, [_f32p, ctypes.c_size_t, _u16p]),
    "tca_f32_to_bf16_host": (_i, [_f32p, ctypes.c_size_t, _u16p]),
    "tca_plan_create": (_i, [_i, _i, _i, _i32p, _i32p, _i, ctypes.POINTER(_vp)]),
    "tca_plan_create_host": (_i, [_i, _i, _i, _i32p, _i32p, _i, ctypes.POINTER(_vp)]),
    "tca_plan_create_transpose": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "tca_plan_destroy": (_i, [_vp]),
    "tca_csr_transpose_device": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tca_csr_transpose_device_noalloc": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "tca_csr_transpose_noalloc_max_scratch": (ctypes.c_size_t, []),
    "tca_plan_get_info": (_i, [_vp, ctypes.POINTER(PlanInfo)]),
    "tca_plan_export": (_i, [_vp, _i32p, _i32p, _i32p, _u16p, _i32p, _i32p]),
    "tca_pack_values": (_i, [_vp, _i, _vp, _vp, _vp]),
    "tca_unpack_values": (_i, [_vp, _i, _vp, _vp, _vp]),
    "tca_spmm": (_i, [_vp, _i, _vp, _i, _vp, _i, _i, _vp, _i, _vp, _vp, _i, _vp]),
    "tca_spmm_dense": (_i, [_vp, _i, _vp, _i, _vp, _i, _i, _vp, _i, _vp, _vp, _i, _vp]),
    "tca_pack_dense": (_i, [_vp, _i, _vp, _vp, _vp]),
    "tca_sddmm": (_i, [_vp, _i, _vp, _i, _i, _vp, _i, _i, _vp, _vp]),
    "tca_bell_from_csr": (_i, [_i, _i, _i32p, _i32p, _u16p, _i, ctypes.POINTER(_i), _i32p, _u16p]),
    "tca_bell_spmm": (_i, [_i, _i, _i, _i, _vp, _i, _vp, _i, _vp, _i, _i, _vp, _i, _i, _i, _vp]),
    "tca_bell_shape": (_i, [_i, _i, _i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "tca_compress_24": (_i, [_i, _i, _i, _vp, _i, _vp, _vp, _vp]),
    "tca_spmm_24": (_i, [_i, _i, _i, _i, _vp, _vp, _i, _vp, _i, _vp, _i, _i, _i, _vp]),
    "tca_quantize_rows_e4m3": (_i, [_i, _vp, _vp, _vp, _vp, _vp]),
    "tca_quantize_cols_e4m3": (_i, [_i, _i, _vp, _i, _vp, _i, _vp, _vp]),
    "tca_batch_create": (_i, [_i, ctypes.POINTER(BatchItem), ctypes.POINTER(_vp)]),
    "tca_sgd_update": (_i, [_i, _vp, _vp, ctypes.c_size_t, ctypes.c_float, _vp]),
    "tca_batch_run": (_i, [_vp, _vp]),
    "tca_batch_run_stage": (_i, [_vp, _i, _vp]),
    "tca_batch_destroy": (_i, [_vp]),
    "tca_batch_get_info": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "tca_rocsparse_spmm": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _dp]),
    "tca_hipblaslt_gemm": (_i, [_i, _i, _i, _vp, _vp, _vp, _i, _dp]),
    "tca_rocsparse_sddmm": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _dp]),
    "tca_rocsparse_spmm_t": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _dp]),
    "tca_hipsparselt_spmm_24": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _i, _dp]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libtca_spmm.so not found at {LIB_PATH}: build it first "
                              "(python -c 'import __graft_entry__ as g; g.build()' or make -C tca-spmm_amd)")
        # One HIP runtime per process: torch bundles libamdhip64 (soname
        # libamdhip64.so.7); loading torch first makes this library bind to that
        # copy instead of pulling in a second runtime from /opt/rocm.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.tca_batch_item_size() != ctypes.sizeof(BatchItem):
            raise ImportError(f"{LIB_PATH}: tca_batch_item is {L.tca_batch_item_size()} bytes in the library, "
                              f"{ctypes.sizeof(BatchItem)} in this binding (stale build? rebuild the library)")
        _lib = L
    return _lib


def check(status, where):
    if status != 0:
        detail = lib().tca_last_error().decode(errors="replace")
        raise TcaError(status, where, detail)


def call(name, *args):
    check(getattr(lib(), name)(*args), name)
