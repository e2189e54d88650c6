"""Seeded inputs, .smtx I/O and CSR helpers (numpy in, numpy out), all through
the C ABI so that C++ callers and Python see the same bits.

Streams are counter-based SplitMix64 sequences: (seed, stream, index) fixes
every value, independent of how many values are drawn.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import PATTERN_BERNOULLI, PATTERN_LOGNORMAL, PATTERN_UNIFORM, call

_i32p = ctypes.POINTER(ctypes.c_int32)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_f32p = ctypes.POINTER(ctypes.c_float)


def rng_u64(seed, stream, index):
    return _lib.lib().tca_rng_u64(seed, stream, index)


def normal_f32(seed, stream, shape, mean=0.0, std=1.0):
    out = np.empty(shape, np.float32)
    call("tca_gen_normal_f32", seed, stream, out.size, mean, std, out.ctypes.data_as(_f32p))
    return out


def uniform_f32(seed, stream, shape, lo=0.0, hi=1.0):
    out = np.empty(shape, np.float32)
    call("tca_gen_uniform_f32", seed, stream, out.size, lo, hi, out.ctypes.data_as(_f32p))
    return out


def f32_to_f16_bits(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.shape, np.uint16)
    call("tca_f32_to_f16_host", x.ctypes.data_as(_f32p), x.size, out.ctypes.data_as(_u16p))
    return out


def f32_to_bf16_bits(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.shape, np.uint16)
    call("tca_f32_to_bf16_host", x.ctypes.data_as(_f32p), x.size, out.ctypes.data_as(_u16p))
    return out


def pattern(kind, n_rows, n_cols, density, param=0.0, seed=0, stream=0):
    """CSR pattern (row_offset int32[n_rows+1], col_idx int32[nnz]) with sorted,
    distinct columns.  kind: 'bernoulli' | 'lognormal' (param = sigma) |
    'uniform' (row length U[(1-param), (1+param)] * density * n_cols)."""
    k = {"bernoulli": PATTERN_BERNOULLI, "lognormal": PATTERN_LOGNORMAL, "uniform": PATTERN_UNIFORM}[kind]
    rp = np.zeros(n_rows + 1, np.int32)
    ci = _i32p()
    nnz = ctypes.c_int()
    call("tca_gen_pattern", k, n_rows, n_cols, density, param, seed, stream, rp.ctypes.data_as(_i32p),
         ctypes.byref(ci), ctypes.byref(nnz))
    cols = np.ctypeslib.as_array(ci, shape=(max(nnz.value, 1),))[:nnz.value].copy()
    _lib.lib().tca_host_free(ci)
    return rp, cols


def smtx_read(path):
    """(n_rows, n_cols, row_offset, col_idx) of a .smtx file."""
    M, K, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rp, ci = _i32p(), _i32p()
    call("tca_smtx_read", path.encode(), ctypes.byref(M), ctypes.byref(K), ctypes.byref(nnz), ctypes.byref(rp),
         ctypes.byref(ci))
    row_offset = np.ctypeslib.as_array(rp, shape=(M.value + 1,)).copy()
    col_idx = np.ctypeslib.as_array(ci, shape=(max(nnz.value, 1),))[:nnz.value].copy()
    _lib.lib().tca_host_free(rp)
    _lib.lib().tca_host_free(ci)
    return M.value, K.value, row_offset, col_idx


def mtx_read(path, one_based=True, pattern_only=False):
    """MatrixMarket coordinate file -> (n_rows, n_cols, row_offset, col_idx, fp16 value bits)."""
    M, K, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rp, ci, vv = _i32p(), _i32p(), _u16p()
    call("tca_mtx_read", path.encode(), int(one_based), int(pattern_only), ctypes.byref(M), ctypes.byref(K),
         ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(vv))
    n = max(nnz.value, 1)
    out = (M.value, K.value, np.ctypeslib.as_array(rp, shape=(M.value + 1,)).copy(),
           np.ctypeslib.as_array(ci, shape=(n,))[:nnz.value].copy(), np.ctypeslib.as_array(vv, shape=(n,))[:nnz.value].copy())
    for ptr in (rp, ci, vv):
        _lib.lib().tca_host_free(ptr)
    return out


def smtx_write(path, n_rows, n_cols, row_offset, col_idx):
    rp = np.ascontiguousarray(row_offset, np.int32)
    ci = np.ascontiguousarray(col_idx, np.int32)
    if ci.size == 0:
        ci = np.zeros(1, np.int32)
    call("tca_smtx_write", path.encode(), n_rows, n_cols, int(rp[-1]), rp.ctypes.data_as(_i32p), ci.ctypes.data_as(_i32p))


def dense_to_csr_f16(dense_bits):
    d = np.ascontiguousarray(dense_bits, np.uint16)
    M, K = d.shape
    nnz = ctypes.c_int()
    call("tca_dense_to_csr_f16", d.ctypes.data_as(_u16p), M, K, None, None, None, 0, ctypes.byref(nnz))
    rp = np.zeros(M + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    v = np.zeros(max(nnz.value, 1), np.uint16)
    call("tca_dense_to_csr_f16", d.ctypes.data_as(_u16p), M, K, rp.ctypes.data_as(_i32p), ci.ctypes.data_as(_i32p),
         v.ctypes.data_as(_u16p), ci.size, ctypes.byref(nnz))
    return rp, ci[:nnz.value], v[:nnz.value]


def csr_to_dense_f16(n_rows, n_cols, row_offset, col_idx, values):
    out = np.zeros((n_rows, n_cols), np.uint16)
    rp = np.ascontiguousarray(row_offset, np.int32)
    ci = np.ascontiguousarray(col_idx, np.int32)
    v = np.ascontiguousarray(values, np.uint16)
    if ci.size == 0:
        ci, v = np.zeros(1, np.int32), np.zeros(1, np.uint16)
    call("tca_csr_to_dense_f16", n_rows, n_cols, rp.ctypes.data_as(_i32p), ci.ctypes.data_as(_i32p),
         v.ctypes.data_as(_u16p), out.ctypes.data_as(_u16p))
    return out


def compare_mat(gold, cmp, tol=1e-3):
    """compare_mat<float> (src/general.cuh:71-123): (pass, failures seen)."""
    g = np.ascontiguousarray(gold, np.float32).ravel()
    c = np.ascontiguousarray(cmp, np.float32).ravel()
    nf = ctypes.c_int()
    ok = _lib.lib().tca_compare_mat_f32(g.ctypes.data_as(_f32p), c.ctypes.data_as(_f32p), g.size, tol, ctypes.byref(nf))
    return bool(ok), nf.value
