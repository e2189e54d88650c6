"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker, never as the computation
being measured or shipped.  The C file oracle/oracle.c restates the reference's
host code; each function there cites the reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int)
_u16p = ctypes.POINTER(ctypes.c_uint16)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_f16_to_f32.restype = ctypes.c_float
        L.orc_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f32_to_f16.argtypes = [ctypes.c_float]
        L.orc_bf16_to_f32.restype = ctypes.c_float
        L.orc_bf16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_f32_to_bf16.restype = ctypes.c_uint16
        L.orc_f32_to_bf16.argtypes = [ctypes.c_float]
        L.orc_e4m3_to_f32.restype = ctypes.c_float
        L.orc_e4m3_to_f32.argtypes = [ctypes.c_uint8]
        L.orc_f32_to_e4m3.restype = ctypes.c_uint8
        L.orc_f32_to_e4m3.argtypes = [ctypes.c_float]
        L.orc_smtx_to_dense.restype = ctypes.c_int
        L.orc_smtx_to_dense.argtypes = [ctypes.c_char_p, _i32p, _i32p, _i32p,
                                        ctypes.POINTER(ctypes.POINTER(ctypes.c_uint16))]
        L.orc_mtx_to_dense.restype = ctypes.c_int
        L.orc_mtx_to_dense.argtypes = [ctypes.c_char_p, ctypes.c_int, _i32p, _i32p, _i32p,
                                       ctypes.POINTER(ctypes.POINTER(ctypes.c_uint16))]
        L.orc_dense_to_csr.restype = ctypes.c_int
        L.orc_dense_to_csr.argtypes = [_u16p, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _u16p]
        L.orc_csr_to_dense.restype = None
        L.orc_csr_to_dense.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _u16p, _u16p]
        L.orc_compare_mat.restype = ctypes.c_int
        L.orc_compare_mat.argtypes = [_f32p, _f32p, ctypes.c_int, ctypes.c_float, _i32p]
        L.orc_spmm_f32.restype = None
        L.orc_spmm_f32.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _u16p, _u16p,
                                   ctypes.c_int, _f32p, ctypes.c_int, ctypes.c_int]
        L.orc_spmm_f64.restype = None
        L.orc_spmm_f64.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, _f64p,
                                   ctypes.c_int, _f64p, ctypes.c_int, ctypes.c_int]
        L.orc_sddmm_f64.restype = None
        L.orc_sddmm_f64.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _f64p, ctypes.c_int,
                                    _f64p, ctypes.c_int, _f64p, ctypes.c_int]
        L.orc_sddmm_f32.restype = None
        L.orc_sddmm_f32.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _u16p, ctypes.c_int,
                                    _u16p, ctypes.c_int, _f32p, ctypes.c_int]
        L.orc_workload_manager.restype = ctypes.c_int
        L.orc_workload_manager.argtypes = [ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int, _i32p,
                                           _i32p, _i32p, _i32p, ctypes.c_int]
        L.orc_tiling_sizes.restype = ctypes.c_int
        L.orc_tiling_sizes.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p, _i32p]
        L.orc_tiling_fill.restype = ctypes.c_int
        L.orc_tiling_fill.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_int,
                                      _i32p, _i32p, _i32p, _u16p, _i32p, _i32p]
        L.orc_transpose_csr.restype = None
        L.orc_transpose_csr.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.orc_free.restype = None
        L.orc_dense_to_bell.argtypes = [_u16p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p, _u16p]
        L.orc_dense_to_bell.restype = ctypes.c_int
        L.orc_bell_spmm_f32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _i32p,
                                        _u16p, _u16p, _f32p]
        L.orc_prune_24_strip.argtypes = [_u16p, ctypes.c_int, ctypes.c_int, _u16p]
        L.orc_prune_24_strip.restype = ctypes.c_int
        L.orc_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


# ---- scalar formats (vectorised over numpy arrays) ------------------------
def f16_bits_to_f32(bits):
    bits = np.ascontiguousarray(bits, dtype=np.uint16)
    return np.array([lib().orc_f16_to_f32(int(b)) for b in bits.ravel()], dtype=np.float32).reshape(bits.shape)


def e4m3_to_f32(codes):
    """Decode OCP E4M3FN codes (vectorised through a 256-entry table)."""
    table = np.array([lib().orc_e4m3_to_f32(i) for i in range(256)], dtype=np.float32)
    return table[np.asarray(codes, dtype=np.uint8)]


def f32_to_e4m3(x):
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float32)
    return np.array([L.orc_f32_to_e4m3(float(v)) for v in x.ravel()], dtype=np.uint8).reshape(x.shape)


def f32_to_e4m3_fast(x):
    """Same rounding as orc_f32_to_e4m3 via a sorted table (for large arrays);
    checked equal to the scalar oracle encode in the oracle tests."""
    table = e4m3_to_f32(np.arange(0x7f, dtype=np.uint8)).astype(np.float64)  # codes 0..0x7e ascending
    x = np.asarray(x, dtype=np.float32)
    a = np.abs(x).astype(np.float64)
    lo = np.clip(np.searchsorted(table, a, side="right") - 1, 0, 0x7e)
    hi = np.minimum(lo + 1, 0x7e)
    d0 = a - table[lo]
    d1 = table[hi] - a
    up = (hi != lo) & ((d1 < d0) | ((d1 == d0) & (lo % 2 == 1)))
    code = np.where(up, hi, lo)
    code = np.where(a >= 448.0, 0x7e, code).astype(np.uint8)
    code = code | (np.signbit(x).astype(np.uint8) << 7)
    code = np.where(np.isnan(x), 0x7f, code)
    return code.astype(np.uint8)


def bf16_bits_to_f32(bits):
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def f32_to_bf16_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16).astype(np.uint16)
    nan = np.isnan(x)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


# ---- .smtx reader / CSR ----------------------------------------------------
def smtx_to_dense(path):
    L = lib()
    M, K, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ptr = ctypes.POINTER(ctypes.c_uint16)()
    rc = L.orc_smtx_to_dense(path.encode(), ctypes.byref(M), ctypes.byref(K), ctypes.byref(nnz), ctypes.byref(ptr))
    if rc != 0:
        return rc, None
    arr = np.ctypeslib.as_array(ptr, shape=(M.value * K.value,)).copy().reshape(M.value, K.value)
    L.orc_free(ptr)
    return 0, (arr, nnz.value)


def mtx_to_dense(path, one_based=True):
    L = lib()
    M, K, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    ptr = ctypes.POINTER(ctypes.c_uint16)()
    rc = L.orc_mtx_to_dense(path.encode(), int(one_based), ctypes.byref(M), ctypes.byref(K), ctypes.byref(nnz),
                            ctypes.byref(ptr))
    if rc != 0:
        return rc, None
    arr = np.ctypeslib.as_array(ptr, shape=(M.value * K.value,)).copy().reshape(M.value, K.value)
    L.orc_free(ptr)
    return 0, (arr, nnz.value)


def dense_to_csr(dense_bits):
    d = np.ascontiguousarray(dense_bits, dtype=np.uint16)
    M, K = d.shape
    L = lib()
    nnz = L.orc_dense_to_csr(_p(d, _u16p), M, K, None, None, None)
    rowptr = np.zeros(M + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.uint16)
    L.orc_dense_to_csr(_p(d, _u16p), M, K, _p(rowptr, _i32p), _p(col, _i32p), _p(val, _u16p))
    return rowptr, col[:nnz], val[:nnz]


def csr_to_dense(M, K, rowptr, col, val):
    out = np.zeros((M, K), np.uint16)
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.uint16)
    lib().orc_csr_to_dense(M, K, _p(rowptr, _i32p), _p(col, _i32p), _p(val, _u16p), _p(out, _u16p))
    return out


def compare_mat(gold, cmp, tol=1e-3):
    g = np.ascontiguousarray(gold, np.float32).ravel()
    c = np.ascontiguousarray(cmp, np.float32).ravel()
    nf = ctypes.c_int()
    ok = lib().orc_compare_mat(_p(g, _f32p), _p(c, _f32p), g.size, tol, ctypes.byref(nf))
    return bool(ok), nf.value


# ---- arithmetic ------------------------------------------------------------
def spmm_f32(M, N, rowptr, col, val_bits, B_bits, nthreads=0):
    """FP32-accumulate CPU SpMM on fp16 bits, ascending column order."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val_bits, np.uint16)
    B = np.ascontiguousarray(B_bits, np.uint16)
    C = np.empty((M, N), np.float32)
    if col.size == 0:
        col = np.zeros(1, np.int32)
        val = np.zeros(1, np.uint16)
    lib().orc_spmm_f32(M, N, _p(rowptr, _i32p), _p(col, _i32p), _p(val, _u16p), _p(B, _u16p),
                       B.shape[1], _p(C, _f32p), N, nthreads)
    return C


def spmm_f64(M, N, rowptr, col, vals, B, nthreads=0):
    """FP64 CSR oracle on exactly-decoded float values."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    v = np.ascontiguousarray(vals, np.float64)
    Bd = np.ascontiguousarray(B, np.float64)
    C = np.empty((M, N), np.float64)
    if col.size == 0:
        col = np.zeros(1, np.int32)
        v = np.zeros(1, np.float64)
    lib().orc_spmm_f64(M, N, _p(rowptr, _i32p), _p(col, _i32p), _p(v, _f64p), _p(Bd, _f64p),
                       Bd.shape[1], _p(C, _f64p), N, nthreads)
    return C


def sddmm_f64(M, N, rowptr, col, G, X, nthreads=0):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    Gd = np.ascontiguousarray(G, np.float64)
    Xd = np.ascontiguousarray(X, np.float64)
    out = np.empty(max(col.size, 1), np.float64)
    lib().orc_sddmm_f64(M, N, _p(rowptr, _i32p), _p(col, _i32p), _p(Gd, _f64p), Gd.shape[1],
                        _p(Xd, _f64p), Xd.shape[1], _p(out, _f64p), nthreads)
    return out[:col.size]


def sddmm_f32(M, N, rowptr, col, G_bits, X_bits, nthreads=0):
    """FP32-accumulate CPU SDDMM on fp16 bits (ascending n): out[e] = G[r] . X[col(e)]."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    Gb = np.ascontiguousarray(G_bits, np.uint16)
    Xb = np.ascontiguousarray(X_bits, np.uint16)
    out = np.empty(max(col.size, 1), np.float32)
    if col.size == 0:
        col = np.zeros(1, np.int32)
    lib().orc_sddmm_f32(M, N, _p(rowptr, _i32p), _p(col, _i32p), _p(Gb, _u16p), Gb.shape[1],
                        _p(Xb, _u16p), Xb.shape[1], _p(out, _f32p), nthreads)
    return out[:int(rowptr[-1])]


def workload_manager(rowptr, N, warps_per_tb=8):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    M = rowptr.size - 1
    cap = 8 * M * 64 + 64
    wpr = np.zeros(max(M, 1), np.int32)
    ridx = np.zeros(cap, np.int32)
    wid = np.zeros(cap, np.int32)
    nb = ctypes.c_int()
    n = lib().orc_workload_manager(M, _p(rowptr, _i32p), N, warps_per_tb, ctypes.byref(nb),
                                   _p(wpr, _i32p), _p(ridx, _i32p), _p(wid, _i32p), cap)
    if n < 0:
        return None
    return nb.value, wpr[:M], ridx[:n], wid[:n]


def tiling(M, K, rowptr, col, rpw):
    """Independent rebuild of the panel tiling; returns a dict of arrays."""
    L = lib()
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    if col.size == 0:
        colp = np.zeros(1, np.int32)
    else:
        colp = col
    G, C = ctypes.c_int(), ctypes.c_int()
    if L.orc_tiling_sizes(M, K, _p(rowptr, _i32p), _p(colp, _i32p), rpw, ctypes.byref(G), ctypes.byref(C)) != 0:
        raise ValueError("bad rows-per-wave")
    G, C = G.value, C.value
    ppg = rpw // 16
    nnz = col.size
    gcp = np.zeros(G + 1, np.int32)
    cols = np.zeros(max(C * 32, 1), np.int32)
    seg = np.zeros(C * ppg + 1, np.int32)
    slot = np.zeros(max(nnz, 1), np.uint16)
    src = np.zeros(max(nnz, 1), np.int32)
    order = np.zeros(max(G, 1), np.int32)
    n = L.orc_tiling_fill(M, K, _p(rowptr, _i32p), _p(colp, _i32p), rpw, _p(gcp, _i32p), _p(cols, _i32p),
                          _p(seg, _i32p), _p(slot, _u16p), _p(src, _i32p), _p(order, _i32p))
    assert n == nnz
    return dict(n_groups=G, n_chunks=C, grp_chunk_ptr=gcp, chunk_cols=cols[:C * 32], seg_ptr=seg,
                entry_slot=slot[:nnz], entry_src=src[:nnz], work_order=order[:G])


def transpose_csr(M, K, rowptr, col):
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    nnz = col.size
    rpT = np.zeros(K + 1, np.int32)
    cT = np.zeros(max(nnz, 1), np.int32)
    perm = np.zeros(max(nnz, 1), np.int32)
    colp = col if nnz else np.zeros(1, np.int32)
    lib().orc_transpose_csr(M, K, _p(rowptr, _i32p), _p(colp, _i32p), _p(rpT, _i32p), _p(cT, _i32p), _p(perm, _i32p))
    return rpT, cT[:nnz], perm[:nnz]


def dense_to_bell(dense_bits, b):
    """Blocked-ELL of a dense FP16 matrix (bits): (col_ind [ceil(M/b), slots] int32,
    values [ceil(M/b)*b, slots*b] uint16, ell_cols)."""
    d = np.ascontiguousarray(dense_bits, np.uint16)
    M, K = d.shape
    L = lib()
    dp = d if d.size else np.zeros(1, np.uint16)
    ell_cols = L.orc_dense_to_bell(_p(dp, _u16p), M, K, b, None, None)
    BR, slots = (M + b - 1) // b, ell_cols // b
    ci = np.zeros(max(BR * slots, 1), np.int32)
    vals = np.zeros(max(BR * b * ell_cols, 1), np.uint16)
    L.orc_dense_to_bell(_p(dp, _u16p), M, K, b, _p(ci, _i32p), _p(vals, _u16p))
    return ci[:BR * slots].reshape(BR, slots), vals[:BR * b * ell_cols].reshape(BR * b, ell_cols), ell_cols


def bell_spmm_f32(M, K, b, col_ind, values, B_bits):
    """C = BELL(A) * B in FP32 (ascending column order)."""
    B = np.ascontiguousarray(B_bits, np.uint16)
    N = B.shape[1]
    ci = np.ascontiguousarray(col_ind, np.int32).reshape(-1)
    v = np.ascontiguousarray(values, np.uint16).reshape(-1)
    ell_cols = values.shape[1] if values.ndim == 2 else 0
    C = np.zeros((M, N), np.float32)
    cip = ci if ci.size else np.zeros(1, np.int32)
    vp = v if v.size else np.zeros(1, np.uint16)
    Bp = B if B.size else np.zeros(1, np.uint16)
    lib().orc_bell_spmm_f32(M, K, N, b, ell_cols, _p(cip, _i32p), _p(vp, _u16p), _p(Bp, _u16p),
                            _p(C, _f32p) if C.size else None)
    return C


def prune_24_strip(A_bits):
    """2:4 STRIP pruning of a dense FP16 matrix (bits), rows of K % 4 == 0."""
    A = np.ascontiguousarray(A_bits, np.uint16)
    M, K = A.shape
    out = np.zeros_like(A)
    if A.size:
        rc = lib().orc_prune_24_strip(_p(A, _u16p), M, K, _p(out, _u16p))
        if rc != 0:
            raise ValueError("K must be a multiple of 4")
    return out
