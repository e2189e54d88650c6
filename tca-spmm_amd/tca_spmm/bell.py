"""Blocked-ELL: host conversion from CSR and the gfx950 SpMM on it.

Replaces perform_cuSPARSE_BlockSpMM (src/spmm_cusparse.cu:181-445): the host
conversion to cuSPARSE's Blocked-ELL arrays (:254-331) and the
cusparseCreateBlockedEll + cusparseSpMM pair (:337-343, :394-398).  Layout:
ceil(M/b) block rows of `slots = ell_cols / b` slots; col_ind[br, j] = block
column of slot j or -1; values (ceil(M/b)*b) x ell_cols row-major.  Both steps
run in libtca_spmm.so (the conversion on the host, the product on the GPU).
"""
import ctypes

import numpy as np

from ._lib import BF16, F16, F32, call


def _i32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _u16p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))


def bell_from_csr(n_rows, n_cols, row_offset, col_idx, values=None, block=32):
    """CSR (values as uint16 bits of F16/BF16, or None for 1.0 everywhere) ->
    (col_ind [ceil(M/b), slots] int32, values [ceil(M/b)*b, ell_cols] uint16, ell_cols)."""
    ro = np.ascontiguousarray(row_offset, np.int32)
    ci = np.ascontiguousarray(col_idx, np.int32)
    nnz = int(ro[n_rows]) if n_rows > 0 else 0
    cip = ci if ci.size else np.zeros(1, np.int32)
    vals = None if values is None else np.ascontiguousarray(values).view(np.uint16)
    if vals is not None and vals.size < nnz:
        raise ValueError("values shorter than the CSR")
    ec = ctypes.c_int(0)
    call("tca_bell_from_csr", n_rows, n_cols, _i32p(ro), _i32p(cip), _u16p(vals) if vals is not None else None, block,
         ctypes.byref(ec), None, None)
    ell_cols = ec.value
    BR, slots = (n_rows + block - 1) // block, ell_cols // block
    col_ind = np.zeros(max(BR * slots, 1), np.int32)
    ell_vals = np.zeros(max(BR * block * ell_cols, 1), np.uint16)
    call("tca_bell_from_csr", n_rows, n_cols, _i32p(ro), _i32p(cip), _u16p(vals) if vals is not None else None, block,
         ctypes.byref(ec), _i32p(col_ind), _u16p(ell_vals))
    return (col_ind[:BR * slots].reshape(BR, slots), ell_vals[:BR * block * ell_cols].reshape(BR * block, ell_cols),
            ell_cols)


def bell_shape(n_rows, n_cols, N, block, tile_n=0):
    """(rows_per_wave, tile_n) that bell_spmm picks for an n_rows x n_cols matrix times N columns."""
    r, t = ctypes.c_int(0), ctypes.c_int(0)
    call("tca_bell_shape", n_rows, n_cols, N, block, tile_n, ctypes.byref(r), ctypes.byref(t))
    return r.value, t.value


class BlockedEll:
    """A Blocked-ELL matrix in device memory (col_ind int32, values float16 /
    bfloat16 torch tensors in the layout above)."""

    def __init__(self, n_rows, n_cols, block, col_ind, values):
        self.n_rows, self.n_cols, self.block = int(n_rows), int(n_cols), int(block)
        self.col_ind, self.values = col_ind, values
        self.ell_cols = int(values.shape[1]) if values.dim() == 2 else 0
        br = (self.n_rows + self.block - 1) // self.block
        if tuple(col_ind.shape) != (br, self.ell_cols // self.block) or values.shape[0] != br * self.block:
            raise ValueError("col_ind / values do not match the Blocked-ELL shape")

    @classmethod
    def from_csr(cls, n_rows, n_cols, row_offset, col_idx, values=None, block=32, dtype=None, device="cuda"):
        """Host conversion, then a copy to `device`.  values: uint16 bits or a
        float16 / bfloat16 tensor or array (None: 1.0 at every entry)."""
        import torch
        if dtype is None:
            dtype = values.dtype if isinstance(values, torch.Tensor) and values.dtype == torch.bfloat16 else torch.float16
        if isinstance(values, torch.Tensor):
            values = values.detach().cpu().contiguous().view(torch.int16).numpy()
        ci, ev, _ = bell_from_csr(n_rows, n_cols, row_offset, col_idx, values, block)
        col_ind = torch.from_numpy(ci).to(device)
        vals = torch.from_numpy(ev.view(np.int16)).view(dtype).to(device)
        return cls(n_rows, n_cols, block, col_ind, vals)


def bell_spmm(bell, B, out=None, tile_n=0, rows_per_wave=0, stream=None):
    """C[M x N] (float32) = bell @ B; B K x N row-major, the values' dtype or float32."""
    import torch
    from .core import _ptr, _stream, tca_dtype
    K, N = B.shape
    if K != bell.n_cols:
        raise ValueError(f"B has {K} rows, the matrix has {bell.n_cols} columns")
    if out is None:
        out = torch.empty((bell.n_rows, N), dtype=torch.float32, device=B.device)
    at = tca_dtype(bell.values)
    if at not in (F16, BF16) or tca_dtype(B) not in (at, F32):
        raise TypeError("bell_spmm: values float16/bfloat16, B the same dtype or float32")
    call("tca_bell_spmm", bell.n_rows, bell.n_cols, bell.block, bell.ell_cols, _ptr(bell.col_ind), at, _ptr(bell.values),
         tca_dtype(B), _ptr(B), B.stride(0), N, _ptr(out), out.stride(0), tile_n, rows_per_wave, _stream(stream))
    return out
