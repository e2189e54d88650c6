"""Host-side mirror of the reference operator for this path, over the C ABI.

`Plan` is the MI355X counterpart of the reference's `workload_manager` tables
(src/our_spmm.cu:347-471): a panel tiling of a CSR pattern.  `spmm` launches the
HIP kernel that replaces `kernel_our_spmm_balanced` (src/our_spmm.cu:290-346).
Torch tensors are used only as device buffers; every computation happens in
libtca_spmm.so.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import BF16, E4M3, F16, F32, call

_TORCH_DTYPE = {torch.float16: F16, torch.bfloat16: BF16, torch.uint8: E4M3, torch.float32: F32}
if hasattr(torch, "float8_e4m3fn"):
    _TORCH_DTYPE[torch.float8_e4m3fn] = E4M3


def tca_dtype(t):
    try:
        return _TORCH_DTYPE[t.dtype]
    except KeyError:
        raise TypeError(f"unsupported dtype {t.dtype}") from None


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _np_i32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _stream(stream=None):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


# Argument checks done before any launch: the C ABI receives raw pointers and
# leading dimensions only, so a column-strided view, an undersized output or a
# tensor on another GPU would otherwise read or write the wrong memory.
def _check_matrix(t, name, rows=None, cols=None, dtypes=None, min_rows=None, min_cols=None):
    if not isinstance(t, torch.Tensor) or t.dim() != 2:
        raise ValueError(f"{name} must be a 2-D tensor")
    if t.shape[1] > 1 and t.stride(1) != 1:
        raise ValueError(f"{name} must have unit stride along its rows (stride(1) == 1), got {t.stride()}")
    if t.stride(0) < t.shape[1]:
        raise ValueError(f"{name} rows overlap (stride(0) {t.stride(0)} < {t.shape[1]} columns)")
    if dtypes is not None and t.dtype not in dtypes:
        raise ValueError(f"{name} has dtype {t.dtype}, expected one of {sorted(str(d) for d in dtypes)}")
    if rows is not None and t.shape[0] != rows:
        raise ValueError(f"{name} has {t.shape[0]} rows, expected {rows}")
    if cols is not None and t.shape[1] != cols:
        raise ValueError(f"{name} has {t.shape[1]} columns, expected {cols}")
    if min_rows is not None and t.shape[0] < min_rows:
        raise ValueError(f"{name} has {t.shape[0]} rows, needs at least {min_rows}")
    if min_cols is not None and t.shape[1] < min_cols:
        raise ValueError(f"{name} has {t.shape[1]} columns, needs at least {min_cols}")


def _check_vector(t, name, min_len, dtypes=None):
    if not isinstance(t, torch.Tensor) or t.dim() != 1 or (t.numel() > 1 and t.stride(0) != 1):
        raise ValueError(f"{name} must be a contiguous 1-D tensor")
    if t.numel() < min_len:
        raise ValueError(f"{name} has {t.numel()} elements, needs at least {min_len}")
    if dtypes is not None and t.dtype not in dtypes:
        raise ValueError(f"{name} has dtype {t.dtype}, expected one of {sorted(str(d) for d in dtypes)}")


def _check_devices(plan, *tensors):
    devs = {t.device for t in tensors if t is not None}
    if len(devs) > 1:
        raise ValueError(f"tensors are on different devices: {sorted(str(d) for d in devs)}")
    if plan.device is None:
        raise ValueError("host-only plan (device=False): launches need a device plan")
    for d in devs:
        if d.type != "cuda" or (d.index if d.index is not None else torch.cuda.current_device()) != plan.device:
            raise ValueError(f"tensor on {d}, the plan lives on cuda:{plan.device}")


_F16ISH = (torch.float16, torch.bfloat16)
_VALUE_DTYPES = tuple(_TORCH_DTYPE)


class Plan:
    """Panel tiling of a CSR pattern (host tables + device copies).

    rows_per_wave in {16, 32, 64}: consecutive rows sharing one condensed
    column list (16-row MFMA panels, 1/2/4 per wave); 128 or 256: one
    workgroup of rows/32 waves (2 panels each) shares every gathered B chunk.
    device=False builds a host-only plan (inspection/export only; launches
    reject it).
    """

    def __init__(self, n_rows, n_cols, row_offset, col_idx, rows_per_wave=16, device=True, _handle=None):
        self._h = None
        if _handle is not None:
            self._h = _handle
        else:
            rp = np.ascontiguousarray(row_offset, dtype=np.int32)
            ci = np.ascontiguousarray(col_idx, dtype=np.int32)
            if rp.size != n_rows + 1:
                raise ValueError("row_offset must have n_rows+1 entries")
            if ci.size == 0:
                ci = np.zeros(1, np.int32)
            h = ctypes.c_void_p()
            fn = "tca_plan_create" if device else "tca_plan_create_host"
            call(fn, n_rows, n_cols, int(rp[-1]), _np_i32p(rp), _np_i32p(ci), rows_per_wave, ctypes.byref(h))
            self._h = h
        info = _lib.PlanInfo()
        call("tca_plan_get_info", self._h, ctypes.byref(info))
        self.n_rows, self.n_cols, self.nnz = info.n_rows, info.n_cols, info.nnz
        self.rows_per_wave = info.rows_per_wave
        self.n_groups, self.n_chunks = info.n_groups, info.n_chunks
        self.max_group_chunks = info.max_group_chunks
        self.transposed = bool(info.transposed)
        self.device_bytes = info.device_bytes
        self.build_ms = info.build_ms
        self.dense_elems = info.dense_elems
        # the GPU holding the device tables (None for a host-only plan)
        self.device = None if info.device_bytes == 0 else torch.cuda.current_device()

    @property
    def handle(self):
        return self._h

    def info(self):
        return dict(n_rows=self.n_rows, n_cols=self.n_cols, nnz=self.nnz, rows_per_wave=self.rows_per_wave,
                    n_groups=self.n_groups, n_chunks=self.n_chunks, max_group_chunks=self.max_group_chunks,
                    transposed=self.transposed, device_bytes=self.device_bytes, build_ms=self.build_ms,
                    dense_elems=self.dense_elems)

    def export(self):
        """Host copies of the tiling tables (numpy)."""
        ppg = self.rows_per_wave // 16
        gcp = np.zeros(self.n_groups + 1, np.int32)
        cols = np.zeros(max(self.n_chunks * 32, 1), np.int32)
        seg = np.zeros(self.n_chunks * ppg + 1, np.int32)
        slot = np.zeros(max(self.nnz, 1), np.uint16)
        src = np.zeros(max(self.nnz, 1), np.int32)
        order = np.zeros(max(self.n_groups, 1), np.int32)
        call("tca_plan_export", self._h, _np_i32p(gcp), _np_i32p(cols), _np_i32p(seg),
             slot.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), _np_i32p(src), _np_i32p(order))
        return dict(n_groups=self.n_groups, n_chunks=self.n_chunks, grp_chunk_ptr=gcp,
                    chunk_cols=cols[:self.n_chunks * 32], seg_ptr=seg, entry_slot=slot[:self.nnz],
                    entry_src=src[:self.nnz], work_order=order[:self.n_groups])

    def transpose(self, rows_per_wave=None):
        """Plan of A^T whose tiled values are read from this plan's tiled values."""
        h = ctypes.c_void_p()
        call("tca_plan_create_transpose", self._h, rows_per_wave or self.rows_per_wave, ctypes.byref(h))
        return Plan(self.n_cols, self.n_rows, None, None, _handle=h)

    def pack(self, values_csr, out=None, stream=None):
        """Tiled-order copy of CSR-ordered device values (any supported dtype)."""
        if out is None:
            out = torch.empty_like(values_csr)
        if self.nnz:
            call("tca_pack_values", self._h, tca_dtype(values_csr), _ptr(values_csr), _ptr(out), _stream(stream))
        return out

    def pack_dense(self, values_tiled, out=None, stream=None):
        """Panel-dense tiles (dense_elems elements, same dtype) of tiled values:
        every (chunk, 16-row panel) as a full 16 x 32 tile in MFMA operand
        order, zeros off the pattern.  For a transposed plan pass the forward
        plan's tiled values.  spmm(..., dense=True) takes the result."""
        if out is None:
            out = torch.empty(max(self.dense_elems, 1), dtype=values_tiled.dtype,
                              device=values_tiled.device)[:self.dense_elems]
        if self.dense_elems:
            call("tca_pack_dense", self._h, tca_dtype(values_tiled), _ptr(values_tiled) if self.nnz else ctypes.c_void_p(0),
                 _ptr(out), _stream(stream))
        return out

    def unpack(self, values_tiled, out=None, stream=None):
        """CSR-ordered copy of tiled-order device values."""
        if out is None:
            out = torch.empty_like(values_tiled)
        if self.nnz:
            call("tca_unpack_values", self._h, tca_dtype(values_tiled), _ptr(values_tiled), _ptr(out), _stream(stream))
        return out

    def close(self):
        if self._h is not None and self._h.value:
            _lib.lib().tca_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlanCache:
    """Plans of fixed sparsity patterns, built once and reused: a training loop
    over pattern-static sparse layers keeps the host preprocessing (tiling,
    transposed tiling) out of its steps.  Keyed by the pattern's shape,
    rows_per_wave and a BLAKE2b digest of row_offset and col_idx."""

    def __init__(self):
        self._plans = {}

    @staticmethod
    def key(n_rows, n_cols, row_offset, col_idx, rows_per_wave):
        import hashlib
        h = hashlib.blake2b(digest_size=16)
        h.update(np.ascontiguousarray(row_offset, dtype=np.int32).tobytes())
        h.update(np.ascontiguousarray(col_idx, dtype=np.int32).tobytes())
        return (int(n_rows), int(n_cols), int(rows_per_wave), h.hexdigest())

    def get(self, n_rows, n_cols, row_offset, col_idx, rows_per_wave=16):
        k = self.key(n_rows, n_cols, row_offset, col_idx, rows_per_wave)
        p = self._plans.get(k)
        if p is None:
            p = self._plans[k] = Plan(n_rows, n_cols, row_offset, col_idx, rows_per_wave=rows_per_wave)
        return p

    def get_transpose(self, n_rows, n_cols, row_offset, col_idx, rows_per_wave=16, rows_per_wave_t=16):
        """(forward plan, transposed plan reading the forward plan's tiled values)."""
        fwd = self.get(n_rows, n_cols, row_offset, col_idx, rows_per_wave)
        k = self.key(n_rows, n_cols, row_offset, col_idx, rows_per_wave) + ("T", int(rows_per_wave_t))
        t = self._plans.get(k)
        if t is None:
            t = self._plans[k] = fwd.transpose(rows_per_wave_t)
        return fwd, t

    def __len__(self):
        return len(self._plans)

    def clear(self):
        self._plans.clear()


def csr_transpose_device(row_offset, col_idx, n_cols, stream=None):
    """CSC of a CSR pattern held in device int32 tensors, built on the device:
    (col_offset[n_cols+1], row_idx[nnz], perm[nnz]) where perm[d] is the CSR
    index of CSC entry d (rows ascending within each column)."""
    n_rows = row_offset.numel() - 1
    nnz = col_idx.numel()
    dev = row_offset.device
    col_offset = torch.empty(n_cols + 1, dtype=torch.int32, device=dev)
    row_idx = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)[:nnz]
    perm = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)[:nnz]
    call("tca_csr_transpose_device", n_rows, n_cols, nnz, _ptr(row_offset), _ptr(col_idx) if nnz else ctypes.c_void_p(0),
         _ptr(col_offset), _ptr(row_idx) if nnz else ctypes.c_void_p(0), _ptr(perm) if nnz else ctypes.c_void_p(0),
         _stream(stream))
    return col_offset, row_idx, perm


def spmm(plan, a_tiled, B, out=None, row_scale=None, col_scale=None, tile_n=0, stream=None, dense=False):
    """C[M x N] (float32) = A (plan + tiled values) @ B.

    a_tiled: tiled values, float16 / bfloat16 / uint8 (E4M3 codes).  For a
    transposed plan pass the forward plan's tiled values.  B: K x N row-major,
    same dtype as A or float32 (converted in-kernel).  row_scale / col_scale
    (float32, may be None) scale the FP32 result per row / column.
    dense=True: a_tiled holds the plan's panel-dense tiles (Plan.pack_dense);
    the result is bit-identical to the tiled-values launch.
    """
    _check_matrix(B, "B", rows=plan.n_cols, dtypes=_VALUE_DTYPES)
    K, N = B.shape
    if out is None:
        out = torch.empty((plan.n_rows, N), dtype=torch.float32, device=B.device)
    _check_matrix(out, "out", dtypes=(torch.float32,), min_rows=plan.n_rows, min_cols=N)
    if dense:
        if a_tiled is None or a_tiled.numel() != plan.dense_elems:
            raise ValueError(f"dense A must hold plan.dense_elems = {plan.dense_elems} elements")
        _check_vector(a_tiled, "dense A", plan.dense_elems, _VALUE_DTYPES)
    elif plan.nnz:
        if a_tiled is None:
            raise ValueError("a_tiled is required for a plan with nonzeros")
        _check_vector(a_tiled, "a_tiled", plan.nnz, _VALUE_DTYPES)
    if row_scale is not None:
        _check_vector(row_scale, "row_scale", plan.n_rows, (torch.float32,))
    if col_scale is not None:
        _check_vector(col_scale, "col_scale", N, (torch.float32,))
    _check_devices(plan, a_tiled, B, out, row_scale, col_scale)
    at = tca_dtype(a_tiled) if a_tiled is not None else F16
    call("tca_spmm_dense" if dense else "tca_spmm", plan.handle, at, _ptr(a_tiled), tca_dtype(B), _ptr(B), B.stride(0), N,
         _ptr(out), out.stride(0), _ptr(row_scale), _ptr(col_scale), tile_n, _stream(stream))
    return out


def sddmm(plan, G, X, out=None, stream=None):
    """out_tiled[e] = sum_n G[row(e), n] X[col(e), n] (float32, tiled order)."""
    _check_matrix(G, "G", rows=plan.n_rows, dtypes=_F16ISH + (torch.float32,))
    N = G.shape[1]
    _check_matrix(X, "X", rows=plan.n_cols, cols=N, dtypes=_F16ISH)
    if out is None:
        out = torch.empty(max(plan.nnz, 1), dtype=torch.float32, device=G.device)[:plan.nnz]
    _check_vector(out, "out", plan.nnz, (torch.float32,))
    _check_devices(plan, G, X, out)
    call("tca_sddmm", plan.handle, tca_dtype(G), _ptr(G), G.stride(0), tca_dtype(X), _ptr(X), X.stride(0), N,
         _ptr(out), _stream(stream))
    return out


def sgd_update(w, g, lr, stream=None):
    """w -= lr * g in place (w float16/bfloat16, g float32, same length and
    order, e.g. tiled weight values and the dW an SDDMM wrote)."""
    if w.numel() != g.numel() or g.dtype != torch.float32 or not (w.is_contiguous() and g.is_contiguous()):
        raise ValueError("sgd_update: w and g must be contiguous, of equal length, g float32")
    if w.device != g.device:
        raise ValueError("sgd_update: w and g must be on one device")
    call("tca_sgd_update", tca_dtype(w), _ptr(w), _ptr(g), w.numel(), float(lr), _stream(stream))
    return w


def quantize_rows_e4m3(row_offset, values, stream=None):
    """Per-row E4M3 codes (uint8) and FP32 scales of CSR values (float16, device)."""
    n_rows = row_offset.numel() - 1
    q = torch.empty(values.shape, dtype=torch.uint8, device=values.device)
    scale = torch.empty(n_rows, dtype=torch.float32, device=values.device)
    call("tca_quantize_rows_e4m3", n_rows, _ptr(row_offset), _ptr(values), _ptr(q), _ptr(scale), _stream(stream))
    return q, scale


def quantize_cols_e4m3(B, stream=None):
    """Per-column E4M3 codes (uint8) and FP32 scales of a dense K x N float16 matrix."""
    _check_matrix(B, "B", dtypes=(torch.float16,))
    K, N = B.shape
    q = torch.empty((K, N), dtype=torch.uint8, device=B.device)
    scale = torch.empty(N, dtype=torch.float32, device=B.device)
    call("tca_quantize_cols_e4m3", K, N, _ptr(B), B.stride(0), _ptr(q), N, _ptr(scale), _stream(stream))
    return q, scale


def rocsparse_spmm(M, N, K, row_offset, col_idx, values, B, out, n_iter=0):
    """rocSPARSE CSR SpMM baseline (f16 -> f32); returns mean ms per call (n_iter > 0)."""
    ms = ctypes.c_double(0.0)
    call("tca_rocsparse_spmm", M, N, K, col_idx.numel(), _ptr(row_offset), _ptr(col_idx), _ptr(values), _ptr(B),
         _ptr(out), n_iter, ctypes.byref(ms))
    return ms.value


def rocsparse_spmm_t(M, N, K, row_offset, col_idx, values, B, out, n_iter=0):
    """rocSPARSE transposed CSR SpMM baseline: out (K x N, f32) = A^T B, B (M x N) f16;
    returns mean ms per call (n_iter > 0)."""
    ms = ctypes.c_double(0.0)
    call("tca_rocsparse_spmm_t", M, N, K, col_idx.numel(), _ptr(row_offset), _ptr(col_idx), _ptr(values), _ptr(B),
         _ptr(out), n_iter, ctypes.byref(ms))
    return ms.value


def hipblaslt_gemm(M, N, K, A_dense, B, out, n_iter=0):
    """hipBLASLt dense GEMM baseline (f16 -> f32); returns mean ms per call (n_iter > 0)."""
    ms = ctypes.c_double(0.0)
    call("tca_hipblaslt_gemm", M, N, K, _ptr(A_dense), _ptr(B), _ptr(out), n_iter, ctypes.byref(ms))
    return ms.value



def rocsparse_sddmm(M, K, N, row_offset, col_idx, G16, X16, out, n_iter=0):
    """rocSPARSE SDDMM baseline: out (CSR order, float32) = (G X^T) on the pattern,
    G (M x N) and X (K x N) float16; returns mean ms per call (n_iter > 0)."""
    ms = ctypes.c_double(0.0)
    call("tca_rocsparse_sddmm", M, K, N, col_idx.numel(), _ptr(row_offset), _ptr(col_idx), _ptr(G16), _ptr(X16),
         _ptr(out), n_iter, ctypes.byref(ms))
    return ms.value
