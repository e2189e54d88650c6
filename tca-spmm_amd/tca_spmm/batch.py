"""One persistent launch over many layers' SpMM / transposed SpMM / SDDMM.

    b = Batch([spmm_item(plan, w_tiled, X, Y), spmm_item(plan_t, w_tiled, G, dX),
               sddmm_item(plan, G, X, dW), ...])
    b.run()            # one kernel; capturable in a HIP graph

Items are fixed at construction (pointers, shapes, tile widths); run() only
launches.  Calls that overlap on one device must use different Batch objects
(each owns its work counter).

Sparse-weight training: spmm_item(plan, w, X, Y, dense="pack", values_order=p)
hands the batch the weight VALUES (in p's tiled order, or CSR order when
values_order is None); the batch owns the item's panel-dense tiles and rebuilds
them from w at the start of every run, so an optimizer that updates w in place
needs no repack launches of its own: run() = one prologue launch (FP16 copies
of FP32 operands, dense tiles) + the persistent launch.  The optimizer step
itself can ride on the SDDMM: sddmm_item(plan, G, X, dW, sgd=(w, lr)) also
applies w <- w - lr * dW to every entry it writes (the arithmetic of
sgd_update, bit for bit), after this run's tiles were packed from w -- so a
training step is one run().
"""
import ctypes

import torch

from . import _lib
from ._lib import OP_SDDMM, OP_SPMM, call
from .core import _F16ISH, _VALUE_DTYPES, _check_devices, _check_matrix, _check_vector, _stream, tca_dtype


def spmm_item(plan, a_tiled, B, out, tile_n=0, dense=False, values_order=None):
    """dense=True: a_tiled holds the plan's panel-dense tiles (Plan.pack_dense).
    dense="pack": a_tiled holds weight values in values_order's tiled order
    (a forward Plan of the same pattern; None = CSR order) and the batch packs
    the plan's dense tiles from them on every run."""
    if dense not in (False, True, "pack"):
        raise ValueError("dense must be False, True or 'pack'")
    if values_order is not None and dense != "pack":
        raise ValueError("values_order applies to dense='pack' items")
    return dict(op=OP_SPMM, plan=plan, a=a_tiled, b=B, out=out, tile_n=tile_n, dense=dense,
                values_order=values_order)


def sddmm_item(plan, G, X, out, sgd=None):
    """sgd=(w, lr): w (F16/BF16 values in the same order as out) is updated
    in place from every dW entry the item writes: w <- round(w - lr * dW)."""
    it = dict(op=OP_SDDMM, plan=plan, a=G, b=X, out=out, tile_n=0)
    if sgd is not None:
        w, lr = sgd
        it["sgd_w"], it["sgd_lr"] = w, float(lr)
    return it


def _check_item(it):
    """The checks tca_spmm / tca_sddmm make, per item (raises ValueError)."""
    plan, a, b, out = it["plan"], it["a"], it["b"], it["out"]
    if it["op"] == OP_SPMM:
        _check_matrix(b, "B", rows=plan.n_cols, dtypes=_VALUE_DTYPES)
        _check_matrix(out, "out", dtypes=(torch.float32,), min_rows=plan.n_rows, min_cols=b.shape[1])
        d = it.get("dense", False)
        if d is True:
            if a.numel() != plan.dense_elems:
                raise ValueError(f"dense A must hold plan.dense_elems = {plan.dense_elems} elements")
            _check_vector(a, "dense A", plan.dense_elems, _VALUE_DTYPES)
        elif plan.nnz:
            _check_vector(a, "values", plan.nnz, _VALUE_DTYPES)
    else:
        _check_matrix(a, "G", rows=plan.n_rows, dtypes=_F16ISH + (torch.float32,))
        _check_matrix(b, "X", rows=plan.n_cols, cols=a.shape[1], dtypes=_F16ISH)
        _check_vector(out, "out", plan.nnz, (torch.float32,))
        if it.get("sgd_w") is not None:
            _check_vector(it["sgd_w"], "sgd weights", plan.nnz, _F16ISH)
            _check_devices(plan, it["sgd_w"])
    if it["op"] == OP_SPMM and it.get("sgd_w") is not None:
        raise ValueError("sgd applies to SDDMM items")
    _check_devices(plan, a, b, out)


class Batch:
    def __init__(self, items):
        for it in items:
            _check_item(it)
        arr = (_lib.BatchItem * len(items))()
        self._keep = []
        for i, it in enumerate(items):
            a, b, out = it["a"], it["b"], it["out"]
            self._keep += [it["plan"], a, b, out]
            e = arr[i]
            e.op = it["op"]
            e.plan = it["plan"].handle
            e.tile_n = it.get("tile_n", 0)
            e.a_dtype = tca_dtype(a)
            e.a = a.data_ptr()
            e.lda = a.stride(0) if a.dim() == 2 else 0
            e.b_dtype = tca_dtype(b)
            e.b = b.data_ptr()
            e.ldb = b.stride(0)
            e.N = b.shape[1]
            e.out = out.data_ptr()
            e.ldc = out.stride(0) if out.dim() == 2 else 0
            d = it.get("dense", False)
            e.a_dense = 2 if d == "pack" else (1 if d else 0)
            vo = it.get("values_order")
            if vo is not None:
                self._keep.append(vo)
                e.values_order = vo.handle
            w = it.get("sgd_w")
            if w is not None:
                self._keep.append(w)
                e.sgd_w = w.data_ptr()
                e.sgd_w_dtype = tca_dtype(w)
                e.sgd_lr = it["sgd_lr"]
        h = ctypes.c_void_p()
        call("tca_batch_create", len(items), arr, ctypes.byref(h))
        self._h = h
        nu, nw = ctypes.c_int(), ctypes.c_int()
        call("tca_batch_get_info", h, ctypes.byref(nu), ctypes.byref(nw))
        self.n_units, self.n_workgroups = nu.value, nw.value

    def run(self, stream=None, stage=3):
        """stage 1: the prologue only (FP16 copies, dense tiles), 2: the
        persistent launch only, 3: both (a full run)."""
        if stage == 3:
            call("tca_batch_run", self._h, _stream(stream))
        else:
            call("tca_batch_run_stage", self._h, stage, _stream(stream))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().tca_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
