"""One persistent launch over many layers' SpMM / transposed SpMM / SDDMM.

    b = Batch([spmm_item(plan, w_tiled, X, Y), spmm_item(plan_t, w_tiled, G, dX),
               sddmm_item(plan, G, X, dW), ...])
    b.run()            # one kernel; capturable in a HIP graph

Items are fixed at construction (pointers, shapes, tile widths); run() only
launches.  Calls that overlap on one device must use different Batch objects
(each owns its work counter).
"""
import ctypes

import torch

from . import _lib
from ._lib import OP_SDDMM, OP_SPMM, call
from .core import _stream, tca_dtype


def spmm_item(plan, a_tiled, B, out, tile_n=0, dense=False):
    """dense=True: a_tiled holds the plan's panel-dense tiles (Plan.pack_dense)."""
    return dict(op=OP_SPMM, plan=plan, a=a_tiled, b=B, out=out, tile_n=tile_n, dense=dense)


def sddmm_item(plan, G, X, out):
    return dict(op=OP_SDDMM, plan=plan, a=G, b=X, out=out, tile_n=0)


class Batch:
    def __init__(self, items):
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
            e.a_dense = 1 if it.get("dense") else 0
        h = ctypes.c_void_p()
        call("tca_batch_create", len(items), arr, ctypes.byref(h))
        self._h = h
        nu, nw = ctypes.c_int(), ctypes.c_int()
        call("tca_batch_get_info", h, ctypes.byref(nu), ctypes.byref(nw))
        self.n_units, self.n_workgroups = nu.value, nw.value

    def run(self, stream=None):
        call("tca_batch_run", self._h, _stream(stream))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().tca_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
