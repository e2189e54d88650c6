"""Sparse-weight training: Y = W X with a CSR weight W, differentiable in W and X.

Backward (hand-written, both on the C ABI):
  dX = W^T G        transposed-plan SpMM
  dW = SDDMM(G, X^T) sampled on W's pattern, FP32, in tiled order
Both read G in the weight's MFMA type: an FP32 G is rounded to it once (round to
nearest even, the same bits the kernels would produce rounding on load), which
halves the bytes both kernels gather.  X may be FP16 / BF16 (the weight's type)
or FP32: the forward rounds an FP32 X on load and the SDDMM gets the same
rounded X.
W's values live in *tiled order* (the plan's entry order) as the parameter,
so no permutation runs per step; `SparseLinear.csr_values()` exports them.
dense=True runs the forward and dX on panel-dense tiles built from the
parameter in each call (one pack each; same bits, faster for large layers).
"""
import numpy as np
import torch

from .core import Plan, sddmm, spmm


class SpmmFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w_tiled, X, plan, plan_t, tile_n, dense=False):
        if X.dim() == 2 and X.shape[1] > 1 and X.stride(1) != 1:
            X = X.contiguous()
        if dense:
            Y = spmm(plan, plan.pack_dense(w_tiled), X, tile_n=tile_n, dense=True)
        else:
            Y = spmm(plan, w_tiled, X, tile_n=tile_n)
        ctx.save_for_backward(w_tiled, X)
        ctx.plan, ctx.plan_t, ctx.tile_n, ctx.dense = plan, plan_t, tile_n, dense
        return Y

    @staticmethod
    def backward(ctx, G):
        w_tiled, X = ctx.saved_tensors
        G = G.contiguous()
        if G.dtype != w_tiled.dtype:
            G = G.to(w_tiled.dtype)
        dW = dX = None
        if ctx.needs_input_grad[0]:
            # an FP32 X entered the forward rounded to the weight's type on load
            # (nearest even); the SDDMM takes X in that type, so round it the same way
            Xw = X if X.dtype == w_tiled.dtype else X.to(w_tiled.dtype)
            dW = sddmm(ctx.plan, G, Xw).to(w_tiled.dtype)
        if ctx.needs_input_grad[1]:
            if ctx.dense:
                dX = spmm(ctx.plan_t, ctx.plan_t.pack_dense(w_tiled), G, tile_n=ctx.tile_n, dense=True).to(X.dtype)
            else:
                dX = spmm(ctx.plan_t, w_tiled, G, tile_n=ctx.tile_n).to(X.dtype)
        return dW, dX, None, None, None, None


def sparse_matmul(w_tiled, X, plan, plan_t, tile_n=0, dense=False):
    """Y (float32) = W X, differentiable in the tiled weight values and X."""
    return SpmmFunction.apply(w_tiled, X, plan, plan_t, tile_n, dense)


class SparseLinear(torch.nn.Module):
    """A CSR-pattern weight (n_out x n_in) applied as Y = W X to X (n_in x N).

    The pattern is fixed at construction; the values are an fp16/bf16
    parameter in tiled order.
    """

    def __init__(self, n_out, n_in, row_offset, col_idx, values_csr, rows_per_wave=16, rows_per_wave_t=None,
                 tile_n=0, device="cuda", dense=False):
        super().__init__()
        self.plan = Plan(n_out, n_in, row_offset, col_idx, rows_per_wave=rows_per_wave)
        self.plan_t = self.plan.transpose(rows_per_wave_t or rows_per_wave)
        v = values_csr if isinstance(values_csr, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(values_csr).view(np.int16)).view(torch.float16)
        self.weight = torch.nn.Parameter(self.plan.pack(v.to(device)))
        self.tile_n = tile_n
        self.dense = dense

    def forward(self, X):
        return sparse_matmul(self.weight, X, self.plan, self.plan_t, self.tile_n, self.dense)

    def csr_values(self, tiled=None):
        return self.plan.unpack((self.weight if tiled is None else tiled).detach())
