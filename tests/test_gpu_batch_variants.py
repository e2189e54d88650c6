"""Every work-unit variant of the batched persistent launch against the FP64
oracle, in each of the launch's three kernel instantiations: a batch holding
tiled-value items (the kernel with every FP16 case), a batch whose SpMM items
are all panel-dense (the leaner FP16 kernel), and a BF16 batch.

Each item reads its own plan and writes its own output: forward SpMM on tiled
values (16-row plans, tile 16/32/64; 32-row, tile 32), transposed SpMM on
tiled values (16-row, tile 16/32/64), panel-dense forward and transposed SpMM
on 16/32/64/128-row plans at every tile width the batch offers, alternately
with caller-packed tiles and tiles the batch packs from the values every run,
and SDDMM.  The shapes are ragged (200 x 168: no multiple of the group or tile
sizes) and the operands 96 wide.

An FP32 operand that the batch's items read at one width gets a 16-bit copy
in the prologue, and the units read the copy; one that they read at two
widths (a 96-wide tensor and its 64-wide view here) gets none, so the units
that read FP32 and round on load run as well.  Both kinds are in every batch.
The oracle takes the 16-bit values (FP16 or BF16) and the FP32 operands rounded
to the A type, in FP64; the kernels accumulate in FP32: rel-L2 <= 1e-5."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_bf16, dev_f16, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen

pytestmark = pytest.mark.gpu

M, K = 200, 168
NW, NN = 96, 64                       # an operand's width, and its narrower view
TILED_FWD = [(16, 16), (16, 32), (16, 64), (32, 32)]
TILED_T = [(16, 16), (16, 32), (16, 64)]
DENSE = [(16, 16), (16, 32), (16, 64), (32, 32), (32, 64), (64, 32), (64, 64), (128, 64)]
DENSE_F32B = [(16, 16), (16, 32), (16, 64), (32, 32)]
TOL = 1e-5


class Inputs:
    """One pattern, its values in the 16-bit type `kind`, the FP16/BF16 X, an
    FP32 X, and two FP32 G: `gc` read at one width (copied), `gx` at two."""

    def __init__(self, kind, cuda):
        self.kind = kind
        self.rp, self.ci = gen.pattern("lognormal", M, K, 0.15, 0.5, seed=11)
        self.nnz = int(self.ci.size)
        to16 = gen.f32_to_f16_bits if kind == "f16" else gen.f32_to_bf16_bits
        self.dec = (lambda b: b.view(np.float16).astype(np.float64)) if kind == "f16" else \
            (lambda b: O.bf16_bits_to_f32(b).astype(np.float64))
        self.rnd = lambda x: self.dec(to16(x))            # FP32 -> the A type -> f64
        self.up = dev_f16 if kind == "f16" else dev_bf16
        w32 = gen.normal_f32(11, 1, (self.nnz,), 0.0, 0.1)
        x32 = gen.normal_f32(11, 2, (K, NW))
        self.w_bits, self.x_bits = to16(w32), to16(x32)
        self.wq, self.xq = self.dec(self.w_bits), self.dec(self.x_bits)
        self.x16 = self.up(self.x_bits)
        self.x32 = dev(gen.normal_f32(11, 3, (K, NW)))
        self.gc = dev(gen.normal_f32(11, 4, (M, NW)))
        self.gx = dev(gen.normal_f32(11, 5, (M, NW)))
        w_dev = self.up(self.w_bits)
        self.plan = {r: T.Plan(M, K, self.rp, self.ci, r) for r in (16, 32, 64, 128)}
        self.tiled = {r: p.pack(w_dev) for r, p in self.plan.items()}
        self.plan_t = {r: self.plan[16].transpose(r) for r in (16, 32, 64, 128)}   # values: plan[16]'s order
        self.rpT, self.cT, perm = O.transpose_csr(M, K, self.rp, self.ci)
        self.wqT = self.wq[perm]
        self.cuda = cuda
        self.items, self.checks = [], []

    def q(self, t):
        """The rounded f64 values of a device operand, as the kernels read it."""
        a = t.float().cpu().numpy() if t.dtype != torch.float32 else t.cpu().numpy()
        return self.rnd(a) if t.dtype == torch.float32 else a.astype(np.float64)

    def _out(self, rows, n):
        return torch.zeros((rows, n), dtype=torch.float32, device=self.cuda)

    def fwd(self, r, tile, B, dense):
        n = B.shape[1]
        out = self._out(M, n)
        p = self.plan[r]
        if dense is True:
            self.items.append(T.spmm_item(p, p.pack_dense(self.tiled[r]), B, out, tile, dense=True))
        elif dense == "pack":
            self.items.append(T.spmm_item(p, self.tiled[r], B, out, tile, dense="pack", values_order=p))
        else:
            self.items.append(T.spmm_item(p, self.tiled[r], B, out, tile))
        ref = O.spmm_f64(M, n, self.rp, self.ci, self.wq, self.q(B))
        self.checks.append((f"fwd rpw {r} tile {tile} B {B.dtype} N {n} dense {dense}", lambda: out.cpu().numpy(), ref))

    def dx(self, r, tile, G, dense):
        n = G.shape[1]
        out = self._out(K, n)
        p = self.plan_t[r]
        if dense is True:
            self.items.append(T.spmm_item(p, p.pack_dense(self.tiled[16]), G, out, tile, dense=True))
        elif dense == "pack":
            self.items.append(T.spmm_item(p, self.tiled[16], G, out, tile, dense="pack", values_order=self.plan[16]))
        else:
            self.items.append(T.spmm_item(p, self.tiled[16], G, out, tile))
        ref = O.spmm_f64(K, n, self.rpT, self.cT, self.wqT, self.q(G))
        self.checks.append((f"dx rpw {r} tile {tile} G {G.dtype} N {n} dense {dense}", lambda: out.cpu().numpy(), ref))

    def dw(self, r, G, X):
        n = G.shape[1]
        out = torch.zeros(max(self.nnz, 1), dtype=torch.float32, device=self.cuda)
        p = self.plan[r]
        self.items.append(T.sddmm_item(p, G, X, out))
        ref = O.sddmm_f64(M, n, self.rp, self.ci, self.q(G), self.q(X))
        self.checks.append((f"dw rpw {r} G {G.dtype} N {n}", lambda: p.unpack(out).cpu().numpy()[:self.nnz], ref))

    def run_and_check(self):
        batch = T.Batch(self.items)
        for rep in range(2):        # the second run starts from the first run's counters and copies
            batch.run()
            torch.cuda.synchronize()
            for name, got, ref in self.checks:
                err = rel_l2(got(), ref)
                assert err <= TOL, f"run {rep}: {name}: rel-L2 {err:.3e}"
        return batch


def _dense_items(I, x32, gx_narrow):
    for k, (r, t) in enumerate(DENSE):
        mode = True if k % 2 == 0 else "pack"
        I.fwd(r, t, I.x16, mode)
        I.dx(r, t, I.gc, "pack" if mode is True else True)
    for k, (r, t) in enumerate(DENSE_F32B):
        I.fwd(r, t, x32, True if k % 2 == 0 else "pack")
        I.dx(r, t, I.gx, "pack" if k % 2 == 0 else True)
    # the narrower views that keep x32 and gx from being copied
    I.fwd(16, 32, x32[:, :NN], True)
    I.dx(32, 32, gx_narrow, "pack")


def test_batch_every_f16_variant_with_tiled_items(cuda):
    I = Inputs("f16", cuda)
    for r, t in TILED_FWD:
        I.fwd(r, t, I.x16, False)
    for r, t in TILED_T:
        I.dx(r, t, I.gc, False)
        I.dx(r, t, I.gx, False)
    _dense_items(I, I.x32, I.gx[:, :NN])
    for r in (16, 64):
        I.dw(r, I.gc, I.x16)
        I.dw(r, I.gx, I.x16)
    I.dw(32, I.gx[:, :NN], I.x16[:, :NN])
    I.run_and_check()


def test_batch_every_f16_variant_panel_dense_only(cuda):
    I = Inputs("f16", cuda)
    _dense_items(I, I.x32, I.gx[:, :NN])
    I.dw(16, I.gc, I.x16)
    I.dw(32, I.gx, I.x16)
    I.run_and_check()


def test_batch_every_bf16_variant(cuda):
    I = Inputs("bf16", cuda)
    _dense_items(I, I.x32, I.gx[:, :NN])
    I.dw(16, I.gc, I.x16)
    I.dw(64, I.gx, I.x16)
    I.dw(32, I.gx[:, :NN], I.x16[:, :NN])
    I.run_and_check()
