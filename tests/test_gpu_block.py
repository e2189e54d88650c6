"""Block kernels on panel-dense tiles (plans of 128 / 256 rows per call; 64 /
128 rows inside the batch): the workgroup's waves share every gathered B chunk
tile and two K groups split the chunk range.  Checked against the FP64 oracle
(rel-L2 <= 1e-5, FP32 accumulation of FP16 products) and against the
tiled-values launch of the same plan (same products, another summation order:
rel-L2 <= 1e-6): forward and transposed plans, F16 / BF16 / an F32 B rounded on
load, tile widths 64 / 128, ragged N and row counts, empty and one-row groups,
groups of one chunk (one K group idle), the 4096 x 4096 x 512 workload (two
K groups per workgroup) and the 4096 x 1024 x 2048 one (one K group: enough
units to fill the chip);
E4M3 A and B with per-row / per-column scales (8-bit tiles, B operands by
transposing 8-bit LDS reads) against the oracle on the quantised inputs and
scales."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_bf16, dev_f16, f16_to_f64, oracle_f32, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads
from tca_spmm.workloads import SparseMatrix

pytestmark = pytest.mark.gpu

WIDE = (128, 256)
TILES = (64, 128)


def _cases():
    yield "lognormal_300x700", *gen.pattern("lognormal", 300, 700, 0.05, 0.5, seed=2, stream=6), 300, 700
    yield "bernoulli_1000x96", *gen.pattern("bernoulli", 1000, 96, 0.3, seed=3, stream=7), 1000, 96
    rp = np.zeros(271, np.int32)
    rp[200:] = 2                                   # one non-empty row: empty groups around it, one chunk
    yield "one_row", rp, np.array([3, 29], np.int32), 270, 30
    yield "empty", np.zeros(150, np.int32), np.zeros(0, np.int32), 149, 64


@pytest.mark.parametrize("rpw", WIDE)
def test_block_f16_against_oracle_and_tiled(cuda, rpw):
    for name, rp, ci, M, K in _cases():
        vals = gen.f32_to_f16_bits(gen.normal_f32(9, rpw, (ci.size,)))
        A = SparseMatrix(name, M, K, rp, ci, vals)
        plan = T.Plan(M, K, rp, ci, rpw)
        tiled = plan.pack(dev_f16(vals if ci.size else np.zeros(1, np.uint16)))
        dense = plan.pack_dense(tiled)
        for N in (8, 136, 512):
            Bb = gen.f32_to_f16_bits(gen.normal_f32(10, N, (K, N)))
            B = dev_f16(Bb)
            ref = oracle_f32(A, Bb)
            for t in TILES:
                got = T.spmm(plan, dense, B, tile_n=t, dense=True).cpu().numpy()
                assert np.isfinite(got).all(), (name, N, t)
                if ci.size == 0:
                    assert not got.any(), (name, N, t)
                    continue
                assert rel_l2(got, ref) <= 1e-5, (name, N, t)
                tv = T.spmm(plan, tiled, B, tile_n=t).cpu().numpy()
                assert rel_l2(got, tv) <= 1e-6, (name, N, t)


def test_block_bf16_and_f32_b(cuda):
    case = workloads.plain_gpu(512, 2048, 0.9)
    A = case.A
    a64 = f16_to_f64(A.values)
    for rpw in WIDE:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        bf = gen.f32_to_bf16_bits(A.values.view(np.float16).astype(np.float32))
        bfb = gen.f32_to_bf16_bits(case.B.view(np.float16).astype(np.float32))
        d_bf = plan.pack_dense(plan.pack(dev_bf16(bf)))
        got = T.spmm(plan, d_bf, dev_bf16(bfb), tile_n=64, dense=True).cpu().numpy()
        ref = O.spmm_f64(A.n_rows, case.N, A.row_offset, A.col_idx, O.bf16_bits_to_f32(bf).astype(np.float64),
                         O.bf16_bits_to_f32(bfb).astype(np.float64))
        assert rel_l2(got, ref) <= 1e-5
        # an F32 B is rounded to the A type on load: same as passing it rounded
        d16 = plan.pack_dense(plan.pack(dev_f16(A.values)))
        B32 = case.B.view(np.float16).astype(np.float32) * np.float32(1.0 + 2.0 ** -13)
        g32 = T.spmm(plan, d16, dev(B32), tile_n=128, dense=True)
        g16 = T.spmm(plan, d16, dev(B32).half(), tile_n=128, dense=True)
        assert torch.equal(g32, g16)
        ref16 = O.spmm_f64(A.n_rows, case.N, A.row_offset, A.col_idx, a64, B32.astype(np.float16).astype(np.float64))
        assert rel_l2(g16.cpu().numpy(), ref16) <= 1e-5


def test_block_transposed_dx(cuda):
    case = workloads.batch12(seed=0)[10]               # 4096 x 1024 at 0.12, N = 128
    W, G = case.A, case.extra["G"]
    fwd = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
    wt = fwd.pack(dev_f16(W.values))
    rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
    g16 = G.astype(np.float16).astype(np.float64)
    ref = O.spmm_f64(W.n_cols, case.N, rpT, cT, f16_to_f64(W.values)[perm], g16)
    for rpw in WIDE:
        tr = fwd.transpose(rpw)
        d = tr.pack_dense(wt)
        for t in TILES:
            got = T.spmm(tr, d, dev(G), tile_n=t, dense=True).cpu().numpy()
            assert rel_l2(got, ref) <= 1e-5, (rpw, t)


def test_block_4096_workload(cuda):
    case = workloads.precision()
    A = case.A
    ref = O.spmm_f64(A.n_rows, case.N, A.row_offset, A.col_idx, f16_to_f64(A.values), f16_to_f64(case.B))
    B = dev_f16(case.B)
    for rpw in WIDE:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        d = plan.pack_dense(plan.pack(dev_f16(A.values)))
        for t in TILES:
            got = T.spmm(plan, d, B, tile_n=t, dense=True).cpu().numpy()
            assert rel_l2(got, ref) <= 1e-5, (rpw, t)
        # two launches give the same bits (K groups add in a fixed order)
        assert torch.equal(T.spmm(plan, d, B, tile_n=64, dense=True), T.spmm(plan, d, B, tile_n=64, dense=True))


def _e4m3_ref(A, q, s, qb, sb, N):
    qa = O.e4m3_to_f32(q.cpu().numpy()).astype(np.float64)
    qbf = O.e4m3_to_f32(qb.cpu().numpy()).astype(np.float64)
    ref = O.spmm_f64(A.n_rows, N, A.row_offset, A.col_idx, qa, qbf)
    return ref * (s.cpu().numpy().astype(np.float64)[:, None] * sb.cpu().numpy().astype(np.float64)[None, :])


@pytest.mark.parametrize("rpw", WIDE)
def test_block_e4m3_against_oracle(cuda, rpw):
    for name, rp, ci, M, K in _cases():
        if ci.size == 0:
            continue
        vals = gen.f32_to_f16_bits(gen.normal_f32(11, rpw, (ci.size,)))
        A = SparseMatrix(name, M, K, rp, ci, vals)
        q, s = T.quantize_rows_e4m3(dev(rp), dev_f16(vals))
        plan = T.Plan(M, K, rp, ci, rpw)
        dense = plan.pack_dense(plan.pack(q))
        for N in (64, 136, 512):
            qb, sb = T.quantize_cols_e4m3(dev_f16(gen.f32_to_f16_bits(gen.normal_f32(12, N, (K, N)))))
            ref = _e4m3_ref(A, q, s, qb, sb, N)
            for t in TILES:
                got = T.spmm(plan, dense, qb, row_scale=s, col_scale=sb, tile_n=t, dense=True).cpu().numpy()
                assert rel_l2(got, ref) <= 1e-5, (name, N, t)


def test_block_e4m3_4096_workload_and_transposed(cuda):
    case = workloads.precision()
    A = case.A
    q, s = T.quantize_rows_e4m3(dev(A.row_offset), dev_f16(A.values))
    qb, sb = T.quantize_cols_e4m3(dev_f16(case.B))
    ref = _e4m3_ref(A, q, s, qb, sb, case.N)
    for rpw in WIDE:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        d = plan.pack_dense(plan.pack(q))
        for t in TILES:
            got = T.spmm(plan, d, qb, row_scale=s, col_scale=sb, tile_n=t, dense=True).cpu().numpy()
            assert rel_l2(got, ref) <= 1e-5, (rpw, t)
    # transposed plan: W^T G with E4M3 values (no scales), against the oracle
    W = workloads.batch12(seed=0)[10].A                 # 4096 x 1024 at 0.12
    fwd = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
    qw, _ = T.quantize_rows_e4m3(dev(W.row_offset), dev_f16(W.values))
    gq, _ = T.quantize_cols_e4m3(dev_f16(gen.f32_to_f16_bits(gen.normal_f32(13, 128, (W.n_rows, 128)))))
    rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
    ref_t = O.spmm_f64(W.n_cols, 128, rpT, cT, O.e4m3_to_f32(qw.cpu().numpy()).astype(np.float64)[perm],
                       O.e4m3_to_f32(gq.cpu().numpy()).astype(np.float64))
    for rpw in WIDE:
        tr = fwd.transpose(rpw)
        got = T.spmm(tr, tr.pack_dense(fwd.pack(qw)), gq, tile_n=64, dense=True).cpu().numpy()
        assert rel_l2(got, ref_t) <= 1e-5, rpw


def test_block_one_k_group_when_units_fill_the_chip(cuda):
    # launches with >= 512 units (group x N-tile) run one K group per
    # workgroup: the 4096 x 1024 x 2048 layer (128-row plans: 1024 / 512
    # units; 256-row plans: 512 with tile 64, 256 -- two K groups -- with
    # tile 128), FP16 and E4M3, against the FP64 oracle
    case = workloads.sparse_linear(seed=0)
    A, N = case.A, case.B.shape[1]
    ref = O.spmm_f64(A.n_rows, N, A.row_offset, A.col_idx, f16_to_f64(A.values), f16_to_f64(case.B))
    B = dev_f16(case.B)
    q, s = T.quantize_rows_e4m3(dev(A.row_offset), dev_f16(A.values))
    qb, sb = T.quantize_cols_e4m3(B)
    ref8 = _e4m3_ref(A, q, s, qb, sb, N)
    for rpw in WIDE:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        d = plan.pack_dense(plan.pack(dev_f16(A.values)))
        d8 = plan.pack_dense(plan.pack(q))
        for t in TILES:
            got = T.spmm(plan, d, B, tile_n=t, dense=True).cpu().numpy()
            assert rel_l2(got, ref) <= 1e-5, (rpw, t)
            got8 = T.spmm(plan, d8, qb, row_scale=s, col_scale=sb, tile_n=t, dense=True).cpu().numpy()
            assert rel_l2(got8, ref8) <= 1e-5, (rpw, t)


@pytest.mark.parametrize("rpw", (64, 128))
def test_batch_block_units_against_oracle(cuda, rpw):
    # batch variants 22 / 23: forward and transposed block units, packed tiles
    layers, items = [], []
    for i in (0, 4, 9, 11):
        case = workloads.batch12(seed=0)[i]
        W = case.A
        pw = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
        pt = pw.transpose(rpw)
        w = pw.pack(dev_f16(W.values))
        L = dict(case=case, pw=pw, w=w, X=dev_f16(case.B), G=dev(case.extra["G"]),
                 Y=torch.full((W.n_rows, case.N), float("nan"), device=cuda),
                 dX=torch.full((W.n_cols, case.N), float("nan"), device=cuda))
        items += [T.spmm_item(pw, w, L["X"], L["Y"], 64, dense="pack", values_order=pw),
                  T.spmm_item(pt, w, L["G"], L["dX"], 64, dense="pack", values_order=pw)]
        layers.append(L)
    bt = T.Batch(items)
    bt.run()
    torch.cuda.synchronize()
    for L in layers:
        W, N = L["case"].A, L["case"].N
        a64 = f16_to_f64(W.values)
        ref_y = O.spmm_f64(W.n_rows, N, W.row_offset, W.col_idx, a64, f16_to_f64(L["case"].B))
        rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
        g16 = L["case"].extra["G"].astype(np.float16).astype(np.float64)
        ref_dx = O.spmm_f64(W.n_cols, N, rpT, cT, a64[perm], g16)
        assert rel_l2(L["Y"].cpu().numpy(), ref_y) <= 1e-5, W.name
        assert rel_l2(L["dX"].cpu().numpy(), ref_dx) <= 1e-5, W.name
