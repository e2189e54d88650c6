"""Panel-dense A tiles (tca_pack_dense / tca_spmm_dense): the tiles hold the
tiled values at the positions the entry slots name (checked element for
element against a host build from the exported tables), and SpMM on them is
bit-identical to SpMM on the tiled values -- forward and transposed plans,
F16 / BF16 / E4M3, every rows_per_wave and tile width, ragged and empty
shapes -- per call and inside the batched launch; both agree with the oracle."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_bf16, dev_f16, f16_to_f64, oracle_f32, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads
from tca_spmm.workloads import SparseMatrix

pytestmark = pytest.mark.gpu

RPWS = (16, 32, 64)
TILES = (16, 32, 64, 128)


def host_dense(plan, tiled_bits, src=None):
    """Dense tiles built on the host from the exported tables: tile
    s = c*PPG + i, element ((k/8)*16 + r)*8 + k%8 for entry slot r<<5 | k."""
    ex = plan.export()
    seg, slot = ex["seg_ptr"], ex["entry_slot"].astype(np.int64)
    out = np.zeros(plan.dense_elems, tiled_bits.dtype)
    tile_of_entry = np.repeat(np.arange(seg.size - 1), np.diff(seg))
    r, k = slot >> 5, slot & 31
    pos = tile_of_entry * 512 + ((k >> 3) * 16 + r) * 8 + (k & 7)
    out[pos] = tiled_bits if src is None else tiled_bits[src]
    return out


def _pattern_cases():
    yield "bernoulli_77x45", *gen.pattern("bernoulli", 77, 45, 0.3, seed=1, stream=5), 77, 45
    yield "lognormal_300x700", *gen.pattern("lognormal", 300, 700, 0.05, 0.5, seed=2, stream=6), 300, 700
    rp = np.zeros(71, np.int32)
    rp[65:] = 2                                    # one non-empty row, empty groups around it
    yield "one_row", rp, np.array([3, 29], np.int32), 70, 30


@pytest.mark.parametrize("rpw", RPWS)
def test_pack_dense_matches_host_build(cuda, rpw):
    for name, rp, ci, M, K in _pattern_cases():
        vals = gen.f32_to_f16_bits(gen.normal_f32(7, rpw, (ci.size,)))
        plan = T.Plan(M, K, rp, ci, rpw)
        tiled = plan.pack(dev_f16(vals))
        dense = plan.pack_dense(tiled)
        assert dense.numel() == plan.dense_elems == plan.n_chunks * (rpw // 16) * 512, name
        want = host_dense(plan, tiled.cpu().numpy().view(np.uint16))
        assert np.array_equal(dense.cpu().numpy().view(np.uint16), want), name


@pytest.mark.parametrize("rpw", RPWS)
def test_dense_spmm_bit_identical_f16(cuda, rpw):
    for name, rp, ci, M, K in _pattern_cases():
        A = SparseMatrix(name, M, K, rp, ci, gen.f32_to_f16_bits(gen.normal_f32(8, rpw, (ci.size,))))
        plan = T.Plan(M, K, rp, ci, rpw)
        tiled = plan.pack(dev_f16(A.values if A.nnz else np.zeros(1, np.uint16)))
        dense = plan.pack_dense(tiled)
        for N in (8, 40, 256):
            Bb = gen.f32_to_f16_bits(gen.normal_f32(8, N, (K, N)))
            B = dev_f16(Bb)
            for t in TILES:
                ref = T.spmm(plan, tiled, B, tile_n=t)
                got = T.spmm(plan, dense, B, tile_n=t, dense=True)
                assert torch.equal(got, ref), (name, N, t)
            assert rel_l2(got.cpu().numpy(), oracle_f32(A, Bb)) <= 1e-5


def test_dense_spmm_bf16_and_f32_b(cuda):
    case = workloads.plain_gpu(512, 2048, 0.9)
    A = case.A
    for rpw in RPWS:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        bf = gen.f32_to_bf16_bits(A.values.view(np.float16).astype(np.float32))
        for vals, Bt in ((dev_bf16(bf), dev_bf16(gen.f32_to_bf16_bits(case.B.view(np.float16).astype(np.float32)))),
                         (dev_f16(A.values), dev(case.B.view(np.float16).astype(np.float32)))):
            tiled = plan.pack(vals)
            dense = plan.pack_dense(tiled)
            for t in (32, 64):
                assert torch.equal(T.spmm(plan, dense, Bt, tile_n=t, dense=True), T.spmm(plan, tiled, Bt, tile_n=t))


def test_dense_spmm_fp8_with_scales(cuda):
    case = workloads.plain_gpu(2048, 512, 0.8)
    A = case.A
    q, s = T.quantize_rows_e4m3(dev(A.row_offset), dev_f16(A.values))
    qb, sb = T.quantize_cols_e4m3(dev_f16(case.B))
    for rpw in RPWS:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        tiled = plan.pack(q)
        dense = plan.pack_dense(tiled)
        assert np.array_equal(dense.cpu().numpy(), host_dense(plan, tiled.cpu().numpy()))
        for t in (16, 64, 128):
            ref = T.spmm(plan, tiled, qb, row_scale=s, col_scale=sb, tile_n=t)
            got = T.spmm(plan, dense, qb, row_scale=s, col_scale=sb, tile_n=t, dense=True)
            assert torch.equal(got, ref)


def test_dense_transposed_plans(cuda):
    case = workloads.sparse_linear(M=1024, K=512, N=256)
    W, G = case.A, case.extra["G"]
    fwd = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
    tiled = fwd.pack(dev_f16(W.values))
    rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
    ref64 = O.spmm_f64(W.n_cols, G.shape[1], rpT, cT, f16_to_f64(W.values)[perm], G.astype(np.float16).astype(np.float64))
    for rpw in RPWS:
        tr = fwd.transpose(rpw)
        dense = tr.pack_dense(tiled)
        want = host_dense(tr, tiled.cpu().numpy().view(np.uint16), tr.export()["entry_src"])
        assert np.array_equal(dense.cpu().numpy().view(np.uint16), want)
        for Gd in (dev(G), dev_f16(G.astype(np.float16).view(np.uint16))):
            for t in (16, 64):
                ref = T.spmm(tr, tiled, Gd, tile_n=t)
                got = T.spmm(tr, dense, Gd, tile_n=t, dense=True)
                assert torch.equal(got, ref), (rpw, t)
        assert rel_l2(got.cpu().numpy(), ref64) <= 1e-5


def test_dense_batch_items_match_tiled_items(cuda):
    """The batch kernel's dense variants (forward and transposed, F32 B) give
    the tiled variants' results: same units and the same split of a unit's
    chunks over the workgroup's waves, so only the order in which units split
    across workgroups add atomically can differ."""
    layers, items_t, items_d = [], [], []
    for case in list(workloads.batch12(seed=0))[:6]:
        W, X, G = case.A, case.B, case.extra["G"]
        fwd = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
        tr = fwd.transpose(16)
        w = fwd.pack(dev_f16(W.values))
        Xd, Gd = dev_f16(X), dev(G)
        L = dict(fwd=fwd, tr=tr, w=w, wd=fwd.pack_dense(w), wtd=tr.pack_dense(w), X=Xd, G=Gd)
        for k in ("Y", "Yd"):
            L[k] = torch.empty((W.n_rows, case.N), dtype=torch.float32, device=cuda)
        for k in ("dX", "dXd"):
            L[k] = torch.empty((W.n_cols, case.N), dtype=torch.float32, device=cuda)
        tile = 64 if case.N >= 64 else 32
        items_t += [T.spmm_item(fwd, w, Xd, L["Y"], tile), T.spmm_item(tr, w, Gd, L["dX"], tile)]
        items_d += [T.spmm_item(fwd, L["wd"], Xd, L["Yd"], tile, dense=True),
                    T.spmm_item(tr, L["wtd"], Gd, L["dXd"], tile, dense=True)]
        layers.append(L)
    bt, bd = T.Batch(items_t), T.Batch(items_d)
    bt.run()
    bd.run()
    torch.cuda.synchronize()
    for L in layers:
        for a, b in (("Y", "Yd"), ("dX", "dXd")):
            assert rel_l2(L[b].cpu().numpy(), L[a].cpu().numpy()) <= 1e-6   # split units add atomically
    for L in layers:
        ref = T.spmm(L["fwd"], L["w"], L["X"], tile_n=64)
        assert rel_l2(L["Yd"].cpu().numpy(), ref.cpu().numpy()) <= 1e-6


def test_dense_rejections(cuda):
    case = workloads.plain_gpu(2048, 512, 0.9)
    A = case.A
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, 16)
    tiled = plan.pack(dev_f16(A.values))
    with pytest.raises(ValueError):
        T.spmm(plan, tiled, dev_f16(case.B), dense=True)            # tiled values, not tiles
    wide = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, 128)
    q, _ = T.quantize_rows_e4m3(dev(A.row_offset), dev_f16(A.values))
    qb, _ = T.quantize_cols_e4m3(dev_f16(case.B))
    with pytest.raises(T.TcaError) as e:                             # wide groups, tiled values: 16-bit only
        T.spmm(wide, wide.pack(q), qb)
    assert e.value.name == "TCA_ERR_UNSUPPORTED"
    with pytest.raises(T.TcaError) as e:                             # E4M3 tiles take an E4M3 B
        T.spmm(wide, wide.pack_dense(wide.pack(q)), dev_f16(case.B), dense=True)
    assert e.value.name == "TCA_ERR_UNSUPPORTED"


@pytest.mark.parametrize("rpw", [128, 256])
def test_dense_wide_plans(cuda, rpw):
    """Wide groups (one workgroup of rows/32 waves sharing each B chunk) on
    panel-dense tiles: forward (F16 and BF16) and transposed (F32 gradient),
    tile_n 64 / 128, against the tiled-values launch.  These launches split
    the few long groups over workgroups that add atomically, so the two agree
    to FP32 summation order, not bit for bit."""
    case = workloads.sparse_linear(M=2048, K=1024, N=256)
    W, X, G = case.A, case.B, case.extra["G"]
    plan = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
    bf = gen.f32_to_bf16_bits(W.values.view(np.float16).astype(np.float32))
    Xbf = dev_bf16(gen.f32_to_bf16_bits(X.view(np.float16).astype(np.float32)))
    for vals, B in ((dev_f16(W.values), dev_f16(X)), (dev_bf16(bf), Xbf)):
        tiled = plan.pack(vals)
        dense = plan.pack_dense(tiled)
        assert np.array_equal(dense.cpu().view(torch.int16).numpy(),
                              host_dense(plan, tiled.cpu().view(torch.int16).numpy()))
        for t in (64, 128):
            got = T.spmm(plan, dense, B, tile_n=t, dense=True).cpu().numpy()
            assert rel_l2(got, T.spmm(plan, tiled, B, tile_n=t).cpu().numpy()) <= 1e-6
    fwd16 = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
    w16 = fwd16.pack(dev_f16(W.values))
    tr = fwd16.transpose(rpw)
    for t in (64, 128):
        got = T.spmm(tr, tr.pack_dense(w16), dev(G), tile_n=t, dense=True).cpu().numpy()
        assert rel_l2(got, T.spmm(tr, w16, dev(G), tile_n=t).cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("rpw,tile,g_f32", [(32, 32, False), (32, 64, False), (64, 32, False), (32, 32, True),
                                            (32, 64, True), (64, 32, True)])
def test_dense_batch_wider_plans(cuda, rpw, tile, g_f32):
    """Dense batch variants on 32- and 64-row plans (forward with an F16 B,
    transposed with an F16 or F32 gradient -- the batch reads an F32 one
    through its F16 copy, so every F16 shape serves it) against per-call dense
    launches."""
    keep, items, refs = [], [], []
    for case in list(workloads.batch12(seed=0))[:4] + list(workloads.batch12(seed=0))[8:10]:
        W, X, G = case.A, case.B, case.extra["G"]
        fwd16 = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
        w16 = fwd16.pack(dev_f16(W.values))
        if g_f32:                                  # transposed product, F32 gradient
            plan = fwd16.transpose(rpw)
            a, B = plan.pack_dense(w16), dev(G)
        else:                                      # forward product, F16 activations
            plan = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
            a, B = plan.pack_dense(plan.pack(dev_f16(W.values))), dev_f16(X)
        out = torch.empty((plan.n_rows, case.N), dtype=torch.float32, device=cuda)
        keep.append((fwd16, plan, a, B, out))
        items.append(T.spmm_item(plan, a, B, out, tile, dense=True))
        refs.append(T.spmm(plan, a, B, tile_n=tile, dense=True))
    bt = T.Batch(items)
    bt.run()
    torch.cuda.synchronize()
    for (_, _, _, _, out), ref in zip(keep, refs):
        assert rel_l2(out.cpu().numpy(), ref.cpu().numpy()) <= 1e-6


def test_dense_empty_plan_and_empty_groups(cuda):
    """nnz = 0: no tiles, C all zeros; a plan whose only entries sit in one
    group: the dense launch equals the tiled one and zero rows stay zero."""
    M, K, N = 70, 30, 48
    B = dev_f16(gen.f32_to_f16_bits(gen.normal_f32(3, 1, (K, N))))
    empty = T.Plan(M, K, np.zeros(M + 1, np.int32), np.zeros(0, np.int32), 32)
    assert empty.dense_elems == 0
    tiles = empty.pack_dense(torch.zeros(1, dtype=torch.float16, device=cuda)[:0])
    C = T.spmm(empty, tiles, B, dense=True)
    torch.cuda.synchronize()
    assert tiles.numel() == 0 and torch.count_nonzero(C).item() == 0
    rp = np.zeros(M + 1, np.int32)
    rp[65:] = 2
    plan = T.Plan(M, K, rp, np.array([3, 29], np.int32), 64)
    tiled = plan.pack(dev_f16(np.array([0x3C00, 0xC000], np.uint16)))
    for t in TILES:
        got = T.spmm(plan, plan.pack_dense(tiled), B, tile_n=t, dense=True)
        assert torch.equal(got, T.spmm(plan, tiled, B, tile_n=t))
        assert torch.count_nonzero(torch.cat([got[:64], got[65:]])).item() == 0


def test_dense_transposed_e4m3_and_bf16(cuda):
    """Transposed plans with E4M3 (scaled) and BF16 values: dense tiles built
    through the entry map give the indirect launch's bits."""
    case = workloads.sparse_linear(M=1024, K=512, N=128)
    W, G = case.A, case.extra["G"]
    fwd = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
    q, s = T.quantize_rows_e4m3(dev(W.row_offset), dev_f16(W.values))
    gq, gs = T.quantize_cols_e4m3(dev_f16(G.astype(np.float16).view(np.uint16)))
    bf = dev_bf16(gen.f32_to_bf16_bits(W.values.view(np.float16).astype(np.float32)))
    gbf = dev_bf16(gen.f32_to_bf16_bits(G))
    for rpw in RPWS:
        tr = fwd.transpose(rpw)
        for vals, Gd in ((fwd.pack(q), gq), (fwd.pack(bf), gbf)):
            dense = tr.pack_dense(vals)
            assert np.array_equal(dense.cpu().view(torch.uint8).numpy() if dense.dtype == torch.uint8 else
                                  dense.cpu().view(torch.int16).numpy(),
                                  host_dense(tr, vals.cpu().view(torch.uint8).numpy() if vals.dtype == torch.uint8 else
                                             vals.cpu().view(torch.int16).numpy(), tr.export()["entry_src"]))
            for t in (32, 64):
                sc = dict(col_scale=gs) if vals.dtype == torch.uint8 else {}
                assert torch.equal(T.spmm(tr, dense, Gd, tile_n=t, dense=True, **sc), T.spmm(tr, vals, Gd, tile_n=t, **sc))
