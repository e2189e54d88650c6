"""Wide-group plans (128 / 256 rows per group, one workgroup of rows/32 waves
sharing every gathered B chunk through LDS, MFMA B operands read with the
transposing LDS load): parity with the CPU oracle on the same shapes and edge
cases as the 16..64-row plans, for the forward product, the transposed
product (dX with an FP32 G) and the SDDMM (dW) on the same plan.  Bars as in
test_gpu_spmm: rel-L2 <= 1e-5 against the FP32-accumulate oracle on the same
inputs, <= 1e-3 against the FP64 oracle; exact cases bit-exact.
"""
import numpy as np
import pytest
import torch

from helpers import dev, dev_bf16, dev_f16, f16_to_f64, oracle_f32, oracle_f64, rel_l2, run_spmm_f16
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads
from tca_spmm.workloads import SparseMatrix

pytestmark = pytest.mark.gpu

WIDE = (128, 256)
TILES = (64, 128)


def _check(A, B, C, tol32=1e-5):
    ref32 = oracle_f32(A, B)
    assert rel_l2(C, ref32) <= tol32, rel_l2(C, ref32)
    assert rel_l2(C, oracle_f64(A, B)) <= 1e-3
    ok, nf = gen.compare_mat(ref32, C)
    assert ok, nf


@pytest.mark.parametrize("rpw", WIDE)
@pytest.mark.parametrize("tile_n", TILES)
def test_wide_plain_cpu(cuda, rpw, tile_n):
    case = workloads.plain_cpu()
    C, plan = run_spmm_f16(case.A, case.B, rpw, tile_n)
    assert plan.rows_per_wave == rpw
    _check(case.A, case.B, C)


def test_wide_plain_gpu_suite(cuda):
    for case in workloads.plain_gpu_suite():
        for rpw in WIDE:
            C, _ = run_spmm_f16(case.A, case.B, rpw, 0)
            _check(case.A, case.B, C)


@pytest.mark.parametrize("N", [8, 24, 40, 200, 264])
def test_wide_ragged_shapes_and_tails(cuda, N):
    # 77 rows: a single, partial group; N not a multiple of the tile
    M, K = 77, 45
    rp, ci = gen.pattern("bernoulli", M, K, 0.3, seed=1, stream=N)
    A = SparseMatrix("ragged", M, K, rp, ci, gen.f32_to_f16_bits(gen.normal_f32(1, 2, (ci.size,))))
    B = gen.f32_to_f16_bits(gen.normal_f32(1, 3, (K, N)))
    for rpw in WIDE:
        for t in TILES:
            C, _ = run_spmm_f16(A, B, rpw, t)
            _check(A, B, C)


def test_wide_exact_cases(cuda):
    K, N = 100, 64
    B = gen.f32_to_f16_bits(gen.normal_f32(2, 1, (K, N)))
    Bf = B.view(np.float16).astype(np.float32)
    ident = SparseMatrix("I", K, K, np.arange(K + 1, dtype=np.int32), np.arange(K, dtype=np.int32),
                         np.full(K, 0x3C00, np.uint16))
    M = 300
    cols = (np.arange(M) * 37) % K
    a = gen.f32_to_f16_bits(gen.normal_f32(2, 2, (M,)))
    single = SparseMatrix("single", M, K, np.arange(M + 1, dtype=np.int32), cols.astype(np.int32), a)
    for rpw in WIDE:
        C, _ = run_spmm_f16(ident, B, rpw, 64)
        assert np.array_equal(C, Bf)
        C, _ = run_spmm_f16(single, B, rpw, 64)
        assert np.array_equal(C, a.view(np.float16).astype(np.float32)[:, None] * Bf[cols])


def test_wide_empty_matrix_and_empty_groups(cuda):
    M, K, N = 600, 30, 48
    B = gen.f32_to_f16_bits(gen.normal_f32(3, 1, (K, N)))
    empty = SparseMatrix("empty", M, K, np.zeros(M + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint16))
    rp = np.zeros(M + 1, np.int32)
    rp[301:] = 2                                  # only row 300 has entries
    one = SparseMatrix("one_row", M, K, rp, np.array([3, 29], np.int32), np.array([0x3C00, 0xC000], np.uint16))
    Bf = B.view(np.float16).astype(np.float32)
    for rpw in WIDE:
        C, plan = run_spmm_f16(empty, B, rpw, 64)
        assert plan.n_chunks == 0 and np.all(C == 0)
        C, _ = run_spmm_f16(one, B, rpw, 64)
        assert np.array_equal(C[300], Bf[3] - 2 * Bf[29]) and np.all(np.delete(C, 300, axis=0) == 0)


def test_wide_long_rows_dense_panels_and_split_groups(cuda):
    # rows of up to 5999 of 6000 columns: panel tiles far above the 64 entries
    # a lane prefetches, and groups long enough to be split across workgroups
    M, K, N = 300, 6000, 128
    rng = np.random.default_rng(4)
    lens = np.concatenate([[5999, 3000, 1], rng.integers(0, 400, M - 3)])
    cols = [np.sort(rng.choice(K, size=l, replace=False)) for l in lens]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate(cols).astype(np.int32)
    A = SparseMatrix("long", M, K, rp, ci, gen.f32_to_f16_bits(gen.normal_f32(4, 1, (ci.size,))))
    B = gen.f32_to_f16_bits(gen.normal_f32(4, 2, (K, N)))
    for rpw in WIDE:
        for t in TILES:
            C, _ = run_spmm_f16(A, B, rpw, t)
            _check(A, B, C)


def test_wide_precision_workload_fp16_and_bf16(cuda):
    case = workloads.precision()
    A, N = case.A, case.N
    for rpw in WIDE:
        C, plan = run_spmm_f16(A, case.B, rpw, 64)
        _check(A, case.B, C)
    abf = gen.f32_to_bf16_bits(A.values.view(np.float16).astype(np.float32))
    bbf = gen.f32_to_bf16_bits(case.B.view(np.float16).astype(np.float32))
    ref_q = O.spmm_f64(A.n_rows, N, A.row_offset, A.col_idx, O.bf16_bits_to_f32(abf).astype(np.float64),
                       O.bf16_bits_to_f32(bbf).astype(np.float64))
    for rpw in WIDE:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        C = T.spmm(plan, plan.pack(dev_bf16(abf)), dev_bf16(bbf), tile_n=128).cpu().numpy()
        assert rel_l2(C, ref_q) <= 1e-5


def test_wide_transposed_dx(cuda):
    case = workloads.sparse_linear(M=1024, K=512, N=256)
    W, G = case.A, case.extra["G"]
    fwd = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
    tiled = fwd.pack(dev_f16(W.values))
    rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
    G16 = G.astype(np.float16).astype(np.float64)
    ref_q = O.spmm_f64(W.n_cols, G.shape[1], rpT, cT, f16_to_f64(W.values)[perm], G16)
    ref = O.spmm_f64(W.n_cols, G.shape[1], rpT, cT, f16_to_f64(W.values)[perm], G.astype(np.float64))
    for rpw in WIDE:
        tr = fwd.transpose(rpw)
        for t in TILES:
            dX = T.spmm(tr, tiled, dev(G), tile_n=t).cpu().numpy()
            assert rel_l2(dX, ref_q) <= 1e-5
            assert rel_l2(dX, ref) <= 1e-3
    # a wide forward plan as the source of the transposed plan's value order
    wide = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 128)
    dX = T.spmm(wide.transpose(16), wide.pack(dev_f16(W.values)), dev(G), tile_n=64).cpu().numpy()
    assert rel_l2(dX, ref_q) <= 1e-5


def test_wide_sddmm_dw(cuda):
    case = workloads.sparse_linear(M=600, K=384, N=256)
    W, X, G = case.A, case.B, case.extra["G"]
    G16 = G.astype(np.float16).astype(np.float64)
    ref_q = O.sddmm_f64(W.n_rows, G.shape[1], W.row_offset, W.col_idx, G16, f16_to_f64(X))
    for rpw in WIDE:
        plan = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
        for Gd in (dev(G), dev_f16(G.astype(np.float16).view(np.uint16))):
            csr = plan.unpack(T.sddmm(plan, Gd, dev_f16(X))).cpu().numpy()
            assert rel_l2(csr, ref_q) <= 1e-5


def test_wide_rejects_fp8_and_small_tiles(cuda):
    case = workloads.plain_cpu()
    A = case.A
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, 128)
    w = plan.pack(dev_f16(A.values))
    with pytest.raises(T.TcaError) as e:
        T.spmm(plan, w, dev_f16(case.B), tile_n=32)
    assert e.value.name == "TCA_ERR_UNSUPPORTED"
    q, s = T.quantize_rows_e4m3(dev(A.row_offset), dev_f16(A.values))
    qb, sb = T.quantize_cols_e4m3(dev_f16(case.B))
    with pytest.raises(T.TcaError) as e:
        T.spmm(plan, plan.pack(q), qb, row_scale=s, col_scale=sb, tile_n=64)
    assert e.value.name == "TCA_ERR_UNSUPPORTED"
