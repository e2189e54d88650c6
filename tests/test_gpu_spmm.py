"""GPU parity of the HIP SpMM path against the CPU oracle, through the C ABI.

Bars: FP16/BF16/FP8 inputs with FP32 accumulation must match the oracle run on
the same (quantised) inputs to rel-L2 <= 1e-5 (accumulation-order noise only),
and the FP64 oracle on the original values to rel-L2 <= 1e-3 (the path's
stated tolerance) for FP16; the reference's compare_mat rule (tol 1e-3) must pass.
Exact cases (identity, single nonzero per row, empty rows) are bit-exact.
"""
import os
import subprocess

import numpy as np
import pytest
import torch

from conftest import BUNDLED_SMTX, PKG
from helpers import dev, dev_bf16, dev_f16, f16_to_f64, oracle_f32, oracle_f64, rel_l2, run_spmm_f16
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads
from tca_spmm.workloads import SparseMatrix

pytestmark = pytest.mark.gpu

RPWS = (16, 32, 64)
TILES = (16, 32, 64, 128)


def _check(A, B, C, tol32=1e-5):
    ref32 = oracle_f32(A, B)
    ref64 = oracle_f64(A, B)
    assert rel_l2(C, ref32) <= tol32, rel_l2(C, ref32)
    assert rel_l2(C, ref64) <= 1e-3
    ok, nf = gen.compare_mat(ref32, C)
    assert ok, nf


@pytest.mark.parametrize("rpw", RPWS)
@pytest.mark.parametrize("tile_n", TILES)
def test_plain_cpu_all_configs(cuda, rpw, tile_n):
    case = workloads.plain_cpu()
    C, _ = run_spmm_f16(case.A, case.B, rpw, tile_n)
    _check(case.A, case.B, C)


def test_plain_gpu_suite(cuda):
    for case in workloads.plain_gpu_suite():
        for rpw in RPWS:
            C, _ = run_spmm_f16(case.A, case.B, rpw, 0)
            _check(case.A, case.B, C)


def test_bundled_smtx_ones_and_values(cuda):
    M, K, rp, ci = gen.smtx_read(BUNDLED_SMTX)
    ones = np.full(ci.size, 0x3C00, np.uint16)
    rnd = gen.f32_to_f16_bits(gen.normal_f32(0, 77, (ci.size,)))
    B = gen.f32_to_f16_bits(gen.normal_f32(0, 78, (K, 256)))
    for vals in (ones, rnd):
        A = SparseMatrix("bundled", M, K, rp, ci, vals)
        for rpw in RPWS:
            C, _ = run_spmm_f16(A, B, rpw, 64)
            _check(A, B, C)
            empty = np.diff(rp) == 0
            assert np.all(C[empty] == 0)


@pytest.mark.parametrize("N", [8, 24, 40, 200, 264])
def test_ragged_shapes_and_tails(cuda, N):
    M, K = 77, 45
    rp, ci = gen.pattern("bernoulli", M, K, 0.3, seed=1, stream=N)
    A = SparseMatrix("ragged", M, K, rp, ci, gen.f32_to_f16_bits(gen.normal_f32(1, 2, (ci.size,))))
    B = gen.f32_to_f16_bits(gen.normal_f32(1, 3, (K, N)))
    for rpw in RPWS:
        for t in TILES:
            C, _ = run_spmm_f16(A, B, rpw, t)
            _check(A, B, C)


def test_exact_cases(cuda):
    K, N = 100, 64
    B = gen.f32_to_f16_bits(gen.normal_f32(2, 1, (K, N)))
    Bf = B.view(np.float16).astype(np.float32)
    ident = SparseMatrix("I", K, K, np.arange(K + 1, dtype=np.int32), np.arange(K, dtype=np.int32),
                         np.full(K, 0x3C00, np.uint16))
    C, _ = run_spmm_f16(ident, B, 16, 32)
    assert np.array_equal(C, Bf)
    # one nonzero per row at a random column: C = a * B[j] exactly
    M = 150
    cols = (np.arange(M) * 37) % K
    a = gen.f32_to_f16_bits(gen.normal_f32(2, 2, (M,)))
    single = SparseMatrix("single", M, K, np.arange(M + 1, dtype=np.int32), cols.astype(np.int32), a)
    for rpw in RPWS:
        C, _ = run_spmm_f16(single, B, rpw, 64)
        assert np.array_equal(C, a.view(np.float16).astype(np.float32)[:, None] * Bf[cols])


def test_empty_matrix_and_empty_groups(cuda):
    M, K, N = 70, 30, 48
    B = gen.f32_to_f16_bits(gen.normal_f32(3, 1, (K, N)))
    empty = SparseMatrix("empty", M, K, np.zeros(M + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint16))
    C, plan = run_spmm_f16(empty, B, 32, 16)
    assert plan.n_chunks == 0 and np.all(C == 0)
    rp = np.zeros(M + 1, np.int32)
    rp[65:] = 2                                   # only row 64 has entries
    A = SparseMatrix("one_row", M, K, rp, np.array([3, 29], np.int32), np.array([0x3C00, 0xC000], np.uint16))
    C, _ = run_spmm_f16(A, B, 64, 32)
    Bf = B.view(np.float16).astype(np.float32)
    assert np.array_equal(C[64], Bf[3] - 2 * Bf[29]) and np.all(np.delete(C, 64, axis=0) == 0)


def test_long_rows_many_chunks(cuda):
    M, K, N = 40, 6000, 128
    lens = np.array([5999, 3000, 1] + [50] * 37)
    rng = np.random.default_rng(4)
    cols = [np.sort(rng.choice(K, size=l, replace=False)) for l in lens]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    ci = np.concatenate(cols).astype(np.int32)
    A = SparseMatrix("long", M, K, rp, ci, gen.f32_to_f16_bits(gen.normal_f32(4, 1, (ci.size,))))
    B = gen.f32_to_f16_bits(gen.normal_f32(4, 2, (K, N)))
    for rpw in RPWS:
        C, _ = run_spmm_f16(A, B, rpw, 128)
        _check(A, B, C)


def test_precision_workload_fp16_full_size(cuda):
    case = workloads.precision()
    C, _ = run_spmm_f16(case.A, case.B, 32, 64)
    _check(case.A, case.B, C)


def test_bf16(cuda):
    case = workloads.plain_gpu(2048, 512, 0.9)
    A = case.A
    va = gen.f32_to_bf16_bits(A.values.view(np.float16).astype(np.float32))
    vb = gen.f32_to_bf16_bits(case.B.view(np.float16).astype(np.float32))
    for rpw in RPWS:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        C = T.spmm(plan, plan.pack(dev_bf16(va)), dev_bf16(vb), tile_n=64).cpu().numpy()
        ref = O.spmm_f64(A.n_rows, case.N, A.row_offset, A.col_idx, O.bf16_bits_to_f32(va).astype(np.float64),
                         O.bf16_bits_to_f32(vb).astype(np.float64))
        assert rel_l2(C, ref) <= 1e-5
        assert rel_l2(C, oracle_f64(A, case.B)) < 1e-2          # format error of bf16 inputs


def _quantize_ref_rows(A):
    v = A.values.view(np.float16).astype(np.float32)
    q = np.zeros(A.nnz, np.uint8)
    s = np.ones(A.n_rows, np.float32)
    for r in range(A.n_rows):
        seg = v[A.row_offset[r]:A.row_offset[r + 1]]
        amax = np.float32(np.abs(seg).max()) if seg.size else np.float32(0)
        s[r] = amax / np.float32(448.0) if amax > 0 else np.float32(1.0)
        q[A.row_offset[r]:A.row_offset[r + 1]] = O.f32_to_e4m3_fast(seg / s[r])
    return q, s


def test_fp8_quantisers_bit_exact(cuda):
    case = workloads.plain_gpu(512, 2048, 0.9)
    A = case.A
    q, s = T.quantize_rows_e4m3(dev(A.row_offset), dev_f16(A.values))
    qr, sr = _quantize_ref_rows(A)
    assert np.array_equal(s.cpu().numpy(), sr) and np.array_equal(q.cpu().numpy(), qr)
    qb, sb = T.quantize_cols_e4m3(dev_f16(case.B))
    Bf = case.B.view(np.float16).astype(np.float32)
    amax = np.abs(Bf).max(axis=0)
    sbr = np.where(amax > 0, amax / np.float32(448.0), np.float32(1.0)).astype(np.float32)
    assert np.array_equal(sb.cpu().numpy(), sbr)
    assert np.array_equal(qb.cpu().numpy(), O.f32_to_e4m3_fast(Bf / sbr[None, :]))
    # column ranges of B as views (row pitch 256): offset, odd-width, unaligned start
    Bd = dev_f16(case.B)
    for lo, hi in ((8, 136), (0, 250), (3, 131)):
        qv, sv = T.quantize_cols_e4m3(Bd[:, lo:hi])
        Bv = Bf[:, lo:hi]
        amax = np.abs(Bv).max(axis=0)
        svr = np.where(amax > 0, amax / np.float32(448.0), np.float32(1.0)).astype(np.float32)
        assert np.array_equal(sv.cpu().numpy(), svr)
        assert np.array_equal(qv.cpu().numpy(), O.f32_to_e4m3_fast(Bv / svr[None, :]))


def test_fp8_spmm_with_scales(cuda):
    case = workloads.precision(M=1024, K=2048, N=256)
    A = case.A
    q, s = T.quantize_rows_e4m3(dev(A.row_offset), dev_f16(A.values))
    qb, sb = T.quantize_cols_e4m3(dev_f16(case.B))
    for rpw in RPWS:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        C = T.spmm(plan, plan.pack(q), qb, row_scale=s, col_scale=sb, tile_n=64).cpu().numpy()
        qa = O.e4m3_to_f32(q.cpu().numpy()).astype(np.float64)
        qbf = O.e4m3_to_f32(qb.cpu().numpy()).astype(np.float64)
        ref = O.spmm_f64(A.n_rows, case.N, A.row_offset, A.col_idx, qa, qbf)
        ref *= s.cpu().numpy().astype(np.float64)[:, None] * sb.cpu().numpy().astype(np.float64)[None, :]
        assert rel_l2(C, ref) <= 1e-5
        assert rel_l2(C, oracle_f64(A, case.B)) < 0.1           # format error of e4m3 inputs


def test_transposed_spmm_dx(cuda):
    case = workloads.sparse_linear(M=1024, K=512, N=256)
    W, G = case.A, case.extra["G"]
    fwd = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
    tiled = fwd.pack(dev_f16(W.values))
    rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
    G16 = G.astype(np.float16).astype(np.float64)
    ref_q = O.spmm_f64(W.n_cols, G.shape[1], rpT, cT, f16_to_f64(W.values)[perm], G16)
    ref = O.spmm_f64(W.n_cols, G.shape[1], rpT, cT, f16_to_f64(W.values)[perm], G.astype(np.float64))
    for rpw in RPWS:
        tr = fwd.transpose(rpw)
        dX = T.spmm(tr, tiled, dev(G), tile_n=64).cpu().numpy()
        assert rel_l2(dX, ref_q) <= 1e-5
        assert rel_l2(dX, ref) <= 1e-3


def test_sddmm_dw(cuda):
    case = workloads.sparse_linear(M=512, K=384, N=256)
    W, X, G = case.A, case.B, case.extra["G"]
    G16 = G.astype(np.float16).astype(np.float64)
    ref_q = O.sddmm_f64(W.n_rows, G.shape[1], W.row_offset, W.col_idx, G16, f16_to_f64(X))
    ref = O.sddmm_f64(W.n_rows, G.shape[1], W.row_offset, W.col_idx, G.astype(np.float64), f16_to_f64(X))
    for rpw in RPWS:
        plan = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
        for Gd in (dev(G), dev_f16(G.astype(np.float16).view(np.uint16))):
            out = T.sddmm(plan, Gd, dev_f16(X))
            assert torch.equal(out, T.sddmm(plan, Gd, dev_f16(X)))     # fixed summation order
            csr = plan.unpack(out).cpu().numpy()
            assert rel_l2(csr, ref_q) <= 1e-5
            assert rel_l2(csr, ref) <= 1e-3


def test_pack_unpack_round_trip(cuda):
    case = workloads.plain_gpu(2048, 512, 0.8)
    A = case.A
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, 32)
    v = dev_f16(A.values)
    t = plan.pack(v)
    assert torch.equal(plan.unpack(t), v)
    src = plan.export()["entry_src"]
    assert np.array_equal(t.cpu().numpy().view(np.uint16), A.values[src])


def test_invalid_arguments_are_rejected(cuda):
    case = workloads.plain_cpu()
    A = case.A
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, 16)
    t = plan.pack(dev_f16(A.values))
    B = dev_f16(case.B)
    with pytest.raises(T.TcaError) as e:
        T.spmm(plan, t, B[:, :12].contiguous())
    assert e.value.name == "TCA_ERR_UNSUPPORTED"
    with pytest.raises(T.TcaError) as e:
        T.spmm(plan, t, B, tile_n=48)
    assert e.value.name == "TCA_ERR_INVALID_ARG"
    with pytest.raises(T.TcaError) as e:
        T.spmm(plan, t, dev(np.zeros((512, 256), np.uint8)))
    assert e.value.name == "TCA_ERR_UNSUPPORTED"
    host_only = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, 16, device=False)
    with pytest.raises(ValueError, match="host-only"):
        T.spmm(host_only, t, B)                 # the wrapper refuses it ...
    import ctypes
    from tca_spmm import _lib
    C = torch.empty((A.n_rows, 256), dtype=torch.float32, device=cuda)
    with pytest.raises(T.TcaError) as e:        # ... and so does the C ABI
        _lib.call("tca_spmm", host_only.handle, T.F16, ctypes.c_void_p(t.data_ptr()), T.F16,
                  ctypes.c_void_p(B.data_ptr()), 256, 256, ctypes.c_void_p(C.data_ptr()), 256, None, None, 0, None)
    assert e.value.name == "TCA_ERR_INVALID_ARG"


def test_dropin_driver_bundled():
    exe = os.path.join(PKG, "tca_test")
    if not os.path.exists(exe):
        pytest.skip("tca_test not built")
    r = subprocess.run([exe, BUNDLED_SMTX], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "preprocess" in r.stdout and "our:" in r.stdout and "validation failed" not in r.stdout


def test_bf16_full_size_standard_plans(cuda):
    # 4096 x 4096 at 0.9, N = 512 in BF16 on the one-wave-per-group plans, tiled
    # values and panel-dense tiles, against the FP64 oracle on the BF16 inputs
    case = workloads.precision()
    A = case.A
    va = gen.f32_to_bf16_bits(A.values.view(np.float16).astype(np.float32))
    vb = gen.f32_to_bf16_bits(case.B.view(np.float16).astype(np.float32))
    ref = O.spmm_f64(A.n_rows, case.N, A.row_offset, A.col_idx, O.bf16_bits_to_f32(va).astype(np.float64),
                     O.bf16_bits_to_f32(vb).astype(np.float64))
    B = dev_bf16(vb)
    for rpw in RPWS:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        t = plan.pack(dev_bf16(va))
        assert rel_l2(T.spmm(plan, t, B, tile_n=64).cpu().numpy(), ref) <= 1e-5, rpw
        assert rel_l2(T.spmm(plan, plan.pack_dense(t), B, tile_n=64, dense=True).cpu().numpy(), ref) <= 1e-5, rpw


def test_fp8_full_size_standard_plans(cuda):
    # 4096 x 4096 at 0.9, N = 512 in E4M3 with per-row / per-column scales,
    # against the FP64 oracle on the quantised inputs and scales
    case = workloads.precision()
    A = case.A
    q, s = T.quantize_rows_e4m3(dev(A.row_offset), dev_f16(A.values))
    qb, sb = T.quantize_cols_e4m3(dev_f16(case.B))
    qa = O.e4m3_to_f32(q.cpu().numpy()).astype(np.float64)
    qbf = O.e4m3_to_f32(qb.cpu().numpy()).astype(np.float64)
    ref = O.spmm_f64(A.n_rows, case.N, A.row_offset, A.col_idx, qa, qbf)
    ref *= s.cpu().numpy().astype(np.float64)[:, None] * sb.cpu().numpy().astype(np.float64)[None, :]
    for rpw in RPWS:
        plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
        t = plan.pack(q)
        C = T.spmm(plan, t, qb, row_scale=s, col_scale=sb, tile_n=64).cpu().numpy()
        assert rel_l2(C, ref) <= 1e-5, rpw
        Cd = T.spmm(plan, plan.pack_dense(t), qb, row_scale=s, col_scale=sb, tile_n=64, dense=True).cpu().numpy()
        assert rel_l2(Cd, ref) <= 1e-5, rpw
    assert rel_l2(C, oracle_f64(A, case.B)) < 0.1                # format error of E4M3 inputs
