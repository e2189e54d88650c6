"""The library baselines bench.py times beside the HIP kernels -- rocSPARSE CSR
SpMM, hipBLASLt dense GEMM (the reference's cuSPARSE / cuBLAS comparisons) and
rocSPARSE SDDMM (the dW counterpart) -- computed on the same inputs agree with
the oracle (SpMM and GEMM to 1e-5; rocSPARSE's FP16 SDDMM to 1e-3, its own
precision), so their times are times of the same product."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_f16, f16_to_f64, oracle_f32, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads

pytestmark = pytest.mark.gpu


def _case():
    A = workloads.sparse("libs_300x200", 300, 200, 0.05, "lognormal", 0.5, 0, 7)
    return A, workloads.dense_f16(0, 7, 0, 3, (200, 64))


def test_rocsparse_spmm_matches_oracle():
    A, B = _case()
    C = torch.zeros((A.n_rows, 64), dtype=torch.float32, device="cuda")
    T.rocsparse_spmm(A.n_rows, 64, A.n_cols, dev(A.row_offset), dev(A.col_idx), dev_f16(A.values), dev_f16(B), C)
    torch.cuda.synchronize()
    assert rel_l2(C.cpu().numpy(), oracle_f32(A, B)) <= 1e-5


def test_hipblaslt_gemm_matches_oracle():
    A, B = _case()
    dense = gen.csr_to_dense_f16(A.n_rows, A.n_cols, A.row_offset, A.col_idx, A.values)
    C = torch.zeros((A.n_rows, 64), dtype=torch.float32, device="cuda")
    T.hipblaslt_gemm(A.n_rows, 64, A.n_cols, dev_f16(dense), dev_f16(B), C)
    torch.cuda.synchronize()
    assert rel_l2(C.cpu().numpy(), oracle_f32(A, B)) <= 1e-5


def test_rocsparse_sddmm_matches_oracle_and_tca_sddmm():
    A, X = _case()                               # W (300 x 200), X (200 x 64)
    G = gen.f32_to_f16_bits(gen.normal_f32(1, 2, (A.n_rows, 64)))
    out = torch.zeros(A.nnz, dtype=torch.float32, device="cuda")
    T.rocsparse_sddmm(A.n_rows, A.n_cols, 64, dev(A.row_offset), dev(A.col_idx), dev_f16(G), dev_f16(X), out)
    torch.cuda.synchronize()
    ref = O.sddmm_f64(A.n_rows, 64, A.row_offset, A.col_idx, f16_to_f64(G), f16_to_f64(X))
    # rocSPARSE's FP16 SDDMM lands ~2e-4 from the FP64 product on this MI355X
    # image (its FP16 path does not keep a full FP32 sum), inside the 1e-3 bar;
    # tca_sddmm accumulates in FP32 and stays at the 1e-5 level
    assert rel_l2(out.cpu().numpy(), ref) <= 1e-3
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, 16)
    ours = plan.unpack(T.sddmm(plan, dev_f16(G), dev_f16(X)))
    assert rel_l2(ours.cpu().numpy(), ref) <= 1e-5


def test_rocsparse_transposed_spmm_matches_oracle():
    A, _ = _case()                               # A (300 x 200): C (200 x 64) = A^T G
    G = gen.f32_to_f16_bits(gen.normal_f32(3, 4, (A.n_rows, 64)))
    C = torch.zeros((A.n_cols, 64), dtype=torch.float32, device="cuda")
    T.rocsparse_spmm_t(A.n_rows, 64, A.n_cols, dev(A.row_offset), dev(A.col_idx), dev_f16(A.values), dev_f16(G), C)
    torch.cuda.synchronize()
    rpT, cT, perm = O.transpose_csr(A.n_rows, A.n_cols, A.row_offset, A.col_idx)
    ref = O.spmm_f32(A.n_cols, 64, rpT, cT, A.values[perm], G)
    assert rel_l2(C.cpu().numpy(), ref) <= 1e-5
