"""Device-side CSR transpose (tca_csr_transpose_device) against the oracle's
host transpose (itself pinned to scipy.sparse in test_oracle), and the
pattern-keyed plan cache."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_f16, f16_to_f64, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads

pytestmark = pytest.mark.gpu


def _check(M, K, rp, ci):
    col_offset, row_idx, perm = T.csr_transpose_device(dev(rp), dev(ci), K)
    rpT, cT, permT = O.transpose_csr(M, K, rp, ci)
    assert np.array_equal(col_offset.cpu().numpy(), rpT)
    assert np.array_equal(row_idx.cpu().numpy(), cT)
    assert np.array_equal(perm.cpu().numpy(), permT)


@pytest.mark.parametrize("shape,density", [((77, 45), 0.3), ((512, 2048), 0.1), ((4096, 1024), 0.1), ((3, 70000), 0.01)])
def test_device_transpose_matches_oracle(cuda, shape, density):
    M, K = shape
    rp, ci = gen.pattern("bernoulli", M, K, density, seed=11, stream=M)
    _check(M, K, rp, ci)


def test_device_transpose_empty_rows_columns_and_nnz0(cuda):
    M, K = 40, 33
    rp = np.array([0] * 20 + [0, 3] + [3] * 19, np.int32)
    ci = np.array([1, 5, 32], np.int32)
    _check(M, K, rp, ci)
    col_offset, row_idx, perm = T.csr_transpose_device(dev(np.zeros(M + 1, np.int32)), dev(np.zeros(0, np.int32)), K)
    assert np.all(col_offset.cpu().numpy() == 0) and row_idx.numel() == 0 and perm.numel() == 0


def test_device_transpose_rejects_bad_arguments(cuda):
    with pytest.raises(T.TcaError) as e:
        T.csr_transpose_device(dev(np.zeros(1, np.int32)), dev(np.zeros(0, np.int32)), 5)   # zero rows
    assert e.value.name == "TCA_ERR_INVALID_ARG"


def test_plan_cache_reuses_pattern_plans(cuda):
    cache = T.PlanCache()
    case = workloads.sparse_linear(M=512, K=384, N=128)
    W, G = case.A, case.extra["G"]
    fwd, tr = cache.get_transpose(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16, 32)
    fwd2, tr2 = cache.get_transpose(W.n_rows, W.n_cols, W.row_offset.copy(), W.col_idx.copy(), 16, 32)
    assert fwd is fwd2 and tr is tr2 and len(cache) == 2
    assert cache.key(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16) != cache.key(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 32)
    # the cached transposed plan computes dX = W^T G
    tiled = fwd.pack(dev_f16(W.values))
    rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
    ref = O.spmm_f64(W.n_cols, G.shape[1], rpT, cT, f16_to_f64(W.values)[perm], G.astype(np.float16).astype(np.float64))
    dX = T.spmm(tr, tiled, dev(G), tile_n=64).cpu().numpy()
    assert rel_l2(dX, ref) <= 1e-5
