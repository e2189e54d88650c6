"""Blocked-ELL on the CPU: the oracle's conversion (the host part of
perform_cuSPARSE_BlockSpMM, src/spmm_cusparse.cu:254-331) pinned by a
hand-worked known answer, the BELL -> dense identity, a brute-force slot count
and sorted block columns; the oracle's BELL product pinned bit for bit against
the CSR oracle (same ascending-column FP32 order); the library's host
converter (tca_bell_from_csr) equal to the oracle element for element."""
import numpy as np
import pytest

from oracle import oracle as O
import tca_spmm as T
from tca_spmm import _lib, gen


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _bell_to_dense(M, K, b, col_ind, values):
    """Dense matrix a Blocked-ELL describes (written independently of the oracle)."""
    out = np.zeros((M, K), np.uint16)
    slots = col_ind.shape[1] if col_ind.ndim == 2 else 0
    for br in range(col_ind.shape[0]):
        for j in range(slots):
            bc = col_ind[br, j]
            if bc < 0:
                continue
            for r in range(br * b, min(M, br * b + b)):
                for c in range(b):
                    k = bc * b + c
                    if k < K and values[r, j * b + c]:
                        out[r, k] = values[r, j * b + c]
    return out


def _random_dense(M, K, density, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((M, K)).astype(np.float32)
    x[rng.random((M, K)) >= density] = 0.0
    return _f16(x)


def test_oracle_bell_known_answer():
    # 5 x 6, b = 2 (worked by hand): block rows {rows 0-1, 2-3, 4}; kept block
    # columns {0, 1}, {1, 2}, {0, 2}; two slots; the pad row 5 is zero.
    A = np.zeros((5, 6), np.float32)
    A[0, 3], A[1, 0], A[2, 5], A[4, 1], A[4, 4], A[3, 2] = 1, 2, 3, 4, 5, 6
    ci, v, ell_cols = O.dense_to_bell(_f16(A), 2)
    assert ell_cols == 4
    assert ci.tolist() == [[0, 1], [1, 2], [0, 2]]
    expect = [[0, 0, 0, 1], [2, 0, 0, 0], [0, 0, 0, 3], [6, 0, 0, 0], [0, 4, 5, 0], [0, 0, 0, 0]]
    assert v.view(np.float16).astype(np.float32).tolist() == expect


@pytest.mark.parametrize("M,K,b,density", [(37, 53, 2, 0.1), (64, 64, 16, 0.02), (100, 75, 16, 0.05),
                                           (70, 130, 32, 0.01), (33, 33, 3, 0.2), (1, 1, 16, 1.0)])
def test_oracle_bell_identity_and_slots(M, K, b, density):
    dense = _random_dense(M, K, density, seed=M * 1000 + K)
    ci, v, ell_cols = O.dense_to_bell(dense, b)
    BR, BC = (M + b - 1) // b, (K + b - 1) // b
    # slots = brute-force maximum of kept blocks per block row
    kept = [[bool(np.any(dense[br * b:br * b + b, bc * b:bc * b + b].view(np.float16) != 0)) for bc in range(BC)]
            for br in range(BR)]
    slots = max((sum(r) for r in kept), default=0)
    assert ell_cols == slots * b and ci.shape == (BR, slots) and v.shape == (BR * b, ell_cols)
    for br in range(BR):
        row = ci[br].tolist()
        n = sum(kept[br])
        assert row[:n] == [bc for bc in range(BC) if kept[br][bc]]      # ascending kept columns
        assert row[n:] == [-1] * (slots - n)                             # -1 after the last
    assert np.array_equal(_bell_to_dense(M, K, b, ci, v), dense)


def test_oracle_bell_spmm_equals_csr_oracle_bitwise():
    for (M, K, b, d) in [(90, 70, 16, 0.05), (64, 200, 32, 0.02), (33, 48, 4, 0.3)]:
        dense = _random_dense(M, K, d, seed=M + K + b)
        B = _f16(np.random.default_rng(b).standard_normal((K, 24)))
        ci, v, _ = O.dense_to_bell(dense, b)
        rp, col, val = O.dense_to_csr(dense)
        ref = O.spmm_f32(M, 24, rp, col, val, B)
        got = O.bell_spmm_f32(M, K, b, ci, v, B)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_oracle_bell_spmm_empty_slots_anywhere():
    # moving -1 slots to the front (with garbage values behind them) keeps the product
    M, K, b = 40, 96, 16
    dense = _random_dense(M, K, 0.03, seed=11)
    B = _f16(np.random.default_rng(3).standard_normal((K, 16)))
    ci, v, ell_cols = O.dense_to_bell(dense, b)
    slots = ci.shape[1]
    ci2 = np.concatenate([np.full((ci.shape[0], 2), -1, np.int32), ci], axis=1)
    v2 = np.concatenate([_f16(np.full((v.shape[0], 2 * b), 7.0)), v], axis=1)
    got = O.bell_spmm_f32(M, K, b, ci2, v2, B)
    ref = dense.view(np.float16).astype(np.float64) @ B.view(np.float16).astype(np.float64)
    assert slots + 2 == ci2.shape[1]
    assert np.allclose(got, ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("M,K,b", [(2048, 512, 32), (300, 700, 16), (77, 45, 64), (129, 257, 128), (0, 10, 16),
                                   (10, 0, 32)])
def test_library_converter_matches_oracle(M, K, b):
    if M and K:
        rp, ci = gen.pattern("lognormal", M, K, 0.05, 0.5, seed=M, stream=K)
    else:
        rp, ci = np.zeros(M + 1, np.int32), np.zeros(0, np.int32)
    vals = _f16(np.random.default_rng(M + K).standard_normal(ci.size))
    vals[vals == 0] = 0x3C00                     # explicit zeros would drop blocks in the dense oracle
    dense = gen.csr_to_dense_f16(M, K, rp, ci, vals)
    o_ci, o_v, o_ec = O.dense_to_bell(dense, b)
    l_ci, l_v, l_ec = T.bell_from_csr(M, K, rp, ci, vals, b)
    assert l_ec == o_ec
    assert np.array_equal(l_ci, o_ci) and np.array_equal(l_v, o_v)


def test_library_converter_pattern_only_is_ones():
    rp, ci = gen.pattern("bernoulli", 50, 40, 0.1, seed=4, stream=4)
    _, v, _ = T.bell_from_csr(50, 40, rp, ci, None, 16)
    assert set(np.unique(v).tolist()) <= {0, 0x3C00} and int((v == 0x3C00).sum()) == ci.size


def test_library_converter_rejections():
    rp, ci = gen.pattern("bernoulli", 20, 20, 0.2, seed=1, stream=1)
    for b in (8, 48, 512):
        with pytest.raises(_lib.TcaError):
            T.bell_from_csr(20, 20, rp, ci, None, b)
    bad = ci.copy()
    bad[0] = 20
    with pytest.raises(_lib.TcaError):
        T.bell_from_csr(20, 20, rp, bad, None, 16)


def test_blocked_ell_class_checks_shapes():
    torch = pytest.importorskip("torch")
    ci, v, ell_cols = T.bell_from_csr(40, 64, *gen.pattern("bernoulli", 40, 64, 0.2, seed=2, stream=2), None, 16)
    vals = torch.from_numpy(v.view(np.int16)).view(torch.float16)
    bell = T.BlockedEll(40, 64, 16, torch.from_numpy(ci), vals)
    assert bell.ell_cols == ell_cols and tuple(bell.col_ind.shape) == (3, ell_cols // 16)
    with pytest.raises(ValueError):
        T.BlockedEll(40, 64, 16, torch.from_numpy(ci[:2]), vals)        # block rows
    with pytest.raises(ValueError):
        T.BlockedEll(40, 64, 16, torch.from_numpy(ci), vals[:32])       # padded rows
    assert T.bell_shape(2048, 512, 256, 32) == (32, 32)                  # >= 512 waves for K = 512
    assert T.bell_shape(512, 2048, 256, 32) == (16, 16)                  # >= 2048 waves for K = 2048
    assert T.bell_shape(8192, 1024, 1024, 32) == (32, 64)
    assert T.bell_shape(100, 64, 8, 16) == (16, 16)                      # rows per wave at most the block
