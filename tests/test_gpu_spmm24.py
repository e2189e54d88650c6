"""2:4 on gfx950's sparse MFMA (tca_compress_24 / tca_spmm_24): the device
pruning + compression rebuilds exactly the oracle's STRIP pruning (ties
included: both take the first pair), and the product equals the oracle's
FP32 product of the pruned matrix (rel-L2 <= 1e-5, compare_mat) for every
rows-per-wave and tile width -- bit-identical across them -- with F16, BF16
and F32 B, ragged rows and strided operands."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_bf16, dev_f16, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import _lib, gen, workloads

pytestmark = pytest.mark.gpu


def _dense(M, K, density, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((M, K)).astype(np.float32)
    x[rng.random((M, K)) >= density] = 0.0
    return gen.f32_to_f16_bits(x)


def _oracle(P_bits, B_bits):
    M = P_bits.shape[0]
    rp, ci, v = O.dense_to_csr(P_bits)
    return O.spmm_f32(M, B_bits.shape[1], rp, ci, v, B_bits)


@pytest.mark.parametrize("M,K,density", [(64, 64, 1.0), (300, 512, 0.6), (2048, 512, 0.3), (37, 128, 0.9)])
def test_compress_matches_oracle_pruning(M, K, density):
    A = _dense(M, K, density, seed=M + K)
    s = T.Sparse24.compress(dev_f16(A))
    torch.cuda.synchronize()
    assert np.array_equal(s.to_dense(), O.prune_24_strip(A))


@pytest.mark.parametrize("M,K,N", [(256, 512, 128), (2048, 512, 256), (512, 2048, 256), (100, 192, 40)])
def test_spmm24_matches_oracle(M, K, N):
    A = _dense(M, K, 0.8, seed=M * 3 + K)
    B = workloads.dense_f16(0, 10, M % 7, 3, (K, N))
    s = T.Sparse24.compress(dev_f16(A))
    C = T.spmm_24(s, dev_f16(B)).cpu().numpy()
    ref = _oracle(O.prune_24_strip(A), B)
    assert rel_l2(C, ref) <= 1e-5
    ok, _ = O.compare_mat(ref, C)
    assert ok


def test_spmm24_every_shape_bit_identical():
    A = _dense(200, 256, 0.7, seed=2)
    B = workloads.dense_f16(0, 10, 9, 3, (256, 72))
    s = T.Sparse24.compress(dev_f16(A))
    ref = _oracle(O.prune_24_strip(A), B)
    base = None
    for rpw in (16, 32, 64):
        for tile in (16, 32, 64):
            C = T.spmm_24(s, dev_f16(B), tile_n=tile, rows_per_wave=rpw).cpu().numpy()
            assert rel_l2(C, ref) <= 1e-5, (rpw, tile)
            base = C if base is None else base
            assert np.array_equal(C.view(np.uint32), base.view(np.uint32)), (rpw, tile)


def test_spmm24_bf16_and_f32_B():
    M, K, N = 128, 256, 64
    Af = np.random.default_rng(4).standard_normal((M, K)).astype(np.float32)
    Ab = gen.f32_to_bf16_bits(Af)
    Bf = gen.normal_f32(5, 6, (K, N))
    Bb = gen.f32_to_bf16_bits(Bf)
    s = T.Sparse24.compress(dev_bf16(Ab))
    P = s.to_dense()
    # the BF16 pruning follows the same rule on BF16 magnitudes: the survivors of
    # every strip are its two largest (their magnitudes sum to the top two), with
    # their bits kept
    Pf = O.bf16_bits_to_f32(P).astype(np.float64)
    a = np.abs(O.bf16_bits_to_f32(Ab).astype(np.float64)).reshape(M, K // 4, 4)
    top2 = np.sort(a, axis=2)[:, :, 2:].sum(axis=2)
    kept = np.abs(Pf).reshape(M, K // 4, 4).sum(axis=2)
    assert np.array_equal(kept, top2)
    assert ((P != 0).reshape(M, K // 4, 4).sum(axis=2) <= 2).all()
    nz = P != 0
    assert np.array_equal(P[nz], Ab[nz])
    C = T.spmm_24(s, dev_bf16(Bb)).cpu().numpy()
    ref = Pf @ O.bf16_bits_to_f32(Bb).astype(np.float64)
    assert rel_l2(C, ref) <= 1e-5
    # F32 B is rounded to BF16 on load
    C32 = T.spmm_24(s, dev(Bf)).cpu().numpy()
    assert np.array_equal(C32, C)


def test_spmm24_strided_and_ragged():
    M, K, N = 70, 128, 24
    A = _dense(M, K, 0.9, seed=8)
    Awide = np.zeros((M, K + 64), np.uint16)
    Awide[:, :K] = A
    dA = dev_f16(Awide)[:, :K]
    s = T.Sparse24.compress(dA)
    B = workloads.dense_f16(0, 10, 3, 3, (K, N))
    Bwide = torch.zeros((K, 40), dtype=torch.float16, device="cuda")
    Bwide[:, :N] = dev_f16(B)
    Cwide = torch.full((M, 48), float("nan"), dtype=torch.float32, device="cuda")
    T.spmm_24(s, Bwide[:, :N], out=Cwide[:, :N])
    got = Cwide.cpu().numpy()
    assert rel_l2(got[:, :N], _oracle(O.prune_24_strip(A), B)) <= 1e-5
    assert np.isnan(got[:, N:]).all()


def test_spmm24_rejections():
    s = T.Sparse24.compress(dev_f16(_dense(32, 64, 1.0, seed=1)))
    with pytest.raises(_lib.TcaError):
        T.Sparse24.compress(dev_f16(_dense(32, 48, 1.0, seed=1)))          # K % 64
    with pytest.raises(_lib.TcaError):
        T.spmm_24(s, torch.ones((64, 12), dtype=torch.float16, device="cuda"))   # N % 8
    with pytest.raises(_lib.TcaError):
        T.spmm_24(s, torch.ones((64, 16), dtype=torch.float16, device="cuda"), tile_n=128)
