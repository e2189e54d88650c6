"""Blocked-ELL SpMM on the GPU (tca_bell_spmm): against the CSR oracle on the
same matrix for every block size, rows-per-wave and tile width (and
bit-identical across those, since every output sums its k steps in the same
order), BF16 and F32 B, strided B / C, empty slots anywhere with non-finite
values behind them, ragged and empty shapes, and the argument checks."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_bf16, dev_f16, oracle_f32, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import _lib, gen, workloads
from tca_spmm.workloads import SparseMatrix

pytestmark = pytest.mark.gpu


def _matrix(M, K, density, seed, kind="lognormal"):
    rp, ci = gen.pattern(kind, M, K, density, 0.5, seed=seed, stream=seed + 1)
    vals = gen.f32_to_f16_bits(gen.normal_f32(seed, 99, (ci.size,)))
    vals[(vals & 0x7FFF) == 0] = 0x3C00
    return SparseMatrix(f"bell_{M}x{K}", M, K, rp, ci, vals)


def _run(A, B_bits, block, tile_n=0, rpw=0, bf16=False):
    bell = T.BlockedEll.from_csr(A.n_rows, A.n_cols, A.row_offset, A.col_idx, A.values, block,
                                 dtype=torch.bfloat16 if bf16 else torch.float16)
    B = dev_bf16(B_bits) if bf16 else dev_f16(B_bits)
    C = T.bell_spmm(bell, B, tile_n=tile_n, rows_per_wave=rpw)
    torch.cuda.synchronize()
    return C.cpu().numpy()


@pytest.mark.parametrize("M,K,N,block", [(200, 300, 40, 16), (2048, 512, 256, 32), (512, 2048, 256, 32),
                                         (130, 257, 64, 64), (300, 520, 24, 128), (77, 45, 8, 256)])
def test_bell_matches_oracle(M, K, N, block):
    A = _matrix(M, K, 0.05, seed=M + K + block)
    B = workloads.dense_f16(0, 8, block, 3, (K, N))
    C = _run(A, B, block)
    ref = oracle_f32(A, B)
    assert rel_l2(C, ref) <= 1e-5
    ok, _ = O.compare_mat(ref, C)
    assert ok


def test_bell_every_shape_bit_identical():
    A = _matrix(333, 410, 0.04, seed=5)
    B = workloads.dense_f16(0, 8, 1, 3, (410, 72))
    ref = oracle_f32(A, B)
    base = None
    for block in (16, 32, 64):
        for rpw in (16, 32, 64):
            if rpw > block:
                continue
            for tile in (16, 32, 64):
                C = _run(A, B, block, tile, rpw)
                assert rel_l2(C, ref) <= 1e-5, (block, rpw, tile)
                if base is None:
                    base = C
                assert np.array_equal(C.view(np.uint32), base.view(np.uint32)), (block, rpw, tile)


def test_bell_suite_matrices_block32():
    for case in workloads.plain_gpu_suite()[::3]:
        A = case.A
        C = _run(A, case.B, 32)
        ref = oracle_f32(A, case.B)
        assert rel_l2(C, ref) <= 1e-5, A.name


def test_bell_bf16_and_f32_B():
    A = _matrix(256, 384, 0.06, seed=9)
    Bf = gen.normal_f32(3, 4, (384, 64))
    # BF16 values and B against the FP64 product of the BF16-rounded operands
    vb = gen.f32_to_bf16_bits(O.f16_bits_to_f32(A.values))
    Ab = SparseMatrix(A.name, A.n_rows, A.n_cols, A.row_offset, A.col_idx, vb)
    Bb = gen.f32_to_bf16_bits(Bf)
    C = _run(Ab, Bb, 32, bf16=True)
    ref = O.spmm_f64(A.n_rows, 64, A.row_offset, A.col_idx, O.bf16_bits_to_f32(vb).astype(np.float64),
                     O.bf16_bits_to_f32(Bb).astype(np.float64))
    assert rel_l2(C, ref) <= 1e-5
    # F32 B is rounded to F16 on load: same as passing the F16-rounded B
    bell = T.BlockedEll.from_csr(A.n_rows, A.n_cols, A.row_offset, A.col_idx, A.values, 32)
    C32 = T.bell_spmm(bell, dev(Bf)).cpu().numpy()
    C16 = T.bell_spmm(bell, dev_f16(gen.f32_to_f16_bits(Bf))).cpu().numpy()
    assert np.array_equal(C32, C16)


def test_bell_strided_B_and_C():
    A = _matrix(150, 200, 0.05, seed=13)
    B = workloads.dense_f16(0, 8, 2, 3, (200, 48))
    bell = T.BlockedEll.from_csr(A.n_rows, A.n_cols, A.row_offset, A.col_idx, A.values, 16)
    Bwide = torch.zeros((200, 64), dtype=torch.float16, device="cuda")
    Bwide[:, :48] = dev_f16(B)
    Cwide = torch.full((150, 80), float("nan"), dtype=torch.float32, device="cuda")
    T.bell_spmm(bell, Bwide[:, :48], out=Cwide[:, :48])
    got = Cwide.cpu().numpy()
    assert rel_l2(got[:, :48], oracle_f32(A, B)) <= 1e-5
    assert np.isnan(got[:, 48:]).all()                 # columns past N untouched


def test_bell_empty_slots_anywhere_are_not_read():
    M, K, b = 96, 160, 16
    A = _matrix(M, K, 0.03, seed=17)
    B = workloads.dense_f16(0, 8, 3, 3, (K, 32))
    ci, v, _ = T.bell_from_csr(M, K, A.row_offset, A.col_idx, A.values, b)
    # two empty slots in front and one in the middle, their values NaN
    nan = np.full((v.shape[0], b), 0x7E00, np.uint16)
    mid = ci.shape[1] // 2
    ci2 = np.concatenate([np.full((ci.shape[0], 2), -1, np.int32), ci[:, :mid], np.full((ci.shape[0], 1), -1, np.int32),
                          ci[:, mid:]], axis=1)
    v2 = np.concatenate([nan, nan, v[:, :mid * b], nan, v[:, mid * b:]], axis=1)
    bell = T.BlockedEll(M, K, b, dev(ci2), dev_f16(v2))
    C = T.bell_spmm(bell, dev_f16(B)).cpu().numpy()
    assert np.isfinite(C).all()
    assert rel_l2(C, oracle_f32(A, B)) <= 1e-5


def test_bell_empty_and_single_row():
    # no entries: ell_cols = 0, every row of C written with zeros
    rp = np.zeros(41, np.int32)
    bell = T.BlockedEll.from_csr(40, 30, rp, np.zeros(0, np.int32), None, 32)
    assert bell.ell_cols == 0
    C = torch.full((40, 16), float("nan"), dtype=torch.float32, device="cuda")
    T.bell_spmm(bell, torch.ones((30, 16), dtype=torch.float16, device="cuda"), out=C)
    assert (C.cpu().numpy() == 0).all()
    # one row, one entry at the last column
    A = SparseMatrix("one", 1, 70, np.array([0, 1], np.int32), np.array([69], np.int32), np.array([0x4000], np.uint16))
    B = workloads.dense_f16(0, 8, 4, 3, (70, 8))
    assert rel_l2(_run(A, B, 64), oracle_f32(A, B)) <= 1e-6


def test_bell_rejections():
    A = _matrix(64, 64, 0.1, seed=1)
    bell = T.BlockedEll.from_csr(64, 64, A.row_offset, A.col_idx, A.values, 16)
    B = torch.ones((64, 12), dtype=torch.float16, device="cuda")
    with pytest.raises(_lib.TcaError):
        T.bell_spmm(bell, B)                                          # N % 8
    B = torch.ones((64, 16), dtype=torch.float16, device="cuda")
    with pytest.raises(_lib.TcaError):
        T.bell_spmm(bell, B, rows_per_wave=32)                        # above the block size
    with pytest.raises(_lib.TcaError):
        T.bell_spmm(bell, B, tile_n=128)
    with pytest.raises(TypeError):
        T.bell_spmm(bell, B.to(torch.bfloat16))


def test_dropin_driver_blocked_ell():
    import os
    import subprocess
    from conftest import BUNDLED_SMTX, PKG
    exe = os.path.join(PKG, "tca_test")
    if not os.path.exists(exe):
        pytest.skip("tca_test not built")
    r = subprocess.run([exe, BUNDLED_SMTX, "256", "0", "32"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Blocked-ELL(32):" in r.stdout and "validation failed" not in r.stdout
