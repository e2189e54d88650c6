"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the oracle (dropped term,
wrong index, transposed operand, wrong rounding) fails it: independent
libraries (numpy float16, torch bfloat16/float8_e4m3fn, scipy.sparse), closed
forms, brute force on tiny inputs and the bundled .smtx fixture's known
structure.  The reference ships no golden outputs for this path.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from conftest import BUNDLED_SMTX
from oracle import oracle as O


# ---------------------------------------------------------------- formats --
def test_f16_decode_all_codes_match_numpy():
    codes = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    ours = O.f16_bits_to_f32(codes)
    ref = codes.view(np.float16).astype(np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(ours), nan)
    assert np.array_equal(ours[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def test_f16_encode_matches_numpy_round_to_nearest_even():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * s for s in (1e-6, 1e-3, 1, 300, 6e4)])
    edge = np.array([0.0, -0.0, 65504, 65519.99, 65520, 1e9, -1e9, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                     2.0 ** -14, 1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, np.inf, -np.inf], np.float32)
    x = np.concatenate([x, edge])
    L = O.lib()
    ours = np.array([L.orc_f32_to_f16(float(v)) for v in x], np.uint16)
    ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)


def test_bf16_encode_decode_match_torch():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * s for s in (1e-30, 1e-3, 1, 1e30)])
    ours = O.f32_to_bf16_bits(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, ref)
    back = O.bf16_bits_to_f32(ours)
    assert np.array_equal(back, torch.from_numpy(ref.view(np.int16)).view(torch.bfloat16).float().numpy())
    L = O.lib()
    assert all(L.orc_f32_to_bf16(float(v)) == r for v, r in zip(x[:500], ref[:500]))


def test_e4m3_decode_all_codes_match_torch_float8():
    codes = np.arange(256, dtype=np.uint8)
    ours = O.e4m3_to_f32(codes)
    ref = torch.from_numpy(codes).view(torch.float8_e4m3fn).float().numpy()
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(ours), nan) and nan.sum() == 2      # 0x7f, 0xff
    assert np.array_equal(ours[~nan], ref[~nan])
    assert O.e4m3_to_f32(np.array([0x7e], np.uint8))[0] == 448.0
    assert O.e4m3_to_f32(np.array([0x01], np.uint8))[0] == 2.0 ** -9


def test_e4m3_encode_nearest_even_and_saturation():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * s for s in (1e-3, 0.05, 1, 60)])
    x = x[np.abs(x) < 440]
    # every decoded value and every midpoint between neighbours
    vals = O.e4m3_to_f32(np.arange(0x7f, dtype=np.uint8)).astype(np.float64)
    mids = ((vals[1:] + vals[:-1]) / 2).astype(np.float32)
    x = np.concatenate([x, vals.astype(np.float32), mids, -mids])
    ours = O.f32_to_e4m3(x)
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(ours, ref)
    # saturation (the format has no infinity; this path clamps to +-448)
    sat = O.f32_to_e4m3(np.array([448, 460, 1e6, -1e6, np.inf], np.float32))
    assert list(sat) == [0x7e, 0x7e, 0x7e, 0xfe, 0x7e]
    assert O.f32_to_e4m3(np.array([np.nan], np.float32))[0] == 0x7f
    # the vectorised encoder agrees with the scalar one
    y = np.concatenate([x, rng.standard_normal(3000).astype(np.float32) * 500])
    assert np.array_equal(O.f32_to_e4m3_fast(y), O.f32_to_e4m3(y))


# ------------------------------------------------------------------ .smtx --
def test_bundled_smtx_structure():
    rc, out = O.smtx_to_dense(BUNDLED_SMTX)
    assert rc == 0
    dense, nnz = out
    assert dense.shape == (2048, 512) and nnz == 1005
    assert int((dense == 0x3C00).sum()) == 1005 and int((dense != 0).sum()) == 1005
    row_nnz = (dense != 0).sum(axis=1)
    assert int((row_nnz > 0).sum()) == 28
    assert row_nnz.max() == 65 and row_nnz[row_nnz > 0].min() == 16
    cols = np.nonzero(dense)[1]
    assert cols.min() == 2 and cols.max() == 511
    # the file itself: header, rows+1 offsets, nnz sorted distinct columns per row
    with open(BUNDLED_SMTX) as f:
        lines = f.read().split("\n")
    assert [int(t) for t in lines[0].split(",")] == [2048, 512, 1005]
    rp = np.array(lines[1].split(), dtype=np.int64)
    ci = np.array(" ".join(lines[2:]).split(), dtype=np.int64)
    assert rp.size == 2049 and rp[0] == 0 and rp[-1] == 1005 and ci.size == 1005
    for r in range(2048):
        seg = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(seg) > 0)


@pytest.mark.parametrize("text,code", [
    ("2, 3, 2\n0 1 2\n0 2\n", 0),
    ("2, 3, 2\n0 1 2 2\n0 2\n", -3),      # extra row offset (the reference overruns its buffer)
    ("2, 3, 2\n0 1\n0 2\n", -3),          # missing row offset
    ("2, 3, 2\n0 1 2\n0\n", -4),          # missing column id
    ("2, 3, 2\n0 1 2\n0 2 1\n", -4),      # extra column id
    ("2, 3, 2\n0 1 2\n0 3\n", -5),        # column out of range
    ("2 3 2\n0 1 2\n0 2\n", -2),          # header not comma separated
    ("2, 3, 2\n0 2 1\n0 2\n", -3),        # decreasing offsets
])
def test_smtx_malformed(tmp_path, text, code):
    p = tmp_path / "m.smtx"
    p.write_text(text)
    rc, _ = O.smtx_to_dense(str(p))
    assert rc == code


def test_smtx_missing_file():
    rc, _ = O.smtx_to_dense("/nonexistent/file.smtx")
    assert rc == -1


# ----------------------------------------------------------- dense <-> CSR --
def test_dense_csr_round_trip_and_brute_force():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((37, 53)).astype(np.float32)
    x[rng.random(x.shape) < 0.8] = 0
    x[0, 0] = -0.0
    x[1, 1] = np.nan
    bits = x.astype(np.float16).view(np.uint16)
    rp, ci, v = O.dense_to_csr(bits)
    rows, cols = np.nonzero((bits & 0x7FFF) != 0)          # +-0 dropped, NaN kept
    assert np.array_equal(ci, cols) and np.array_equal(np.diff(rp), np.bincount(rows, minlength=37))
    assert (0, 0) not in set(zip(rows, cols)) and (1, 1) in set(zip(rows, cols))
    assert np.array_equal(v, bits[rows, cols])
    back = O.csr_to_dense(37, 53, rp, ci, v)
    assert np.array_equal(back, np.where((bits & 0x7FFF) != 0, bits, 0))


# ------------------------------------------------------------- compare_mat --
def test_compare_mat_rule():
    ok, nf = O.compare_mat([1.0], [1.0005])
    assert ok and nf == 0                       # rel 5e-4 < tol
    ok, nf = O.compare_mat([1e-4], [5e-4])
    assert ok                                   # rel 4 >= tol but abs 4e-4 < tol
    ok, nf = O.compare_mat([1.0], [1.01])
    assert not ok and nf == 1
    ok, nf = O.compare_mat([0.0], [0.0])
    assert ok                                   # 0/0 is NaN, never >= tol
    ok, nf = O.compare_mat([0.0], [0.5])
    assert not ok
    ok, nf = O.compare_mat(np.ones(50), np.full(50, 2.0))
    assert not ok and nf == 10                  # early return after the 10th failure


# --------------------------------------------------------------- SpMM f32 --
def _rand_csr(M, K, density, seed):
    rng = np.random.default_rng(seed)
    mask = rng.random((M, K)) < density
    vals = rng.uniform(-1, 1, (M, K)).astype(np.float16)
    dense = np.where(mask, vals, np.float16(0))
    dense[dense == 0] = 0
    rp, ci, v = O.dense_to_csr(dense.view(np.uint16))
    return dense, rp, ci, v


def test_spmm_f32_against_dense_f64():
    M, K, N = 45, 70, 33
    dense, rp, ci, v = _rand_csr(M, K, 0.15, 5)
    B = np.random.default_rng(6).standard_normal((K, N)).astype(np.float16)
    C = O.spmm_f32(M, N, rp, ci, v, B.view(np.uint16))
    ref = dense.astype(np.float64) @ B.astype(np.float64)
    assert np.max(np.abs(C - ref)) < 1e-5 * max(1.0, np.abs(ref).max())
    # an A^T or B^T mix-up changes the result far beyond rounding
    assert np.max(np.abs(C - (dense.astype(np.float64) @ B.astype(np.float64)))) < 1e-4


def test_spmm_f32_closed_forms():
    K, N = 16, 24
    B = np.random.default_rng(7).standard_normal((K, N)).astype(np.float16)
    # identity pattern: C == B exactly
    rp = np.arange(K + 1, dtype=np.int32)
    ci = np.arange(K, dtype=np.int32)
    one = np.full(K, 0x3C00, np.uint16)
    assert np.array_equal(O.spmm_f32(K, N, rp, ci, one, B.view(np.uint16)), B.astype(np.float32))
    # single-nnz rows: C = a * B[j], exact (f16 x f16 is exact in f32)
    a = np.array([0.3, -2.5, 7.0], np.float16)
    rp = np.array([0, 1, 2, 3], np.int32)
    ci = np.array([5, 0, 15], np.int32)
    C = O.spmm_f32(3, N, rp, ci, a.view(np.uint16), B.view(np.uint16))
    exp = a.astype(np.float32)[:, None] * B.astype(np.float32)[ci]
    assert np.array_equal(C, exp)
    # all-ones values (the reference driver's A): row sums of gathered rows
    rp = np.array([0, 3, 3, 5], np.int32)
    ci = np.array([1, 4, 9, 0, 15], np.int32)
    C = O.spmm_f32(3, N, rp, ci, np.full(5, 0x3C00, np.uint16), B.view(np.uint16))
    Bf = B.astype(np.float32)
    assert np.allclose(C[0], Bf[1] + Bf[4] + Bf[9], atol=1e-6)
    assert np.all(C[1] == 0)                    # empty row
    assert np.allclose(C[2], Bf[0] + Bf[15], atol=1e-6)


def test_spmm_f64_and_f32_agree():
    M, K, N = 64, 96, 40
    dense, rp, ci, v = _rand_csr(M, K, 0.3, 8)
    B = np.random.default_rng(9).standard_normal((K, N)).astype(np.float16)
    C32 = O.spmm_f32(M, N, rp, ci, v, B.view(np.uint16))
    C64 = O.spmm_f64(M, N, rp, ci, v.view(np.float16).astype(np.float64), B.astype(np.float64))
    ref = dense.astype(np.float64) @ B.astype(np.float64)
    assert np.max(np.abs(C64 - ref)) < 1e-12
    rel = np.linalg.norm(C32 - C64) / np.linalg.norm(C64)
    assert rel < 1e-6


def test_sddmm_f64_against_dense():
    M, K, N = 30, 41, 64
    dense, rp, ci, v = _rand_csr(M, K, 0.2, 10)
    rng = np.random.default_rng(11)
    G = rng.standard_normal((M, N))
    X = rng.standard_normal((K, N))
    out = O.sddmm_f64(M, N, rp, ci, G, X)
    full = G @ X.T
    rows = np.repeat(np.arange(M), np.diff(rp))
    assert np.allclose(out, full[rows, ci], rtol=1e-12, atol=1e-12)
    assert not np.allclose(out, (X @ G.T)[ci % M, rows % K]) if M != K else True


def test_sddmm_f32_exact_on_small_integers_and_close_to_f64():
    # small-integer fp16 operands: every product and partial sum is exact in
    # FP32, so the result must equal numpy's dense product sampled on the
    # pattern exactly; a transposed operand or a dropped n would not
    M, K, N = 23, 37, 96
    dense, rp, ci, v = _rand_csr(M, K, 0.3, 12)
    rng = np.random.default_rng(13)
    G = rng.integers(-8, 9, (M, N)).astype(np.float16)
    X = rng.integers(-8, 9, (K, N)).astype(np.float16)
    out = O.sddmm_f32(M, N, rp, ci, G.view(np.uint16), X.view(np.uint16))
    rows = np.repeat(np.arange(M), np.diff(rp))
    full = G.astype(np.float64) @ X.astype(np.float64).T
    assert np.array_equal(out.astype(np.float64), full[rows, ci])
    short = G.astype(np.float64)[:, :-1] @ X.astype(np.float64)[:, :-1].T
    assert not np.array_equal(out.astype(np.float64), short[rows, ci])
    # random normal operands: FP32 accumulation stays within 1e-6 of FP64
    Gn = rng.standard_normal((M, N)).astype(np.float16)
    Xn = rng.standard_normal((K, N)).astype(np.float16)
    o32 = O.sddmm_f32(M, N, rp, ci, Gn.view(np.uint16), Xn.view(np.uint16))
    o64 = O.sddmm_f64(M, N, rp, ci, Gn.astype(np.float64), Xn.astype(np.float64))
    assert np.linalg.norm(o32 - o64) / np.linalg.norm(o64) < 1e-6


# ------------------------------------------------------ workload_manager --
def test_workload_manager_known_answer_bundled():
    rc, out = O.smtx_to_dense(BUNDLED_SMTX)
    rp, ci, v = O.dense_to_csr(out[0])
    nb, wpr, ridx, wid = O.workload_manager(rp, 256, 8)
    assert nb == 24
    assert int((ridx == -1).sum()) == 2                      # idle warp slots
    # no row packed into two different blocks
    blocks = {}
    for slot, r in enumerate(ridx):
        if r >= 0:
            blocks.setdefault(int(r), set()).add(slot // 8)
    assert all(len(b) == 1 for r, b in blocks.items() if wpr[r] <= 8)
    # every nonempty row gets exactly its warp demand
    for r in np.nonzero(np.diff(rp))[0]:
        assert int((ridx == r).sum()) == wpr[r]
    assert O.workload_manager(rp, 256, 4) is None             # warps/TB outside [8, 32]


def test_workload_manager_demand_rule():
    # rows of 0, 8, 9, 16, 17, 32, 33, 64, 65, 200 nnz at N=256
    lens = [0, 8, 9, 16, 17, 32, 33, 64, 65, 200]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    nb, wpr, ridx, wid = O.workload_manager(rp, 256, 8)
    assert list(wpr) == [0, 1, 2, 2, 4, 4, 8, 8, 16, 32]      # >64: 8*min(ceil(nnz/64), N/64)


# -------------------------------------------------------------- tiling --
@pytest.mark.parametrize("rpw", [16, 32, 64])
def test_tiling_rebuild_reconstructs_csr(rpw):
    M, K = 150, 300
    dense, rp, ci, v = _rand_csr(M, K, 0.05, 12 + rpw)
    t = O.tiling(M, K, rp, ci, rpw)
    rows_of = np.repeat(np.arange(M), np.diff(rp))
    ppg = rpw // 16
    assert np.array_equal(np.sort(t["entry_src"]), np.arange(ci.size))   # a permutation
    for g in range(t["n_groups"]):
        r0, r1 = g * rpw, min(M, (g + 1) * rpw)
        union = np.unique(ci[rp[r0]:rp[r1]])
        c0, c1 = t["grp_chunk_ptr"][g], t["grp_chunk_ptr"][g + 1]
        assert c1 - c0 == (union.size + 31) // 32
        cols = t["chunk_cols"][c0 * 32:c1 * 32]
        assert np.array_equal(cols[:union.size], union) and np.all(cols[union.size:] == -1)
        for c in range(c0, c1):
            for i in range(ppg):
                e0, e1 = t["seg_ptr"][c * ppg + i], t["seg_ptr"][c * ppg + i + 1]
                for e in range(e0, e1):
                    s = int(t["entry_slot"][e])
                    src = t["entry_src"][e]
                    assert rows_of[src] == g * rpw + 16 * i + (s >> 5)
                    assert ci[src] == t["chunk_cols"][c * 32 + (s & 31)]
    counts = np.diff(t["grp_chunk_ptr"])
    assert np.all(np.diff(counts[t["work_order"]]) <= 0)                 # heavy groups first


def test_transpose_matches_scipy():
    M, K = 40, 25
    dense, rp, ci, v = _rand_csr(M, K, 0.2, 13)
    rpT, cT, perm = O.transpose_csr(M, K, rp, ci)
    A = sp.csr_matrix((np.arange(1, ci.size + 1, dtype=np.float64), ci, rp), shape=(M, K))
    AT = A.T.tocsr()
    AT.sort_indices()
    assert np.array_equal(rpT, AT.indptr) and np.array_equal(cT, AT.indices)
    assert np.array_equal(perm + 1, AT.data.astype(np.int64))


# ------------------------------------------------------------ MatrixMarket --
def test_mtx_reader_brute_force(tmp_path):
    p = tmp_path / "a.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% comment\n3 4 5\n"
                 "1 1 0.5\n3 4 -2\n2 2\n1 1 0.25\n  2 3 1.5\n")
    rc, (dense, nnz) = O.mtx_to_dense(str(p), one_based=True)
    assert rc == 0 and nnz == 5
    exp = np.zeros((3, 4), np.float16)
    exp[0, 0] = 0.25          # the later duplicate overwrites
    exp[2, 3] = -2
    exp[1, 1] = 1.0           # missing value -> 1.0
    exp[1, 2] = 1.5
    assert np.array_equal(dense, exp.view(np.uint16))
    p.write_text("2 2 2\n0 1 3\n1 0 4\n")
    rc, (dense, _) = O.mtx_to_dense(str(p), one_based=False)
    assert rc == 0 and dense.view(np.float16)[0, 1] == 3 and dense.view(np.float16)[1, 0] == 4
    p.write_text("2 2 3\n1 1 1\n")
    assert O.mtx_to_dense(str(p))[0] == -5                  # count differs from the header
    p.write_text("2 2 1\n3 1 1\n")
    assert O.mtx_to_dense(str(p))[0] == -4                  # index out of range
