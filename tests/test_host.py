"""CPU tests of the product library's host side, checked against the oracle:
the C ABI loads and exports every declared symbol, the .smtx reader, CSR
conversion, compare_mat, the seeded generators and the panel tiling (bit-exact
against the oracle's independent rebuild).  No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import BUNDLED_SMTX, ROOT
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import _lib, gen, workloads


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tca_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared("tca_spmm.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers all of them
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)
    assert L.tca_version() == 101
    assert L.tca_batch_item_size() == ctypes.sizeof(_lib.BatchItem)


def test_status_strings():
    L = _lib.lib()
    for s in range(8):
        assert L.tca_status_string(s)


# ------------------------------------------------------------------ .smtx --
def test_smtx_read_matches_oracle_bundled():
    M, K, rp, ci = gen.smtx_read(BUNDLED_SMTX)
    rc, (dense, nnz) = O.smtx_to_dense(BUNDLED_SMTX)
    orp, oci, _ = O.dense_to_csr(dense)
    assert (M, K, nnz) == (2048, 512, 1005)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)


def test_smtx_write_read_round_trip(tmp_path):
    rp, ci = gen.pattern("lognormal", 300, 200, 0.05, 0.5, seed=3, stream=7)
    p = str(tmp_path / "x.smtx")
    gen.smtx_write(p, 300, 200, rp, ci)
    M, K, rp2, ci2 = gen.smtx_read(p)
    assert (M, K) == (300, 200) and np.array_equal(rp, rp2) and np.array_equal(ci, ci2)
    rc, (dense, nnz) = O.smtx_to_dense(p)          # the oracle reads the same file
    assert rc == 0 and nnz == ci.size and int((dense == 0x3C00).sum()) == ci.size


@pytest.mark.parametrize("text,status", [
    ("2, 3, 2\n0 1 2 2\n0 2\n", "TCA_ERR_FORMAT"),
    ("2, 3, 2\n0 1\n0 2\n", "TCA_ERR_FORMAT"),
    ("2, 3, 2\n0 1 2\n0\n", "TCA_ERR_FORMAT"),
    ("2, 3, 2\n0 1 2\n0 2 1\n", "TCA_ERR_FORMAT"),
    ("2, 3, 2\n0 1 2\n0 3\n", "TCA_ERR_FORMAT"),
    ("2 3 2\n0 1 2\n0 2\n", "TCA_ERR_FORMAT"),
    ("2, 3, x\n0 1 2\n0 2\n", "TCA_ERR_FORMAT"),
])
def test_smtx_malformed_rejected(tmp_path, text, status):
    p = tmp_path / "bad.smtx"
    p.write_text(text)
    with pytest.raises(T.TcaError) as e:
        gen.smtx_read(str(p))
    assert e.value.name == status


def test_smtx_missing_file():
    with pytest.raises(T.TcaError) as e:
        gen.smtx_read("/nonexistent.smtx")
    assert e.value.name == "TCA_ERR_IO"


# ------------------------------------------------------- CSR / compare --
def test_dense_to_csr_matches_oracle():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((64, 80)).astype(np.float32)
    x[rng.random(x.shape) < 0.85] = 0
    x[3, 3] = -0.0
    x[4, 4] = np.nan
    bits = x.astype(np.float16).view(np.uint16)
    a = gen.dense_to_csr_f16(bits)
    b = O.dense_to_csr(bits)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert np.array_equal(gen.csr_to_dense_f16(64, 80, *a), O.csr_to_dense(64, 80, *b))


def test_compare_mat_matches_oracle():
    rng = np.random.default_rng(2)
    g = rng.standard_normal(1000).astype(np.float32)
    for noise in (1e-5, 1e-3, 1e-2, 1.0):
        c = g + noise * rng.standard_normal(1000).astype(np.float32)
        assert gen.compare_mat(g, c) == O.compare_mat(g, c)


# ------------------------------------------------------------ generators --
def test_generators_deterministic_and_counter_based():
    a = gen.normal_f32(5, 9, (1000,))
    b = gen.normal_f32(5, 9, (2000,))
    assert np.array_equal(a, b[:1000])                    # element i independent of the count
    assert not np.array_equal(a, gen.normal_f32(5, 10, (1000,)))
    assert abs(a.mean()) < 0.1 and abs(a.std() - 1) < 0.05
    u = gen.uniform_f32(1, 1, (100000,), -1, 1)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01


def test_f16_bf16_host_rounding_matches_oracle():
    x = np.random.default_rng(3).standard_normal(5000).astype(np.float32) * 100
    x = np.concatenate([x, np.array([65519.9, 65520, 2.0 ** -25, 3 * 2.0 ** -26, 1e-8], np.float32)])
    assert np.array_equal(gen.f32_to_f16_bits(x), x.astype(np.float16).view(np.uint16))
    assert np.array_equal(gen.f32_to_bf16_bits(x), O.f32_to_bf16_bits(x))


def test_patterns_follow_their_distributions():
    rp, ci = gen.pattern("bernoulli", 512, 512, 0.1, seed=0, stream=1)
    lens = np.diff(rp)
    assert abs(lens.mean() - 51.2) < 1.5 and 4 < lens.var() ** 0.5 < 9      # Binomial(512, 0.1): sd 6.8
    rp, ci = gen.pattern("lognormal", 2048, 512, 0.1, 0.5, seed=0, stream=2)
    assert abs(ci.size / (2048 * 512) - 0.1) < 0.002
    lens = np.diff(rp)
    assert 0.35 < np.log(lens[lens > 0]).std() < 0.65                        # sigma 0.5
    rp, ci = gen.pattern("uniform", 1000, 4096, 0.1, 0.2, seed=0, stream=3)
    lens = np.diff(rp)
    assert lens.min() >= int(409.6 * 0.8) - 1 and lens.max() <= int(409.6 * 1.2) + 1
    for r in range(0, 1000, 97):                                             # sorted, distinct, in range
        seg = ci[rp[r]:rp[r + 1]]
        assert np.all(np.diff(seg) > 0) and seg.min() >= 0 and seg.max() < 4096


def test_workload_shapes():
    c = workloads.plain_cpu()
    assert c.A.n_rows == 512 and c.B.shape == (512, 256) and abs(c.A.sparsity - 0.9) < 0.01
    suite = workloads.plain_gpu_suite()
    assert len(suite) == 10
    for case in suite:
        assert case.B.shape == (case.A.n_cols, 256)
    layers = workloads.batch12(with_grad=False)
    assert len(layers) == 12 and {l.N for l in layers} <= {128, 384, 1024}
    assert all(0.45 < l.A.sparsity < 0.99 for l in layers)


# --------------------------------------------------------------- tiling --
def _cases():
    yield "bundled", gen.smtx_read(BUNDLED_SMTX)
    for (M, K) in ((2048, 512), (512, 2048)):
        for s in (0.7, 0.98):
            c = workloads.plain_gpu(M, K, s)
            yield c.A.name, (M, K, c.A.row_offset, c.A.col_idx)
    c = workloads.plain_cpu()
    yield c.A.name, (512, 512, c.A.row_offset, c.A.col_idx)
    yield "ragged_77x45", (77, 45, *gen.pattern("bernoulli", 77, 45, 0.3, seed=4, stream=4))
    yield "empty_rows", (40, 33, np.array([0] * 20 + [0, 3] + [3] * 19, np.int32), np.array([1, 5, 32], np.int32))


@pytest.mark.parametrize("rpw", [16, 32, 64, 128, 256])
def test_tiling_bit_exact_against_oracle_rebuild(rpw):
    for name, (M, K, rp, ci) in _cases():
        plan = T.Plan(M, K, rp, ci, rows_per_wave=rpw, device=False)
        got = plan.export()
        ref = O.tiling(M, K, rp, ci, rpw)
        for k in ("grp_chunk_ptr", "chunk_cols", "seg_ptr", "entry_slot", "entry_src", "work_order"):
            assert np.array_equal(got[k], ref[k]), (name, k)
        assert plan.n_chunks == ref["n_chunks"]


def test_transpose_plan_maps_into_forward_tiled_order():
    M, K = 96, 70
    rp, ci = gen.pattern("bernoulli", M, K, 0.2, seed=5, stream=5)
    fwd = T.Plan(M, K, rp, ci, 16, device=False)
    tr = fwd.transpose(32)
    assert (tr.n_rows, tr.n_cols, tr.transposed) == (K, M, True)
    f, t = fwd.export(), tr.export()
    rows_of = np.repeat(np.arange(M), np.diff(rp))
    # transposed entry e sits at (row j, col i) of A^T and reads forward tiled value t.entry_src[e]
    for e in range(0, ci.size, 7):
        g, panel_row = None, None
        src_csr = f["entry_src"][t["entry_src"][e]]
        s = int(t["entry_slot"][e])
        c = np.searchsorted(t["seg_ptr"], e, side="right") - 1       # segment index = chunk*ppg + panel
        chunk, panel = divmod(c, 2)
        g = np.searchsorted(t["grp_chunk_ptr"], chunk, side="right") - 1
        row_T = g * 32 + 16 * panel + (s >> 5)
        col_T = t["chunk_cols"][chunk * 32 + (s & 31)]
        assert row_T == ci[src_csr] and col_T == rows_of[src_csr]


def test_plan_rejects_bad_csr():
    with pytest.raises(T.TcaError) as e:
        T.Plan(3, 4, np.array([0, 2, 1, 3], np.int32), np.array([0, 1, 2], np.int32), device=False)
    assert e.value.name == "TCA_ERR_FORMAT"
    with pytest.raises(T.TcaError):
        T.Plan(2, 4, np.array([0, 2, 3], np.int32), np.array([1, 1, 2], np.int32), device=False)   # duplicate column
    with pytest.raises(T.TcaError) as e:
        T.Plan(2, 4, np.array([0, 1, 2], np.int32), np.array([0, 4], np.int32), device=False)      # out of range
    assert e.value.name == "TCA_ERR_FORMAT"
    with pytest.raises(T.TcaError) as e:
        T.Plan(2, 4, np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32), rows_per_wave=8, device=False)
    assert e.value.name == "TCA_ERR_INVALID_ARG"
    # an offset past nnz ahead of a decrease: rejected before any column is read
    with pytest.raises(T.TcaError) as e:
        T.Plan(2, 4, np.array([0, 100, 5], np.int32), np.array([0, 1, 2, 3, 1], np.int32), device=False)
    assert e.value.name == "TCA_ERR_FORMAT"


# ------------------------------------------------------------ MatrixMarket --
def test_mtx_read_matches_oracle(tmp_path):
    rng = np.random.default_rng(9)
    M, K = 60, 45
    rows, cols = np.nonzero(rng.random((M, K)) < 0.1)
    perm = rng.permutation(rows.size)                       # file order is not sorted
    vals = rng.standard_normal(rows.size).astype(np.float32)
    lines = ["%%MatrixMarket matrix coordinate real general", "% generated", f"{M} {K} {rows.size + 1}"]
    lines += [f"{rows[i] + 1} {cols[i] + 1} {vals[i]:.6f}" for i in perm]
    lines.append(f"{rows[perm[0]] + 1} {cols[perm[0]] + 1} 7.5")   # duplicate: last wins
    p = str(tmp_path / "r.mtx")
    open(p, "w").write("\n".join(lines) + "\n")
    M2, K2, rp, ci, v = gen.mtx_read(p, one_based=True)
    rc, (dense, nnz) = O.mtx_to_dense(p, one_based=True)
    assert (M2, K2) == (M, K) and rc == 0
    assert np.array_equal(gen.csr_to_dense_f16(M, K, rp, ci, v), dense)
    assert ci.size == rows.size                             # the duplicate is merged
    for r in range(M):
        assert np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0)
    _, _, rp2, ci2, v2 = gen.mtx_read(p, one_based=True, pattern_only=True)
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci) and np.all(v2 == 0x3C00)


@pytest.mark.parametrize("text", ["2 2 3\n1 1 1\n", "2 2 1\n3 1 1\n", "x y z\n1 1 1\n", ""])
def test_mtx_read_rejects_malformed(tmp_path, text):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(T.TcaError) as e:
        gen.mtx_read(str(p))
    assert e.value.name == "TCA_ERR_FORMAT"
