"""MatrixMarket files through the HIP path: a pattern-only one-based file and
a real-valued file with comments and a repeated entry (the last value wins, as
the reference's dense reader overwrites, src/file_input.cu:8-134), read by
tca_mtx_read and multiplied by the SpMM kernels (tiled values and panel-dense
tiles), against the oracle's own reading of the same file."""
import numpy as np
import pytest
import torch

from helpers import dev_f16, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen

pytestmark = pytest.mark.gpu


def _write(path, M, K, entries, pattern, comments=True):
    lines = ["%%MatrixMarket matrix coordinate " + ("pattern" if pattern else "real") + " general"]
    if comments:
        lines.append("% written by the test")
    lines.append(f"{M} {K} {len(entries)}")
    for r, c, v in entries:
        lines.append(f"{r + 1} {c + 1}" if pattern else f"{r + 1} {c + 1} {v:.6g}")
    path.write_text("\n".join(lines) + "\n")


def _run(path, pattern, cuda):
    M, K, rp, ci, vals = gen.mtx_read(str(path), one_based=True, pattern_only=pattern)
    rc, (dense, nnz) = O.mtx_to_dense(str(path), one_based=True)
    assert rc == 0
    orp, oci, ov = O.dense_to_csr(dense)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(vals, ov)
    N = 256
    Bb = gen.f32_to_f16_bits(gen.normal_f32(4, 1, (K, N)))
    ref = O.spmm_f64(M, N, rp, ci, vals.view(np.float16).astype(np.float64), Bb.view(np.float16).astype(np.float64))
    B = dev_f16(Bb)
    for rpw in (16, 32, 64, 128):
        plan = T.Plan(M, K, rp, ci, rpw)
        t = plan.pack(dev_f16(vals))
        assert rel_l2(T.spmm(plan, t, B, tile_n=64).cpu().numpy(), ref) <= 1e-5, rpw
        assert rel_l2(T.spmm(plan, plan.pack_dense(t), B, tile_n=64, dense=True).cpu().numpy(), ref) <= 1e-5, rpw
    return M, K, rp, ci, vals


def test_mtx_pattern_one_based_through_hip(cuda, tmp_path):
    rp, ci = gen.pattern("lognormal", 700, 300, 0.05, 0.5, seed=6, stream=2)
    rows = np.repeat(np.arange(700), np.diff(rp))
    entries = list(zip(rows.tolist(), ci.tolist(), [1.0] * ci.size))
    rng = np.random.default_rng(0)
    rng.shuffle(entries)                         # coordinate files need not be sorted
    p = tmp_path / "pattern.mtx"
    _write(p, 700, 300, entries, pattern=True)
    M, K, rp2, ci2, vals = _run(p, True, cuda)
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
    assert np.all(vals == 0x3C00)                # pattern entries are 1.0


def test_mtx_real_values_and_repeats_through_hip(cuda, tmp_path):
    rp, ci = gen.pattern("bernoulli", 300, 500, 0.1, seed=7, stream=3)
    rows = np.repeat(np.arange(300), np.diff(rp))
    v = np.random.default_rng(1).standard_normal(ci.size).astype(np.float32)
    entries = list(zip(rows.tolist(), ci.tolist(), v.tolist()))
    entries.append((int(rows[0]), int(ci[0]), 2.5))       # a repeat: the last value wins
    p = tmp_path / "real.mtx"
    _write(p, 300, 500, entries, pattern=False)
    M, K, rp2, ci2, vals = _run(p, False, cuda)
    assert np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
    assert vals.view(np.float16)[0] == np.float16(2.5)
