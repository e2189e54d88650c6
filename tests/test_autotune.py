"""The autotuner's table: record() / record_batch() entries come back from
choice_for() for the same shape and the nearest density, per A layout and
per op, batch entries only for batch=True lookups; without an entry the
density rule (per call) or the batch rule applies.  The GPU test checks that
tune() times every candidate and returns them fastest first."""
import json

import numpy as np
import pytest

from tca_spmm import autotune, workloads
from tca_spmm.workloads import SparseMatrix


@pytest.fixture
def fresh_table(monkeypatch, tmp_path):
    monkeypatch.setattr(autotune, "_table", {"entries": []})
    return tmp_path / "table.json"


def _matrix(M, K, nnz):
    rp = np.linspace(0, nnz, M + 1).astype(np.int32)
    return SparseMatrix("m", M, K, rp, np.zeros(nnz, np.int32), np.zeros(nnz, np.uint16))


def test_record_and_choice_round_trip(fresh_table):
    A = _matrix(2048, 512, 104858)                # density 0.1
    res = [{"rows_per_wave": 64, "tile_n": 32, "us": 4.0}, {"rows_per_wave": 16, "tile_n": 64, "us": 5.0}]
    autotune.record(A, 256, res, path=str(fresh_table), layout="dense")
    assert autotune.choice_for(2048, 512, 104858, 256, dense=True) == (64, 32)
    # nearest density within the table's reach still matches, the tiled layout does not
    assert autotune.choice_for(2048, 512, 90000, 256, dense=True) == (64, 32)
    assert autotune.choice_for(2048, 512, 104858, 256) == autotune.heuristic(2048, 512, 104858, 256)
    # another shape falls back to the density rule
    assert autotune.choice_for(1024, 512, 104858, 256, dense=True) == autotune.heuristic(1024, 512, 104858, 256)
    # transposed entries are keyed by the transposed shape
    autotune.record(A, 256, [{"rows_per_wave": 32, "tile_n": 16, "us": 3.0}], op="spmm_t", path=str(fresh_table))
    assert autotune.choice_for(512, 2048, 104858, 256, op="spmm_t") == (32, 16)
    saved = json.loads(fresh_table.read_text())
    assert len(saved["entries"]) == 2


def test_record_batch_is_read_back_for_batch_lookups_only(fresh_table):
    cases = workloads.batch12(seed=0)[:2]
    choices = [{"fwd": (128, 64), "dx": (64, 64)}, {"fwd": (16, 32), "dx": (32, 32)}]
    autotune.record_batch(cases, choices, 123.0, path=str(fresh_table))
    for c, ch in zip(cases, choices):
        A, N = c.A, c.N
        assert autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N, batch=True, dense=True) == ch["fwd"]
        assert autotune.choice_for(A.n_cols, A.n_rows, A.nnz, N, op="spmm_t", batch=True, dense=True) == ch["dx"]
        # per-call lookups ignore batch entries
        assert autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N, dense=True) == autotune.heuristic(
            A.n_rows, A.n_cols, A.nnz, N)
    # a later per-call record of the same shape keeps the batch entries
    A = cases[0].A
    autotune.record(A, cases[0].N, [{"rows_per_wave": 16, "tile_n": 16, "us": 1.0}], path=str(fresh_table),
                    layout="dense")
    assert autotune.choice_for(A.n_rows, A.n_cols, A.nnz, cases[0].N, batch=True, dense=True) == (128, 64)
    # every recorded batch choice is a variant the batch kernel has
    for e in json.loads(fresh_table.read_text())["entries"]:
        if e.get("batch"):
            assert (e["rows_per_wave"], e["tile_n"]) in autotune.BATCH_DENSE_CANDIDATES


def test_batch_rule_without_entries(fresh_table):
    assert autotune.choice_for(512, 2048, 100000, 128, batch=True, dense=True) == (32, 64)
    assert autotune.choice_for(512, 2048, 100000, 16, batch=True) == (16, 16)


def test_committed_table_batch_entries_are_batch_variants():
    with open(autotune.TABLE_PATH) as f:
        tab = json.load(f)
    for e in tab["entries"]:
        if e.get("batch"):
            assert (e["rows_per_wave"], e["tile_n"]) in autotune.BATCH_DENSE_CANDIDATES
        else:
            assert (e["rows_per_wave"], e["tile_n"]) in autotune.ALL_CANDIDATES


@pytest.mark.gpu
def test_tune_returns_every_candidate_sorted(cuda):
    import torch
    case = workloads.plain_gpu(2048, 512, 0.9)
    B = torch.from_numpy(case.B.view(np.int16)).view(torch.float16).to(cuda)
    for dense in (False, True):
        cands = [(16, 32), (32, 64), (64, 32), (128, 64)]
        res = autotune.tune(case.A, B, dense=dense, candidates=cands, iters=5)
        assert sorted((r["rows_per_wave"], r["tile_n"]) for r in res) == sorted(cands)
        us = [r["us"] for r in res]
        assert us == sorted(us) and all(u > 0 for u in us)
