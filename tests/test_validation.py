"""Argument checks of the torch-facing wrappers, on the CPU (host-only plans
and CPU tensors): every check fires before anything reaches the C ABI, so a
column-strided view, an undersized or mistyped output, a short value array or
a host-only plan raises ValueError instead of reading or writing the wrong
memory."""
import numpy as np
import pytest
import torch

import tca_spmm as T
from tca_spmm import gen


@pytest.fixture(scope="module")
def hplan():
    rp, ci = gen.pattern("bernoulli", 40, 24, 0.2, seed=1, stream=1)
    return T.Plan(40, 24, rp, ci, 16, device=False)


def test_spmm_rejects_bad_arguments(hplan):
    p = hplan
    a = torch.zeros(p.nnz, dtype=torch.float16)
    B = torch.zeros(24, 32, dtype=torch.float16)
    with pytest.raises(ValueError, match="unit stride"):
        T.spmm(p, a, torch.zeros(24, 64, dtype=torch.float16)[:, ::2])
    with pytest.raises(ValueError, match="rows"):
        T.spmm(p, a, torch.zeros(23, 32, dtype=torch.float16))
    with pytest.raises(ValueError, match="dtype"):
        T.spmm(p, a, B, out=torch.zeros(40, 32, dtype=torch.float16))
    with pytest.raises(ValueError, match="at least 40"):
        T.spmm(p, a, B, out=torch.zeros(39, 32))
    with pytest.raises(ValueError, match="at least 32"):
        T.spmm(p, a, B, out=torch.zeros(40, 16))
    with pytest.raises(ValueError, match="needs at least"):
        T.spmm(p, a[:-1], B)
    with pytest.raises(ValueError, match="dense_elems"):
        T.spmm(p, a, B, dense=True)
    with pytest.raises(ValueError, match="row_scale"):
        T.spmm(p, a, B, row_scale=torch.ones(39))
    with pytest.raises(ValueError, match="host-only"):
        T.spmm(p, a, B)


def test_sddmm_rejects_bad_arguments(hplan):
    p = hplan
    G = torch.zeros(40, 32, dtype=torch.float16)
    X = torch.zeros(24, 32, dtype=torch.float16)
    with pytest.raises(ValueError, match="unit stride"):
        T.sddmm(p, torch.zeros(40, 64, dtype=torch.float16)[:, ::2], X)
    with pytest.raises(ValueError, match="columns"):
        T.sddmm(p, G, torch.zeros(24, 16, dtype=torch.float16))
    with pytest.raises(ValueError, match="dtype"):
        T.sddmm(p, G, torch.zeros(24, 32, dtype=torch.float32))
    with pytest.raises(ValueError, match="needs at least"):
        T.sddmm(p, G, X, out=torch.zeros(p.nnz - 1))
    with pytest.raises(ValueError, match="host-only"):
        T.sddmm(p, G, X)


def test_batch_items_are_checked(hplan):
    p = hplan
    a = torch.zeros(p.nnz, dtype=torch.float16)
    with pytest.raises(ValueError, match="at least 40"):
        T.Batch([T.spmm_item(p, a, torch.zeros(24, 32, dtype=torch.float16), torch.zeros(10, 32))])
    with pytest.raises(ValueError, match="needs at least"):
        T.Batch([T.sddmm_item(p, torch.zeros(40, 32), torch.zeros(24, 32, dtype=torch.float16), torch.zeros(3))])
