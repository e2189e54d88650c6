"""A fresh plan's FIRST tca_spmm and tca_sddmm can be captured in a HIP graph:
every table a launch reads (SpMM unit tables of every split factor, the SDDMM
tiling and its unit tables) is built when the plan is created, so the
launches allocate and copy nothing.  The graph is replayed several times with
eager work in between, each replay checked against the oracle."""
import numpy as np
import pytest
import torch

from helpers import dev_f16, f16_to_f64, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rpw", (16, 32, 64, 128))
def test_first_calls_of_a_fresh_plan_are_capturable(cuda, rpw):
    case = workloads.plain_gpu(512, 2048, 0.9)
    A = case.A
    N = case.N
    vals = dev_f16(A.values)
    B = dev_f16(case.B)
    Gb = gen.f32_to_f16_bits(gen.normal_f32(3, rpw, (A.n_rows, N)))
    G = dev_f16(Gb)
    ref_c = O.spmm_f64(A.n_rows, N, A.row_offset, A.col_idx, f16_to_f64(A.values), f16_to_f64(case.B))
    ref_dw = O.sddmm_f64(A.n_rows, N, A.row_offset, A.col_idx, f16_to_f64(Gb), f16_to_f64(case.B))
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)      # fresh: no launch has run on it
    w = plan.pack(vals)
    C = torch.empty((A.n_rows, N), dtype=torch.float32, device=cuda)
    dW = torch.empty(A.nnz, dtype=torch.float32, device=cuda)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        T.spmm(plan, w, B, out=C, tile_n=64)
        T.sddmm(plan, G, B, out=dW)
    for rep in range(3):
        C.fill_(float("nan"))                # eager work between replays
        dW.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert rel_l2(C.cpu().numpy(), ref_c) <= 1e-5, rep
        assert rel_l2(plan.unpack(dW).cpu().numpy(), ref_dw) <= 1e-5, rep
