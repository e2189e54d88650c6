"""Batched persistent launch over the twelve uneven layers (forward SpMM,
transposed SpMM dX and SDDMM dW of every layer in ONE kernel): the outputs
must match the per-call kernels (SpMM to FP32 summation-order noise: the two
launches split a group's chunks differently; SDDMM bit for bit), match the
FP64 oracle, and repeat across HIP-graph replays (the launch resets its own
work counter): the SDDMM bit for bit, the SpMM outputs to FP32
summation-order noise (units split across workgroups add atomically)."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_f16, f16_to_f64, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import workloads

pytestmark = pytest.mark.gpu


def _layers(cuda):
    out = []
    for case in workloads.batch12(seed=0):
        W, X, G = case.A, case.B, case.extra["G"]
        rpw, tile = T.autotune.choice_for(W.n_rows, W.n_cols, W.nnz, case.N, batch=True)
        rpw_t, tile_t = T.autotune.choice_for(W.n_cols, W.n_rows, W.nnz, case.N, op="spmm_t", batch=True)
        plan = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
        plan_t = plan.transpose(rpw_t)
        w = plan.pack(dev_f16(W.values))
        Xd, Gd = dev_f16(X), dev(G)
        out.append(dict(case=case, plan=plan, plan_t=plan_t, w=w, X=Xd, G=Gd, tile=tile, tile_t=tile_t,
                        Y=torch.zeros((W.n_rows, case.N), dtype=torch.float32, device=cuda),
                        dX=torch.zeros((W.n_cols, case.N), dtype=torch.float32, device=cuda),
                        dW=torch.zeros(W.nnz, dtype=torch.float32, device=cuda)))
    return out


def _items(layers):
    items = []
    for L in layers:
        items.append(T.spmm_item(L["plan"], L["w"], L["X"], L["Y"], L["tile"]))
        items.append(T.spmm_item(L["plan_t"], L["w"], L["G"], L["dX"], L["tile_t"]))
        items.append(T.sddmm_item(L["plan"], L["G"], L["X"], L["dW"]))
    return items


def test_batch_equals_per_call_and_oracle(cuda):
    layers = _layers(cuda)
    batch = T.Batch(_items(layers))
    assert batch.n_units > 0 and batch.n_workgroups > 0
    batch.run()
    torch.cuda.synchronize()
    for L in layers:
        Y = T.spmm(L["plan"], L["w"], L["X"], tile_n=L["tile"])
        dX = T.spmm(L["plan_t"], L["w"], L["G"], tile_n=L["tile_t"])
        dW = T.sddmm(L["plan"], L["G"], L["X"])
        assert torch.equal(L["dW"], dW)
        for got, ref in ((L["Y"], Y), (L["dX"], dX)):
            assert rel_l2(got.cpu().numpy(), ref.cpu().numpy()) <= 1e-6
    for L in layers[::5]:
        W, X, G = L["case"].A, L["case"].B, L["case"].extra["G"]
        ref = O.spmm_f64(W.n_rows, L["case"].N, W.row_offset, W.col_idx, f16_to_f64(W.values), f16_to_f64(X))
        assert rel_l2(L["Y"].cpu().numpy(), ref) <= 1e-5
        g16 = G.astype(np.float16).astype(np.float64)
        ref_dw = O.sddmm_f64(W.n_rows, L["case"].N, W.row_offset, W.col_idx, g16, f16_to_f64(X))
        assert rel_l2(L["plan"].unpack(L["dW"]).cpu().numpy(), ref_dw) <= 1e-5


def test_batch_graph_replays(cuda):
    layers = _layers(cuda)
    batch = T.Batch(_items(layers))
    batch.run()
    torch.cuda.synchronize()
    ref = [(L["Y"].clone(), L["dX"].clone(), L["dW"].clone()) for L in layers]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        batch.run()                      # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    with torch.cuda.graph(g):
        batch.run()
    for rep in range(4):
        for L in layers:                 # eager work between replays
            L["Y"].zero_()
            L["dX"].zero_()
            L["dW"].zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for L, (y, dx, dw) in zip(layers, ref):
            # units split across workgroups add atomically (order may vary between
            # replays); the SDDMM is never split and must repeat bit for bit
            assert torch.equal(L["dW"], dw), rep
            for got, want in ((L["Y"], y), (L["dX"], dx)):
                assert rel_l2(got.cpu().numpy(), want.cpu().numpy()) <= 1e-6, rep


def test_batch_rejects_unsupported(cuda):
    case = workloads.plain_cpu()
    A = case.A
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, 16)
    w = plan.pack(dev_f16(A.values)).to(torch.bfloat16)
    B = dev_f16(case.B).to(torch.bfloat16)
    C = torch.empty((A.n_rows, 256), dtype=torch.float32, device=cuda)
    with pytest.raises(T.TcaError) as e:
        T.Batch([T.spmm_item(plan, w, B, C)])
    assert e.value.name == "TCA_ERR_UNSUPPORTED"


def test_batch_with_fp16_gradient_matches_fp32_gradient(cuda):
    # G rounded once to FP16 gives the same MFMA inputs as an FP32 G rounded on load
    layers = _layers(cuda)
    batch = T.Batch(_items(layers))
    batch.run()
    torch.cuda.synchronize()
    ref = [(L["dX"].clone(), L["dW"].clone()) for L in layers]
    items = []
    for L in layers:
        L["G16"] = L["G"].half()
        L["dX"].zero_()
        L["dW"].zero_()
        items.append(T.spmm_item(L["plan_t"], L["w"], L["G16"], L["dX"], L["tile_t"]))
        items.append(T.sddmm_item(L["plan"], L["G16"], L["X"], L["dW"]))
    T.Batch(items).run()
    torch.cuda.synchronize()
    for L, (dx, dw) in zip(layers, ref):
        assert torch.equal(L["dX"], dx) and torch.equal(L["dW"], dw)


def test_batch_rounds_fp32_operands_on_every_run(cuda):
    # FP32 G is rounded into the batch's FP16 copy at the start of every run, so
    # a G rewritten between runs is what the next run reads
    layers = _layers(cuda)[:4]
    batch = T.Batch(_items(layers))
    batch.run()
    torch.cuda.synchronize()
    for L in layers:
        L["G"].mul_(-0.75).add_(0.125)
    batch.run()
    torch.cuda.synchronize()
    for L in layers:
        dW = T.sddmm(L["plan"], L["G"], L["X"])
        dX = T.spmm(L["plan_t"], L["w"], L["G"], tile_n=L["tile_t"])
        assert torch.equal(L["dW"], dW)
        assert rel_l2(L["dX"].cpu().numpy(), dX.cpu().numpy()) <= 1e-6
