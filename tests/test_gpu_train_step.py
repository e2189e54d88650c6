"""The twelve-layer sparse-weight training step as the bench runs it: one batch
whose forward and dX items hand the batch the weight VALUES (dense="pack",
in the SDDMM plan's tiled order, the order dW comes back in), so every run
packs the panel-dense tiles from the current weights (prologue launch) and
then runs forward + dX + dW of all layers in one persistent launch.

Checked on EVERY layer against the FP64 oracle (forward on the FP16 inputs,
dX and dW on G rounded to FP16, the rounding the kernels apply), after an
in-place weight update (the tiles follow the weights), with values in CSR
order, across HIP-graph replays with weight updates between them, and a
values_order of another pattern is rejected; with the optimizer step fused into
the SDDMM (sddmm_item(sgd=...)) the weights after a run equal a separate
tca_sgd_update on that run's dW, bit for bit."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_bf16, dev_f16, f16_to_f64, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-5   # rel-L2 against the FP64 oracle (FP32 accumulation of FP16 products)


def _setup(cuda, csr_order=False, layers=None):
    out = []
    cases = workloads.batch12(seed=0)
    for i, case in enumerate(cases):
        if layers is not None and i not in layers:
            continue
        W, X, G = case.A, case.B, case.extra["G"]
        p16 = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
        v_csr = dev_f16(W.values)
        w = v_csr.clone() if csr_order else p16.pack(v_csr)
        rpw, tile = T.autotune.choice_for(W.n_rows, W.n_cols, W.nnz, case.N, batch=True, dense=True)
        rpw_t, tile_t = T.autotune.choice_for(W.n_cols, W.n_rows, W.nnz, case.N, op="spmm_t", batch=True, dense=True)
        pf = p16 if rpw == 16 else T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
        pt = p16.transpose(rpw_t)
        out.append(dict(case=case, p16=p16, pf=pf, pt=pt, w=w, X=dev_f16(X), G=dev(G), tile=tile, tile_t=tile_t,
                        Y=torch.full((W.n_rows, case.N), float("nan"), device=cuda),
                        dX=torch.full((W.n_cols, case.N), float("nan"), device=cuda),
                        dW=torch.full((W.nnz,), float("nan"), device=cuda)))
    return out


def _batch(layers, csr_order=False, sgd_lr=None):
    items = []
    for L in layers:
        vo = None if csr_order else L["p16"]
        items.append(T.spmm_item(L["pf"], L["w"], L["X"], L["Y"], L["tile"], dense="pack", values_order=vo))
        items.append(T.spmm_item(L["pt"], L["w"], L["G"], L["dX"], L["tile_t"], dense="pack", values_order=vo))
        sgd = None if sgd_lr is None else (L["w"], sgd_lr)
        items.append(T.sddmm_item(L["p16"], L["G"], L["X"], L["dW"], sgd=sgd))
    return T.Batch(items)


def _check_oracle(L, csr_order=False):
    W, N = L["case"].A, L["case"].N
    w = L["w"].cpu().numpy().view(np.uint16)
    v = w if csr_order else L["p16"].unpack(L["w"]).cpu().numpy().view(np.uint16)
    a64 = f16_to_f64(v)
    x64 = f16_to_f64(L["case"].B)
    g64 = L["G"].cpu().numpy().astype(np.float16).astype(np.float64)
    ref_y = O.spmm_f64(W.n_rows, N, W.row_offset, W.col_idx, a64, x64)
    rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
    ref_dx = O.spmm_f64(W.n_cols, N, rpT, cT, a64[perm], g64)
    ref_dw = O.sddmm_f64(W.n_rows, N, W.row_offset, W.col_idx, g64, x64)
    dw = L["p16"].unpack(L["dW"]).cpu().numpy()
    errs = (rel_l2(L["Y"].cpu().numpy(), ref_y), rel_l2(L["dX"].cpu().numpy(), ref_dx), rel_l2(dw, ref_dw))
    assert max(errs) <= TOL, (W.name, errs)
    return errs


def test_train_step_every_layer_against_oracle(cuda):
    layers = _setup(cuda)
    bt = _batch(layers)
    bt.run()
    torch.cuda.synchronize()
    for L in layers:
        _check_oracle(L)


def test_train_step_tiles_follow_weight_updates(cuda):
    layers = _setup(cuda)
    bt = _batch(layers)
    bt.run()
    torch.cuda.synchronize()
    for L in layers:                       # an SGD-like in-place update of the weights
        L["w"].add_(L["dW"].half(), alpha=-0.01)
        L["G"].mul_(0.5)
    bt.run()
    torch.cuda.synchronize()
    for L in layers:
        _check_oracle(L)
        # the per-call launches on freshly packed tiles agree to FP32 summation order
        wf = L["w"] if L["pf"] is L["p16"] else L["pf"].pack(L["p16"].unpack(L["w"]))
        y = T.spmm(L["pf"], L["pf"].pack_dense(wf), L["X"], tile_n=L["tile"], dense=True)
        dx = T.spmm(L["pt"], L["pt"].pack_dense(L["w"]), L["G"], tile_n=L["tile_t"], dense=True)
        assert rel_l2(L["Y"].cpu().numpy(), y.cpu().numpy()) <= 1e-6
        assert rel_l2(L["dX"].cpu().numpy(), dx.cpu().numpy()) <= 1e-6
        assert torch.equal(L["dW"], T.sddmm(L["p16"], L["G"], L["X"]))


def test_train_step_values_in_csr_order(cuda):
    layers = _setup(cuda, csr_order=True, layers=(0, 5, 9, 11))
    bt = _batch(layers, csr_order=True)
    bt.run()
    torch.cuda.synchronize()
    for L in layers:
        _check_oracle(L, csr_order=True)


def test_train_step_stages(cuda):
    # stage 1 (prologue) then stage 2 (persistent launch) = a full run
    layers = _setup(cuda, layers=(1, 8, 10))
    bt = _batch(layers)
    bt.run()
    torch.cuda.synchronize()
    ref = [(L["Y"].clone(), L["dW"].clone()) for L in layers]
    for L in layers:
        L["Y"].fill_(float("nan"))
        L["dW"].fill_(float("nan"))
    bt.run(stage=1)
    bt.run(stage=2)
    torch.cuda.synchronize()
    for L, (y, dw) in zip(layers, ref):
        assert rel_l2(L["Y"].cpu().numpy(), y.cpu().numpy()) <= 1e-6
        assert torch.equal(L["dW"], dw)


def test_train_step_graph_replays_follow_weights(cuda):
    layers = _setup(cuda, layers=(0, 4, 8, 11))
    bt = _batch(layers)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        bt.run()                          # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        bt.run()
    for rep in range(3):
        for L in layers:                  # eager work between replays: new weights
            L["w"].mul_(-0.75)
            L["Y"].fill_(float("nan"))
            L["dX"].fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = [(L["Y"].clone(), L["dX"].clone(), L["dW"].clone()) for L in layers]
        bt.run()                          # eager run on the same weights
        torch.cuda.synchronize()
        for L, (y, dx, dw) in zip(layers, got):
            assert torch.equal(dw, L["dW"]), rep
            assert rel_l2(y.cpu().numpy(), L["Y"].cpu().numpy()) <= 1e-6, rep
            assert rel_l2(dx.cpu().numpy(), L["dX"].cpu().numpy()) <= 1e-6, rep
    _check_oracle(layers[-1])


def test_train_step_fused_sgd(cuda):
    # the SDDMM items also apply w <- round(w - lr * dW) (sddmm_item(sgd=...)):
    # every run computes forward / dX / dW with the weights its prologue packed,
    # then leaves the updated weights -- bit for bit what a separate
    # tca_sgd_update on the run's dW gives; the next run's products follow them
    lr = 0.0625   # large enough that the updates change most FP16 weights
    layers = _setup(cuda, layers=(0, 5, 9, 11))
    bt = _batch(layers, sgd_lr=lr)
    for run in range(2):
        before = [L["w"].clone() for L in layers]
        bt.run()
        torch.cuda.synchronize()
        for L, w0 in zip(layers, before):
            after = L["w"].clone()
            L["w"].copy_(w0)               # the oracle check uses the weights the run multiplied with
            _check_oracle(L)
            want = w0.clone()
            T.sgd_update(want, L["dW"], lr)
            torch.cuda.synchronize()
            assert torch.equal(after.view(torch.int16), want.view(torch.int16)), (L["case"].A.name, run)
            assert not torch.equal(after, w0), L["case"].A.name
            L["w"].copy_(after)


def test_fused_sgd_bf16_weights_and_rejections(cuda):
    L = _setup(cuda, layers=(3,))[0]
    W = L["case"].A
    wb = (torch.arange(W.nnz, device=cuda, dtype=torch.float32) % 97 - 48.0).div_(64.0).to(torch.bfloat16)
    want = wb.clone()
    T.Batch([T.sddmm_item(L["p16"], L["G"], L["X"], L["dW"], sgd=(wb, 0.5))]).run()
    torch.cuda.synchronize()
    T.sgd_update(want, L["dW"], 0.5)
    torch.cuda.synchronize()
    assert torch.equal(wb.view(torch.int16), want.view(torch.int16))
    with pytest.raises(ValueError):   # SpMM items take no weight update
        T.Batch([dict(T.spmm_item(L["pf"], L["w"], L["X"], L["Y"], L["tile"], dense="pack", values_order=L["p16"]),
                      sgd_w=L["w"], sgd_lr=0.1)])
    with pytest.raises(ValueError):   # too short
        T.Batch([T.sddmm_item(L["p16"], L["G"], L["X"], L["dW"], sgd=(wb[:-1], 0.5))])
    with pytest.raises(T.TcaError) as e:
        T.Batch([T.sddmm_item(L["p16"], L["G"], L["X"], L["dW"], sgd=(wb, float("nan")))])
    assert e.value.name == "TCA_ERR_INVALID_ARG"
    # an SpMM item reading the updated weights directly (tiled values) would race
    with pytest.raises(T.TcaError) as e:
        T.Batch([T.spmm_item(L["p16"], L["w"], L["X"], L["Y"], 64),
                 T.sddmm_item(L["p16"], L["G"], L["X"], L["dW"], sgd=(L["w"], 0.5))])
    assert e.value.name == "TCA_ERR_INVALID_ARG"
    # so would two updates of overlapping weights
    big = torch.zeros(W.nnz + 64, dtype=torch.float16, device=cuda)
    dW2 = torch.empty_like(L["dW"])
    with pytest.raises(T.TcaError) as e:
        T.Batch([T.sddmm_item(L["p16"], L["G"], L["X"], L["dW"], sgd=(big[:W.nnz], 0.5)),
                 T.sddmm_item(L["p16"], L["G"], L["X"], dW2, sgd=(big[8:W.nnz + 8], 0.5))])
    assert e.value.name == "TCA_ERR_INVALID_ARG"


def test_train_step_bf16(cuda):
    # the same training step in BF16: BF16 weights (tiled order) and
    # activations, FP32 G rounded once per run into a batch-owned BF16 copy,
    # the fused update on BF16 weights; against the FP64 oracle on the BF16
    # inputs, and the weights against a separate tca_sgd_update bit for bit
    lr = 0.0625
    cases = workloads.batch12(seed=0)
    layers, items = [], []
    for i in (0, 4, 9, 11):
        c = cases[i]
        W, N = c.A, c.N
        wb = gen.f32_to_bf16_bits(W.values.view(np.float16).astype(np.float32))
        xb = gen.f32_to_bf16_bits(c.B.view(np.float16).astype(np.float32))
        p16 = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
        rpw, tile = T.autotune.choice_for(W.n_rows, W.n_cols, W.nnz, N, batch=True, dense=True)
        rpw_t, tile_t = T.autotune.choice_for(W.n_cols, W.n_rows, W.nnz, N, op="spmm_t", batch=True, dense=True)
        pf = p16 if rpw == 16 else T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
        pt = p16.transpose(rpw_t)
        L = dict(case=c, p16=p16, w=p16.pack(dev_bf16(wb)), X=dev_bf16(xb), G=dev(c.extra["G"]), xb=xb,
                 Y=torch.full((W.n_rows, N), float("nan"), device=cuda),
                 dX=torch.full((W.n_cols, N), float("nan"), device=cuda),
                 dW=torch.full((W.nnz,), float("nan"), device=cuda))
        items += [T.spmm_item(pf, L["w"], L["X"], L["Y"], tile, dense="pack", values_order=p16),
                  T.spmm_item(pt, L["w"], L["G"], L["dX"], tile_t, dense="pack", values_order=p16),
                  T.sddmm_item(p16, L["G"], L["X"], L["dW"], sgd=(L["w"], lr))]
        layers.append(L)
    bt = T.Batch(items)
    for run in range(2):
        before = [L["w"].clone() for L in layers]
        bt.run()
        torch.cuda.synchronize()
        for L, w0 in zip(layers, before):
            W, N = L["case"].A, L["case"].N
            a64 = O.bf16_bits_to_f32(L["p16"].unpack(w0).view(torch.int16).cpu().numpy().view(np.uint16)).astype(np.float64)
            x64 = O.bf16_bits_to_f32(L["xb"]).astype(np.float64)
            g64 = O.bf16_bits_to_f32(O.f32_to_bf16_bits(L["case"].extra["G"])).astype(np.float64)
            rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
            errs = (rel_l2(L["Y"].cpu().numpy(), O.spmm_f64(W.n_rows, N, W.row_offset, W.col_idx, a64, x64)),
                    rel_l2(L["dX"].cpu().numpy(), O.spmm_f64(W.n_cols, N, rpT, cT, a64[perm], g64)),
                    rel_l2(L["p16"].unpack(L["dW"]).cpu().numpy(), O.sddmm_f64(W.n_rows, N, W.row_offset, W.col_idx, g64, x64)))
            assert max(errs) <= TOL, (W.name, run, errs)
            want = w0.clone()
            T.sgd_update(want, L["dW"], lr)
            torch.cuda.synchronize()
            assert torch.equal(L["w"].view(torch.int16), want.view(torch.int16)), (W.name, run)
    # F16 and BF16 items in one batch: rejected
    L = layers[0]
    w16 = torch.zeros_like(L["w"], dtype=torch.float16)
    with pytest.raises(T.TcaError) as e:
        T.Batch([T.spmm_item(L["p16"], w16, L["X"].to(torch.float16), L["Y"], 64, dense="pack", values_order=L["p16"]),
                 T.sddmm_item(L["p16"], L["G"], L["X"], L["dW"])])
    assert e.value.name == "TCA_ERR_UNSUPPORTED"


def test_train_step_rejects_foreign_values_order(cuda):
    a, b = workloads.batch12(seed=0)[0].A, workloads.batch12(seed=0)[1].A
    pa = T.Plan(a.n_rows, a.n_cols, a.row_offset, a.col_idx, 32)
    pb = T.Plan(b.n_rows, b.n_cols, b.row_offset, b.col_idx, 16)
    w = torch.zeros(max(a.nnz, b.nnz), dtype=torch.float16, device=cuda)   # long enough for either pattern
    X = torch.zeros((a.n_cols, 128), dtype=torch.float16, device=cuda)
    Y = torch.empty((a.n_rows, 128), dtype=torch.float32, device=cuda)
    with pytest.raises(T.TcaError) as e:
        T.Batch([T.spmm_item(pa, w, X, Y, 64, dense="pack", values_order=pb)])
    assert e.value.name == "TCA_ERR_INVALID_ARG"


@pytest.mark.parametrize("n", [8, 1003, 4096 * 9 + 5])
def test_sgd_update_matches_fma_rounding(cuda, n):
    # w <- round(w - lr * g) with one FP32 rounding of the fused multiply-add,
    # then round to nearest even into the weight type; the tail (n % 8) included
    rng = np.random.default_rng(n)
    lr = np.float32(0.0375)
    g = rng.standard_normal(n).astype(np.float32)
    for wt, npt in ((torch.float16, np.float16), (torch.bfloat16, None)):
        w32 = rng.standard_normal(n).astype(np.float32)
        if npt is not None:
            w_host = w32.astype(np.float16)
            w = torch.from_numpy(w_host.view(np.int16)).view(torch.float16).to(cuda)
            x = w_host.astype(np.float64)
        else:
            w = torch.from_numpy(w32).to(cuda).to(torch.bfloat16)
            x = w.float().cpu().numpy().astype(np.float64)
        want32 = (x - np.float64(lr) * g.astype(np.float64)).astype(np.float32)
        T.sgd_update(w, dev(g), float(lr))
        torch.cuda.synchronize()
        if npt is not None:
            want = want32.astype(np.float16).view(np.uint16)
            got = w.cpu().numpy().view(np.uint16)
        else:
            want = torch.from_numpy(want32).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
            got = w.view(torch.int16).cpu().numpy().view(np.uint16)
        diff = np.abs(got.astype(np.int64) - want.astype(np.int64))
        assert diff.max() <= 1 and (diff == 0).mean() >= 0.9999, (wt, n)
