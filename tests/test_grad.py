"""Sparse-linear differentiation at full size (W 4096x1024 at 0.9, X 1024x2048,
G 4096x2048): the FP64 gradient formula is pinned by central differences
(h = 1e-3, 256 sampled W entries and 256 sampled X entries, exact for this
linear loss up to roundoff); the GPU gradients (SDDMM dW, transposed-SpMM dX)
are checked against the dense FP64 gradient at rel-L2 <= 1e-3, and their FP32
outputs against the FP64 oracle on the same fp16-rounded inputs at 1e-5.
The two finite-difference tests need no GPU and run in the CPU suite."""
import numpy as np
import pytest
import torch

from helpers import dev, dev_bf16, dev_f16, f16_to_f64, rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import gen, workloads

@pytest.fixture(scope="module")
def problem():
    case = workloads.sparse_linear()
    W, X, G = case.A, case.B, case.extra["G"]
    M, K, N = W.n_rows, W.n_cols, X.shape[1]
    w64, x64, g64 = f16_to_f64(W.values), f16_to_f64(X), G.astype(np.float64)
    dW = O.sddmm_f64(M, N, W.row_offset, W.col_idx, g64, x64)
    rpT, cT, perm = O.transpose_csr(M, K, W.row_offset, W.col_idx)
    dX = O.spmm_f64(K, N, rpT, cT, w64[perm], g64)
    return dict(W=W, X=X, G=G, M=M, K=K, N=N, w64=w64, x64=x64, g64=g64, dW=dW, dX=dX, rpT=rpT, cT=cT, perm=perm)


def _row_loss(p, i, vals_row):
    """sum_n G[i,n] * (W[i,:] X)[i,n] for row i with the given values."""
    rp = np.array([0, vals_row.size], np.int32)
    cols = p["W"].col_idx[p["W"].row_offset[i]:p["W"].row_offset[i + 1]]
    y = O.spmm_f64(1, p["N"], rp, cols, vals_row, p["x64"], nthreads=1)
    return float(np.dot(p["g64"][i], y[0]))


def test_fd_pins_fp64_dw(problem):
    p = problem
    h = 1e-3
    rng = np.random.default_rng(0)
    rows_of = np.repeat(np.arange(p["M"]), np.diff(p["W"].row_offset))
    for e in rng.choice(p["W"].nnz, 256, replace=False):
        i = rows_of[e]
        r0, r1 = p["W"].row_offset[i], p["W"].row_offset[i + 1]
        base = p["w64"][r0:r1].copy()
        plus, minus = base.copy(), base.copy()
        plus[e - r0] += h
        minus[e - r0] -= h
        fd = (_row_loss(p, i, plus) - _row_loss(p, i, minus)) / (2 * h)
        assert abs(fd - p["dW"][e]) <= 1e-6 * (1 + abs(p["dW"][e])), (e, fd, p["dW"][e])


def test_fd_pins_fp64_dx(problem):
    p = problem
    h = 1e-3
    rng = np.random.default_rng(1)
    W = p["W"]
    for _ in range(256):
        j, n = int(rng.integers(p["K"])), int(rng.integers(p["N"]))
        col = p["x64"][:, n].copy()

        def loss(c):
            y = O.spmm_f64(p["M"], 1, W.row_offset, W.col_idx, p["w64"], c[:, None], nthreads=8)
            return float(np.dot(p["g64"][:, n], y[:, 0]))
        cp, cm = col.copy(), col.copy()
        cp[j] += h
        cm[j] -= h
        fd = (loss(cp) - loss(cm)) / (2 * h)
        assert abs(fd - p["dX"][j, n]) <= 1e-6 * (1 + abs(p["dX"][j, n])), (j, n, fd, p["dX"][j, n])


@pytest.mark.gpu
def test_gpu_gradients_against_dense_fp64(cuda, problem):
    p = problem
    W = p["W"]
    layer = T.SparseLinear(p["M"], p["K"], W.row_offset, W.col_idx, W.values, rows_per_wave=16)
    X = dev_f16(p["X"]).requires_grad_(True)
    G = dev(p["G"])
    Y = layer(X)
    # forward against the oracle
    Yref = O.spmm_f64(p["M"], p["N"], W.row_offset, W.col_idx, p["w64"], p["x64"])
    assert rel_l2(Y.detach().cpu().numpy(), Yref) <= 1e-5
    (Y * G).sum().backward()
    dW = layer.csr_values(layer.weight.grad).float().cpu().numpy()
    dX = X.grad.float().cpu().numpy()
    assert rel_l2(dW, p["dW"]) <= 1e-3
    assert rel_l2(dX, p["dX"]) <= 1e-3
    # raw FP32 kernel outputs vs FP64 on the fp16-rounded G they consume
    g16 = p["G"].astype(np.float16).astype(np.float64)
    dW32 = layer.plan.unpack(T.sddmm(layer.plan, G, X.detach())).cpu().numpy()
    ref_dw = O.sddmm_f64(p["M"], p["N"], W.row_offset, W.col_idx, g16, p["x64"])
    assert rel_l2(dW32, ref_dw) <= 1e-5
    dX32 = T.spmm(layer.plan_t, layer.weight.detach(), G).cpu().numpy()
    ref_dx = O.spmm_f64(p["K"], p["N"], p["rpT"], p["cT"], p["w64"][p["perm"]], g16)
    assert rel_l2(dX32, ref_dx) <= 1e-5


@pytest.mark.gpu
def test_gpu_gradients_dense_tiles_match_tiled(cuda, problem):
    """SparseLinear(dense=True) -- forward and dX on panel-dense tiles packed
    from the parameter in each call -- gives the tiled-values layer's Y, dW and
    dX bit for bit."""
    p = problem
    W = p["W"]
    out = []
    for dense in (False, True):
        layer = T.SparseLinear(p["M"], p["K"], W.row_offset, W.col_idx, W.values, rows_per_wave=32,
                               rows_per_wave_t=16, tile_n=64, dense=dense)
        X = dev_f16(p["X"]).requires_grad_(True)
        Y = layer(X)
        (Y * dev(p["G"])).sum().backward()
        out.append((Y.detach(), layer.weight.grad, X.grad))
    for a, b in zip(*out):
        assert torch.equal(a, b)


@pytest.mark.gpu
def test_gpu_gradients_fp32_activations(cuda, problem):
    """An FP32 X: the forward rounds it to FP16 on load and dW = SDDMM(G, X)
    takes the same rounded X; W.grad and X.grad against the FP64 formula on the
    FP16-rounded X (rel-L2 <= 1e-3, G rounded to FP16 by the kernels)."""
    p = problem
    W = p["W"]
    rng = np.random.default_rng(5)
    X32 = (p["X"].view(np.float16).astype(np.float32) + rng.standard_normal(p["X"].shape).astype(np.float32) * 1e-3)
    x64 = X32.astype(np.float16).astype(np.float64)
    layer = T.SparseLinear(p["M"], p["K"], W.row_offset, W.col_idx, W.values, rows_per_wave=16)
    X = dev(X32).requires_grad_(True)
    Y = layer(X)
    Yref = O.spmm_f64(p["M"], p["N"], W.row_offset, W.col_idx, p["w64"], x64)
    assert rel_l2(Y.detach().cpu().numpy(), Yref) <= 1e-5
    (Y * dev(p["G"])).sum().backward()
    assert X.grad.dtype == torch.float32
    ref_dw = O.sddmm_f64(p["M"], p["N"], W.row_offset, W.col_idx, p["g64"], x64)
    dW = layer.csr_values(layer.weight.grad).float().cpu().numpy()
    assert rel_l2(dW, ref_dw) <= 1e-3
    assert rel_l2(X.grad.cpu().numpy(), p["dX"]) <= 1e-3


@pytest.mark.gpu
def test_gpu_gradients_bf16(cuda, problem):
    """BF16 weights and activations: Y and the FP32 kernel outputs of dW and dX
    (G rounded to BF16, as the layer rounds it) against the FP64 oracle on the
    BF16 inputs at 1e-5; W.grad and X.grad, BF16 like their tensors, against the
    same at 4e-3 (one BF16 rounding of each gradient: ~2^-9 relative)."""
    p = problem
    W, M, K, N = p["W"], p["M"], p["K"], p["N"]
    wb = gen.f32_to_bf16_bits(W.values.view(np.float16).astype(np.float32))
    xb = gen.f32_to_bf16_bits(p["X"].view(np.float16).astype(np.float32))
    gb = gen.f32_to_bf16_bits(p["G"])
    w64, x64, g64 = (O.bf16_bits_to_f32(b).astype(np.float64) for b in (wb, xb, gb))
    layer = T.SparseLinear(M, K, W.row_offset, W.col_idx, dev_bf16(wb), rows_per_wave=16)
    X = dev_bf16(xb).requires_grad_(True)
    Y = layer(X)
    assert rel_l2(Y.detach().cpu().numpy(), O.spmm_f64(M, N, W.row_offset, W.col_idx, w64, x64)) <= 1e-5
    (Y * dev(p["G"])).sum().backward()
    ref_dw = O.sddmm_f64(M, N, W.row_offset, W.col_idx, g64, x64)
    ref_dx = O.spmm_f64(K, N, p["rpT"], p["cT"], w64[p["perm"]], g64)
    assert layer.weight.grad.dtype == torch.bfloat16 and X.grad.dtype == torch.bfloat16
    assert rel_l2(layer.csr_values(layer.weight.grad).float().cpu().numpy(), ref_dw) <= 4e-3
    assert rel_l2(X.grad.float().cpu().numpy(), ref_dx) <= 4e-3
    Gb = dev_bf16(gb)
    assert rel_l2(layer.plan.unpack(T.sddmm(layer.plan, Gb, X.detach())).cpu().numpy(), ref_dw) <= 1e-5
    assert rel_l2(T.spmm(layer.plan_t, layer.weight.detach(), Gb).cpu().numpy(), ref_dx) <= 1e-5
