"""The synthetic workloads of this path, generated from seeded streams.

  plain_cpu      512x512 CSR, row nnz ~ Binomial(512, 0.1), values U[-1,1],
                 B 512x256 N(0,1)  (the reference's CPU-runnable case)
  plain_gpu      FFN-shaped 2048x512 and 512x2048 at sparsity
                 {0.7, 0.8, 0.9, 0.95, 0.98}, lognormal(0.5) row lengths,
                 values N(0, 0.02), B K x 256 N(0,1)
  precision      4096x4096 at 0.9, row lengths uniform +-20%, values
                 N(0, 0.02), B 4096x512 N(0,1)
  sparse_linear  W 4096x1024 at 0.9 (lognormal rows), X 1024x2048 N(0,1),
                 G 4096x2048 float32 N(0,1)
  batch12        twelve uneven layers (4x 512x2048, 4x 2048x512, 2x 1024^2,
                 2x 4096x1024), sparsity U[0.5, 0.98], N in {128, 384, 1024}
All matrices are host numpy arrays (fp16 as uint16 bits).
"""
from dataclasses import dataclass, field

import numpy as np

from . import gen


@dataclass
class SparseMatrix:
    name: str
    n_rows: int
    n_cols: int
    row_offset: np.ndarray
    col_idx: np.ndarray
    values: np.ndarray          # fp16 bits, CSR order

    @property
    def nnz(self):
        return int(self.col_idx.size)

    @property
    def sparsity(self):
        return 1.0 - self.nnz / float(self.n_rows * self.n_cols)


@dataclass
class SpmmCase:
    A: SparseMatrix
    B: np.ndarray               # fp16 bits, K x N
    extra: dict = field(default_factory=dict)

    @property
    def N(self):
        return int(self.B.shape[1])


def _stream(cfg, layer, tensor):
    return cfg * 1000 + layer * 10 + tensor


def _values(seed, stream, n, kind, a, b):
    if kind == "uniform":
        return gen.f32_to_f16_bits(gen.uniform_f32(seed, stream, (n,), a, b))
    return gen.f32_to_f16_bits(gen.normal_f32(seed, stream, (n,), a, b))


def sparse(name, M, K, density, kind, param, seed, cfg, layer=0, value_kind="normal", va=0.0, vb=0.02):
    rp, ci = gen.pattern(kind, M, K, density, param, seed=seed, stream=_stream(cfg, layer, 1))
    vals = _values(seed, _stream(cfg, layer, 2), ci.size, value_kind, va, vb)
    return SparseMatrix(name, M, K, rp, ci, vals)


def dense_f16(seed, cfg, layer, tensor, shape):
    return gen.f32_to_f16_bits(gen.normal_f32(seed, _stream(cfg, layer, tensor), shape))


def plain_cpu(seed=0, N=256):
    A = sparse("plain_cpu_512x512_s0.9", 512, 512, 0.1, "bernoulli", 0.0, seed, 1, value_kind="uniform", va=-1.0, vb=1.0)
    return SpmmCase(A, dense_f16(seed, 1, 0, 3, (512, N)))


PLAIN_GPU_SHAPES = ((2048, 512), (512, 2048))
PLAIN_GPU_SPARSITIES = (0.7, 0.8, 0.9, 0.95, 0.98)


def plain_gpu(M, K, sparsity, seed=0, N=256):
    layer = PLAIN_GPU_SHAPES.index((M, K)) * 10 + PLAIN_GPU_SPARSITIES.index(sparsity)
    A = sparse(f"plain_gpu_{M}x{K}_s{sparsity}", M, K, 1.0 - sparsity, "lognormal", 0.5, seed, 2, layer)
    return SpmmCase(A, dense_f16(seed, 2, layer, 3, (K, N)))


def plain_gpu_suite(seed=0, N=256):
    return [plain_gpu(M, K, s, seed, N) for (M, K) in PLAIN_GPU_SHAPES for s in PLAIN_GPU_SPARSITIES]


def precision(seed=0, M=4096, K=4096, N=512):
    A = sparse(f"precision_{M}x{K}_s0.9", M, K, 0.1, "uniform", 0.2, seed, 3)
    return SpmmCase(A, dense_f16(seed, 3, 0, 3, (K, N)))


def sparse_linear(seed=0, M=4096, K=1024, N=2048):
    W = sparse(f"sparse_linear_{M}x{K}_s0.9", M, K, 0.1, "lognormal", 0.5, seed, 4)
    X = dense_f16(seed, 4, 0, 3, (K, N))
    G = gen.normal_f32(seed, _stream(4, 0, 4), (M, N))
    return SpmmCase(W, X, extra={"G": G})


BATCH12_SHAPES = [(512, 2048)] * 4 + [(2048, 512)] * 4 + [(1024, 1024)] * 2 + [(4096, 1024)] * 2
BATCH12_NS = (128, 384, 1024)


def batch12(seed=0, with_grad=True):
    """Twelve uneven layers: (W, X, G) per layer, sparsity and N drawn from streams."""
    cases = []
    s_draw = gen.uniform_f32(seed, _stream(5, 99, 1), (len(BATCH12_SHAPES),), 0.5, 0.98)
    n_draw = gen.uniform_f32(seed, _stream(5, 99, 2), (len(BATCH12_SHAPES),), 0.0, 3.0)
    for i, (M, K) in enumerate(BATCH12_SHAPES):
        s = float(s_draw[i])
        N = BATCH12_NS[min(int(n_draw[i]), 2)]
        W = sparse(f"batch12_l{i}_{M}x{K}_s{s:.3f}", M, K, 1.0 - s, "lognormal", 0.5, seed, 5, i)
        X = dense_f16(seed, 5, i, 3, (K, N))
        extra = {"G": gen.normal_f32(seed, _stream(5, i, 4), (M, N))} if with_grad else {}
        cases.append(SpmmCase(W, X, extra=extra))
    return cases


def alg_bytes_fwd(M, K, nnz, N, sv=2, sb=2, scales=False):
    """Format-independent bytes of one forward SpMM: CSR (value + int32 column),
    row offsets, B once, C (f32) once; FP8 adds the row and column scales."""
    b = (sv + 4) * nnz + 4 * (M + 1) + sb * K * N + 4 * M * N
    if scales:
        b += 4 * M + 4 * N
    return b


def alg_bytes_dx(M, K, nnz, N):
    """dX = W^T G: W (fp16 + index), K+1 offsets, G (f32) once, dX (f32) once."""
    return 6 * nnz + 4 * (K + 1) + 4 * M * N + 4 * K * N


def alg_bytes_dw(M, K, nnz, N):
    """dW = SDDMM(G, X^T): index, offsets, G (f32), X (fp16), dW (f32)."""
    return 4 * nnz + 4 * (M + 1) + 4 * M * N + 2 * K * N + 4 * nnz
