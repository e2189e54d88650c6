"""2:4 structured sparsity (replaces perform_cuSPARSELT_24sparsity,
src/spmm_cusparse.cu:2-179).

`Sparse24.compress` prunes (STRIP rule) and compresses a dense matrix on the GPU
and `spmm_24` multiplies with gfx950's sparse MFMA -- both kernels of
libtca_spmm.so.  `hipsparselt_spmm_24` is the library leg: PyTorch bundles an
older hipSPARSELt, and a process that imported torch binds the library's
hipSPARSELt calls to that copy (tca_hipsparselt_spmm_24 refuses such a
process), so it runs in a child process of its own -- the drop-in driver's
`--sparse24` mode -- with the operands passed as files.
"""
import os
import subprocess
import tempfile

import numpy as np

_EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tca_test")


def hipsparselt_spmm_24(A_bits, B_bits, n_iter=20, timeout=120):
    """A (M x K) and B (K x N) as F16 bits -> (pruned A bits, C = pruned A @ B as F16
    bits, mean ms per call over n_iter calls).  The library leg is switched off
    for now (tca_hipsparselt_spmm_24 returns TCA_ERR_UNSUPPORTED, so this raises
    RuntimeError); Sparse24 / spmm_24 are the 2:4 path."""
    A = np.ascontiguousarray(A_bits, np.uint16)
    B = np.ascontiguousarray(B_bits, np.uint16)
    (M, K), N = A.shape, B.shape[1]
    if B.shape[0] != K:
        raise ValueError("B must have K rows")
    if not os.path.exists(_EXE):
        raise RuntimeError(f"{_EXE} is not built (make -C tca-spmm_amd tca_test)")
    with tempfile.TemporaryDirectory() as d:
        A.tofile(os.path.join(d, "A.bin"))
        B.tofile(os.path.join(d, "B.bin"))
        r = subprocess.run([_EXE, "--sparse24", str(M), str(N), str(K), str(n_iter), d], capture_output=True, text=True,
                           timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError(f"tca_test --sparse24 failed ({r.returncode}): {r.stderr.strip() or r.stdout.strip()}")
        P = np.fromfile(os.path.join(d, "P.bin"), np.uint16).reshape(M, K)
        C = np.fromfile(os.path.join(d, "C.bin"), np.uint16).reshape(M, N)
    ms = float(r.stdout.split("hipSPARSELt 2:4 ms")[1].split()[0])
    return P, C, ms


class Sparse24:
    """A compressed 2:4 matrix in device memory: values (n_rows x n_cols/2, the
    dense dtype) and meta (n_rows x n_cols/16 int16 words of 2-bit positions)."""

    def __init__(self, n_rows, n_cols, values, meta):
        self.n_rows, self.n_cols, self.values, self.meta = int(n_rows), int(n_cols), values, meta

    @classmethod
    def compress(cls, dense, stream=None):
        """Prune a dense float16 / bfloat16 device matrix (STRIP rule) and compress it."""
        import torch
        from ._lib import call
        from .core import _ptr, _stream, tca_dtype
        M, K = dense.shape
        values = torch.empty((M, K // 2), dtype=dense.dtype, device=dense.device)
        meta = torch.empty((M, K // 16), dtype=torch.int16, device=dense.device)
        call("tca_compress_24", M, K, tca_dtype(dense), _ptr(dense), dense.stride(0), _ptr(values), _ptr(meta),
             _stream(stream))
        return cls(M, K, values, meta)

    def to_dense(self):
        """The pruned dense matrix, rebuilt on the host (uint16 bits)."""
        import torch
        v = self.values.cpu().contiguous().view(torch.int16).numpy().view(np.uint16)
        m = self.meta.cpu().numpy().view(np.uint16).astype(np.int64)
        out = np.zeros((self.n_rows, self.n_cols), np.uint16)
        j = np.arange(8)
        pos = (m[:, :, None] >> (2 * j)) & 3                               # [M, K/16, 8]
        col = np.arange(self.n_cols // 16)[None, :, None] * 16 + 4 * (j // 2) + pos
        rows = np.repeat(np.arange(self.n_rows), self.n_cols // 2)
        out[rows, col.reshape(-1)] = v.reshape(-1)
        return out


def spmm_24(s24, B, out=None, tile_n=0, rows_per_wave=0, stream=None):
    """C[M x N] (float32) = s24 @ B with the sparse MFMA; B K x N, the values' dtype or float32."""
    import torch
    from ._lib import call
    from .core import _ptr, _stream, tca_dtype
    K, N = B.shape
    if K != s24.n_cols:
        raise ValueError(f"B has {K} rows, the matrix has {s24.n_cols} columns")
    if out is None:
        out = torch.empty((s24.n_rows, N), dtype=torch.float32, device=B.device)
    call("tca_spmm_24", s24.n_rows, s24.n_cols, N, tca_dtype(s24.values), _ptr(s24.values), _ptr(s24.meta),
         tca_dtype(B), _ptr(B), B.stride(0), _ptr(out), out.stride(0), tile_n, rows_per_wave, _stream(stream))
    return out
