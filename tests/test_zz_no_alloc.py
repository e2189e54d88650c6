"""tca_csr_transpose_device_noalloc (the transpose on the library's fixed
per-GPU scratch array) against tca_csr_transpose_device (the allocating entry
point), both called through the C ABI: bit-identical column offsets, row
indices and permutation on two calls in a row, the second on a smaller
pattern (the array keeps the first call's contents), with and without a
caller-provided permutation; a pattern beyond the array's size is rejected on
the host before anything is launched."""
import ctypes

import numpy as np
import pytest
import torch

from tca_spmm import _lib, gen

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _transpose(fn, n_rows, n_cols, rp, ci, with_perm, cuda):
    nnz = ci.numel()
    co = torch.full((n_cols + 1,), -7, dtype=torch.int32, device=cuda)
    ri = torch.full((max(nnz, 1),), -7, dtype=torch.int32, device=cuda)
    perm = torch.full((max(nnz, 1),), -7, dtype=torch.int32, device=cuda) if with_perm else None
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.call(fn, n_rows, n_cols, nnz, _vp(rp), _vp(ci), _vp(co), _vp(ri),
              _vp(perm) if with_perm else ctypes.c_void_p(0), stream)
    torch.cuda.synchronize()
    return co.cpu().numpy(), ri.cpu().numpy()[:nnz], (perm.cpu().numpy()[:nnz] if with_perm else None)


def _pattern(seed, n_rows, n_cols, density):
    rp, ci = gen.pattern("lognormal", n_rows, n_cols, density, 0.5, seed=seed, stream=3)
    return rp, ci


def test_noalloc_matches_allocating_transpose_twice(cuda):
    shapes = [(4096, 1024, 0.1, 1), (700, 3000, 0.02, 2)]   # second call smaller
    for with_perm in (True, False):
        for n_rows, n_cols, d, seed in shapes:
            rp, ci = _pattern(seed, n_rows, n_cols, d)
            rpd = torch.from_numpy(rp).to(cuda)
            cid = torch.from_numpy(ci).to(cuda)
            ref = _transpose("tca_csr_transpose_device", n_rows, n_cols, rpd, cid, with_perm, cuda)
            got = _transpose("tca_csr_transpose_device_noalloc", n_rows, n_cols, rpd, cid, with_perm, cuda)
            for r, g in zip(ref, got):
                if r is not None:
                    assert np.array_equal(r, g), (n_rows, n_cols, with_perm)
            # and the CSC is the transpose of the CSR (scipy-free check)
            co, ri, _ = got
            rows = np.repeat(np.arange(n_rows), np.diff(rp))
            order = np.lexsort((rows, ci))
            assert np.array_equal(ri, rows[order])
            assert np.array_equal(co, np.searchsorted(ci[order], np.arange(n_cols + 1)))


def test_noalloc_rejects_oversized_pattern(cuda):
    cap = _lib.lib().tca_csr_transpose_noalloc_max_scratch()
    assert cap == 64 << 20
    nnz = cap // 16 + 1024          # more than the fixed array holds (16 B per entry before the sort temp)
    n_rows = 2
    rp = torch.tensor([0, nnz // 2, nnz], dtype=torch.int32, device=cuda)
    ci = torch.zeros(nnz, dtype=torch.int32, device=cuda)
    co = torch.empty(9, dtype=torch.int32, device=cuda)
    ri = torch.empty(nnz, dtype=torch.int32, device=cuda)
    with pytest.raises(_lib.TcaError) as e:
        _lib.call("tca_csr_transpose_device_noalloc", n_rows, 8, nnz, _vp(rp), _vp(ci), _vp(co), _vp(ri),
                  ctypes.c_void_p(0), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert e.value.name == "TCA_ERR_UNSUPPORTED"
