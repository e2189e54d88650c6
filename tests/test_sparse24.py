"""2:4 STRIP pruning on the CPU (the oracle's restatement of the rule
perform_cuSPARSELT_24sparsity asks the library for, src/spmm_cusparse.cu:95-96):
hand-worked strips, the two largest magnitudes by brute force on tie-free rows,
idempotence, at most two survivors per strip with their bits kept."""
import numpy as np
import pytest

from oracle import oracle as O


def _f16(x):
    return np.asarray(x, np.float32).astype(np.float16).view(np.uint16)


def _tie_free(M, K, seed):
    """Rows whose magnitudes are distinct inside every strip (so the rule has one answer)."""
    rng = np.random.default_rng(seed)
    mags = np.stack([rng.permutation(4) for _ in range(M * K // 4)]).reshape(M, K).astype(np.float32)
    mags = 0.5 + mags + rng.integers(0, 8, (M, K // 4)).repeat(4, axis=1) * 4.0
    sign = np.where(rng.random((M, K)) < 0.5, -1.0, 1.0)
    return _f16(mags * sign)


def test_prune_known_answer():
    A = _f16([[1, -3, 2, 0.5, 0, 0, 0, 1, 4, 4, 1, 1, -0.25, 0.5, -0.75, 1]])
    got = O.prune_24_strip(A).view(np.float16).astype(np.float32)
    # L1 of each pair, worked by hand: (1,2)=5; zeros tie, (0,3) first; (0,1)=8; (2,3)=1.75
    assert got.tolist() == [[0, -3, 2, 0, 0, 0, 0, 1, 4, 4, 0, 0, 0, 0, -0.75, 1]]


def test_prune_keeps_two_largest_on_tie_free_rows():
    A = _tie_free(64, 128, seed=3)
    P = O.prune_24_strip(A)
    a = np.abs(A.view(np.float16).astype(np.float32)).reshape(64, 32, 4)
    keep = np.zeros_like(a, bool)
    top2 = np.argsort(-a, axis=2)[:, :, :2]
    np.put_along_axis(keep, top2, True, axis=2)
    expect = np.where(keep.reshape(64, 128), A, 0).astype(np.uint16)
    assert np.array_equal(P, expect)


def test_prune_idempotent_and_at_most_two_per_strip():
    rng = np.random.default_rng(5)
    A = _f16(rng.standard_normal((33, 64)) * (rng.random((33, 64)) < 0.7))
    P = O.prune_24_strip(A)
    assert np.array_equal(O.prune_24_strip(P), P)
    nz = (P.reshape(33, 16, 4) & 0x7FFF) != 0
    assert nz.sum(axis=2).max() <= 2
    kept = P != 0
    assert np.array_equal(P[kept], A[kept])        # survivors keep their bits


def test_prune_rejects_ragged_k():
    with pytest.raises(ValueError):
        O.prune_24_strip(np.zeros((2, 6), np.uint16))


def test_hipsparselt_refuses_a_foreign_copy():
    """In a process that imported torch, the hipSPARSELt calls bind to PyTorch's
    bundled copy; the C ABI refuses before touching the device."""
    import ctypes
    import tca_spmm  # noqa: F401  (imports torch first)
    from tca_spmm import _lib
    L = _lib.lib()
    ms = ctypes.c_double()
    maps = open("/proc/self/maps").read()
    if not any("torch/lib/libhipsparselt" in line for line in maps.splitlines()):
        pytest.skip("this torch build bundles no hipSPARSELt")
    # placeholder pointers: the call must return before using any of them
    s = L.tca_hipsparselt_spmm_24(64, 64, 64, ctypes.c_void_p(16), None, ctypes.c_void_p(32), ctypes.c_void_p(48), 0,
                                  ctypes.byref(ms))
    assert s == 2 and b"another copy" in L.tca_last_error()
