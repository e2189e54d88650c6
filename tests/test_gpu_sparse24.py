"""hipSPARSELt 2:4 baseline (tca_hipsparselt_spmm_24, run by the drop-in
driver's --sparse24 mode in a process of its own): the library's STRIP pruning
equals the oracle's on tie-free rows, and its F16 product equals the oracle's
FP32 product of the pruned A rounded to F16 (rel-L2 <= 1e-3 and the reference's
compare_mat rule)."""
import numpy as np
import pytest

from helpers import rel_l2
from oracle import oracle as O
import tca_spmm as T
from tca_spmm import workloads
from test_sparse24 import _tie_free

pytestmark = pytest.mark.gpu


@pytest.mark.skip(reason="hipSPARSELt path switched off: its call sequence ended in a host segmentation fault "
                         "(cause not found); tca_hipsparselt_spmm_24 returns TCA_ERR_UNSUPPORTED")
@pytest.mark.parametrize("M,K,N", [(256, 512, 128), (2048, 512, 256)])
def test_hipsparselt_prune_and_product(M, K, N):
    A = _tie_free(M, K, seed=M + K)
    B = workloads.dense_f16(0, 9, 0, 3, (K, N))
    P, C, ms = T.hipsparselt_spmm_24(A, B, n_iter=2)
    assert ms > 0
    assert np.array_equal(P, O.prune_24_strip(A))
    rp, ci, v = O.dense_to_csr(P)
    ref = O.spmm_f32(M, N, rp, ci, v, B)
    got = C.view(np.float16).astype(np.float32)
    assert rel_l2(got, ref.astype(np.float16).astype(np.float32)) <= 1e-3
    ok, _ = O.compare_mat(ref, got)
    assert ok
