"""2:4 structured-sparsity baseline on hipSPARSELt (replaces
perform_cuSPARSELT_24sparsity, src/spmm_cusparse.cu:2-179).

PyTorch bundles an older hipSPARSELt, and a process that imported torch binds
the library's hipSPARSELt calls to that copy (tca_hipsparselt_spmm_24 refuses
such a process).  The baseline therefore runs in a child process of its own --
the drop-in driver's `--sparse24` mode -- with the operands passed as files.
"""
import os
import subprocess
import tempfile

import numpy as np

_EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tca_test")


def hipsparselt_spmm_24(A_bits, B_bits, n_iter=20, timeout=120):
    """A (M x K) and B (K x N) as F16 bits -> (pruned A bits, C = pruned A @ B as F16
    bits, mean ms per call over n_iter calls)."""
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
