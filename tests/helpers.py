"""Shared helpers for the GPU tests: host<->device copies and error metrics."""
import numpy as np
import torch

import tca_spmm as T
from oracle import oracle as O


def dev_f16(bits, device="cuda"):
    return torch.from_numpy(np.ascontiguousarray(bits).view(np.int16)).view(torch.float16).to(device)


def dev_bf16(bits, device="cuda"):
    return torch.from_numpy(np.ascontiguousarray(bits).view(np.int16)).view(torch.bfloat16).to(device)


def dev(x, device="cuda"):
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def rel_l2(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(got - ref) / den) if den > 0 else float(np.linalg.norm(got))


def f16_to_f64(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float64)


def run_spmm_f16(A, B_bits, rpw=16, tile_n=0, device="cuda"):
    """GPU C = A @ B through the C ABI (values packed to tiled order)."""
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rows_per_wave=rpw)
    vals = dev_f16(A.values if A.nnz else np.zeros(1, np.uint16), device)
    tiled = plan.pack(vals)
    C = T.spmm(plan, tiled, dev_f16(B_bits, device), tile_n=tile_n)
    torch.cuda.synchronize()
    return C.cpu().numpy(), plan


def oracle_f32(A, B_bits):
    return O.spmm_f32(A.n_rows, B_bits.shape[1], A.row_offset, A.col_idx, A.values, B_bits)


def oracle_f64(A, B_bits):
    return O.spmm_f64(A.n_rows, B_bits.shape[1], A.row_offset, A.col_idx, f16_to_f64(A.values), f16_to_f64(B_bits))
