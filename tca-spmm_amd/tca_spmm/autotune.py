"""Autotuner for the two launch parameters of the panel SpMM:
rows_per_wave (16/32/64 rows sharing one condensed column list per wave, or
128/256 rows per workgroup sharing each gathered B chunk) and tile_n
(16/32/64/128 output columns; 64/128 for the 128/256-row groups).

`tune()` times every candidate on the device with HIP events (launches queued
behind a sleep so host enqueue cost stays off the clock) and returns them
fastest first.  Results recorded with `record()` go to tuning_table.json next
to this file; `choice_for()` looks up the nearest recorded shape (same M, K,
N class, closest density) and falls back to a density rule when none exists.
Entries are kept per A layout: tiled values, or panel-dense tiles
(`layout="dense"`).
"""
import json
import math
import os

RPW_CANDIDATES = (16, 32, 64)
TILE_CANDIDATES = (16, 32, 64, 128)
WIDE_RPW_CANDIDATES = (128, 256)
WIDE_TILE_CANDIDATES = (64, 128)
ALL_CANDIDATES = ([(r, t) for r in RPW_CANDIDATES for t in TILE_CANDIDATES]
                  + [(r, t) for r in WIDE_RPW_CANDIDATES for t in WIDE_TILE_CANDIDATES])
TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_table.json")

_table = None


def _load():
    global _table
    if _table is None:
        try:
            with open(TABLE_PATH) as f:
                _table = json.load(f)
        except (OSError, ValueError):
            _table = {"entries": []}
    return _table


def heuristic(M, K, nnz, N):
    d = nnz / float(max(1, M * K))
    rpw = 32 if d >= 0.05 else 16
    tile = 64 if N >= 128 else (32 if N >= 32 else 16)
    return rpw, tile


# (rows_per_wave -> tile_n values) compiled into the persistent batch kernel
BATCH_VARIANTS = {"spmm": {16: (16, 32, 64), 32: (32,)}, "spmm_t": {16: (16, 32, 64)}}


def choice_for(M, K, nnz, N, op="spmm", batch=False, dense=False):
    """(rows_per_wave, tile_n) for an M x K matrix (op "spmm") or for the
    product with the transpose of a K x M matrix whose transpose is M x K
    (op "spmm_t"), with nnz nonzeros and N dense columns.  dense=True: for A
    as panel-dense tiles.

    batch=True: a variant the persistent batch kernel has.  There the
    parallelism comes from the many items, and every work unit pays a fixed
    claim + pipeline-start cost, so few, long units win: tiled values on
    16 rows x 64 columns; panel-dense tiles on 32 rows (each gathered B row
    serves two MFMA panels) x 64 columns, forward and transposed (a batch
    reads an F32 gradient through its own F16 copy, which has the 64-column
    variant).  Measured faster than the per-call tuned choices on the
    ten-product suite and the twelve-layer forward + backward step (there the
    transposed products at 32 x 64 took the step from 193 to 164 us; 32 x 32
    had been the best while the F32 gradient was gathered directly)."""
    if batch:
        t64 = 64 if N >= 64 else (32 if N >= 32 else 16)
        if dense and N >= 32:
            return 32, t64
        return 16, t64
    d = nnz / float(max(1, M * K))
    layout = "dense" if dense else "tiled"
    best, best_dist = None, None
    for e in _load()["entries"]:
        if e.get("op", "spmm") != op or e["M"] != M or e["K"] != K or e.get("layout", "tiled") != layout:
            continue
        dist = abs(math.log(max(e["density"], 1e-6)) - math.log(max(d, 1e-6))) + abs(math.log(e["N"] / N))
        if best is None or dist < best_dist:
            best, best_dist = e, dist
    if best is not None and best_dist < 0.7:
        return best["rows_per_wave"], best["tile_n"]
    return heuristic(M, K, nnz, N)


def tune(A, B_dev, values_dev=None, candidates=None, iters=30, device="cuda", transposed=False, dense=False):
    """Time each (rows_per_wave, tile_n) on the device.  A: workloads.SparseMatrix
    (fp16 values), B_dev: K x N device tensor (for transposed=True: the
    M x N float32 gradient, and the product is A^T B).  dense=True times A as
    panel-dense tiles.  Returns a list of dicts sorted by time (us per launch)."""
    import numpy as np
    import torch

    from .core import Plan, spmm

    if values_dev is None:
        values_dev = torch.from_numpy(A.values.view(np.int16)).view(torch.float16).to(device)
    N = B_dev.shape[1]
    cands = candidates or ALL_CANDIDATES

    out = torch.empty((A.n_cols if transposed else A.n_rows, N), dtype=torch.float32, device=device)
    stream = torch.cuda.current_stream()
    res = []
    plans = {}
    fwd = Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rows_per_wave=16) if transposed else None
    fwd_tiled = fwd.pack(values_dev) if transposed else None
    for rpw, tile in cands:
        if rpw not in plans:
            if transposed:
                pt = fwd.transpose(rpw)
                plans[rpw] = (pt, pt.pack_dense(fwd_tiled) if dense else fwd_tiled)
            else:
                p = Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rows_per_wave=rpw)
                t = p.pack(values_dev)
                plans[rpw] = (p, p.pack_dense(t) if dense else t)
        p, tiled = plans[rpw]
        for _ in range(3):
            spmm(p, tiled, B_dev, out=out, tile_n=tile, dense=dense)
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        if hasattr(torch.cuda, "_sleep"):
            torch.cuda._sleep(1_000_000)
        s.record(stream)
        for _ in range(iters):
            spmm(p, tiled, B_dev, out=out, tile_n=tile, dense=dense)
        e.record(stream)
        torch.cuda.synchronize()
        res.append({"rows_per_wave": rpw, "tile_n": tile, "us": s.elapsed_time(e) * 1e3 / iters})
    res.sort(key=lambda r: r["us"])
    return res


def record(A, N, results, op="spmm", path=None, layout="tiled"):
    """Store the fastest candidate for this shape (and A layout) in the tuning table."""
    tab = _load()
    d = A.nnz / float(A.n_rows * A.n_cols)
    best = results[0]
    M, K = (A.n_cols, A.n_rows) if op == "spmm_t" else (A.n_rows, A.n_cols)
    tab["entries"] = [e for e in tab["entries"]
                      if not (e.get("op", "spmm") == op and e["M"] == M and e["K"] == K and e["N"] == N
                              and abs(e["density"] - d) < 1e-3 and e.get("layout", "tiled") == layout)]
    tab["entries"].append({"op": op, "layout": layout, "M": M, "K": K, "N": N, "density": round(d, 5),
                           "rows_per_wave": best["rows_per_wave"], "tile_n": best["tile_n"], "us": best["us"],
                           "candidates": results})
    with open(path or TABLE_PATH, "w") as f:
        json.dump(tab, f, indent=1)


def main(argv=None):
    """Tune the built-in workloads on the current device and write a table.

        python -m tca_spmm.autotune --out tuning_table.json [--workloads plain_gpu,batch12,...]
    """
    import argparse

    import numpy as np
    import torch

    from . import workloads

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE_PATH)
    ap.add_argument("--workloads", default="plain_gpu,batch12,precision,sparse_linear")
    ap.add_argument("--iters", type=int, default=30)
    ap.add_argument("--layouts", default="tiled,dense")
    args = ap.parse_args(argv)
    global _table
    _table = {"entries": []}
    dev = torch.device("cuda:0")

    def f16(bits):
        return torch.from_numpy(np.ascontiguousarray(bits).view(np.int16)).view(torch.float16).to(dev)

    jobs = []
    sel = set(args.workloads.split(","))
    if "plain_gpu" in sel:
        jobs += [(c.A, c.B, None) for c in workloads.plain_gpu_suite()]
    if "batch12" in sel:
        jobs += [(c.A, c.B, c.extra["G"]) for c in workloads.batch12()]
    if "precision" in sel:
        c = workloads.precision()
        jobs.append((c.A, c.B, None))
    if "sparse_linear" in sel:
        c = workloads.sparse_linear()
        jobs.append((c.A, c.B, c.extra["G"]))
    for layout in args.layouts.split(","):
        dense = layout == "dense"
        for A, B, G in jobs:
            r = tune(A, f16(B), iters=args.iters, dense=dense)
            record(A, B.shape[1], r, op="spmm", path=args.out, layout=layout)
            print(f"{layout:5s} spmm   {A.name:40s} N={B.shape[1]:5d} best rpw={r[0]['rows_per_wave']} "
                  f"tile={r[0]['tile_n']} {r[0]['us']:.1f} us (worst {r[-1]['us']:.1f})", flush=True)
            if G is not None:
                rt = tune(A, torch.from_numpy(G).to(dev), iters=args.iters, transposed=True, dense=dense)
                record(A, G.shape[1], rt, op="spmm_t", path=args.out, layout=layout)
                print(f"{layout:5s} spmm_t {A.name:40s} N={G.shape[1]:5d} best rpw={rt[0]['rows_per_wave']} "
                      f"tile={rt[0]['tile_n']} {rt[0]['us']:.1f} us (worst {rt[-1]['us']:.1f})", flush=True)


if __name__ == "__main__":
    main()
