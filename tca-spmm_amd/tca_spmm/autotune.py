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

`tune_batch()` chooses, layer by layer, the batched variants of the fused
training step (forward and dX of every layer inside one persistent launch) by
coordinate descent on the measured time of the whole batch; `record_batch()`
stores them as entries marked "batch", which `choice_for(batch=True,
dense=True)` returns for that layer and op.
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


# (rows_per_wave, tile_n) of the persistent batch kernel's panel-dense SpMM
# variants with an F16 B (an F32 B is read through the batch's F16 copy): one
# wave per 16 / 32 / 64 rows, or block units (the workgroup's four waves share
# each B chunk tile) on 64- and 128-row plans with 64 columns
BATCH_DENSE_CANDIDATES = ((16, 16), (16, 32), (16, 64), (32, 32), (32, 64), (64, 32), (64, 64), (128, 64))


def _nearest(entries, op, M, K, N, d, layout, batch):
    best, best_dist = None, None
    for e in entries:
        if (e.get("op", "spmm") != op or e["M"] != M or e["K"] != K or e.get("layout", "tiled") != layout
                or bool(e.get("batch", False)) != batch):
            continue
        dist = abs(math.log(max(e["density"], 1e-6)) - math.log(max(d, 1e-6))) + abs(math.log(e["N"] / N))
        if best is None or dist < best_dist:
            best, best_dist = e, dist
    return best if best is not None and best_dist < 0.7 else None


def choice_for(M, K, nnz, N, op="spmm", batch=False, dense=False):
    """(rows_per_wave, tile_n) for an M x K matrix (op "spmm") or for the
    product with the transpose of a K x M matrix whose transpose is M x K
    (op "spmm_t"), with nnz nonzeros and N dense columns.  dense=True: for A
    as panel-dense tiles.

    batch=True: a variant the persistent batch kernel has.  With dense=True
    the layer's own entry recorded by tune_batch() (nearest density, same M,
    K, N), else the rule: there the parallelism comes from the many items,
    and every work unit pays a fixed claim + pipeline-start cost, so few,
    long units win: tiled values on 16 rows x 64 columns; panel-dense tiles
    on 32 rows (each gathered B row serves two MFMA panels) x 64 columns,
    forward and transposed."""
    d = nnz / float(max(1, M * K))
    layout = "dense" if dense else "tiled"
    if batch:
        if dense:
            e = _nearest(_load()["entries"], op, M, K, N, d, layout, True)
            if e is not None:
                return e["rows_per_wave"], e["tile_n"]
        t64 = 64 if N >= 64 else (32 if N >= 32 else 16)
        if dense and N >= 32:
            return 32, t64
        return 16, t64
    e = _nearest(_load()["entries"], op, M, K, N, d, layout, False)
    if e is not None:
        return e["rows_per_wave"], e["tile_n"]
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
                      if e.get("batch") or not (e.get("op", "spmm") == op and e["M"] == M and e["K"] == K and e["N"] == N
                                                and abs(e["density"] - d) < 1e-3 and e.get("layout", "tiled") == layout)]
    tab["entries"].append({"op": op, "layout": layout, "M": M, "K": K, "N": N, "density": round(d, 5),
                           "rows_per_wave": best["rows_per_wave"], "tile_n": best["tile_n"], "us": best["us"],
                           "candidates": results})
    with open(path or TABLE_PATH, "w") as f:
        json.dump(tab, f, indent=1)


def tune_batch(cases, candidates=BATCH_DENSE_CANDIDATES, iters=20, passes=1, device="cuda", log=None,
               start="table", sgd_lr=1e-6):
    """Per-layer choice of the fused training step's batched SpMM variants:
    forward (op "spmm") and dX (op "spmm_t") of every layer, by coordinate
    descent on the measured time of the whole batch run (prologue packing the
    dense tiles from the weight values + the persistent launch over forward,
    dX and dW of all layers, the SDDMM items applying the SGD update with
    sgd_lr; None: no update).  cases: workloads.SpmmCase with extra["G"].
    start: "table" = the choices choice_for(batch=True, dense=True) gives now
    (recorded ones or the rule), "rule" = 32 rows x 64 columns everywhere.
    Returns (choices, us): choices[i] = {"fwd": (rpw, tile), "dx": (rpw, tile)}
    and the step's batch time with them; every candidate tried is timed."""
    import numpy as np
    import torch

    from .batch import Batch, sddmm_item, spmm_item
    from .core import Plan

    layers = []
    for c in cases:
        W = c.A
        v = torch.from_numpy(W.values.view(np.int16)).view(torch.float16).to(device)
        p16 = Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
        layers.append(dict(case=c, v=v, plans={16: p16}, tplans={},
                           X=torch.from_numpy(c.B.view(np.int16)).view(torch.float16).to(device),
                           G=torch.from_numpy(c.extra["G"]).to(device),
                           Y=torch.empty((W.n_rows, c.N), dtype=torch.float32, device=device),
                           dX=torch.empty((W.n_cols, c.N), dtype=torch.float32, device=device)))

    def plan(L, r):
        if r not in L["plans"]:
            W = L["case"].A
            L["plans"][r] = Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, r)
        return L["plans"][r]

    def tplan(L, r):
        if r not in L["tplans"]:
            L["tplans"][r] = L["plans"][16].transpose(r)
        return L["tplans"][r]

    def measure(choice):
        items, keep = [], []
        for L, ch in zip(layers, choice):
            pw = plan(L, ch["fwd"][0])
            pt = tplan(L, ch["dx"][0])
            w = pw.pack(L["v"])
            dW = torch.empty(max(L["case"].A.nnz, 1), dtype=torch.float32, device=device)
            keep += [w, dW]
            items += [spmm_item(pw, w, L["X"], L["Y"], ch["fwd"][1], dense="pack", values_order=pw),
                      spmm_item(pt, w, L["G"], L["dX"], ch["dx"][1], dense="pack", values_order=pw),
                      sddmm_item(pw, L["G"], L["X"], dW, sgd=None if sgd_lr is None else (w, sgd_lr))]
        b = Batch(items)
        for _ in range(3):
            b.run()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        if hasattr(torch.cuda, "_sleep"):
            torch.cuda._sleep(2_000_000)
        s.record()
        for _ in range(iters):
            b.run()
        e.record()
        torch.cuda.synchronize()
        b.close()
        return s.elapsed_time(e) * 1e3 / iters

    choice = []
    for L in layers:
        W, N = L["case"].A, L["case"].N
        if start == "table":
            choice.append({"fwd": tuple(choice_for(W.n_rows, W.n_cols, W.nnz, N, batch=True, dense=True)),
                           "dx": tuple(choice_for(W.n_cols, W.n_rows, W.nnz, N, op="spmm_t", batch=True, dense=True))})
        else:
            t = 64 if N >= 64 else 32
            choice.append({"fwd": (32, t), "dx": (32, t)})
    best = measure(choice)
    if log:
        log(f"batch start {best:.1f} us")
    for _ in range(passes):
        for i in range(len(layers)):
            for op in ("fwd", "dx"):
                for cand in candidates:
                    if tuple(cand) == tuple(choice[i][op]) or cand[1] > max(16, layers[i]["case"].N):
                        continue
                    trial = [dict(c) for c in choice]
                    trial[i][op] = tuple(cand)
                    us = measure(trial)
                    if us < best * 0.995:
                        best, choice = us, trial
                if log:
                    log(f"layer {i} {op}: {choice[i][op]} -> {best:.1f} us")
    return choice, best


def record_batch(cases, choices, us, path=None):
    """Store the batch choices (one entry per layer and op, marked batch) in the
    tuning table; choice_for(batch=True, dense=True) reads them back."""
    tab = _load()
    for c, ch in zip(cases, choices):
        A, N = c.A, c.N
        d = A.nnz / float(A.n_rows * A.n_cols)
        for op, key in (("spmm", "fwd"), ("spmm_t", "dx")):
            M, K = (A.n_cols, A.n_rows) if op == "spmm_t" else (A.n_rows, A.n_cols)
            tab["entries"] = [e for e in tab["entries"]
                              if not (e.get("batch") and e.get("op", "spmm") == op and e["M"] == M and e["K"] == K
                                      and e["N"] == N and abs(e["density"] - d) < 1e-3)]
            tab["entries"].append({"op": op, "layout": "dense", "batch": True, "M": M, "K": K, "N": N,
                                   "density": round(d, 5), "rows_per_wave": int(ch[key][0]), "tile_n": int(ch[key][1]),
                                   "us": us})
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
    ap.add_argument("--batch", action="store_true", help="also tune the batched training step of batch12")
    ap.add_argument("--batch-only", action="store_true", help="only tune the batched training step (keeps the table)")
    ap.add_argument("--passes", type=int, default=1, help="coordinate-descent passes of the batch tuning")
    ap.add_argument("--update", action="store_true",
                    help="start from the current table and replace only the entries tuned now")
    args = ap.parse_args(argv)
    global _table
    dev = torch.device("cuda:0")
    if args.batch_only:
        _table = None
        _load()
        cases = workloads.batch12()
        ch, us = tune_batch(cases, passes=args.passes, log=lambda m: print(m, flush=True))
        record_batch(cases, ch, us, path=args.out)
        return
    if args.update:
        _table = None
        _load()
    else:
        _table = {"entries": []}

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
    if args.batch:
        cases = workloads.batch12()
        ch, us = tune_batch(cases, passes=args.passes, log=lambda m: print(m, flush=True))
        record_batch(cases, ch, us, path=args.out)


if __name__ == "__main__":
    main()
