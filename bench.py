#!/usr/bin/env python3
"""Benchmark of the row-balanced CSR x dense FP16 SpMM on one MI355X.

Workload (one "step"): the FFN-shaped GPU-forward suite -- CSR A of 2048x512
and 512x2048 at sparsity 0.7/0.8/0.9/0.95/0.98 (lognormal(0.5) row lengths,
values N(0,0.02)), B = K x 256 N(0,1), FP16 inputs, FP32 accumulate and output:
ten SpMMs per step, inputs resident in HBM (A as panel-dense 16 x 32 tiles of
32-row plans, built once like the tiled values), run as ONE persistent batched
launch (the ten products' work units drain through one grid), K launches back
to back.  value = sum(2*nnz*N) / step time (TFLOP/s).

Also reported: the HBM roofline of that launch (the ten products' algorithmic
bytes / its measured duration vs 8 TB/s), the ten tuned per-call launches
(dense tiles: graph and eager; tiled values: graph), per-matrix warm and cold
launch times, rocSPARSE and hipBLASLt on the same inputs, and the CPU oracle
(FP32-accumulate C port of naive_gemm restricted to nonzeros) timed on the
host cores.
Prints one JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "tca-spmm_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
METRIC = "SpMM eff. TFLOP/s (2·nnz·N) + %HBM roofline; fwd+bwd ms vs plain FP16"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_traffic(name):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE),
    read from the committed summary profiles/<name>.txt ('traffic_bytes_per_launch: X')."""
    path = os.path.join(ROOT, "profiles", name + ".txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def time_graph(torch, fn, iters):
    """Mean ms of fn() captured in a HIP graph and replayed back to back
    (event pair on the capture-free current stream, queue backlog first)."""
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    for _ in range(5):
        g.replay()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if hasattr(torch.cuda, "_sleep"):
        torch.cuda._sleep(5_000_000)
    s.record(st)
    for _ in range(iters):
        g.replay()
    e.record(st)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / iters


def time_eager(torch, fn, iters, backlog=False):
    """Mean ms of fn() launched eagerly: host launch cost included when the
    GPU outruns the host, unless `backlog` queues a spin kernel first so the
    launches are all enqueued before the GPU reaches them."""
    for _ in range(3):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if backlog and hasattr(torch.cuda, "_sleep"):
        torch.cuda._sleep(20_000_000)
    s.record(st)
    for _ in range(iters):
        fn()
    e.record(st)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / iters


def bench_fwd_bwd(torch, np, T, O, dev, libs=True):
    """Twelve uneven layers, forward + dX + dW: per-call path (36 launches,
    tuned tiled-values launches and tuned panel-dense launches) vs one
    persistent launch (panel-dense forward and dX, SDDMM dW) replayed back to
    back and from a HIP graph.  Weights are fixed over the timed steps, so
    their tiled values and dense tiles are built before timing; `repack_ms`
    is what rebuilding them after a weight update costs per step."""
    from tca_spmm import workloads
    from tca_spmm.workloads import alg_bytes_dw, alg_bytes_dx, alg_bytes_fwd
    layers, items, flops, nbytes = [], [], 0, 0
    for case in workloads.batch12(seed=0):
        W, X, G = case.A, case.B, case.extra["G"]
        w_vals = torch.from_numpy(W.values.view(np.int16)).view(torch.float16).to(dev)
        Xd = torch.from_numpy(X.view(np.int16)).view(torch.float16).to(dev)
        Gd = torch.from_numpy(G).to(dev)
        L = dict(X=Xd, G=Gd, w_vals=w_vals, row_offset=W.row_offset, col_idx=W.col_idx,
                 Y=torch.empty((W.n_rows, case.N), dtype=torch.float32, device=dev),
                 dX=torch.empty((W.n_cols, case.N), dtype=torch.float32, device=dev),
                 dW=torch.empty(max(W.nnz, 1), dtype=torch.float32, device=dev)[:W.nnz])
        fwd16 = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, 16)
        L["w16"] = fwd16.pack(w_vals)
        L["sddmm_plan"] = fwd16
        # per-call paths: the tuned launch of each product, tiled values and dense tiles
        for lay, dense in (("t", False), ("d", True)):
            rpw, tile = T.autotune.choice_for(W.n_rows, W.n_cols, W.nnz, case.N, dense=dense)
            rpw_t, tile_t = T.autotune.choice_for(W.n_cols, W.n_rows, W.nnz, case.N, op="spmm_t", dense=dense)
            plan = fwd16 if rpw == 16 else T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw)
            w = L["w16"] if rpw == 16 else plan.pack(w_vals)
            plan_t = fwd16.transpose(rpw_t)
            L[lay] = dict(plan=plan, plan_t=plan_t, tile=tile, tile_t=tile_t, dense=dense,
                          a=plan.pack_dense(w) if dense else w,
                          at=plan_t.pack_dense(L["w16"]) if dense else L["w16"])
        # batched path: panel-dense forward and dX on the batch's choices, SDDMM dW
        rpw_b, tile_b = T.autotune.choice_for(W.n_rows, W.n_cols, W.nnz, case.N, batch=True, dense=True)
        rpw_bt, tile_bt = T.autotune.choice_for(W.n_cols, W.n_rows, W.nnz, case.N, op="spmm_t", batch=True, dense=True)
        pb = fwd16 if rpw_b == 16 else T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, rpw_b)
        wb = L["w16"] if rpw_b == 16 else pb.pack(w_vals)
        pbt = fwd16.transpose(rpw_bt)
        L["b"] = dict(plan=pb, plan_t=pbt, wb=wb, a=pb.pack_dense(wb), at=pbt.pack_dense(L["w16"]))
        items += [T.spmm_item(pb, L["b"]["a"], Xd, L["Y"], tile_b, dense=True),
                  T.spmm_item(pbt, L["b"]["at"], Gd, L["dX"], tile_bt, dense=True),
                  T.sddmm_item(fwd16, Gd, Xd, L["dW"])]
        flops += 3 * 2 * W.nnz * case.N
        nbytes += (alg_bytes_fwd(W.n_rows, W.n_cols, W.nnz, case.N) + alg_bytes_dx(W.n_rows, W.n_cols, W.nnz, case.N)
                   + alg_bytes_dw(W.n_rows, W.n_cols, W.nnz, case.N))
        layers.append(L)

    def per_call(lay, op=None):
        def run():
            for L in layers:
                c = L[lay]
                if op in (None, "fwd"):
                    T.spmm(c["plan"], c["a"], L["X"], out=L["Y"], tile_n=c["tile"], dense=c["dense"])
                if op in (None, "dx"):
                    T.spmm(c["plan_t"], c["at"], L["G"], out=L["dX"], tile_n=c["tile_t"], dense=c["dense"])
                if op in (None, "dw"):
                    T.sddmm(L["sddmm_plan"], L["G"], L["X"], out=L["dW"])
        return run

    def repack():
        for L in layers:
            L["sddmm_plan"].pack(L["w_vals"], out=L["w16"])
            if L["b"]["wb"] is not L["w16"]:
                L["b"]["plan"].pack(L["w_vals"], out=L["b"]["wb"])
            L["b"]["plan"].pack_dense(L["b"]["wb"], out=L["b"]["a"])
            L["b"]["plan_t"].pack_dense(L["w16"], out=L["b"]["at"])

    bt = T.Batch(items)
    # parity gate: the fused launch against the per-call dense-tile launches
    bt.run()
    torch.cuda.synchronize()
    fused_out = [(L["Y"].clone(), L["dX"].clone(), L["dW"].clone()) for L in layers]
    per_call("d")()
    torch.cuda.synchronize()
    for L, (y, dx, dw) in zip(layers, fused_out):
        for got, ref in ((y, L["Y"]), (dx, L["dX"]), (dw, L["dW"])):
            rel = float(torch.linalg.norm(got - ref) / max(float(torch.linalg.norm(ref)), 1e-30))
            if rel > 1e-6:
                raise RuntimeError(f"fused forward + backward differs from the per-call launches (rel-L2 {rel:.2e})")
    plain_eager = time_eager(torch, per_call("t"), 20)
    plain_graph = time_graph(torch, per_call("t"), 20)
    dense_graph = time_graph(torch, per_call("d"), 20)
    fused = time_eager(torch, bt.run, 50, backlog=True)
    fused_graph = time_graph(torch, bt.run, 50)
    repack_ms = time_graph(torch, repack, 20)
    per_op = {lay: {op: time_graph(torch, per_call(lay, op), 20) for op in ("fwd", "dx", "dw")} for lay in ("t", "d")}
    best_plain = min(plain_graph, dense_graph)
    # library counterpart of the dW step: rocSPARSE SDDMM per layer (G rounded to FP16
    # beforehand, the rounding the kernels apply on load)
    lib_dw = lib_dx = None
    if libs:
        try:
            lib_dx = 0.0
            for L in layers:
                W = L["sddmm_plan"]
                ro = torch.from_numpy(L["row_offset"]).to(dev)
                ci = torch.from_numpy(L["col_idx"]).to(dev)
                out = torch.empty((W.n_cols, L["X"].shape[1]), dtype=torch.float32, device=dev)
                lib_dx += T.rocsparse_spmm_t(W.n_rows, L["X"].shape[1], W.n_cols, ro, ci, L["w_vals"], L["G"].half(), out,
                                             n_iter=10)
        except Exception as ex:  # report, do not hide
            lib_dx = float("nan")
            log("rocSPARSE transposed SpMM failed:", ex)
        try:
            lib_dw = 0.0
            for L in layers:
                W = L["sddmm_plan"]
                ro = torch.from_numpy(L["row_offset"]).to(dev)
                ci = torch.from_numpy(L["col_idx"]).to(dev)
                out = torch.empty(max(W.nnz, 1), dtype=torch.float32, device=dev)
                lib_dw += T.rocsparse_sddmm(W.n_rows, W.n_cols, L["X"].shape[1], ro, ci, L["G"].half(), L["X"], out,
                                            n_iter=10)
        except Exception as ex:  # report, do not hide
            lib_dw = float("nan")
            log("rocSPARSE SDDMM failed:", ex)
    return {"plain_per_call_eager_ms": plain_eager, "plain_per_call_graph_ms": plain_graph,
            "dense_per_call_graph_ms": dense_graph, "fused_ms": fused, "fused_graph_ms": fused_graph,
            "speedup_vs_graph": best_plain / fused, "speedup_vs_tiled_graph": plain_graph / fused,
            "speedup_vs_eager": plain_eager / fused, "repack_ms": repack_ms,
            "fused_tflops": flops / (fused * 1e-3) / 1e12, "fused_hbm_frac": nbytes / (fused * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "per_call_graph_ms_by_op": {"tiled": per_op["t"], "dense": per_op["d"]},
            "rocsparse_dw_ms": lib_dw, "rocsparse_dx_ms": lib_dx,
            "launches_plain": 3 * len(layers), "units": bt.n_units, "workgroups": bt.n_workgroups}


def bench_precision(torch, np, T, O, dev):
    """4096x4096 at 0.9, N=512: FP16, BF16 and FP8-E4M3 (row/column scales)
    throughput, with rel-L2 / max-abs against the FP64 oracle on the original
    values (format + kernel error) and on the quantised values (kernel error)."""
    from tca_spmm import workloads
    from tca_spmm.workloads import alg_bytes_fwd
    case = workloads.precision(seed=0)
    A, N = case.A, case.N
    rp, ci = A.row_offset, A.col_idx
    a16 = A.values.view(np.float16)
    b16 = case.B.view(np.float16)
    ref = O.spmm_f64(A.n_rows, N, rp, ci, a16.astype(np.float64), b16.astype(np.float64))
    rpw, tile = T.autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N)
    plan = T.Plan(A.n_rows, A.n_cols, rp, ci, rpw)
    out = {}

    def err(C, r):
        d = C.astype(np.float64) - r
        return float(np.linalg.norm(d) / np.linalg.norm(r)), float(np.abs(d).max())

    va = torch.from_numpy(A.values.view(np.int16)).view(torch.float16).to(dev)
    vb = torch.from_numpy(case.B.view(np.int16)).view(torch.float16).to(dev)
    C = torch.empty((A.n_rows, N), dtype=torch.float32, device=dev)
    rpw_d, tile_d = T.autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N, dense=True)
    plan_d = plan if rpw_d == rpw else T.Plan(A.n_rows, A.n_cols, rp, ci, rpw_d)
    for name, cast in (("f16", torch.float16), ("bf16", torch.bfloat16)):
        ta, tb = plan.pack(va.to(cast)), vb.to(cast)
        da = plan_d.pack_dense(plan_d.pack(va.to(cast)))
        ms_d = time_graph(torch, lambda: T.spmm(plan_d, da, tb, out=C, tile_n=tile_d, dense=True), 50)
        Cd = C.cpu().numpy()
        ms = time_graph(torch, lambda: T.spmm(plan, ta, tb, out=C, tile_n=tile), 50)
        Cn = C.cpu().numpy()
        rq = O.spmm_f64(A.n_rows, N, rp, ci, plan.unpack(ta).float().cpu().numpy().astype(np.float64),
                        tb.float().cpu().numpy().astype(np.float64))
        e_all, m_all = err(Cn, ref)
        e_q, m_q = err(Cn, rq)
        out[name] = {"us": ms * 1e3, "dense_us": ms_d * 1e3, "tflops": 2 * A.nnz * N / (ms * 1e-3) / 1e12,
                     "dense_tflops": 2 * A.nnz * N / (ms_d * 1e-3) / 1e12, "dense_rel_l2_vs_fp64": err(Cd, ref)[0],
                     "hbm_frac": alg_bytes_fwd(A.n_rows, A.n_cols, A.nnz, N) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "rel_l2_vs_fp64": e_all, "max_abs_vs_fp64": m_all,
                     "rel_l2_vs_fp64_on_quantised": e_q, "max_abs_vs_fp64_on_quantised": m_q}
    rpd = torch.from_numpy(rp).to(dev)
    qa, sa = T.quantize_rows_e4m3(rpd, va)
    qb, sb = T.quantize_cols_e4m3(vb)
    ta = plan.pack(qa)
    da = plan_d.pack_dense(plan_d.pack(qa))
    ms_d = time_graph(torch, lambda: T.spmm(plan_d, da, qb, out=C, row_scale=sa, col_scale=sb, tile_n=tile_d, dense=True), 50)
    Cd = C.cpu().numpy()
    ms = time_graph(torch, lambda: T.spmm(plan, ta, qb, out=C, row_scale=sa, col_scale=sb, tile_n=tile), 50)
    Cn = C.cpu().numpy()
    rq = O.spmm_f64(A.n_rows, N, rp, ci, O.e4m3_to_f32(qa.cpu().numpy()).astype(np.float64),
                    O.e4m3_to_f32(qb.cpu().numpy()).astype(np.float64))
    rq *= sa.cpu().numpy().astype(np.float64)[:, None] * sb.cpu().numpy().astype(np.float64)[None, :]
    e_all, m_all = err(Cn, ref)
    e_q, m_q = err(Cn, rq)
    qms = time_graph(torch, lambda: T.quantize_cols_e4m3(vb), 20)
    out["e4m3"] = {"us": ms * 1e3, "dense_us": ms_d * 1e3, "tflops": 2 * A.nnz * N / (ms * 1e-3) / 1e12,
                   "dense_tflops": 2 * A.nnz * N / (ms_d * 1e-3) / 1e12, "dense_rel_l2_vs_fp64": err(Cd, ref)[0],
                   "hbm_frac": alg_bytes_fwd(A.n_rows, A.n_cols, A.nnz, N, 1, 1, True) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "rel_l2_vs_fp64": e_all, "max_abs_vs_fp64": m_all,
                   "rel_l2_vs_fp64_on_quantised": e_q, "max_abs_vs_fp64_on_quantised": m_q,
                   "quantize_cols_us": qms * 1e3}
    out["config"] = {"rows_per_wave": rpw, "tile_n": tile, "dense_rows_per_wave": rpw_d, "dense_tile_n": tile_d,
                     "nnz": A.nnz, "N": N}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sections", default="fwdbwd,precision",
                    help="comma list of extra sections: fwdbwd, precision")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-libs", action="store_true", help="skip rocSPARSE / hipBLASLt")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy 10x (HBM counter calibration)")
    args = ap.parse_args()
    args.sections = set(s for s in args.sections.split(",") if s)
    if args.copy_calibration:
        import torch
        a = torch.empty(256 << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": 256 << 20, "launches": 10}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the path would shard by row groups (B replicated, dX reduce-scattered), which is not run here.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import tca_spmm as T
    from tca_spmm import workloads
    from tca_spmm.workloads import alg_bytes_fwd

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N = 256
    suite = workloads.plain_gpu_suite(seed=0, N=N)

    def f16(bits):
        return torch.from_numpy(np.ascontiguousarray(bits).view(np.int16)).view(torch.float16).to(dev)

    entries = []
    for case in suite:
        A = case.A
        vals = f16(A.values)
        e = dict(case=case, B=f16(case.B), C=torch.empty((A.n_rows, N), dtype=torch.float32, device=dev),
                 flops=2 * A.nnz * N, bytes=alg_bytes_fwd(A.n_rows, A.n_cols, A.nnz, N))
        # per-call launches: each product's tuned (rows_per_wave, tile_n), with A
        # as tiled values ("t") and as panel-dense tiles ("d")
        for lay, dense in (("t", False), ("d", True)):
            rpw, tile = T.autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N, dense=dense)
            plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rows_per_wave=rpw)
            a = plan.pack(vals)
            e[lay] = dict(plan=plan, a=plan.pack_dense(a) if dense else a, tile=tile, rpw=rpw, dense=dense)
        # the step's batch item: panel-dense tiles of a 32-row plan, 64-column tiles
        rpw_b, tile_b = T.autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N, batch=True, dense=True)
        pb = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rows_per_wave=rpw_b)
        e["b"] = dict(plan=pb, a=pb.pack_dense(pb.pack(vals)), tile=tile_b, rpw=rpw_b)
        entries.append(e)
    total_flops = sum(e["flops"] for e in entries)
    total_bytes = sum(e["bytes"] for e in entries)

    def launch(e, lay):
        c = e[lay]
        T.spmm(c["plan"], c["a"], e["B"], out=e["C"], tile_n=c["tile"], dense=c["dense"])

    def per_call(lay):
        def run():
            for e in entries:
                launch(e, lay)
        return run

    # The step: the ten products as ONE persistent batched launch (their work
    # units drain through one grid), the fastest way found to run them.
    batch = T.Batch([T.spmm_item(e["b"]["plan"], e["b"]["a"], e["B"], e["C"], e["b"]["tile"], dense=True)
                     for e in entries])

    def step():
        batch.run()

    # parity gate before timing: every product of the step, and of the per-call
    # launches, against the CPU oracle
    from oracle import oracle as O
    refs = []
    for e in entries:
        A = e["case"].A
        refs.append(O.spmm_f32(A.n_rows, N, A.row_offset, A.col_idx, A.values, e["case"].B))
    worst = 0.0
    for run in (step, per_call("t"), per_call("d")):
        for e in entries:
            e["C"].fill_(float("nan"))
        run()
        torch.cuda.synchronize()
        for e, ref in zip(entries, refs):
            rel = float(np.linalg.norm(e["C"].cpu().numpy() - ref) / np.linalg.norm(ref))
            worst = max(worst, rel)
            assert rel <= 1e-5, f"{e['case'].A.name}: rel-L2 {rel} vs oracle"
    log(f"parity ok: worst rel-L2 vs CPU oracle {worst:.2e}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    # timed region: K steps back to back on one stream, an event pair around
    # them, behind a queue backlog so the host's enqueue cost stays off the GPU
    # clock.  A step is one kernel launch, so no graph is needed to hide launch
    # cost; the same step replayed from a HIP graph is reported beside it.
    K = args.steps
    t_all0, t_all1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    if hasattr(torch.cuda, "_sleep"):
        torch.cuda._sleep(20_000_000)
    wall0 = time.perf_counter()
    t_all0.record(stream)
    for _ in range(K):
        step()
    t_all1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_per_step = t_all0.elapsed_time(t_all1) / K
    value = total_flops / (ms_per_step * 1e-3) / 1e12
    # roofline of the step's kernel (batch_kernel, one launch per step): the
    # ten products' algorithmic bytes / the launch's duration
    launch_us = ms_per_step * 1e3
    achieved = total_bytes / (ms_per_step * 1e-3) / 1e9
    graph_ms = time_graph(torch, step, K)

    # the ten tuned per-call launches, captured in one graph (one stream) and
    # eager, with dense tiles and with tiled values
    per_call_graph_ms = time_graph(torch, per_call("d"), K)
    eager_ms = time_eager(torch, per_call("d"), K)
    tiled_graph_ms = time_graph(torch, per_call("t"), K)

    extra = {"graph_replay_ms_per_step": graph_ms, "wall_s_timed": wall,
             "batch_units": batch.n_units, "batch_workgroups": batch.n_workgroups,
             "per_call_graph_ms_per_step": per_call_graph_ms, "per_call_eager_ms_per_step": eager_ms,
             "per_call_tiled_graph_ms_per_step": tiled_graph_ms,
             "speedup_vs_per_call_graph": per_call_graph_ms / ms_per_step,
             "configs": [{"name": e["case"].A.name, "nnz": e["case"].A.nnz,
                          "per_call_dense": [e["d"]["rpw"], e["d"]["tile"]], "per_call_tiled": [e["t"]["rpw"], e["t"]["tile"]],
                          "batch": [e["b"]["rpw"], e["b"]["tile"]]} for e in entries]}

    # per-matrix times of the tuned per-call launch (dense tiles; tiled values
    # beside it): 50 back-to-back launches after a queue backlog
    per = []
    for e in entries:
        us = time_eager(torch, lambda: launch(e, "d"), 50, backlog=True) * 1e3
        us_t = time_eager(torch, lambda: launch(e, "t"), 50, backlog=True) * 1e3
        per.append({"name": e["case"].A.name, "us": us, "tiled_us": us_t, "tflops": e["flops"] / (us * 1e-6) / 1e12,
                    "hbm_frac": e["bytes"] / (us * 1e-6) / 1e9 / HBM_PEAK_GBS})
    # cold: a 512 MiB write between launches evicts L2 and the Infinity Cache, so
    # every operand comes from HBM (event pair around the launch alone)
    flush = torch.empty(512 << 20, dtype=torch.uint8, device=dev)
    for e, row in zip(entries, per):
        cold = []
        for _ in range(10):
            flush.fill_(1)
            s, t = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(stream)
            launch(e, "d")
            t.record(stream)
            torch.cuda.synchronize()
            cold.append(s.elapsed_time(t) * 1e3)
        row["cold_us"] = float(np.median(cold))
    del flush
    extra["per_matrix"] = per

    # Blocked-ELL (the reference's block-SpMM baseline, a gfx950 kernel here) on
    # the same ten products, blocks of 16 and 32: the format pads every block
    # row to the longest one and every kept block to b x b, so it moves and
    # multiplies `fill` times the nonzeros
    bell = {"rows": [], "ms_per_step": {}}
    for b in (16, 32):
        tot = 0.0
        for i, e in enumerate(entries):
            A = e["case"].A
            M_ = T.BlockedEll.from_csr(A.n_rows, A.n_cols, A.row_offset, A.col_idx, A.values, b)
            Cb = torch.empty_like(e["C"])
            us = time_eager(torch, lambda: T.bell_spmm(M_, e["B"], out=Cb), 50, backlog=True) * 1e3
            launch(e, "d")
            torch.cuda.synchronize()
            rel = float(torch.linalg.norm(Cb - e["C"]) / torch.linalg.norm(e["C"]))
            if b == 16:
                bell["rows"].append({"name": A.name})
            bell["rows"][i][f"b{b}_us"] = us
            bell["rows"][i][f"b{b}_fill"] = M_.values.numel() / max(A.nnz, 1)
            bell["rows"][i][f"b{b}_rel_vs_spmm"] = rel
            tot += us
        bell["ms_per_step"][f"b{b}"] = tot * 1e-3
    extra["blocked_ell"] = bell

    # 2:4 (the reference's cuSPARSELt comparison, here gfx950's sparse MFMA): each
    # product's A densified, pruned to 2:4 and compressed outside the timed
    # region -- a different, denser product than the SpMM's
    s24 = {"rows": []}
    tot = 0.0
    for e in entries:
        A = e["case"].A
        dense = torch.zeros((A.n_rows, A.n_cols), dtype=torch.float16, device=dev)
        rows = torch.repeat_interleave(torch.arange(A.n_rows, device=dev), torch.from_numpy(np.diff(A.row_offset)).to(dev))
        dense[rows, torch.from_numpy(A.col_idx).to(dev).long()] = f16(A.values)
        sp = T.Sparse24.compress(dense)
        C24 = torch.empty_like(e["C"])
        us = time_eager(torch, lambda: T.spmm_24(sp, e["B"], out=C24), 50, backlog=True) * 1e3
        kept = int((sp.values != 0).sum())
        s24["rows"].append({"name": A.name, "us": us, "kept_nnz_frac": kept / max(A.nnz, 1)})
        tot += us
        del dense, sp
    s24["ms_per_step"] = tot * 1e-3
    extra["sparse24"] = s24

    # achievable HBM bandwidth on this box: a 1 GiB device-to-device copy
    src = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    src.fill_(1)
    copy_ms = time_eager(torch, lambda: dst.copy_(src), 10, backlog=True)
    copy_gbs = 2 * src.numel() / (copy_ms * 1e-3) / 1e9      # bytes read + written
    del src, dst

    # library baselines on the same inputs
    if not args.no_libs:
        lib_ms = {"rocsparse": 0.0, "hipblaslt": 0.0}
        for e in entries:
            A = e["case"].A
            rp = torch.from_numpy(A.row_offset).to(dev)
            ci = torch.from_numpy(A.col_idx).to(dev)
            vv = f16(A.values)
            Cl = torch.empty_like(e["C"])
            try:
                lib_ms["rocsparse"] += T.rocsparse_spmm(A.n_rows, N, A.n_cols, rp, ci, vv, e["B"], Cl, n_iter=20)
            except Exception as ex:  # report, do not hide
                lib_ms["rocsparse"] = float("nan")
                log("rocSPARSE failed:", ex)
            dense = torch.zeros((A.n_rows, A.n_cols), dtype=torch.float16, device=dev)
            rows = torch.repeat_interleave(torch.arange(A.n_rows, device=dev), torch.from_numpy(np.diff(A.row_offset)).to(dev))
            dense[rows, ci.long()] = vv
            try:
                lib_ms["hipblaslt"] += T.hipblaslt_gemm(A.n_rows, N, A.n_cols, dense, e["B"], Cl, n_iter=20)
            except Exception as ex:
                lib_ms["hipblaslt"] = float("nan")
                log("hipBLASLt failed:", ex)
        extra["rocsparse_ms_per_step"] = lib_ms["rocsparse"]
        extra["hipblaslt_ms_per_step"] = lib_ms["hipblaslt"]
        extra["speedup_vs_rocsparse"] = lib_ms["rocsparse"] / ms_per_step if lib_ms["rocsparse"] else None
        extra["speedup_vs_hipblaslt"] = lib_ms["hipblaslt"] / ms_per_step if lib_ms["hipblaslt"] else None

    if "fwdbwd" in args.sections:
        extra["fwd_bwd_12_layers"] = bench_fwd_bwd(torch, np, T, O, dev, libs=not args.no_libs)
    if "precision" in args.sections:
        extra["precision_4096"] = bench_precision(torch, np, T, O, dev)

    cpu = None
    if not args.no_cpu:
        try:
            cores = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            cores = min(16, os.cpu_count() or 1)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            for e in entries:
                A = e["case"].A
                O.spmm_f32(A.n_rows, N, A.row_offset, A.col_idx, A.values, e["case"].B, cores)
            reps += 1
        dt = time.perf_counter() - t0
        # one thread on the first matrix, for the scaling of the CPU figure
        A0 = entries[0]["case"].A
        reps1, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < 2.0:
            O.spmm_f32(A0.n_rows, N, A0.row_offset, A0.col_idx, A0.values, entries[0]["case"].B, 1)
            reps1 += 1
        dt1 = time.perf_counter() - t1
        extra["cpu_one_thread"] = {"matrix": A0.name, "ms": dt1 / reps1 * 1e3,
                                   "tflops": entries[0]["flops"] * reps1 / dt1 / 1e12}
        cpu = {"value": total_flops * reps / dt / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
               "sample": f"the same ten-matrix suite (N=256) repeated {reps}x in {dt:.1f}s "
                         f"({dt / reps * 1e3:.2f} ms per suite), FP32 accumulate, OpenMP over rows"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic",
        "config": {"workload": "plain_gpu: CSR 2048x512 and 512x2048 at sparsity 0.7/0.8/0.9/0.95/0.98 "
                               "(lognormal(0.5) rows), B Kx256, FP16 in, FP32 accumulate/out; 10 SpMMs per step "
                               "as one persistent batched launch, A as panel-dense 16x32 tiles",
                   "N": N, "matrices": len(entries), "nnz_total": int(sum(e["case"].A.nnz for e in entries)),
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("batch_plain_gpu_suite_traffic"),
                     "kernel": "batch_kernel (the step's one launch: ten products' algorithmic bytes / the "
                               "launch's duration, event pair around K back-to-back launches)",
                     "alg_bytes_per_step": total_bytes, "launch_us_mean": launch_us,
                     "copy_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
