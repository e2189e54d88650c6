#!/usr/bin/env python3
"""Benchmark of the row-balanced CSR x dense SpMM path on one MI355X.

Headline workload (one "step"): the twelve-layer sparse-weight training step --
twelve uneven CSR layers (4x 512x2048, 4x 2048x512, 2x 1024x1024,
2x 4096x1024, sparsity U[0.5, 0.98], lognormal(0.5) rows, N in
{128, 384, 1024}), FP16 weights and activations, FP32 upstream gradient G.
A step is: forward Y = W X, dX = W^T G and dW = SDDMM(G, X^T) of all twelve
layers with the current weights, and the SGD update of every weight from its
dW (w -= lr * dW).  The fused path runs it as one batch run: the prologue
launch (FP16 copy of every G, panel-dense tiles of every W rebuilt from the
current values) + ONE persistent launch over all forward / dX / dW work
units, whose SDDMM units also apply the weight update to every entry they
write.  value = sum over layers of 3 * 2*nnz*N (forward, dX, dW) / step
time, in TFLOP/s.

Beside it, in the same run: the plain per-call FP16 path (36 tuned launches of
tca_spmm / tca_sddmm on tiled values, with its own repack and the same SGD
update, one HIP graph) and the per-call panel-dense path with its repack; the
roofline of the persistent launch (the step's algorithmic bytes / its measured
duration vs 8 TB/s); the CPU oracle on the same step (all host cores); the
plain GPU suite (ten FFN-shaped products) as one batched launch, per call,
Blocked-ELL, 2:4, rocSPARSE and hipBLASLt; the 4096^2 x 512 precision sweep
with hipBLASLt on the densified matrix; the 512x512 CPU-runnable case on the
GPU and on the CPU.  Prints one JSON line on stdout.
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
MFMA_F16_PEAK_TFLOPS = 2500.0
METRIC = "SpMM eff. TFLOP/s (2·nnz·N) + %HBM roofline; fwd+bwd ms vs plain FP16"
SGD_LR = 1e-6


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


def time_cold(torch, np, fn, reps, flush):
    """Median ms of fn() with every cache level flushed first: a write of
    `flush` (>= 512 MiB, beyond the L2s and the 256 MB Infinity Cache)
    between runs, each run bracketed by its own event pair."""
    fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    out = []
    for _ in range(reps):
        flush.fill_(1)
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record(st)
        fn()
        e.record(st)
        torch.cuda.synchronize()
        out.append(s.elapsed_time(e))
    return float(np.median(out))


def rel_l2(np, got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))


# --------------------------------------------------------------------------
# the twelve-layer training step
# --------------------------------------------------------------------------
class TrainStep:
    """The step run three ways, each path with the weights in its own
    preferred order, its own FP16 weight buffer and FP32 dW buffer (all layers
    in one buffer each, so a per-call path's SGD update is one launch):

      "f"  fused: weights in the tiled order of the batch's forward plan (the
           order its SDDMM writes dW in); the batch packs the forward and
           transposed dense tiles from them in its prologue launch, and its
           SDDMM items update them from dW (sddmm_item(sgd=...));
      "t"  plain per call: weights in the tiled order of each layer's tuned
           forward plan; forward, dX (transposed plan reading the same values)
           and dW are 36 tca_spmm / tca_sddmm launches, no repack needed;
      "d"  per call on panel-dense tiles: as "t" with the tuned dense plans,
           plus the repack of both dense tile sets after every update."""

    PATHS = ("f", "t", "d")

    def __init__(self, torch, np, T, dev, seed=0):
        from tca_spmm import workloads
        from tca_spmm.workloads import alg_bytes_dw, alg_bytes_dx, alg_bytes_fwd
        self.torch, self.np, self.T = torch, np, T
        cases = workloads.batch12(seed=seed)
        offs, tot = [], 0
        for c in cases:
            offs.append(tot)
            tot += (c.A.nnz + 7) // 8 * 8            # 16-byte aligned views
        self.w_all = {p: torch.zeros(max(tot, 8), dtype=torch.float16, device=dev) for p in self.PATHS}
        self.dw_all = {p: torch.zeros(max(tot, 8), dtype=torch.float32, device=dev) for p in self.PATHS}
        self.layers, self.flops, self.alg_bytes, self.dense_flops = [], 0, 0, 0
        items = []
        for c, off in zip(cases, offs):
            W, N = c.A, c.N
            L = dict(case=c, N=N,
                     X=torch.from_numpy(c.B.view(np.int16)).view(torch.float16).to(dev),
                     G=torch.from_numpy(c.extra["G"]).to(dev),
                     Y=torch.empty((W.n_rows, N), dtype=torch.float32, device=dev),
                     dX=torch.empty((W.n_cols, N), dtype=torch.float32, device=dev))
            v_csr = torch.from_numpy(W.values.view(np.int16)).view(torch.float16).to(dev)
            plans, tplans = {}, {}
            for path in self.PATHS:
                if path == "f":
                    r, t = T.autotune.choice_for(W.n_rows, W.n_cols, W.nnz, N, batch=True, dense=True)
                    r_t, t_t = T.autotune.choice_for(W.n_cols, W.n_rows, W.nnz, N, op="spmm_t", batch=True, dense=True)
                else:
                    dense = path == "d"
                    r, t = T.autotune.choice_for(W.n_rows, W.n_cols, W.nnz, N, dense=dense)
                    r_t, t_t = T.autotune.choice_for(W.n_cols, W.n_rows, W.nnz, N, op="spmm_t", dense=dense)
                if r not in plans:
                    plans[r] = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, r)
                plan = plans[r]
                if (r, r_t) not in tplans:
                    tplans[(r, r_t)] = plan.transpose(r_t)
                plan_t = tplans[(r, r_t)]
                w = self.w_all[path][off:off + W.nnz]
                w.copy_(plan.pack(v_csr))
                P = dict(plan=plan, plan_t=plan_t, tile=t, tile_t=t_t, w=w, dW=self.dw_all[path][off:off + W.nnz],
                         choice=[[r, t], [r_t, t_t]])
                if path == "d":
                    P["a"], P["at"] = plan.pack_dense(w), plan_t.pack_dense(w)
                L[path] = P
            f = L["f"]
            items += [T.spmm_item(f["plan"], f["w"], L["X"], L["Y"], f["tile"], dense="pack", values_order=f["plan"]),
                      T.spmm_item(f["plan_t"], f["w"], L["G"], L["dX"], f["tile_t"], dense="pack", values_order=f["plan"]),
                      T.sddmm_item(f["plan"], L["G"], L["X"], f["dW"], sgd=(f["w"], SGD_LR))]
            self.layers.append(L)
            self.flops += 3 * 2 * W.nnz * N
            self.dense_flops += 3 * 2 * W.n_rows * W.n_cols * N
            self.alg_bytes += (alg_bytes_fwd(W.n_rows, W.n_cols, W.nnz, N) + alg_bytes_dx(W.n_rows, W.n_cols, W.nnz, N)
                               + alg_bytes_dw(W.n_rows, W.n_cols, W.nnz, N))
        self.items = items   # per layer: forward, dX, dW (+ update)
        self.batch = T.Batch(items)

    def sgd(self, path):
        self.T.sgd_update(self.w_all[path], self.dw_all[path], SGD_LR)

    def repack(self, path):
        """What a path rebuilds after a weight update: the dense tiles of "d"
        (the fused path packs inside its batch run; "t" reads its values)."""
        if path == "d":
            for L in self.layers:
                P = L["d"]
                P["plan"].pack_dense(P["w"], out=P["a"])
                P["plan_t"].pack_dense(P["w"], out=P["at"])

    def products(self, path, op=None):
        T = self.T
        dense = path == "d"
        for L in self.layers:
            P = L[path]
            if op in (None, "fwd"):
                T.spmm(P["plan"], P["a"] if dense else P["w"], L["X"], out=L["Y"], tile_n=P["tile"], dense=dense)
            if op in (None, "dx"):
                T.spmm(P["plan_t"], P["at"] if dense else P["w"], L["G"], out=L["dX"], tile_n=P["tile_t"], dense=dense)
            if op in (None, "dw"):
                T.sddmm(P["plan"], L["G"], L["X"], out=P["dW"])

    def fused_step(self):
        self.batch.run()   # products with the current weights, then their update

    def per_call_step(self, path):
        def run():
            self.sgd(path)
            self.repack(path)
            self.products(path)
        return run

    def parity_gate(self, O):
        """Fused outputs of every layer against the per-call tiled launches
        (FP32 summation order; dW mapped through CSR order) on two steps in a
        row -- the second multiplies with the weights the first step's fused
        update left, the per-call path applies the same update in between --
        and two layers of the first step against the FP64 oracle."""
        torch, np = self.torch, self.np

        def weights(path):   # every layer's weights in CSR order
            return torch.cat([L[path]["plan"].unpack(L[path]["w"]) for L in self.layers]).view(torch.int16)

        worst = 0.0
        for step in range(2):
            w_before = weights("f")
            if step == 1:    # both paths multiply with the same updated weights
                if not torch.equal(weights("t"), w_before):
                    raise RuntimeError("per-call weights differ from the fused step's after one update")
                log(f"parity gate: {int((w_before != w_first).sum())} of {w_before.numel()} weights changed by "
                    "the first step's update")
            w_first = w_before
            self.batch.run()
            torch.cuda.synchronize()
            got = [(L["Y"].clone(), L["dX"].clone(), L["f"]["plan"].unpack(L["f"]["dW"])) for L in self.layers]
            if step == 0:
                fused = got
            self.products("t")
            torch.cuda.synchronize()
            for L, (y, dx, dw) in zip(self.layers, got):
                for g, ref in ((y, L["Y"]), (dx, L["dX"]), (dw, L["t"]["plan"].unpack(L["t"]["dW"]))):
                    r = float(torch.linalg.norm(g - ref) / max(float(torch.linalg.norm(ref)), 1e-30))
                    worst = max(worst, r)
                    if r > 1e-6:
                        raise RuntimeError(f"{L['case'].A.name}: fused step {step} differs from the per-call "
                                           f"launches ({r:.2e})")
            self.sgd("t")
        oracle_err = {}
        for i in (0, len(self.layers) - 1):
            L = self.layers[i]
            W, N = L["case"].A, L["N"]
            a64 = W.values.view(np.float16).astype(np.float64)
            x64 = L["case"].B.view(np.float16).astype(np.float64)
            g64 = L["case"].extra["G"].astype(np.float16).astype(np.float64)
            ref_y = O.spmm_f64(W.n_rows, N, W.row_offset, W.col_idx, a64, x64)
            rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
            ref_dx = O.spmm_f64(W.n_cols, N, rpT, cT, a64[perm], g64)
            ref_dw = O.sddmm_f64(W.n_rows, N, W.row_offset, W.col_idx, g64, x64)
            y, dx, dw = fused[i]
            e = [rel_l2(np, y.cpu().numpy(), ref_y), rel_l2(np, dx.cpu().numpy(), ref_dx),
                 rel_l2(np, dw.cpu().numpy(), ref_dw)]
            if max(e) > 1e-5:
                raise RuntimeError(f"{W.name}: fused step vs FP64 oracle {e}")
            oracle_err[W.name] = {"fwd": e[0], "dx": e[1], "dw": e[2]}
        return worst, oracle_err

    def cpu_step(self, O, threads):
        """The same step's products on the CPU oracle (FP32 accumulation of FP16
        products): forward, dX on the transposed CSR, dW, every layer."""
        np = self.np
        for L in self.layers:
            W, N = L["case"].A, L["N"]
            if "cpu" not in L:
                rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
                g16 = L["case"].extra["G"].astype(np.float16).view(np.uint16)
                L["cpu"] = (rpT, cT, np.ascontiguousarray(W.values[perm]), g16)
            rpT, cT, vT, g16 = L["cpu"]
            O.spmm_f32(W.n_rows, N, W.row_offset, W.col_idx, W.values, L["case"].B, threads)
            O.spmm_f32(W.n_cols, N, rpT, cT, vT, g16, threads)
            O.sddmm_f32(W.n_rows, N, W.row_offset, W.col_idx, g16, L["case"].B, threads)


def bench_train_step(torch, np, T, O, dev, args):
    ts = TrainStep(torch, np, T, dev)
    worst, oracle_err = ts.parity_gate(O)
    log(f"train step parity ok: worst rel-L2 vs per-call {worst:.2e}, oracle {oracle_err}")
    for _ in range(args.warmup):
        ts.fused_step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    K = args.steps
    # timed region: K steps back to back on one stream, an event pair around
    # them, behind a queue backlog so the host's enqueue cost stays off the clock
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    if hasattr(torch.cuda, "_sleep"):
        torch.cuda._sleep(20_000_000)
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(K):
        ts.fused_step()
    t1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_per_step = t0.elapsed_time(t1) / K
    # the persistent launch alone (the step's dominant kernel), K back to back
    ts.batch.run(stage=1)
    main_ms = time_eager(torch, lambda: ts.batch.run(stage=2), K, backlog=True)
    prologue_ms = time_eager(torch, lambda: ts.batch.run(stage=1), K, backlog=True)
    # the persistent launch over one kind of item at a time (the same items:
    # every layer's forward, dX, or dW with its update), to show its composition
    main_by_op = {}
    for k, op in enumerate(("fwd", "dx", "dw") if "composition" in args.sections else ()):
        sub = T.Batch(ts.items[k::3])
        sub.run(stage=1)
        main_by_op[op] = time_eager(torch, lambda sub=sub: sub.run(stage=2), K, backlog=True)
        del sub
    # the separate update kernel the per-call paths launch (the fused path
    # applies it inside the SDDMM units)
    sgd_ms = time_eager(torch, lambda: ts.sgd("f"), K, backlog=True)
    graph_ms = time_graph(torch, ts.fused_step, K)
    # cold: every operand from HBM (a 512 MiB write evicts L2 and the Infinity Cache)
    flush = torch.empty(512 << 20, dtype=torch.uint8, device=dev)
    cold_ms = time_cold(torch, np, ts.fused_step, 10, flush)
    plain_cold_ms = time_cold(torch, np, ts.per_call_step("t"), 10, flush)
    del flush
    # the per-call paths, each with its own repack and the same SGD update
    plain_graph = time_graph(torch, ts.per_call_step("t"), 20)
    dense_graph = time_graph(torch, ts.per_call_step("d"), 20)
    plain_eager = time_eager(torch, ts.per_call_step("t"), 10)
    repack_d = time_graph(torch, lambda: ts.repack("d"), 20)
    per_op = {lay: {op: time_graph(torch, (lambda lay=lay, op=op: ts.products(lay, op)), 20)
                    for op in ("fwd", "dx", "dw")} for lay in ("t", "d")}
    res = {
        "ms_per_step": ms_per_step, "wall_s_timed": wall, "flops_per_step": ts.flops,
        "dense_equivalent_flops_per_step": ts.dense_flops, "alg_bytes_per_step": ts.alg_bytes,
        "main_launch_ms": main_ms, "main_launch_ms_one_op_at_a_time": main_by_op, "prologue_ms": prologue_ms, "sgd_kernel_ms": sgd_ms, "graph_replay_ms_per_step": graph_ms,
        "cold_ms_per_step": cold_ms, "plain_per_call_cold_ms": plain_cold_ms,
        "plain_per_call_graph_ms": plain_graph, "plain_per_call_eager_ms": plain_eager,
        "dense_per_call_graph_ms": dense_graph, "dense_repack_ms": repack_d,
        "speedup_vs_plain_per_call": plain_graph / ms_per_step,
        "speedup_vs_dense_per_call": dense_graph / ms_per_step,
        "speedup_vs_plain_per_call_eager": plain_eager / ms_per_step,
        "per_call_graph_ms_by_op": {"tiled": per_op["t"], "dense": per_op["d"]},
        "units": ts.batch.n_units, "workgroups": ts.batch.n_workgroups,
        "parity": {"worst_rel_l2_vs_per_call": worst, "rel_l2_vs_fp64_oracle": oracle_err},
        "layers": [{"name": L["case"].A.name, "nnz": L["case"].A.nnz, "N": L["N"], "fused": L["f"]["choice"],
                    "per_call_tiled": L["t"]["choice"], "per_call_dense": L["d"]["choice"]} for L in ts.layers],
    }
    if not args.no_libs:
        lib_dx = lib_dw = 0.0
        try:
            for L in ts.layers:
                W = L["case"].A
                ro = torch.from_numpy(W.row_offset).to(dev)
                ci = torch.from_numpy(W.col_idx).to(dev)
                vals = L["t"]["plan"].unpack(L["t"]["w"])
                g16 = L["G"].half()
                out = torch.empty((W.n_cols, L["N"]), dtype=torch.float32, device=dev)
                lib_dx += T.rocsparse_spmm_t(W.n_rows, L["N"], W.n_cols, ro, ci, vals, g16, out, n_iter=5)
                outw = torch.empty(max(W.nnz, 1), dtype=torch.float32, device=dev)
                lib_dw += T.rocsparse_sddmm(W.n_rows, W.n_cols, L["N"], ro, ci, g16, L["X"], outw, n_iter=5)
        except Exception as ex:  # report, do not hide
            log("rocSPARSE backward legs failed:", ex)
            lib_dx = lib_dw = float("nan")
        res["rocsparse_dx_ms"] = lib_dx
        res["rocsparse_dw_ms"] = lib_dw
    res["bf16_step"] = bench_train_step_bf16(torch, np, T, O, dev, K)
    return ts, res


def bench_train_step_bf16(torch, np, T, O, dev, K):
    """The same twelve-layer training step in BF16 (weights and activations
    BF16, G FP32 copied to BF16 by the batch, the fused update on BF16
    weights): one batch run per step, timed like the FP16 step; the first
    run's outputs of two layers against the FP64 oracle on the BF16 inputs."""
    from tca_spmm import gen, workloads
    cases = workloads.batch12(seed=0)
    items, keep, check = [], [], []
    for i, c in enumerate(cases):
        W, N = c.A, c.N
        wb = gen.f32_to_bf16_bits(W.values.view(np.float16).astype(np.float32))
        xb = gen.f32_to_bf16_bits(c.B.view(np.float16).astype(np.float32))
        r, t = T.autotune.choice_for(W.n_rows, W.n_cols, W.nnz, N, batch=True, dense=True)
        r_t, t_t = T.autotune.choice_for(W.n_cols, W.n_rows, W.nnz, N, op="spmm_t", batch=True, dense=True)
        plan = T.Plan(W.n_rows, W.n_cols, W.row_offset, W.col_idx, r)
        plan_t = plan.transpose(r_t)
        w = plan.pack(torch.from_numpy(wb.view(np.int16)).view(torch.bfloat16).to(dev))
        X = torch.from_numpy(xb.view(np.int16)).view(torch.bfloat16).to(dev)
        G = torch.from_numpy(c.extra["G"]).to(dev)
        Y = torch.empty((W.n_rows, N), dtype=torch.float32, device=dev)
        dX = torch.empty((W.n_cols, N), dtype=torch.float32, device=dev)
        dW = torch.empty(max(W.nnz, 1), dtype=torch.float32, device=dev)
        items += [T.spmm_item(plan, w, X, Y, t, dense="pack", values_order=plan),
                  T.spmm_item(plan_t, w, G, dX, t_t, dense="pack", values_order=plan),
                  T.sddmm_item(plan, G, X, dW, sgd=(w, SGD_LR))]
        keep += [plan, plan_t, w, X, G, Y, dX, dW]
        if i in (0, len(cases) - 1):
            check.append((c, plan, wb, xb, Y, dX, dW))
    batch = T.Batch(items)
    batch.run()
    torch.cuda.synchronize()
    errs = {}
    for c, plan, wb, xb, Y, dX, dW in check:
        W, N = c.A, c.N
        a64 = O.bf16_bits_to_f32(wb).astype(np.float64)
        x64 = O.bf16_bits_to_f32(xb).astype(np.float64)
        g64 = O.bf16_bits_to_f32(O.f32_to_bf16_bits(c.extra["G"])).astype(np.float64)
        rpT, cT, perm = O.transpose_csr(W.n_rows, W.n_cols, W.row_offset, W.col_idx)
        e = [rel_l2(np, Y.cpu().numpy(), O.spmm_f64(W.n_rows, N, W.row_offset, W.col_idx, a64, x64)),
             rel_l2(np, dX.cpu().numpy(), O.spmm_f64(W.n_cols, N, rpT, cT, a64[perm], g64)),
             rel_l2(np, plan.unpack(dW).cpu().numpy(), O.sddmm_f64(W.n_rows, N, W.row_offset, W.col_idx, g64, x64))]
        if max(e) > 1e-5:
            raise RuntimeError(f"{W.name}: BF16 step vs FP64 oracle {e}")
        errs[W.name] = {"fwd": e[0], "dx": e[1], "dw": e[2]}
    ms = time_eager(torch, batch.run, K, backlog=True)
    flops = sum(3 * 2 * c.A.nnz * c.N for c in cases)
    out = {"ms_per_step": ms, "tflops": flops / (ms * 1e-3) / 1e12,
           "main_launch_ms": time_eager(torch, lambda: batch.run(stage=2), K, backlog=True),
           "rel_l2_vs_fp64_oracle": errs}
    batch.close()
    return out


# --------------------------------------------------------------------------
# extra sections
# --------------------------------------------------------------------------
def bench_plain_gpu_suite(torch, np, T, O, dev, args):
    """The ten FFN-shaped products (CSR 2048x512 and 512x2048 at sparsity
    0.7..0.98, B K x 256): one batched launch, tuned per-call launches,
    per-matrix warm/cold, Blocked-ELL, 2:4, rocSPARSE and hipBLASLt."""
    from tca_spmm import workloads
    from tca_spmm.workloads import alg_bytes_fwd
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
        for lay, dense in (("t", False), ("d", True)):
            rpw, tile = T.autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N, dense=dense)
            plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rows_per_wave=rpw)
            a = plan.pack(vals)
            e[lay] = dict(plan=plan, a=plan.pack_dense(a) if dense else a, tile=tile, rpw=rpw, dense=dense)
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

    batch = T.Batch([T.spmm_item(e["b"]["plan"], e["b"]["a"], e["B"], e["C"], e["b"]["tile"], dense=True)
                     for e in entries])
    refs = [O.spmm_f32(e["case"].A.n_rows, N, e["case"].A.row_offset, e["case"].A.col_idx, e["case"].A.values,
                       e["case"].B) for e in entries]
    worst = 0.0
    for run in (batch.run, per_call("t"), per_call("d")):
        for e in entries:
            e["C"].fill_(float("nan"))
        run()
        torch.cuda.synchronize()
        for e, ref in zip(entries, refs):
            r = rel_l2(np, e["C"].cpu().numpy(), ref)
            worst = max(worst, r)
            assert r <= 1e-5, f"{e['case'].A.name}: rel-L2 {r} vs oracle"
    batch_ms = time_eager(torch, batch.run, 50, backlog=True)
    out = {"batch_ms": batch_ms, "tflops": total_flops / (batch_ms * 1e-3) / 1e12,
           "hbm_frac": total_bytes / (batch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "alg_bytes": total_bytes,
           "worst_rel_l2_vs_oracle": worst,
           "per_call_graph_ms": time_graph(torch, per_call("d"), 20),
           "per_call_tiled_graph_ms": time_graph(torch, per_call("t"), 20),
           "per_call_eager_ms": time_eager(torch, per_call("d"), 20),
           "batch_units": batch.n_units}
    per = []
    for e in entries:
        us = time_eager(torch, lambda: launch(e, "d"), 50, backlog=True) * 1e3
        us_t = time_eager(torch, lambda: launch(e, "t"), 50, backlog=True) * 1e3
        per.append({"name": e["case"].A.name, "us": us, "tiled_us": us_t, "tflops": e["flops"] / (us * 1e-6) / 1e12,
                    "hbm_frac": e["bytes"] / (us * 1e-6) / 1e9 / HBM_PEAK_GBS,
                    "per_call_dense": [e["d"]["rpw"], e["d"]["tile"]], "per_call_tiled": [e["t"]["rpw"], e["t"]["tile"]]})
    # cold: a 512 MiB write between launches evicts L2 and the Infinity Cache
    flush = torch.empty(512 << 20, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()
    for e, row in zip(entries, per):
        cold = []
        for _ in range(5):
            flush.fill_(1)
            s, t = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record(stream)
            launch(e, "d")
            t.record(stream)
            torch.cuda.synchronize()
            cold.append(s.elapsed_time(t) * 1e3)
        row["cold_us"] = float(np.median(cold))
    del flush
    out["per_matrix"] = per
    # Blocked-ELL (the reference's block-SpMM baseline, a gfx950 kernel here)
    bell = {}
    for b in (16, 32):
        tot = 0.0
        for e in entries:
            A = e["case"].A
            M_ = T.BlockedEll.from_csr(A.n_rows, A.n_cols, A.row_offset, A.col_idx, A.values, b)
            Cb = torch.empty_like(e["C"])
            tot += time_eager(torch, lambda: T.bell_spmm(M_, e["B"], out=Cb), 30, backlog=True)
        bell[f"b{b}_ms"] = tot
    out["blocked_ell"] = bell
    # 2:4 on the sparse MFMA: each A pruned to 2:4 (a different, denser product)
    tot, kept = 0.0, []
    for e in entries:
        A = e["case"].A
        dense = torch.zeros((A.n_rows, A.n_cols), dtype=torch.float16, device=dev)
        rows = torch.repeat_interleave(torch.arange(A.n_rows, device=dev), torch.from_numpy(np.diff(A.row_offset)).to(dev))
        dense[rows, torch.from_numpy(A.col_idx).to(dev).long()] = f16(A.values)
        sp = T.Sparse24.compress(dense)
        C24 = torch.empty_like(e["C"])
        tot += time_eager(torch, lambda: T.spmm_24(sp, e["B"], out=C24), 30, backlog=True)
        kept.append(int((sp.values != 0).sum()) / max(A.nnz, 1))
        del dense, sp
    out["sparse24"] = {"ms": tot, "kept_nnz_frac_min": min(kept)}
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
        out["rocsparse_ms"] = lib_ms["rocsparse"]
        out["hipblaslt_ms"] = lib_ms["hipblaslt"]
    return out


def bench_precision(torch, np, T, O, dev, args):
    """4096x4096 at 0.9, N=512: FP16, BF16 and FP8-E4M3 (row/column scales)
    per call, with rel-L2 / max-abs against the FP64 oracle on the original
    values (format + kernel error) and on the quantised values (kernel error);
    hipBLASLt dense FP16 GEMM and rocSPARSE on the same matrix.  Every time is
    the mean of 50 back-to-back launches on one stream behind a queue backlog,
    the convention of the library baselines' own timing."""
    from tca_spmm import workloads
    from tca_spmm.workloads import alg_bytes_fwd
    case = workloads.precision(seed=0)
    A, N = case.A, case.N
    rp, ci = A.row_offset, A.col_idx
    ref = O.spmm_f64(A.n_rows, N, rp, ci, A.values.view(np.float16).astype(np.float64),
                     case.B.view(np.float16).astype(np.float64))
    out = {}

    def err(C, r):
        d = C.astype(np.float64) - r
        return float(np.linalg.norm(d) / np.linalg.norm(r)), float(np.abs(d).max())

    va = torch.from_numpy(A.values.view(np.int16)).view(torch.float16).to(dev)
    vb = torch.from_numpy(case.B.view(np.int16)).view(torch.float16).to(dev)
    C = torch.empty((A.n_rows, N), dtype=torch.float32, device=dev)
    rpw, tile = T.autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N)
    rpw_d, tile_d = T.autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N, dense=True)
    plan = T.Plan(A.n_rows, A.n_cols, rp, ci, rpw)
    plan_d = plan if rpw_d == rpw else T.Plan(A.n_rows, A.n_cols, rp, ci, rpw_d)
    rpw0, tile0, rpw_d0, tile_d0 = rpw, tile, rpw_d, tile_d
    nbytes = alg_bytes_fwd(A.n_rows, A.n_cols, A.nnz, N)
    for name, cast in (("f16", torch.float16), ("bf16", torch.bfloat16)):
        ta, tb = plan.pack(va.to(cast)), vb.to(cast)
        da = plan_d.pack_dense(plan_d.pack(va.to(cast)))
        ms_d = time_eager(torch, lambda: T.spmm(plan_d, da, tb, out=C, tile_n=tile_d, dense=True), 50, backlog=True)
        Cd = C.cpu().numpy()
        ms = time_eager(torch, lambda: T.spmm(plan, ta, tb, out=C, tile_n=tile), 50, backlog=True)
        Cn = C.cpu().numpy()
        rq = O.spmm_f64(A.n_rows, N, rp, ci, plan.unpack(ta).float().cpu().numpy().astype(np.float64),
                        tb.float().cpu().numpy().astype(np.float64))
        e_all, m_all = err(Cn, ref)
        e_q, m_q = err(Cn, rq)
        best = min(ms, ms_d)
        out[name] = {"us": ms * 1e3, "dense_us": ms_d * 1e3, "best_us": best * 1e3,
                     "tflops": 2 * A.nnz * N / (best * 1e-3) / 1e12,
                     "hbm_frac": nbytes / (best * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "dense_rel_l2_vs_fp64_on_quantised": err(Cd, rq)[0],
                     "rel_l2_vs_fp64": e_all, "max_abs_vs_fp64": m_all,
                     "rel_l2_vs_fp64_on_quantised": e_q, "max_abs_vs_fp64_on_quantised": m_q}
    rpd = torch.from_numpy(rp).to(dev)
    qa, sa = T.quantize_rows_e4m3(rpd, va)
    qb, sb = T.quantize_cols_e4m3(vb)
    # E4M3 tiled values run on the one-wave-per-group kernels (16 / 32 / 64
    # rows; the wide kernels of 128 / 256-row plans take 16-bit values); the
    # dense tiles run on the block kernels like FP16
    if rpw >= 128:
        rpw, tile = 32, 64
        plan = T.Plan(A.n_rows, A.n_cols, rp, ci, rpw)
    ta = plan.pack(qa)
    da = plan_d.pack_dense(plan_d.pack(qa))
    ms_d = time_eager(torch, lambda: T.spmm(plan_d, da, qb, out=C, row_scale=sa, col_scale=sb, tile_n=tile_d, dense=True), 50, backlog=True)
    ms = time_eager(torch, lambda: T.spmm(plan, ta, qb, out=C, row_scale=sa, col_scale=sb, tile_n=tile), 50, backlog=True)
    Cn = C.cpu().numpy()
    rq = O.spmm_f64(A.n_rows, N, rp, ci, O.e4m3_to_f32(qa.cpu().numpy()).astype(np.float64),
                    O.e4m3_to_f32(qb.cpu().numpy()).astype(np.float64))
    rq *= sa.cpu().numpy().astype(np.float64)[:, None] * sb.cpu().numpy().astype(np.float64)[None, :]
    e_all, m_all = err(Cn, ref)
    e_q, m_q = err(Cn, rq)
    best = min(ms, ms_d)
    out["e4m3"] = {"us": ms * 1e3, "dense_us": ms_d * 1e3, "best_us": best * 1e3,
                   "tflops": 2 * A.nnz * N / (best * 1e-3) / 1e12,
                   "hbm_frac": alg_bytes_fwd(A.n_rows, A.n_cols, A.nnz, N, 1, 1, True) / (best * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "rel_l2_vs_fp64": e_all, "max_abs_vs_fp64": m_all,
                   "rel_l2_vs_fp64_on_quantised": e_q, "max_abs_vs_fp64_on_quantised": m_q,
                   "quantize_cols_us": time_graph(torch, lambda: T.quantize_cols_e4m3(vb), 20) * 1e3}
    out["e4m3"]["plans"] = {"rows_per_wave": rpw, "tile_n": tile, "dense_rows_per_wave": rpw_d, "dense_tile_n": tile_d}
    out["config"] = {"rows_per_wave": rpw0, "tile_n": tile0, "dense_rows_per_wave": rpw_d0, "dense_tile_n": tile_d0,
                     "nnz": A.nnz, "N": N, "alg_bytes_f16": nbytes}
    if not args.no_libs:
        dense = torch.zeros((A.n_rows, A.n_cols), dtype=torch.float16, device=dev)
        rows = torch.repeat_interleave(torch.arange(A.n_rows, device=dev), torch.from_numpy(np.diff(rp)).to(dev))
        dense[rows, torch.from_numpy(ci).to(dev).long()] = va
        try:
            ms_blt = T.hipblaslt_gemm(A.n_rows, N, A.n_cols, dense, vb, C, n_iter=50)
            out["hipblaslt_dense_us"] = ms_blt * 1e3
            out["speedup_vs_hipblaslt"] = ms_blt * 1e3 / out["f16"]["best_us"]
        except Exception as ex:
            log("hipBLASLt failed:", ex)
        try:
            ms_rs = T.rocsparse_spmm(A.n_rows, N, A.n_cols, rpd, torch.from_numpy(ci).to(dev), va, vb, C, n_iter=10)
            out["rocsparse_us"] = ms_rs * 1e3
        except Exception as ex:
            log("rocSPARSE failed:", ex)
        del dense
    return out


def bench_linear(torch, np, T, O, dev, args):
    """One sparse linear layer at full size: W 4096x1024 at 0.9 (lognormal
    rows), X 1024x2048 FP16, upstream gradient G 4096x2048 FP32.  Forward, dX
    and dW per call (tuned shapes, panel-dense tiles; G passed as FP32, which
    the kernels round to FP16 on load),
    the same three as one batch run with the fused SGD update (a one-layer
    training step: prologue + persistent launch), and SparseLinear's forward +
    backward through torch.autograd; each checked against the FP64 oracle
    (rel-L2 <= 1e-5), each time the mean of back-to-back launches behind a
    queue backlog; hipBLASLt's dense FP16 GEMM of the densified W beside the
    forward."""
    from tca_spmm import workloads
    from tca_spmm.workloads import alg_bytes_dw, alg_bytes_dx, alg_bytes_fwd
    case = workloads.sparse_linear(seed=0)
    W, N = case.A, case.B.shape[1]
    M, K, nnz = W.n_rows, W.n_cols, W.nnz
    G32 = case.extra["G"]
    a64 = W.values.view(np.float16).astype(np.float64)
    x64 = case.B.view(np.float16).astype(np.float64)
    g64 = G32.astype(np.float64)
    g64q = G32.astype(np.float16).astype(np.float64)   # the batch reads G rounded to FP16
    rpT, cT, perm = O.transpose_csr(M, K, W.row_offset, W.col_idx)
    ref_y = O.spmm_f64(M, N, W.row_offset, W.col_idx, a64, x64)
    ref = {"fwd": ref_y, "dx": O.spmm_f64(K, N, rpT, cT, a64[perm], g64),
           "dw": O.sddmm_f64(M, N, W.row_offset, W.col_idx, g64, x64)}
    ref_q = {"fwd": ref_y, "dx": O.spmm_f64(K, N, rpT, cT, a64[perm], g64q),
             "dw": O.sddmm_f64(M, N, W.row_offset, W.col_idx, g64q, x64)}
    X = torch.from_numpy(case.B.view(np.int16)).view(torch.float16).to(dev)
    G = torch.from_numpy(G32).to(dev)
    v_csr = torch.from_numpy(W.values.view(np.int16)).view(torch.float16).to(dev)
    Y = torch.empty((M, N), dtype=torch.float32, device=dev)
    dX = torch.empty((K, N), dtype=torch.float32, device=dev)
    p16 = T.Plan(M, K, W.row_offset, W.col_idx, 16)
    w16 = p16.pack(v_csr)
    dW = torch.empty(nnz, dtype=torch.float32, device=dev)
    alg = {"fwd": alg_bytes_fwd(M, K, nnz, N), "dx": alg_bytes_dx(M, K, nnz, N), "dw": alg_bytes_dw(M, K, nnz, N)}
    flops = 2 * nnz * N
    out = {"config": {"M": M, "K": K, "nnz": nnz, "N": N, "alg_bytes": alg}}

    def check(tag, got, key, refs):
        r = rel_l2(np, got, refs[key])
        if r > 1e-5:
            raise RuntimeError(f"sparse linear {tag} {key}: rel-L2 {r:.2e} vs the FP64 oracle")
        return r

    def entry(ms, key, err):
        return {"us": ms * 1e3, "tflops": flops / (ms * 1e-3) / 1e12,
                "hbm_frac": alg[key] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "rel_l2_vs_fp64": err}

    # per call
    r_f, t_f = T.autotune.choice_for(M, K, nnz, N, dense=True)
    r_t, t_t = T.autotune.choice_for(K, M, nnz, N, op="spmm_t", dense=True)
    pf = p16 if r_f == 16 else T.Plan(M, K, W.row_offset, W.col_idx, r_f)
    pt = p16.transpose(r_t)
    df = pf.pack_dense(pf.pack(v_csr))
    dt = pt.pack_dense(w16)
    calls = {"fwd": lambda: T.spmm(pf, df, X, out=Y, tile_n=t_f, dense=True),
             "dx": lambda: T.spmm(pt, dt, G, out=dX, tile_n=t_t, dense=True),
             "dw": lambda: T.sddmm(p16, G, X, out=dW)}
    per_call = {}
    for key, fn in calls.items():
        fn()
        torch.cuda.synchronize()
        got = {"fwd": Y, "dx": dX, "dw": None}[key]
        got = p16.unpack(dW).cpu().numpy() if key == "dw" else got.cpu().numpy()
        e = check("per call", got, key, ref_q)   # the kernels round an FP32 G to FP16 on load
        per_call[key] = entry(time_eager(torch, fn, 20, backlog=True), key, e)
    per_call["shapes"] = {"fwd": [r_f, t_f], "dx": [r_t, t_t]}
    out["per_call"] = per_call
    # one batch run: the layer's training step (dense tiles packed from the
    # weights, fused SGD update with a zero rate so that the weights stay put)
    rb_f, tb_f = T.autotune.choice_for(M, K, nnz, N, batch=True, dense=True)
    rb_t, tb_t = T.autotune.choice_for(K, M, nnz, N, op="spmm_t", batch=True, dense=True)
    bf = p16 if rb_f == 16 else T.Plan(M, K, W.row_offset, W.col_idx, rb_f)
    bt = p16.transpose(rb_t)
    wb = bf.pack(v_csr)
    dWb = torch.empty(nnz, dtype=torch.float32, device=dev)
    batch = T.Batch([T.spmm_item(bf, wb, X, Y, tb_f, dense="pack", values_order=bf),
                     T.spmm_item(bt, wb, G, dX, tb_t, dense="pack", values_order=bf),
                     T.sddmm_item(bf, G, X, dWb, sgd=(wb, 0.0))])
    batch.run()
    torch.cuda.synchronize()
    errs = {"fwd": check("batch", Y.cpu().numpy(), "fwd", ref_q), "dx": check("batch", dX.cpu().numpy(), "dx", ref_q),
            "dw": check("batch", bf.unpack(dWb).cpu().numpy(), "dw", ref_q)}
    ms_b = time_eager(torch, batch.run, 20, backlog=True)
    out["batch_step"] = {"us": ms_b * 1e3, "tflops": 3 * flops / (ms_b * 1e-3) / 1e12,
                         "hbm_frac": sum(alg.values()) / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "main_launch_us": time_eager(torch, lambda: batch.run(stage=2), 20, backlog=True) * 1e3,
                         "prologue_us": time_eager(torch, lambda: batch.run(stage=1), 20, backlog=True) * 1e3,
                         "rel_l2_vs_fp64_on_fp16_g": errs, "shapes": {"fwd": [rb_f, tb_f], "dx": [rb_t, tb_t]},
                         "per_call_sum_us": sum(per_call[k]["us"] for k in ("fwd", "dx", "dw"))}
    batch.close()
    # SparseLinear through torch.autograd (forward, then backward with G)
    lin = T.SparseLinear(M, K, W.row_offset, W.col_idx, W.values, rows_per_wave=r_f, rows_per_wave_t=r_t,
                         tile_n=t_f, device=dev, dense=True)
    Xg = X.clone().requires_grad_(True)

    res = {}

    def fwd_bwd():
        lin.weight.grad = None
        Xg.grad = None
        y = lin(Xg)
        y.backward(G)
        res["y"] = y

    fwd_bwd()
    torch.cuda.synchronize()
    # the gradients come back in the FP16 of the weights and of X (G rounded to
    # FP16 on the way in): their rounding alone is ~3e-4 rel-L2
    e_y = check("autograd", res["y"].detach().cpu().numpy(), "fwd", ref)
    e_g = {k: rel_l2(np, v, ref_q[k]) for k, v in
           (("dx", Xg.grad.float().cpu().numpy()), ("dw", lin.plan.unpack(lin.weight.grad).float().cpu().numpy()))}
    if max(e_g.values()) > 1e-3:
        raise RuntimeError(f"sparse linear autograd gradients vs the FP64 oracle: {e_g}")
    out["autograd_fwd_bwd"] = {"us": time_eager(torch, fwd_bwd, 10) * 1e3, "rel_l2_fwd_vs_fp64": e_y,
                               "rel_l2_fp16_grads_vs_fp64": e_g, "grad_dtype": str(lin.weight.grad.dtype)}
    if not args.no_libs:
        dense = torch.zeros((M, K), dtype=torch.float16, device=dev)
        rows = torch.repeat_interleave(torch.arange(M, device=dev), torch.from_numpy(np.diff(W.row_offset)).to(dev))
        dense[rows, torch.from_numpy(W.col_idx).to(dev).long()] = v_csr
        try:
            out["hipblaslt_dense_fwd_us"] = T.hipblaslt_gemm(M, N, K, dense, X, Y, n_iter=20) * 1e3
        except Exception as ex:
            log("hipBLASLt failed:", ex)
        del dense
    return out


def bench_plain_cpu(torch, np, T, O, dev, cores):
    """512x512 at 0.9 (Binomial rows), N=256: the reference's CPU-runnable
    case, on the GPU (tuned per-call launch) and on the CPU oracle (all
    cores and one thread), same inputs."""
    from tca_spmm import workloads
    case = workloads.plain_cpu()
    A, N = case.A, case.N
    rpw, tile = T.autotune.choice_for(A.n_rows, A.n_cols, A.nnz, N, dense=True)
    plan = T.Plan(A.n_rows, A.n_cols, A.row_offset, A.col_idx, rpw)
    va = torch.from_numpy(A.values.view(np.int16)).view(torch.float16).to(dev)
    da = plan.pack_dense(plan.pack(va))
    B = torch.from_numpy(case.B.view(np.int16)).view(torch.float16).to(dev)
    C = torch.empty((A.n_rows, N), dtype=torch.float32, device=dev)
    T.spmm(plan, da, B, out=C, tile_n=tile, dense=True)
    ref = O.spmm_f32(A.n_rows, N, A.row_offset, A.col_idx, A.values, case.B, cores)
    r = rel_l2(np, C.cpu().numpy(), ref)
    assert r <= 1e-5, f"plain_cpu rel-L2 {r}"
    gpu_us = time_eager(torch, lambda: T.spmm(plan, da, B, out=C, tile_n=tile, dense=True), 100, backlog=True) * 1e3
    res = {"gpu_us": gpu_us, "rel_l2_gpu_vs_cpu": r, "flops": 2 * A.nnz * N}
    for th, key in ((cores, "cpu_all_cores"), (1, "cpu_one_thread")):
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 1.0:
            O.spmm_f32(A.n_rows, N, A.row_offset, A.col_idx, A.values, case.B, th)
            reps += 1
        dt = (time.perf_counter() - t0) / reps
        res[key] = {"us": dt * 1e6, "threads": th, "gflops": 2 * A.nnz * N / dt / 1e9}
    res["gpu_speedup_vs_cpu_all_cores"] = res["cpu_all_cores"]["us"] / gpu_us
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sections", default="composition,suite,precision,linear,plain_cpu",
                    help="comma list of extra sections: composition (the persistent launch over each kind of "
                         "item alone), suite, precision, linear, plain_cpu")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-libs", action="store_true", help="skip rocSPARSE / hipBLASLt")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
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
    from oracle import oracle as O

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1

    ts, step = bench_train_step(torch, np, T, O, dev, args)
    log(f"train step {step['ms_per_step'] * 1e3:.1f} us, plain per-call {step['plain_per_call_graph_ms'] * 1e3:.1f} us")
    ms_per_step = step["ms_per_step"]
    value = ts.flops / (ms_per_step * 1e-3) / 1e12
    main_ms = step["main_launch_ms"]
    achieved = ts.alg_bytes / (main_ms * 1e-3) / 1e9

    # achievable HBM bandwidth on this box: a 1 GiB device-to-device copy
    src = torch.empty(1 << 30, dtype=torch.uint8, device=dev)
    dst = torch.empty_like(src)
    src.fill_(1)
    copy_ms = time_eager(torch, lambda: dst.copy_(src), 10, backlog=True)
    copy_gbs = 2 * src.numel() / (copy_ms * 1e-3) / 1e9      # bytes read + written
    del src, dst

    extra = {"train_step": step}
    if "suite" in args.sections:
        extra["plain_gpu_suite"] = bench_plain_gpu_suite(torch, np, T, O, dev, args)
    if "precision" in args.sections:
        extra["precision_4096"] = bench_precision(torch, np, T, O, dev, args)
    if "linear" in args.sections:
        extra["sparse_linear_4096x1024"] = bench_linear(torch, np, T, O, dev, args)
    if "plain_cpu" in args.sections:
        extra["plain_cpu_512"] = bench_plain_cpu(torch, np, T, O, dev, cores)

    cpu = None
    if not args.no_cpu:
        ts.cpu_step(O, cores)                         # warm-up (and the transposes)
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < args.cpu_seconds:
            ts.cpu_step(O, cores)
            reps += 1
        dt = time.perf_counter() - t0
        cpu = {"value": ts.flops * reps / dt / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
               "sample": f"the full twelve-layer step's products (forward, dX on the transposed CSR, dW; FP32 "
                         f"accumulation of FP16 products, OpenMP over rows, {cores} threads = the process's "
                         f"affinity set) repeated {reps}x in {dt:.1f}s ({dt / reps * 1e3:.1f} ms per step)"}
        L0 = ts.layers[0]
        W0 = L0["case"].A
        reps1, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < 2.0:
            O.spmm_f32(W0.n_rows, L0["N"], W0.row_offset, W0.col_idx, W0.values, L0["case"].B, 1)
            reps1 += 1
        dt1 = time.perf_counter() - t1
        extra["cpu_one_thread"] = {"product": W0.name + " forward", "ms": dt1 / reps1 * 1e3,
                                   "tflops": 2 * W0.nnz * L0["N"] * reps1 / dt1 / 1e12}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16",
        "data": "synthetic",
        "config": {"workload": "twelve-layer sparse-weight training step: 4x 512x2048, 4x 2048x512, 2x 1024x1024, "
                               "2x 4096x1024 CSR layers, sparsity U[0.5,0.98], lognormal(0.5) rows, N in "
                               "{128,384,1024}; forward + dX + dW + SGD update per step (fused: prologue launch "
                               "packing dense tiles from the current weights + one persistent launch whose "
                               "SDDMM units apply the update)",
                   "layers": len(ts.layers), "nnz_total": int(sum(L["case"].A.nnz for L in ts.layers)),
                   "flops_per_step": ts.flops, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("batch_train_step_traffic"),
                     "kernel": "batch_kernel (the step's persistent launch: the twelve layers' forward + dX + dW "
                               "algorithmic bytes / the launch's mean duration, K back-to-back launches)",
                     "alg_bytes_per_launch": ts.alg_bytes, "launch_us_mean": main_ms * 1e3,
                     "copy_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs,
                     "mfma_dense_equivalent_tflops": ts.dense_flops / (main_ms * 1e-3) / 1e12,
                     "mfma_peak_tflops": MFMA_F16_PEAK_TFLOPS},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
