// Batched persistent launch: the host builds one descriptor per item and one
// record per work unit (heaviest first), uploads both once, and every run is a
// single kernel launch that drains the unit list (see batch_kernel).
#include "tca_internal.hpp"

#include <algorithm>
#include <numeric>

using tca::fail;

struct tca_batch {
    int n_items = 0, n_units = 0, grid = 0, device = 0, n_zero = 0;
    int mode = 0;   // 1: every item F16 and every SpMM item panel-dense (the leaner kernel); 2: the same in BF16
    void *d_mem = nullptr;  // items | units | zero list | counters
    tca::BatchItemDev *d_items = nullptr;
    tca::BatchUnitDev *d_units = nullptr;
    tca::ZeroDesc *d_zero = nullptr;
    int *d_counters = nullptr;
    int *d_part_off = nullptr;   // first unit of each partition (kBatchParts + 1)
    int nparts = 1;
    // FP32 operands (SpMM B, SDDMM G) rounded once per run into FP16 copies that
    // the items read instead: the gathers then move half the bytes, with the
    // same bits as rounding on load
    void *d_shadow = nullptr;    // conversion descriptors | FP16 copies
    tca::ConvDesc *d_conv = nullptr;
    int n_conv = 0;
    long conv8 = 0;
    // panel-dense tiles of a_dense = 2 items, rebuilt from the weight values on
    // every run (the same prologue launch as the FP16 copies)
    void *d_pack_mem = nullptr;  // pack descriptors | dense tiles | value maps
    tca::PackDesc *d_pack = nullptr;
    int n_pack = 0;
    long pack_blocks = 0;
};

namespace {

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

extern "C" size_t tca_batch_item_size(void) { return sizeof(tca_batch_item); }

extern "C" tca_status tca_batch_create(int n_items, const tca_batch_item *items, tca_batch **batch)
{
    if (!batch || n_items <= 0 || !items) return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: bad argument");
    *batch = nullptr;
    // FP32 operands read by several items or one: one 16-bit copy per buffer, in
    // the type the reading items round it to on load (an SPMM item's A type, an
    // SDDMM item's X type: FP16 or BF16), used when every item reading it
    // agrees on its shape and that type
    std::vector<tca_batch_item> its(items, items + n_items);
    struct Operand {
        const void *src;
        int rows, N, ld;
        bool ok;
        size_t off;
        tca_dtype to;
    };
    std::vector<Operand> ops;
    auto operand_of = [&](int i, bool is_b) -> std::pair<const void *, Operand> {
        const tca_batch_item &it = its[i];
        const tca_plan *p = it.plan;
        if (!p) return {nullptr, {}};
        if (it.op == TCA_OP_SPMM && is_b && it.b_dtype == TCA_F32 && it.b && (it.a_dtype == TCA_F16 || it.a_dtype == TCA_BF16))
            return {it.b, {it.b, p->host.n_cols, it.N, it.ldb, true, 0, it.a_dtype}};
        if (it.op == TCA_OP_SDDMM && !is_b && it.a_dtype == TCA_F32 && it.a && (it.b_dtype == TCA_F16 || it.b_dtype == TCA_BF16))
            return {it.a, {it.a, p->host.n_rows, it.N, it.lda, true, 0, it.b_dtype}};
        return {nullptr, {}};
    };
    for (int i = 0; i < n_items; i++)
        for (bool is_b : {true, false}) {
            const auto found = operand_of(i, is_b);
            const void *ptr = found.first;
            const Operand &o = found.second;
            if (!ptr || o.N <= 0 || o.N % 8 || o.ld < o.N || o.ld % 8 || o.rows <= 0) continue;
            auto f = std::find_if(ops.begin(), ops.end(), [&](const Operand &x) { return x.src == ptr; });
            if (f == ops.end()) ops.push_back(o);
            else if (f->rows != o.rows || f->N != o.N || f->ld != o.ld || f->to != o.to) f->ok = false;
        }
    ops.erase(std::remove_if(ops.begin(), ops.end(), [](const Operand &x) { return !x.ok; }), ops.end());
    if (ops.size() > (size_t)tca::kMaxConv) ops.resize(tca::kMaxConv);   // the rest stay FP32
    const size_t conv_bytes = (sizeof(tca::ConvDesc) * std::max<size_t>(ops.size(), 1) + 255) / 256 * 256;
    size_t shadow_bytes = conv_bytes;
    for (Operand &o : ops) {
        o.off = shadow_bytes;
        shadow_bytes += ((size_t)o.rows * o.N * 2 + 255) / 256 * 256;
    }
    struct DevGuard {
        void *p = nullptr;
        ~DevGuard()
        {
            if (p) (void)hipFree(p);
        }
    } shadow;
    std::vector<tca::ConvDesc> conv;
    long conv8 = 0;
    if (!ops.empty()) {
        if (hipMalloc(&shadow.p, shadow_bytes) != hipSuccess) return fail(TCA_ERR_ALLOC, "tca_batch_create: FP16 copies");
        char *base = static_cast<char *>(shadow.p);
        for (const Operand &o : ops) {
            tca::ConvDesc c{};
            c.src = static_cast<const float *>(o.src);
            c.dst = reinterpret_cast<uint16_t *>(base + o.off);
            c.rows = o.rows;
            c.N = o.N;
            c.ld = o.ld;
            c.bf16 = o.to == TCA_BF16 ? 1 : 0;
            c.first8 = conv8;
            conv8 += ((long)o.rows * o.N / 8 + 255) / 256 * 256;   // next operand on a workgroup boundary
            conv.push_back(c);
        }
        for (tca_batch_item &it : its) {
            for (const tca::ConvDesc &c : conv) {
                const tca_dtype to = c.bf16 ? TCA_BF16 : TCA_F16;
                if (it.op == TCA_OP_SPMM && it.b_dtype == TCA_F32 && it.b == c.src) {
                    it.b = c.dst;
                    it.ldb = c.N;
                    it.b_dtype = to;
                }
                if (it.op == TCA_OP_SDDMM && it.a_dtype == TCA_F32 && it.a == c.src) {
                    it.a = c.dst;
                    it.lda = c.N;
                    it.a_dtype = to;
                }
            }
        }
        if (hipMemcpy(shadow.p, conv.data(), sizeof(tca::ConvDesc) * conv.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(TCA_ERR_HIP, "tca_batch_create: upload failed");
    }
    // Items whose weight values the batch packs into panel-dense tiles on every
    // run: map each entry of the item's plan to its value's index in `a`.
    struct PackJob {
        int item;
        std::vector<int32_t> map;   // empty: identity
        size_t tiles_off = 0, map_off = 0;
        long n_tiles = 0;
    };
    std::vector<PackJob> jobs;
    size_t pack_bytes = 0;
    for (int i = 0; i < n_items; i++) {
        const tca_batch_item &it = its[i];
        if (it.a_dense != 2) continue;
        const tca_plan *p = it.plan;
        const std::string who = "tca_batch_create: item " + std::to_string(i);
        if (it.op != TCA_OP_SPMM || !p || !p->d_base)
            return fail(TCA_ERR_INVALID_ARG, who + ": a_dense = 2 needs an SPMM item with a device plan");
        if (it.a_dtype != TCA_F16 && it.a_dtype != TCA_BF16)
            return fail(TCA_ERR_UNSUPPORTED, who + ": a_dense = 2 needs F16 or BF16 values");
        const tca_plan *q = it.values_order;
        if (q && (q->transposed || q->pattern_key != p->pattern_key || q->host.nnz != p->host.nnz))
            return fail(TCA_ERR_INVALID_ARG, who + ": values_order must be a forward plan of the item's pattern");
        if (p->host.nnz > 0 && !it.a) return fail(TCA_ERR_INVALID_ARG, who + ": NULL values");
        PackJob j;
        j.item = i;
        j.n_tiles = (long)p->host.n_chunks * (p->host.rpw / 16);
        if (q != p) {
            const int nnz = p->host.nnz;
            const std::vector<int32_t> &csr = p->transposed ? p->entry_csr : p->host.entry_src;
            j.map.resize(nnz);
            if (q) {
                std::vector<int32_t> tiled_of_csr(nnz);
                for (int e = 0; e < nnz; e++) tiled_of_csr[q->host.entry_src[e]] = e;
                for (int e = 0; e < nnz; e++) j.map[e] = tiled_of_csr[csr[e]];
            } else {
                for (int e = 0; e < nnz; e++) j.map[e] = csr[e];
            }
        }
        jobs.push_back(std::move(j));
    }
    if ((int)jobs.size() > tca::kMaxPack) return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: too many packed items");
    std::vector<tca::PackDesc> packs;
    long pack_blocks = 0;
    struct DevGuard2 {
        void *p = nullptr;
        ~DevGuard2()
        {
            if (p) (void)hipFree(p);
        }
    } packmem;
    if (!jobs.empty()) {
        pack_bytes = (sizeof(tca::PackDesc) * jobs.size() + 255) / 256 * 256;
        for (PackJob &j : jobs) {
            j.tiles_off = pack_bytes;
            pack_bytes += ((size_t)j.n_tiles * 1024 + 255) / 256 * 256;
            j.map_off = pack_bytes;
            pack_bytes += (j.map.size() * 4 + 255) / 256 * 256;
        }
        if (hipMalloc(&packmem.p, pack_bytes) != hipSuccess) return fail(TCA_ERR_ALLOC, "tca_batch_create: dense tiles");
        char *base = static_cast<char *>(packmem.p);
        for (PackJob &j : jobs) {
            tca_batch_item &it = its[j.item];
            const tca_plan *p = it.plan;
            if (!j.map.empty() &&
                hipMemcpy(base + j.map_off, j.map.data(), j.map.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
                return fail(TCA_ERR_HIP, "tca_batch_create: upload failed");
            tca::PackDesc d{};
            d.seg_ptr = p->d_seg_ptr;
            d.entry_slot = p->d_entry_slot;
            d.map = j.map.empty() ? nullptr : reinterpret_cast<const int32_t *>(base + j.map_off);
            d.src = static_cast<const uint16_t *>(it.a);
            d.dst = reinterpret_cast<uint16_t *>(base + j.tiles_off);
            d.n_tiles = j.n_tiles;
            d.first_block = pack_blocks;
            pack_blocks += (j.n_tiles + tca::kPackTilesPerBlock - 1) / tca::kPackTilesPerBlock;
            packs.push_back(d);
            it.a = d.dst;    // the items read the batch's tiles
            it.a_dense = 1;
        }
        if (hipMemcpy(packmem.p, packs.data(), sizeof(tca::PackDesc) * packs.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(TCA_ERR_HIP, "tca_batch_create: upload failed");
    }
    items = its.data();
    // A fused weight update must not race with a unit that reads the weights
    // directly: SpMM items on tiled values (the packed and the caller's dense
    // tiles are other memory) may not overlap any SDDMM item's sgd_w, and no
    // two SDDMM items may update overlapping weights.
    for (int i = 0; i < n_items; i++) {
        const tca_batch_item &wi = items[i];
        if (wi.op != TCA_OP_SDDMM || !wi.sgd_w || !wi.plan) continue;
        const char *w0 = static_cast<const char *>(wi.sgd_w), *w1 = w0 + (size_t)wi.plan->host.nnz * 2;
        for (int k = i + 1; k < n_items; k++) {
            const tca_batch_item &oi = items[k];
            if (oi.op != TCA_OP_SDDMM || !oi.sgd_w || !oi.plan) continue;
            const char *o0 = static_cast<const char *>(oi.sgd_w), *o1 = o0 + (size_t)oi.plan->host.nnz * 2;
            if (o0 < w1 && w0 < o1)
                return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: SDDMM items " + std::to_string(i) + " and " +
                                                     std::to_string(k) + " update overlapping weights (sgd_w)");
        }
        for (int k = 0; k < n_items; k++) {
            const tca_batch_item &ri = items[k];
            if (ri.op != TCA_OP_SPMM || ri.a_dense || !ri.a || !ri.plan) continue;
            const char *a0 = static_cast<const char *>(ri.a);
            const char *a1 = a0 + (size_t)ri.plan->host.nnz * (ri.a_dtype == TCA_E4M3 ? 1 : 2);
            if (a0 < w1 && w0 < a1)
                return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: SPMM item " + std::to_string(k) + " reads the weights SDDMM item " +
                                                     std::to_string(i) + " updates (sgd_w) in the same launch");
        }
    }
    std::vector<tca::BatchItemDev> dev(n_items);
    std::vector<tca::BatchUnitDev> units;
    std::vector<double> cost;
    for (int i = 0; i < n_items; i++) {
        const tca_batch_item &it = items[i];
        const tca_plan *p = it.plan;
        if (!p || !p->d_base) return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: item without a device plan");
        const tca::PlanTables &t = p->host;
        const int ppg = t.rpw / 16;
        tca::BatchItemDev &d = dev[i];
        d = tca::BatchItemDev{};
        d.n_rows = t.n_rows;
        d.nnz = t.nnz;
        d.n_chunks = t.n_chunks;
        d.N = it.N;
        d.rpw = t.rpw;
        d.grp_chunk_ptr = p->d_grp_chunk_ptr;
        d.chunk_cols = p->d_chunk_cols;
        d.seg_ptr = p->d_seg_ptr;
        d.entry_slot = p->d_entry_slot;
        d.entry_lds = p->d_entry_lds;
        d.entry_src = p->d_entry_src;
        d.a = it.a;
        d.b = it.b;
        d.out = it.out;
        if (it.op == TCA_OP_SPMM) {
            if (it.sgd_w)
                return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: SPMM item " + std::to_string(i) + ": sgd_w applies to SDDMM items");
            const int tile = it.tile_n ? it.tile_n : tca::default_tile_n(p, it.N);
            if (it.N <= 0 || it.N % 8 || it.ldb < it.N || it.ldc < it.N || it.ldb % 8 || it.ldc % 4 || !it.b || !it.out ||
                !aligned16(it.b) || !aligned16(it.out) || (t.nnz > 0 && !it.a))
                return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: SPMM item " + std::to_string(i) + " has bad shapes/buffers");
            d.variant = tca::batch_variant(TCA_OP_SPMM, t.rpw, tile, it.a_dtype, it.b_dtype, p->transposed != 0,
                                           it.a_dense != 0);
            if (d.variant < 0)
                return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: SPMM item " + std::to_string(i) +
                                                     " (batched SpMM: f16 A with f16 B forward on rows_per_wave 16 with tile_n "
                                                     "16/32/64 or rows_per_wave 32 with tile_n 32; f16 A with f32 or f16 B "
                                                     "transposed on rows_per_wave 16 with tile_n 16/32/64; panel-dense f16 "
                                                     "or bf16 A with a B of the same type on rows_per_wave 16 (tile_n "
                                                     "16/32/64), 32 (32/64), 64 (32/64) or 128 (64), or f32 B on "
                                                     "rows_per_wave 16 (16/32/64) or 32 (32))");
            if (it.a_dense && tca::dense_bytes(t, it.a_dtype) >= (1L << 31))
                return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: SPMM item " + std::to_string(i) + " dense tiles reach 2 GiB");
            d.ldb = it.ldb;
            d.ldc = it.ldc;
            d.n_tiles = (it.N + tile - 1) / tile;
            d.b_bytes = tca::operand_bytes(t.n_cols, it.N, it.ldb, it.b_dtype == TCA_F32 ? 4 : 2);   // B: f32 or f16
            if (d.b_bytes == 0 || (long)t.nnz * 4 >= (1L << 31))
                return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: SPMM item " + std::to_string(i) + " reaches 2 GiB");
            for (int gi = 0; gi < t.n_groups; gi++) {
                const int g = t.work_order[gi];
                const int nch = t.grp_chunk_ptr[g + 1] - t.grp_chunk_ptr[g];
                for (int tt = 0; tt < d.n_tiles; tt++) {
                    units.push_back({i, g, tt, t.grp_chunk_ptr[g], t.grp_chunk_ptr[g + 1], 0, 0, 0});
                    // estimated ns: unit start + per chunk (index + B rows + MFMAs; an
                    // FP32 B gathers twice the bytes; tiled values add the entry work)
                    // (block units: the four waves share each chunk, about half the
                    // cost per panel of a wave per 32 rows)
                    const bool block = d.variant == 22 || d.variant == 23 || d.variant == 34 || d.variant == 35;
                    const double chunk = (100.0 + (it.b_dtype == TCA_F32 ? 280.0 : 150.0) * ppg * tile / 64.0) *
                                         (it.a_dense ? 1.0 : 1.6) * (block ? 0.5 : 1.0);
                    cost.push_back(1500.0 + nch * chunk);
                }
            }
        } else if (it.op == TCA_OP_SDDMM) {
            if (p->transposed) return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: SDDMM needs a forward plan");
            if (it.N <= 0 || it.N % 32 || it.lda < it.N || it.ldb < it.N || it.lda % 8 || it.ldb % 8 || !it.a || !it.b ||
                !aligned16(it.a) || !aligned16(it.b) || (t.nnz > 0 && !it.out))
                return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: SDDMM item " + std::to_string(i) + " has bad shapes/buffers");
            d.variant = tca::batch_variant(TCA_OP_SDDMM, t.rpw, 0, it.a_dtype, it.b_dtype, false, false);
            if (d.variant < 0)
                return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: SDDMM item " + std::to_string(i) +
                                                     " (batched SDDMM: f16 X with f32 or f16 G, bf16 X with f32 or bf16 G)");
            // the SDDMM runs on the plan's 64-row tiling and writes through dst
            const tca_plan *aux = nullptr;
            const int32_t *dst = nullptr;
            tca_status s = tca::sddmm_tiling(p, &aux, &dst);
            if (s != TCA_OK) return s;
            const tca::PlanTables &ta = aux->host;
            d.grp_chunk_ptr = aux->d_grp_chunk_ptr;
            d.chunk_cols = aux->d_chunk_cols;
            d.seg_ptr = aux->d_seg_ptr;
            d.entry_slot = aux->d_entry_slot;
            d.entry_lds = aux->d_entry_lds;
            d.entry_src = aux->d_entry_src;
            d.dst = dst;
            if (it.sgd_w) {
                if (it.sgd_w_dtype != TCA_F16 && it.sgd_w_dtype != TCA_BF16)
                    return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: SDDMM item " + std::to_string(i) + ": sgd_w must be F16 or BF16");
                if (!(it.sgd_lr == it.sgd_lr))
                    return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: SDDMM item " + std::to_string(i) + ": sgd_lr is NaN");
                d.sgd_w = static_cast<uint16_t *>(it.sgd_w);
                d.sgd_lr = it.sgd_lr;
                d.sgd_bf16 = it.sgd_w_dtype == TCA_BF16 ? 1 : 0;
            }
            d.rpw = ta.rpw;
            d.n_chunks = ta.n_chunks;
            d.ldb = it.ldb;   // X
            d.ldc = it.lda;   // G
            d.a_bytes = tca::operand_bytes(t.n_rows, it.N, it.lda, it.a_dtype == TCA_F32 ? 4 : 2);
            d.b_bytes = tca::operand_bytes(t.n_cols, it.N, it.ldb, 2);
            if (d.a_bytes == 0 || d.b_bytes == 0 || (long)t.nnz * 4 >= (1L << 31))
                return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: SDDMM item " + std::to_string(i) + " reaches 2 GiB");
            for (int gi = 0; gi < ta.n_groups; gi++) {
                const int g = ta.work_order[gi];
                for (int c = ta.grp_chunk_ptr[g]; c < ta.grp_chunk_ptr[g + 1]; c += tca::kBatchSddmmCB) {
                    const int nc = std::min(tca::kBatchSddmmCB, ta.grp_chunk_ptr[g + 1] - c);
                    units.push_back({i, g, c, 0, 0, 0, 0, 0});
                    // estimated ns: unit start and sampling + per 32-wide contraction step
                    cost.push_back(4000.0 + 700.0 * (it.N / 32.0) * nc / tca::kBatchSddmmCB);
                }
            }
        } else {
            return fail(TCA_ERR_INVALID_ARG, "tca_batch_create: unknown op");
        }
    }
    // Split SpMM units much longer than the average share of a workgroup into
    // chunk-range pieces (>= 16 chunks, 4 per wave of the workgroup); their
    // items accumulate into a zeroed output.
    int dev_id = 0, cus = 0;
    TCA_HIP_TRY(hipGetDevice(&dev_id));
    TCA_HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id));
    bool dense_only = true, any_bf16 = false, any_f16 = false;
    for (int i = 0; i < n_items; i++) {
        if (items[i].op == TCA_OP_SPMM && !items[i].a_dense) dense_only = false;
        const tca_dtype t16 = items[i].op == TCA_OP_SPMM ? items[i].a_dtype : items[i].b_dtype;   // A / X type
        (t16 == TCA_BF16 ? any_bf16 : any_f16) = true;
    }
    if (any_bf16 && any_f16)
        return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: a batch takes F16 or BF16 items, not both");
    const int mode = any_bf16 ? 2 : dense_only ? 1 : 0;
    const int workers = std::max(cus, 1) * tca::batch_blocks_per_cu(mode);
    double total_cost = 0.0;
    for (double c : cost) total_cost += c;
    const double share = total_cost / workers;
    std::vector<char> item_split(n_items, 0);
    {
        std::vector<tca::BatchUnitDev> u2;
        std::vector<double> c2;
        for (size_t k = 0; k < units.size(); k++) {
            const tca::BatchUnitDev &un = units[k];
            const bool is_spmm = items[un.item].op == TCA_OP_SPMM;
            const int nch = un.c_hi - un.c_lo;
            int P = 1;
            if (is_spmm && cost[k] > 2.0 * share) P = std::min((int)(cost[k] / share), nch / 16);
            if (P <= 1) {
                u2.push_back(un);
                c2.push_back(cost[k]);
                continue;
            }
            item_split[un.item] = 1;
            for (int q = 0; q < P; q++) {
                tca::BatchUnitDev v = un;
                v.c_lo = un.c_lo + (int)((long)nch * q / P);
                v.c_hi = un.c_lo + (int)((long)nch * (q + 1) / P);
                u2.push_back(v);
                c2.push_back(cost[k] / P);
            }
        }
        units.swap(u2);
        cost.swap(c2);
    }
    std::vector<tca::ZeroDesc> zero;
    for (int i = 0; i < n_items; i++)
        if (item_split[i]) zero.push_back({static_cast<float *>(items[i].out), dev[i].n_rows, items[i].N, items[i].ldc, 0});
    for (auto &un : units)
        if (item_split[un.item]) un.atomic = 1;

    // Partitions (one per XCD): the units of one (item, group) -- its N-tiles,
    // chunk pieces or SDDMM blocks, which read the same A tiles or G rows --
    // form a bundle; bundles in item order (a layer's forward, dX and dW
    // items are adjacent, sharing X and G), then group order, are cut into
    // kBatchParts runs of equal estimated cost, so that an XCD's L2 fetches a
    // group's tiles once and an item's B operand once per XCD its run spans.
    // Inside a partition the units run heaviest first; workgroups whose
    // partition is exhausted take units from the others (the kernel), which
    // absorbs the cost model's errors at the tail.
    const int grid = std::max(1, std::min((int)units.size(), workers));
    const int nparts = std::min(tca::kBatchParts, grid);
    std::vector<int> order(units.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return units[a].item != units[b].item ? units[a].item < units[b].item : units[a].g < units[b].g;
    });
    std::vector<std::vector<int>> part(nparts);
    {
        double total = 0.0;
        for (double c : cost) total += c;
        const double target = std::max(total / nparts, 1e-9);
        double cum = 0.0;
        for (size_t k = 0; k < order.size();) {
            size_t e = k;
            double c = 0.0;
            while (e < order.size() && units[order[e]].item == units[order[k]].item &&
                   units[order[e]].g == units[order[k]].g)
                c += cost[order[e++]];
            const int x = std::min(nparts - 1, (int)((cum + 0.5 * c) / target));
            for (size_t q = k; q < e; q++) part[x].push_back(order[q]);
            cum += c;
            k = e;
        }
    }
    std::vector<tca::BatchUnitDev> sorted;
    std::vector<int> part_off(tca::kBatchParts + 1, 0);
    for (int x = 0; x < nparts; x++) {
        std::stable_sort(part[x].begin(), part[x].end(), [&](int a, int b) { return cost[a] > cost[b]; });
        part_off[x] = (int)sorted.size();
        for (int k : part[x]) sorted.push_back(units[k]);
    }
    for (int x = nparts; x <= tca::kBatchParts; x++) part_off[x] = (int)sorted.size();

    tca_batch *b = new (std::nothrow) tca_batch();
    if (!b) return fail(TCA_ERR_ALLOC, "tca_batch_create: host allocation");
    b->n_items = n_items;
    b->n_units = (int)sorted.size();
    b->n_zero = (int)zero.size();
    b->device = dev_id;
    b->mode = mode;
    b->grid = grid;
    b->nparts = nparts;
    b->n_conv = (int)conv.size();
    b->conv8 = conv8;
    b->d_conv = static_cast<tca::ConvDesc *>(shadow.p);
    b->d_shadow = shadow.p;
    shadow.p = nullptr;   // owned by the batch from here
    b->d_pack_mem = packmem.p;
    b->d_pack = static_cast<tca::PackDesc *>(packmem.p);
    b->n_pack = (int)packs.size();
    b->pack_blocks = pack_blocks;
    packmem.p = nullptr;
    const size_t bi = sizeof(tca::BatchItemDev) * dev.size();
    const size_t bu = sizeof(tca::BatchUnitDev) * std::max<size_t>(sorted.size(), 1);
    const size_t bz = sizeof(tca::ZeroDesc) * std::max<size_t>(zero.size(), 1);
    const size_t bc = (bi + bu + bz + 255) / 256 * 256;   // counters start on a 256-byte line
    const size_t total = bc + tca::kBatchCounterBytes + 256;   // + partition offsets
    if (hipMalloc(&b->d_mem, total) != hipSuccess) {
        b->d_mem = nullptr;
        tca_batch_destroy(b);
        return fail(TCA_ERR_ALLOC, "tca_batch_create: device allocation");
    }
    char *m = static_cast<char *>(b->d_mem);
    b->d_items = reinterpret_cast<tca::BatchItemDev *>(m);
    b->d_units = reinterpret_cast<tca::BatchUnitDev *>(m + bi);
    b->d_zero = reinterpret_cast<tca::ZeroDesc *>(m + bi + bu);
    b->d_counters = reinterpret_cast<int *>(m + bc);
    b->d_part_off = reinterpret_cast<int *>(m + bc + tca::kBatchCounterBytes);
    if (hipMemcpy(b->d_items, dev.data(), bi, hipMemcpyHostToDevice) != hipSuccess ||
        (!sorted.empty() && hipMemcpy(b->d_units, sorted.data(), sizeof(tca::BatchUnitDev) * sorted.size(),
                                      hipMemcpyHostToDevice) != hipSuccess) ||
        (!zero.empty() && hipMemcpy(b->d_zero, zero.data(), sizeof(tca::ZeroDesc) * zero.size(),
                                    hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(b->d_part_off, part_off.data(), sizeof(int) * part_off.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(b->d_counters, 0, tca::kBatchCounterBytes) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        tca_batch_destroy(b);
        return fail(TCA_ERR_HIP, "tca_batch_create: upload failed");
    }
    *batch = b;
    return TCA_OK;
}

extern "C" tca_status tca_batch_run_stage(tca_batch *batch, int stage, void *stream)
{
    if (!batch) return fail(TCA_ERR_INVALID_ARG, "tca_batch_run: NULL batch");
    if (stage < 1 || stage > 3) return fail(TCA_ERR_INVALID_ARG, "tca_batch_run_stage: stage must be 1, 2 or 3");
    if ((stage & 1) && (batch->n_conv || batch->n_pack)) {
        const tca_status s = tca::launch_prologue(batch->d_conv, batch->n_conv, batch->conv8, batch->d_pack,
                                                  batch->n_pack, batch->pack_blocks, static_cast<hipStream_t>(stream));
        if (s != TCA_OK) return s;
    }
    if (!(stage & 2)) return TCA_OK;
    return tca::launch_batch(batch->d_items, batch->d_units, batch->n_units, batch->d_counters, batch->d_part_off,
                             batch->nparts, batch->grid, batch->d_zero, batch->n_zero, batch->mode,
                             static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_batch_run(tca_batch *batch, void *stream) { return tca_batch_run_stage(batch, 3, stream); }

extern "C" tca_status tca_batch_destroy(tca_batch *batch)
{
    if (!batch) return TCA_OK;
    if (batch->d_mem) (void)hipFree(batch->d_mem);
    if (batch->d_shadow) (void)hipFree(batch->d_shadow);
    if (batch->d_pack_mem) (void)hipFree(batch->d_pack_mem);
    delete batch;
    return TCA_OK;
}

extern "C" tca_status tca_batch_get_info(const tca_batch *batch, int *n_units, int *n_workgroups)
{
    if (!batch) return fail(TCA_ERR_INVALID_ARG, "tca_batch_get_info: NULL batch");
    if (n_units) *n_units = batch->n_units;
    if (n_workgroups) *n_workgroups = batch->grid;
    return TCA_OK;
}
