// C ABI: plan lifetime, value packing, SpMM / SDDMM / quantisation entry
// points.  Argument checking happens here; kernels assume valid arguments.
#include "tca_internal.hpp"

#include <vector>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

using tca::fail;

namespace {

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Element index of panel-tile slot (r16 << 5) | k in the SpMM kernels' LDS A
// tile: 16 rows x 32 k stored as octets, two rows per 128-byte line, octet
// slot XOR-swizzled by the row pair (the device function a_oct); precomputed
// here so that the kernels scatter entries without address arithmetic.
uint16_t a_tile_index(uint16_t slot)
{
    const int r = slot >> 5, q = (slot >> 3) & 3, m = r >> 1;
    return (uint16_t)((m * 8 + ((((r & 1) << 2) | q) ^ (m & 3))) * 8 + (slot & 7));
}

tca_status upload(tca_plan *p)
{
    const tca::PlanTables &t = p->host;
    const int ppg = t.rpw / 16;
    std::vector<uint16_t> lds(t.entry_slot.size());
    for (size_t e = 0; e < lds.size(); e++) lds[e] = a_tile_index(t.entry_slot[e]);
    const size_t sz[7] = {
        sizeof(int32_t) * (size_t)(t.n_groups + 1),
        sizeof(int32_t) * (size_t)t.n_chunks * 32,
        sizeof(int32_t) * ((size_t)t.n_chunks * ppg + 1),
        sizeof(uint16_t) * (size_t)t.nnz,
        sizeof(int32_t) * (size_t)t.nnz,
        sizeof(int32_t) * (size_t)t.n_groups,
        sizeof(uint16_t) * (size_t)t.nnz,
    };
    size_t off[7], total = 0;
    for (int i = 0; i < 7; i++) {
        off[i] = total;
        total = align_up(total + std::max<size_t>(sz[i], 4), 256);
    }
    TCA_HIP_TRY(hipGetDevice(&p->device));
    void *base = nullptr;
    if (hipMalloc(&base, total) != hipSuccess) return fail(TCA_ERR_ALLOC, "plan: device allocation failed");
    p->d_base = base;
    p->d_bytes = total;
    char *b = static_cast<char *>(base);
    const void *src[7] = {t.grp_chunk_ptr.data(), t.chunk_cols.data(), t.seg_ptr.data(), t.entry_slot.data(),
                          t.entry_src.data(),     t.work_order.data(), lds.data()};
    for (int i = 0; i < 7; i++)
        if (sz[i]) TCA_HIP_TRY(hipMemcpy(b + off[i], src[i], sz[i], hipMemcpyHostToDevice));
    p->d_grp_chunk_ptr = reinterpret_cast<const int32_t *>(b + off[0]);
    p->d_chunk_cols = reinterpret_cast<const int32_t *>(b + off[1]);
    p->d_seg_ptr = reinterpret_cast<const int32_t *>(b + off[2]);
    p->d_entry_slot = reinterpret_cast<const uint16_t *>(b + off[3]);
    p->d_entry_src = reinterpret_cast<const int32_t *>(b + off[4]);
    p->d_work_order = reinterpret_cast<const int32_t *>(b + off[5]);
    p->d_entry_lds = reinterpret_cast<const uint16_t *>(b + off[6]);
    return TCA_OK;
}

int dtype_bytes(tca_dtype d)
{
    switch (d) {
    case TCA_F16:
    case TCA_BF16: return 2;
    case TCA_E4M3: return 1;
    case TCA_F32: return 4;
    }
    return 0;
}

bool aligned(const void *p, size_t a) { return (reinterpret_cast<uintptr_t>(p) % a) == 0; }

// FNV-1a over the shape and the CSR arrays: identifies a pattern across plans
uint64_t pattern_digest(int n_rows, int n_cols, int nnz, const int32_t *row_offset, const int32_t *col_idx)
{
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *p, size_t n) {
        const unsigned char *c = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; i++) h = (h ^ c[i]) * 1099511628211ull;
    };
    const int dims[3] = {n_rows, n_cols, nnz};
    mix(dims, sizeof(dims));
    mix(row_offset, sizeof(int32_t) * ((size_t)n_rows + 1));
    if (nnz > 0) mix(col_idx, sizeof(int32_t) * (size_t)nnz);
    return h;
}

}  // namespace

namespace {

// Every table a launch on the plan reads: the SpMM unit tables and, for a
// forward plan, its SDDMM tiling with the SDDMM unit tables.  Built here so
// that tca_spmm / tca_sddmm / tca_batch_create never allocate or change a plan
// (capturable from the first call, safe from several threads).
tca_status prepare_launch_tables(tca_plan *p)
{
    tca_status s = tca::prepare_spmm_units(p);
    if (s != TCA_OK || p->transposed) return s;
    const tca_plan *aux = nullptr;
    const int32_t *dst = nullptr;
    s = tca::sddmm_tiling(p, &aux, &dst);
    for (int cb = 4; s == TCA_OK && cb >= 1; cb >>= 1) {
        const int32_t *units = nullptr;
        int n = 0;
        s = tca::sddmm_units(aux, cb, &units, &n);
    }
    return s;
}

tca_status plan_create(int n_rows, int n_cols, int nnz, const int32_t *row_offset, const int32_t *col_idx,
                       int rows_per_wave, tca_plan **plan, bool device)
{
    if (!plan) return fail(TCA_ERR_INVALID_ARG, "tca_plan_create: NULL plan");
    *plan = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    tca_plan *p = new (std::nothrow) tca_plan();
    if (!p) return fail(TCA_ERR_ALLOC, "tca_plan_create: host allocation");
    tca_status s = tca::build_tables(n_rows, n_cols, nnz, row_offset, col_idx, rows_per_wave, p->host);
    if (s != TCA_OK) { delete p; return s; }
    p->csr_row_offset.assign(row_offset, row_offset + n_rows + 1);
    p->csr_col_idx.assign(col_idx, col_idx + nnz);
    p->pattern_key = pattern_digest(n_rows, n_cols, nnz, row_offset, col_idx);
    p->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (device) {
        s = upload(p);
        if (s == TCA_OK) s = prepare_launch_tables(p);
        if (s != TCA_OK) { tca_plan_destroy(p); return s; }
    }
    *plan = p;
    return TCA_OK;
}

}  // namespace

extern "C" tca_status tca_plan_create_host(int n_rows, int n_cols, int nnz, const int32_t *row_offset,
                                           const int32_t *col_idx, int rows_per_wave, tca_plan **plan)
{
    return plan_create(n_rows, n_cols, nnz, row_offset, col_idx, rows_per_wave, plan, false);
}

extern "C" tca_status tca_plan_create(int n_rows, int n_cols, int nnz, const int32_t *row_offset, const int32_t *col_idx,
                                      int rows_per_wave, tca_plan **plan)
{
    return plan_create(n_rows, n_cols, nnz, row_offset, col_idx, rows_per_wave, plan, true);
}

extern "C" tca_status tca_plan_create_transpose(const tca_plan *fwd, int rows_per_wave, tca_plan **plan_t)
{
    if (!fwd || !plan_t) return fail(TCA_ERR_INVALID_ARG, "tca_plan_create_transpose: NULL argument");
    if (fwd->transposed) return fail(TCA_ERR_INVALID_ARG, "tca_plan_create_transpose: plan is already a transpose");
    *plan_t = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    const int M = fwd->host.n_rows, K = fwd->host.n_cols, nnz = fwd->host.nnz;
    // CSC of the forward CSR: rows of the transpose, with the CSR index of each entry
    std::vector<int32_t> rpT(K + 1, 0), cT(nnz), perm(nnz);
    const std::vector<int32_t> &rp = fwd->csr_row_offset, &ci = fwd->csr_col_idx;
    for (int e = 0; e < nnz; e++) rpT[ci[e] + 1]++;
    for (int j = 0; j < K; j++) rpT[j + 1] += rpT[j];
    std::vector<int32_t> fill(rpT.begin(), rpT.end() - 1);
    for (int r = 0; r < M; r++)
        for (int e = rp[r]; e < rp[r + 1]; e++) {
            const int d = fill[ci[e]]++;
            cT[d] = r;
            perm[d] = e;
        }
    tca_plan *p = new (std::nothrow) tca_plan();
    if (!p) return fail(TCA_ERR_ALLOC, "tca_plan_create_transpose: host allocation");
    tca_status s = tca::build_tables(K, M, nnz, rpT.data(), cT.data(), rows_per_wave, p->host);
    if (s != TCA_OK) { delete p; return s; }
    // entry_src: transpose CSR index -> forward CSR index -> forward tiled index
    std::vector<int32_t> fwd_tiled_of_csr(nnz);
    for (int e = 0; e < nnz; e++) fwd_tiled_of_csr[fwd->host.entry_src[e]] = e;
    p->entry_csr.resize(nnz);
    for (int e = 0; e < nnz; e++) {
        p->entry_csr[e] = perm[p->host.entry_src[e]];
        p->host.entry_src[e] = fwd_tiled_of_csr[p->entry_csr[e]];
    }
    p->transposed = 1;
    p->pattern_key = fwd->pattern_key;
    p->csr_row_offset = std::move(rpT);
    p->csr_col_idx = std::move(cT);
    p->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (fwd->d_base) {  // a host-only forward plan gives a host-only transpose
        s = upload(p);
        if (s == TCA_OK) s = prepare_launch_tables(p);
        if (s != TCA_OK) { tca_plan_destroy(p); return s; }
    }
    *plan_t = p;
    return TCA_OK;
}

extern "C" tca_status tca_plan_destroy(tca_plan *plan)
{
    if (!plan) return TCA_OK;
    if (plan->sddmm_aux) tca_plan_destroy(plan->sddmm_aux);
    int cur = 0;
    bool switched = false;
    if ((plan->d_base || plan->d_sddmm_dst || plan->d_sddmm_units[0] || plan->d_sddmm_units[1] ||
         plan->d_sddmm_units[2] || !plan->unit_tables.empty()) &&
        hipGetDevice(&cur) == hipSuccess && cur != plan->device)
        switched = hipSetDevice(plan->device) == hipSuccess;
    if (plan->d_base) (void)hipFree(plan->d_base);
    if (plan->d_sddmm_dst) (void)hipFree(plan->d_sddmm_dst);
    for (void *u : plan->d_sddmm_units)
        if (u) (void)hipFree(u);
    for (auto &u : plan->unit_tables)
        if (u.dev) (void)hipFree(u.dev);
    if (switched) (void)hipSetDevice(cur);
    delete plan;
    return TCA_OK;
}

namespace tca {

tca_status sddmm_tiling(const tca_plan *cp, const tca_plan **aux, const int32_t **dst)
{
    tca_plan *p = const_cast<tca_plan *>(cp);
    if (p->host.rpw == kSddmmRows) {
        *aux = p;
        *dst = nullptr;
        return TCA_OK;
    }
    if (!p->sddmm_aux) {
        const PlanTables &t = p->host;
        tca_plan *a = new (std::nothrow) tca_plan();
        if (!a) return fail(TCA_ERR_ALLOC, "sddmm: host allocation");
        tca_status s = build_tables(t.n_rows, t.n_cols, t.nnz, p->csr_row_offset.data(), p->csr_col_idx.data(),
                                    kSddmmRows, a->host);
        if (s == TCA_OK) s = upload(a);
        if (s != TCA_OK) {
            tca_plan_destroy(a);
            return s;
        }
        // entry of the 64-row tiling -> CSR index -> this plan's tiled index
        std::vector<int32_t> tiled_of_csr(t.nnz), d(t.nnz);
        for (int e = 0; e < t.nnz; e++) tiled_of_csr[t.entry_src[e]] = e;
        for (int e = 0; e < t.nnz; e++) d[e] = tiled_of_csr[a->host.entry_src[e]];
        if (t.nnz > 0) {
            if (hipMalloc(&p->d_sddmm_dst, sizeof(int32_t) * t.nnz) != hipSuccess) {
                tca_plan_destroy(a);
                return fail(TCA_ERR_ALLOC, "sddmm: device allocation");
            }
            if (hipMemcpy(p->d_sddmm_dst, d.data(), sizeof(int32_t) * t.nnz, hipMemcpyHostToDevice) != hipSuccess) {
                tca_plan_destroy(a);
                return fail(TCA_ERR_HIP, "sddmm: upload");
            }
        }
        p->sddmm_aux = a;
    }
    *aux = p->sddmm_aux;
    *dst = p->d_sddmm_dst;
    return TCA_OK;
}

tca_status sddmm_units(const tca_plan *cp, int cb, const int32_t **units, int *n)
{
    tca_plan *p = const_cast<tca_plan *>(cp);
    const int slot = cb == 4 ? 0 : cb == 2 ? 1 : 2;
    if (!p->d_sddmm_units[slot]) {
        const PlanTables &t = p->host;
        std::vector<int32_t> ug, uc;
        for (int gi = 0; gi < t.n_groups; gi++) {
            const int g = t.work_order[gi];
            for (int c = t.grp_chunk_ptr[g]; c < t.grp_chunk_ptr[g + 1]; c += cb) {
                ug.push_back(g);
                uc.push_back(c);
            }
        }
        const size_t m = ug.size();
        ug.insert(ug.end(), uc.begin(), uc.end());
        if (hipMalloc(&p->d_sddmm_units[slot], sizeof(int32_t) * std::max<size_t>(ug.size(), 2)) != hipSuccess)
            return fail(TCA_ERR_ALLOC, "sddmm: unit table allocation");
        if (!ug.empty())
            TCA_HIP_TRY(hipMemcpy(p->d_sddmm_units[slot], ug.data(), sizeof(int32_t) * ug.size(), hipMemcpyHostToDevice));
        p->n_sddmm_units[slot] = (int)m;
    }
    *units = static_cast<const int32_t *>(p->d_sddmm_units[slot]);
    *n = p->n_sddmm_units[slot];
    return TCA_OK;
}

}  // namespace tca

extern "C" tca_status tca_plan_get_info(const tca_plan *plan, tca_plan_info *info)
{
    if (!plan || !info) return fail(TCA_ERR_INVALID_ARG, "tca_plan_get_info: NULL argument");
    const tca::PlanTables &t = plan->host;
    info->n_rows = t.n_rows;
    info->n_cols = t.n_cols;
    info->nnz = t.nnz;
    info->rows_per_wave = t.rpw;
    info->n_groups = t.n_groups;
    info->n_chunks = t.n_chunks;
    info->max_group_chunks = t.max_group_chunks;
    info->transposed = plan->transposed;
    info->device_bytes = plan->d_bytes;
    info->build_ms = plan->build_ms;
    info->dense_elems = (size_t)tca::dense_elems(t);
    return TCA_OK;
}

extern "C" tca_status tca_plan_export(const tca_plan *plan, int32_t *grp_chunk_ptr, int32_t *chunk_cols, int32_t *seg_ptr,
                                      uint16_t *entry_slot, int32_t *entry_src, int32_t *work_order)
{
    if (!plan) return fail(TCA_ERR_INVALID_ARG, "tca_plan_export: NULL plan");
    const tca::PlanTables &t = plan->host;
    auto cp = [](auto *dst, const auto &v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(grp_chunk_ptr, t.grp_chunk_ptr);
    cp(chunk_cols, t.chunk_cols);
    cp(seg_ptr, t.seg_ptr);
    cp(entry_slot, t.entry_slot);
    cp(entry_src, t.entry_src);
    cp(work_order, t.work_order);
    return TCA_OK;
}

extern "C" tca_status tca_pack_values(const tca_plan *plan, tca_dtype dtype, const void *d_csr, void *d_tiled, void *stream)
{
    if (!plan || (plan->host.nnz > 0 && (!d_csr || !d_tiled))) return fail(TCA_ERR_INVALID_ARG, "tca_pack_values: NULL argument");
    if (!plan->d_base) return fail(TCA_ERR_INVALID_ARG, "tca_pack_values: host-only plan");
    if (plan->transposed) return fail(TCA_ERR_INVALID_ARG, "tca_pack_values: use the forward plan");
    const int eb = dtype_bytes(dtype);
    if (!eb) return fail(TCA_ERR_INVALID_ARG, "tca_pack_values: dtype");
    return tca::launch_gather(plan->d_entry_src, plan->host.nnz, eb, d_csr, d_tiled, false, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_unpack_values(const tca_plan *plan, tca_dtype dtype, const void *d_tiled, void *d_csr, void *stream)
{
    if (!plan || (plan->host.nnz > 0 && (!d_csr || !d_tiled))) return fail(TCA_ERR_INVALID_ARG, "tca_unpack_values: NULL argument");
    if (!plan->d_base) return fail(TCA_ERR_INVALID_ARG, "tca_unpack_values: host-only plan");
    if (plan->transposed) return fail(TCA_ERR_INVALID_ARG, "tca_unpack_values: use the forward plan");
    const int eb = dtype_bytes(dtype);
    if (!eb) return fail(TCA_ERR_INVALID_ARG, "tca_unpack_values: dtype");
    return tca::launch_gather(plan->d_entry_src, plan->host.nnz, eb, d_tiled, d_csr, true, static_cast<hipStream_t>(stream));
}

static tca_status spmm_entry(const tca_plan *plan, tca_dtype a_dtype, const void *d_a_tiled, tca_dtype b_dtype,
                             const void *d_B, int ldb, int N, float *d_C, int ldc, const float *d_row_scale,
                             const float *d_col_scale, int tile_n, bool dense, void *stream)
{
    if (!plan) return fail(TCA_ERR_INVALID_ARG, "tca_spmm: NULL plan");
    if (!plan->d_base) return fail(TCA_ERR_INVALID_ARG, "tca_spmm: host-only plan");
    if (N <= 0) return fail(TCA_ERR_INVALID_ARG, "tca_spmm: N must be positive");
    if (N % 8) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm: N must be a multiple of 8");
    if (ldb < N || ldc < N || ldb % 8 || ldc % 4) return fail(TCA_ERR_INVALID_ARG, "tca_spmm: ldb/ldc too small or misaligned");
    if (!d_B || !d_C || (plan->host.nnz > 0 && !d_a_tiled)) return fail(TCA_ERR_INVALID_ARG, "tca_spmm: NULL buffer");
    if (!aligned(d_B, 16) || !aligned(d_C, 16)) return fail(TCA_ERR_INVALID_ARG, "tca_spmm: B and C must be 16-byte aligned");
    if (tile_n != 0 && tile_n != 16 && tile_n != 32 && tile_n != 64 && tile_n != 128)
        return fail(TCA_ERR_INVALID_ARG, "tca_spmm: tile_n must be 0, 16, 32, 64 or 128");
    if (b_dtype == TCA_F32 && a_dtype == TCA_E4M3) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm: quantise B to E4M3 first");
    if (b_dtype != a_dtype && b_dtype != TCA_F32) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm: B dtype must equal A dtype or be F32");
    return tca::launch_spmm(plan, a_dtype, d_a_tiled, b_dtype, d_B, ldb, N, d_C, ldc, d_row_scale, d_col_scale, tile_n,
                            dense, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_spmm(const tca_plan *plan, tca_dtype a_dtype, const void *d_a_tiled, tca_dtype b_dtype,
                               const void *d_B, int ldb, int N, float *d_C, int ldc, const float *d_row_scale,
                               const float *d_col_scale, int tile_n, void *stream)
{
    return spmm_entry(plan, a_dtype, d_a_tiled, b_dtype, d_B, ldb, N, d_C, ldc, d_row_scale, d_col_scale, tile_n, false,
                      stream);
}

extern "C" tca_status tca_spmm_dense(const tca_plan *plan, tca_dtype a_dtype, const void *d_a_dense, tca_dtype b_dtype,
                                     const void *d_B, int ldb, int N, float *d_C, int ldc, const float *d_row_scale,
                                     const float *d_col_scale, int tile_n, void *stream)
{
    if (!aligned(d_a_dense, 16)) return fail(TCA_ERR_INVALID_ARG, "tca_spmm_dense: tiles must be 16-byte aligned");
    return spmm_entry(plan, a_dtype, d_a_dense, b_dtype, d_B, ldb, N, d_C, ldc, d_row_scale, d_col_scale, tile_n, true,
                      stream);
}

extern "C" tca_status tca_pack_dense(const tca_plan *plan, tca_dtype dtype, const void *d_tiled, void *d_dense,
                                     void *stream)
{
    if (!plan || (plan->host.nnz > 0 && !d_tiled) || (plan->host.n_chunks > 0 && !d_dense))
        return fail(TCA_ERR_INVALID_ARG, "tca_pack_dense: NULL argument");
    if (!plan->d_base) return fail(TCA_ERR_INVALID_ARG, "tca_pack_dense: host-only plan");
    if (dtype != TCA_F16 && dtype != TCA_BF16 && dtype != TCA_E4M3) return fail(TCA_ERR_INVALID_ARG, "tca_pack_dense: dtype");
    if (!aligned(d_dense, 16)) return fail(TCA_ERR_INVALID_ARG, "tca_pack_dense: tiles must be 16-byte aligned");
    return tca::launch_pack_dense(plan, dtype == TCA_E4M3 ? 1 : 2, d_tiled, d_dense, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_bell_spmm(int n_rows, int n_cols, int block, int ell_cols, const int32_t *d_col_ind,
                                    tca_dtype a_dtype, const void *d_values, tca_dtype b_dtype, const void *d_B, int ldb,
                                    int N, float *d_C, int ldc, int tile_n, int rows_per_wave, void *stream)
{
    if (n_rows < 0 || n_cols < 0 || N <= 0) return fail(TCA_ERR_INVALID_ARG, "tca_bell_spmm: bad shape");
    if (block < 16 || block > 256 || (block & (block - 1)))
        return fail(TCA_ERR_UNSUPPORTED, "tca_bell_spmm: block must be 16, 32, 64, 128 or 256");
    if (ell_cols < 0 || ell_cols % block) return fail(TCA_ERR_INVALID_ARG, "tca_bell_spmm: ell_cols must be a multiple of block");
    if (N % 8) return fail(TCA_ERR_UNSUPPORTED, "tca_bell_spmm: N must be a multiple of 8");
    if (ldb < N || ldc < N || ldb % 8 || ldc % 4) return fail(TCA_ERR_INVALID_ARG, "tca_bell_spmm: ldb/ldc too small or misaligned");
    if (!d_B || !d_C || (n_rows > 0 && ell_cols > 0 && (!d_col_ind || !d_values)))
        return fail(TCA_ERR_INVALID_ARG, "tca_bell_spmm: NULL buffer");
    if (!aligned(d_B, 16) || !aligned(d_C, 16) || !aligned(d_values, 16) || !aligned(d_col_ind, 4))
        return fail(TCA_ERR_INVALID_ARG, "tca_bell_spmm: values, B and C must be 16-byte aligned");
    if (tile_n != 0 && tile_n != 16 && tile_n != 32 && tile_n != 64)
        return fail(TCA_ERR_INVALID_ARG, "tca_bell_spmm: tile_n must be 0, 16, 32 or 64");
    if (rows_per_wave != 0 && rows_per_wave != 16 && rows_per_wave != 32 && rows_per_wave != 64)
        return fail(TCA_ERR_INVALID_ARG, "tca_bell_spmm: rows_per_wave must be 0, 16, 32 or 64");
    if (a_dtype != TCA_F16 && a_dtype != TCA_BF16) return fail(TCA_ERR_UNSUPPORTED, "tca_bell_spmm: values must be F16 or BF16");
    if (b_dtype != a_dtype && b_dtype != TCA_F32) return fail(TCA_ERR_UNSUPPORTED, "tca_bell_spmm: B dtype must equal A dtype or be F32");
    return tca::launch_bell_spmm(n_rows, n_cols, N, block, ell_cols, d_col_ind, a_dtype, d_values, b_dtype, d_B, ldb, d_C,
                                 ldc, tile_n, rows_per_wave, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_compress_24(int n_rows, int n_cols, tca_dtype dtype, const void *d_dense, int ld,
                                      void *d_values, uint16_t *d_meta, void *stream)
{
    if (n_rows < 0 || n_cols < 0 || ld < n_cols || ld % 8) return fail(TCA_ERR_INVALID_ARG, "tca_compress_24: bad shape or ld");
    if (n_cols % 64) return fail(TCA_ERR_UNSUPPORTED, "tca_compress_24: n_cols must be a multiple of 64");
    if (dtype != TCA_F16 && dtype != TCA_BF16) return fail(TCA_ERR_UNSUPPORTED, "tca_compress_24: dtype must be F16 or BF16");
    if ((long)n_rows * n_cols > 0 && (!d_dense || !d_values || !d_meta)) return fail(TCA_ERR_INVALID_ARG, "tca_compress_24: NULL buffer");
    if (!aligned(d_dense, 16) || !aligned(d_values, 16) || !aligned(d_meta, 2))
        return fail(TCA_ERR_INVALID_ARG, "tca_compress_24: dense and values must be 16-byte aligned");
    return tca::launch_compress24(n_rows, n_cols, dtype, d_dense, ld, d_values, d_meta, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_spmm_24(int n_rows, int n_cols, int N, tca_dtype a_dtype, const void *d_values,
                                  const uint16_t *d_meta, tca_dtype b_dtype, const void *d_B, int ldb, float *d_C,
                                  int ldc, int tile_n, int rows_per_wave, void *stream)
{
    if (n_rows < 0 || n_cols < 0 || N <= 0) return fail(TCA_ERR_INVALID_ARG, "tca_spmm_24: bad shape");
    if (n_cols % 64) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm_24: n_cols must be a multiple of 64");
    if (N % 8) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm_24: N must be a multiple of 8");
    if (ldb < N || ldc < N || ldb % 8 || ldc % 4) return fail(TCA_ERR_INVALID_ARG, "tca_spmm_24: ldb/ldc too small or misaligned");
    if (!d_B || !d_C || ((long)n_rows * n_cols > 0 && (!d_values || !d_meta)))
        return fail(TCA_ERR_INVALID_ARG, "tca_spmm_24: NULL buffer");
    if (!aligned(d_B, 16) || !aligned(d_C, 16) || !aligned(d_values, 16) || !aligned(d_meta, 2))
        return fail(TCA_ERR_INVALID_ARG, "tca_spmm_24: values, B and C must be 16-byte aligned");
    if (tile_n != 0 && tile_n != 16 && tile_n != 32 && tile_n != 64)
        return fail(TCA_ERR_INVALID_ARG, "tca_spmm_24: tile_n must be 0, 16, 32 or 64");
    if (rows_per_wave != 0 && rows_per_wave != 16 && rows_per_wave != 32 && rows_per_wave != 64)
        return fail(TCA_ERR_INVALID_ARG, "tca_spmm_24: rows_per_wave must be 0, 16, 32 or 64");
    if (a_dtype != TCA_F16 && a_dtype != TCA_BF16) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm_24: values must be F16 or BF16");
    if (b_dtype != a_dtype && b_dtype != TCA_F32) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm_24: B dtype must equal A dtype or be F32");
    return tca::launch_spmm24(n_rows, n_cols, N, a_dtype, d_values, d_meta, b_dtype, d_B, ldb, d_C, ldc, tile_n,
                              rows_per_wave, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_bell_shape(int n_rows, int n_cols, int N, int block, int tile_n, int *rows_per_wave,
                                     int *tile)
{
    if (!rows_per_wave || !tile || block < 16 || (block & (block - 1)))
        return fail(TCA_ERR_INVALID_ARG, "tca_bell_shape: bad argument");
    tca::bell_shape(n_rows, n_cols, N, block, tile_n, rows_per_wave, tile);
    return TCA_OK;
}

extern "C" tca_status tca_sddmm(const tca_plan *plan, tca_dtype g_dtype, const void *d_G, int ldg, tca_dtype x_dtype,
                                const void *d_X, int ldx, int N, float *d_out_tiled, void *stream)
{
    if (!plan) return fail(TCA_ERR_INVALID_ARG, "tca_sddmm: NULL plan");
    if (!plan->d_base) return fail(TCA_ERR_INVALID_ARG, "tca_sddmm: host-only plan");
    if (N <= 0 || N % 32) return fail(TCA_ERR_UNSUPPORTED, "tca_sddmm: N must be a positive multiple of 32");
    if (ldg < N || ldx < N || ldg % 8 || ldx % 8) return fail(TCA_ERR_INVALID_ARG, "tca_sddmm: ldg/ldx too small or misaligned");
    if (!d_G || !d_X || (plan->host.nnz > 0 && !d_out_tiled)) return fail(TCA_ERR_INVALID_ARG, "tca_sddmm: NULL buffer");
    if (!aligned(d_G, 16) || !aligned(d_X, 16)) return fail(TCA_ERR_INVALID_ARG, "tca_sddmm: G and X must be 16-byte aligned");
    return tca::launch_sddmm(plan, g_dtype, d_G, ldg, x_dtype, d_X, ldx, N, d_out_tiled, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_quantize_rows_e4m3(int n_rows, const int32_t *d_row_offset, const uint16_t *d_values, uint8_t *d_q,
                                             float *d_row_scale, void *stream)
{
    if (n_rows < 0 || (n_rows > 0 && (!d_row_offset || !d_row_scale))) return fail(TCA_ERR_INVALID_ARG, "tca_quantize_rows_e4m3: bad argument");
    return tca::launch_quantize_rows(n_rows, d_row_offset, d_values, d_q, d_row_scale, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_quantize_cols_e4m3(int K, int N, const uint16_t *d_B, int ldb, uint8_t *d_q, int ldq,
                                             float *d_col_scale, void *stream)
{
    if (K < 0 || N < 0 || ldb < N || ldq < N || (K > 0 && N > 0 && (!d_B || !d_q || !d_col_scale)))
        return fail(TCA_ERR_INVALID_ARG, "tca_quantize_cols_e4m3: bad argument");
    if ((K + 63) / 64 > 65535) return fail(TCA_ERR_UNSUPPORTED, "tca_quantize_cols_e4m3: K above 4194240");
    return tca::launch_quantize_cols(K, N, d_B, ldb, d_q, ldq, d_col_scale, static_cast<hipStream_t>(stream));
}

extern "C" tca_status tca_sgd_update(tca_dtype w_dtype, void *d_w, const float *d_g, size_t n, float lr, void *stream)
{
    if (w_dtype != TCA_F16 && w_dtype != TCA_BF16) return fail(TCA_ERR_UNSUPPORTED, "tca_sgd_update: weights must be F16 or BF16");
    if (n == 0) return TCA_OK;
    if (!d_w || !d_g) return fail(TCA_ERR_INVALID_ARG, "tca_sgd_update: NULL buffer");
    if (!aligned(d_w, 16) || !aligned(d_g, 16)) return fail(TCA_ERR_INVALID_ARG, "tca_sgd_update: w and g must be 16-byte aligned");
    return tca::launch_sgd(w_dtype == TCA_BF16, d_w, d_g, (long)n, lr, static_cast<hipStream_t>(stream));
}
