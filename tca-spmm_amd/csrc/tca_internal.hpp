// Internal declarations shared by the host code, the kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include "../../include/tca_spmm.h"

namespace tca {

// Thread-local error detail behind tca_last_error().
void set_error(const std::string &msg);
tca_status fail(tca_status s, const std::string &msg);

#define TCA_HIP_TRY(expr)                                                              \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess)                                                          \
            return ::tca::fail(TCA_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Panel tiling of one sparse matrix (see tca_plan_create).
struct PlanTables {
    int n_rows = 0, n_cols = 0, nnz = 0, rpw = 16;
    int n_groups = 0, n_chunks = 0, max_group_chunks = 0;
    std::vector<int32_t> grp_chunk_ptr;  // n_groups + 1
    std::vector<int32_t> chunk_cols;     // 32 * n_chunks, -1 padded
    std::vector<int32_t> seg_ptr;        // n_chunks * ppg + 1
    std::vector<uint16_t> entry_slot;    // nnz: (r16 << 5) | k
    std::vector<int32_t> entry_src;      // nnz: CSR index (or forward tiled index)
    std::vector<int32_t> work_order;     // n_groups, heavy first
};

// Parses .smtx text into a malloc'ed CSR pattern.
tca_status smtx_parse(const std::string &text, int *n_rows, int *n_cols, int *nnz, int32_t **row_offset,
                      int32_t **col_idx);

// Builds the tables from a CSR pattern (sorted, distinct columns per row).
tca_status build_tables(int n_rows, int n_cols, int nnz, const int32_t *row_offset, const int32_t *col_idx,
                        int rpw, PlanTables &t);

}  // namespace tca

struct tca_plan {
    tca::PlanTables host;
    int transposed = 0;
    int device = 0;
    double build_ms = 0.0;
    // device tables (one allocation)
    void *d_base = nullptr;
    size_t d_bytes = 0;
    const int32_t *d_grp_chunk_ptr = nullptr;
    const int32_t *d_chunk_cols = nullptr;
    const int32_t *d_seg_ptr = nullptr;
    const uint16_t *d_entry_slot = nullptr;
    const uint16_t *d_entry_lds = nullptr;   // entry_slot as an index into the SpMM LDS A tile
    const int32_t *d_entry_src = nullptr;
    const int32_t *d_work_order = nullptr;
    // SDDMM: the 64-row tiling its kernel runs on (this plan itself when it has
    // 64-row groups) and, for each of that tiling's entries, its index in this
    // plan's tiled order (null for this plan itself); built with the plan
    tca_plan *sddmm_aux = nullptr;
    int32_t *d_sddmm_dst = nullptr;
    // SDDMM work units (group, first chunk) of this plan as an SDDMM tiling,
    // for 4-, 2- and 1-chunk blocks, built with the plan
    void *d_sddmm_units[3] = {nullptr, nullptr, nullptr};
    int n_sddmm_units[3] = {0, 0, 0};
    // SpMM unit tables (group, pieces, c_lo, c_hi), one per split factor, built
    // with the plan (prepare_spmm_units)
    struct UnitTable {
        int pieces = 1, count = 0;
        bool split = false;
        void *dev = nullptr;
    };
    std::vector<UnitTable> unit_tables;
    // CSR pattern kept for building the transpose
    std::vector<int32_t> csr_row_offset, csr_col_idx;
    // transposed plans: CSR index (in the forward matrix) of each entry; forward
    // plans use host.entry_src
    std::vector<int32_t> entry_csr;
    // digest of the forward matrix's pattern (shape, row_offset, col_idx), the
    // same for a forward plan and every transpose made from it
    uint64_t pattern_key = 0;
};

namespace tca {

// Kernel launchers (tca_kernels.hip).
tca_status launch_spmm(const tca_plan *p, tca_dtype a_dtype, const void *a, tca_dtype b_dtype, const void *B,
                       int ldb, int N, float *C, int ldc, const float *row_scale, const float *col_scale,
                       int tile_n, bool dense, hipStream_t stream);
// Panel-dense tiles of a plan: one 16 x 32 tile per (chunk, 16-row panel).
inline long dense_elems(const PlanTables &t) { return (long)t.n_chunks * (t.rpw / 16) * 512; }
inline long dense_bytes(const PlanTables &t, tca_dtype dt) { return dense_elems(t) * (dt == TCA_E4M3 ? 1 : 2); }
tca_status launch_pack_dense(const tca_plan *p, int elem_bytes, const void *tiled, void *dense, hipStream_t stream);
tca_status launch_sddmm(const tca_plan *p, tca_dtype g_dtype, const void *G, int ldg, tca_dtype x_dtype,
                        const void *X, int ldx, int N, float *out, hipStream_t stream);
tca_status launch_gather(const int32_t *idx, int n, int elem_bytes, const void *src, void *dst, bool scatter,
                         hipStream_t stream);
tca_status launch_sgd(bool bf16, void *w, const float *g, long n, float lr, hipStream_t stream);
tca_status launch_quantize_rows(int n_rows, const int32_t *row_offset, const uint16_t *vals, uint8_t *q,
                                float *scale, hipStream_t stream);
tca_status launch_quantize_cols(int K, int N, const uint16_t *B, int ldb, uint8_t *q, int ldq, float *scale,
                                hipStream_t stream);
int default_tile_n(const tca_plan *p, int N);
// Unit tables of every split factor a standalone SpMM launch can pick, built
// at plan creation (launches only look them up).
tca_status prepare_spmm_units(tca_plan *p);
tca_status launch_warm_up(hipStream_t stream);
// The SDDMM tiling of a forward plan (tca_capi.cpp): *aux has 64-row groups,
// *dst maps its entries to p's tiled order (null when *aux == p).  Built, with
// its unit tables, when the forward plan is created; later calls look it up.
constexpr int kSddmmRows = 64;
tca_status sddmm_tiling(const tca_plan *p, const tca_plan **aux, const int32_t **dst);
// SDDMM work units (group, first chunk) of a 64-row tiling for blocks of cb (4,
// 2 or 1) chunks, heavy groups first: n pairs at *units (groups, then chunks).
tca_status sddmm_units(const tca_plan *aux, int cb, const int32_t **units, int *n);
// Bytes from the first to the last element of a rows x N row-major operand with
// leading dimension ld; 0 when that is not below 2^31 (the kernels address
// operands with 32-bit byte offsets).
uint32_t operand_bytes(long rows, long N, long ld, int elem_bytes);

// Blocked-ELL SpMM (tca_bell.hip): shape choice and launch.
void bell_shape(int M, int K, int N, int block, int tile_n, int *rpw, int *nt);
// 2:4 sparse MFMA path (tca_sparse24.hip).
tca_status launch_compress24(int M, int K, tca_dtype dtype, const void *dense, int ld, void *values, uint16_t *meta,
                             hipStream_t stream);
tca_status launch_spmm24(int M, int K, int N, tca_dtype a_dtype, const void *values, const uint16_t *meta,
                         tca_dtype b_dtype, const void *B, int ldb, float *C, int ldc, int tile_n, int rows_per_wave,
                         hipStream_t stream);
tca_status launch_bell_spmm(int M, int K, int N, int block, int ell_cols, const int32_t *col_ind, tca_dtype a_dtype,
                            const void *values, tca_dtype b_dtype, const void *B, int ldb, float *C, int ldc,
                            int tile_n, int rows_per_wave, hipStream_t stream);

// Batched persistent launch: one descriptor per item, one record per work unit.
struct BatchItemDev {
    int32_t variant;   // kernel variant id (batch_variant)
    int32_t n_rows, N, ldb, ldc, n_tiles, rpw, nnz, n_chunks;
    uint32_t a_bytes, b_bytes;   // SPMM: -, B; SDDMM: G, X (bytes from first to last used element)
    const int32_t *grp_chunk_ptr, *chunk_cols, *seg_ptr;   // SDDMM: of the 64-row tiling
    const uint16_t *entry_slot, *entry_lds;
    const int32_t *entry_src;
    const int32_t *dst;          // SDDMM: output index of each entry (null: identity)
    const void *a, *b;
    float *out;
    uint16_t *sgd_w;             // SDDMM: weights updated from each written entry (null: none)
    float sgd_lr;
    int32_t sgd_bf16;
};
struct BatchUnitDev {
    int32_t item, g, t;        // SPMM: (group, N-tile); SDDMM: (group, first chunk)
    int32_t c_lo, c_hi;        // SPMM: chunk range of the group
    int32_t atomic, pad0, pad1;
};
struct ZeroDesc {              // an output that split units accumulate into
    float *C;
    int32_t rows, N, ldc, pad;
};
struct ConvDesc {              // an FP32 operand a batch rounds to an FP16 / BF16 copy on every run
    const float *src;
    uint16_t *dst;             // rows x N, row pitch N
    int32_t rows, N, ld, bf16; // bf16: the copy is BF16 (else FP16)
    int64_t first8;            // its first 8-element group in the conversion launch (a multiple of 256)
};
// FP16 copies: every operand of a ConvDesc list is rounded (nearest even, the
// rounding the kernels apply on load) into its copy, total8 groups of 8
// elements over all of them, at most kMaxConv operands (launch_prologue).
constexpr int kMaxConv = 256;
struct PackDesc {              // panel-dense tiles a batch rebuilds from weight values on every run
    const int32_t *seg_ptr;    // the plan's (chunk, panel) entry ranges: tile t = seg_ptr[t]..seg_ptr[t+1]
    const uint16_t *entry_slot;
    const int32_t *map;        // index of each entry's value in src (null: the entry's own index)
    const uint16_t *src;       // weight values (16-bit)
    uint16_t *dst;             // n_tiles x 512 elements
    int64_t n_tiles;
    int64_t first_block;       // its first workgroup in the prologue's pack range
};
constexpr int kMaxPack = 512;
constexpr int kPackTilesPerBlock = 4;
// One launch: every FP16 copy (conversion groups of 8, blocks [0, conv_blocks))
// and every dense-tile pack (blocks from conv_blocks on, kPackTilesPerBlock
// tiles each).  total8 as for the conversion; pack_blocks = the packs' blocks.
tca_status launch_prologue(const ConvDesc *conv, int n_conv, long total8, const PackDesc *pack, int n_pack,
                           long pack_blocks, hipStream_t stream);
constexpr int kBatchSddmmCB = 4;   // chunks per SDDMM unit in a batch
// Work counters of a batch launch: the unit list is cut into up to
// kBatchParts partitions (one per XCD under round-robin workgroup dispatch),
// each with its own counter 4 KiB from the next (distinct memory channels),
// then one count of finished workgroups.
constexpr int kBatchParts = 8;
constexpr int kCounterStride = 1024;   // ints
constexpr size_t kBatchCounterBytes = (size_t)(kBatchParts + 1) * kCounterStride * sizeof(int);
// -1 when the combination has no batched variant.
int batch_variant(tca_op op, int rpw, int tile_n, tca_dtype a_dtype, tca_dtype b_dtype, bool transposed, bool dense);
// resident one-wave workgroups of batch_kernel per compute unit
// mode: 0 = every F16 case, 1 = F16 panel-dense SpMM + SDDMM, 2 = the BF16 ones
int batch_blocks_per_cu(int mode);
tca_status launch_batch(const BatchItemDev *items, const BatchUnitDev *units, int n_units, int *counters,
                        const int *part_off, int nparts, int grid, const ZeroDesc *zero, int n_zero, int mode,
                        hipStream_t stream);

}  // namespace tca
