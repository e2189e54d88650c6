/*
 * tca_spmm.h -- C ABI of the MI355X (gfx950) row-balanced CSR x dense SpMM.
 *
 * Plain C: pointers, sizes, status codes.  Device pointers are HIP device
 * addresses; `stream` is a hipStream_t passed as void* (NULL = default stream).
 * Every entry point returns a tca_status; tca_last_error() gives the detail of
 * the last failure on the calling thread.  Nothing here allocates on the
 * caller's behalf except where stated (plans, batches, tca_smtx_read).
 *
 * Each declaration names the reference interface it replaces
 * (mlsys-lab-sogang/TCA-SpMM, paths relative to its root).
 */
#ifndef TCA_SPMM_H
#define TCA_SPMM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCA_VERSION 101

typedef enum tca_status {
    TCA_OK = 0,
    TCA_ERR_INVALID_ARG = 1,   /* bad shape, NULL pointer, misaligned buffer   */
    TCA_ERR_UNSUPPORTED = 2,   /* valid but not implemented (dtype pair, N%8)   */
    TCA_ERR_HIP = 3,           /* a HIP runtime call failed                     */
    TCA_ERR_ALLOC = 4,         /* host or device allocation failed              */
    TCA_ERR_IO = 5,            /* file could not be opened / written            */
    TCA_ERR_FORMAT = 6,        /* malformed .smtx / inconsistent CSR            */
    TCA_ERR_LIBRARY = 7        /* rocSPARSE / hipBLASLt reported an error       */
} tca_status;

typedef enum tca_dtype {
    TCA_F16 = 0,   /* IEEE binary16 */
    TCA_BF16 = 1,  /* bfloat16 */
    TCA_E4M3 = 2,  /* OCP FP8 E4M3FN (gfx950 native) */
    TCA_F32 = 3    /* binary32 (dense operand only; converted on load) */
} tca_dtype;

const char *tca_status_string(tca_status s);
const char *tca_last_error(void);
int tca_version(void);
void tca_host_free(void *p);

/* ------------------------------------------------------------------------ */
/* .smtx I/O.  Replaces smtx_input_to_Dense (src/file_input.cu:135-193,      */
/* decl src/file_input.cuh:8).  Returns the pattern as CSR (malloc'ed,        */
/* release with tca_host_free) and rejects files the reference would overrun  */
/* (token counts, non-monotone offsets, out-of-range columns).                */
/* ------------------------------------------------------------------------ */
tca_status tca_smtx_read(const char *path, int *n_rows, int *n_cols, int *nnz,
                         int32_t **row_offset, int32_t **col_idx);
tca_status tca_smtx_write(const char *path, int n_rows, int n_cols, int nnz,
                          const int32_t *row_offset, const int32_t *col_idx);

/* MatrixMarket coordinate text.  Replaces mtx_input_to_Dense / mtx_input_to_CSR
 * (src/file_input.cu:8-134, flags from src/arg_parse.h:82-89): lines starting
 * with '%' are skipped, then "rows cols nnz", then one "r c [v]" per entry.
 * one_based subtracts 1 from r and c; pattern_only (or a missing value) gives
 * 1.0.  The CSR comes back with columns sorted per row; a repeated (r, c)
 * keeps the last value, as the dense reader's overwrite does.  values are
 * fp16 bits (round to nearest even).  All arrays malloc'ed. */
tca_status tca_mtx_read(const char *path, int one_based, int pattern_only, int *n_rows, int *n_cols, int *nnz,
                        int32_t **row_offset, int32_t **col_idx, uint16_t **values);

/* ------------------------------------------------------------------------ */
/* Dense <-> CSR (fp16 bits).  Replaces convert_dense_to_csr /               */
/* convert_csr_to_dense (src/miscalleneous.cu:29-65, :66-82).  With          */
/* col_idx == NULL only counts (row_offset may also be NULL).                */
/* ------------------------------------------------------------------------ */
tca_status tca_dense_to_csr_f16(const uint16_t *dense, int n_rows, int n_cols, int32_t *row_offset,
                                int32_t *col_idx, uint16_t *values, int capacity, int *nnz);
tca_status tca_csr_to_dense_f16(int n_rows, int n_cols, const int32_t *row_offset, const int32_t *col_idx,
                                const uint16_t *values, uint16_t *dense);

/* compare_mat<float> (src/general.cuh:71-123): 1 = pass.  *n_fail counts
 * failures up to the reference's early exit after the 10th. */
int tca_compare_mat_f32(const float *gold, const float *cmp, int n, float tol, int *n_fail);

/* ------------------------------------------------------------------------ */
/* Seeded synthetic inputs (counter-based SplitMix64 streams).  Replace       */
/* prepare_cpu_matrix (src/general.cuh:140-174), which is unseeded.           */
/* ------------------------------------------------------------------------ */
typedef enum tca_pattern_kind {
    TCA_PATTERN_BERNOULLI = 0,   /* each (r,c) kept with prob. density         */
    TCA_PATTERN_LOGNORMAL = 1,   /* row nnz ~ lognormal(sigma=param), rescaled */
    TCA_PATTERN_UNIFORM = 2      /* row nnz ~ U[(1-param),(1+param)] * mean    */
} tca_pattern_kind;
uint64_t tca_rng_u64(uint64_t seed, uint64_t stream, uint64_t index);
tca_status tca_gen_pattern(tca_pattern_kind kind, int n_rows, int n_cols, double density, double param,
                           uint64_t seed, uint64_t stream, int32_t *row_offset /* n_rows+1 */,
                           int32_t **col_idx /* malloc'ed */, int *nnz);
tca_status tca_gen_normal_f32(uint64_t seed, uint64_t stream, size_t n, double mean, double std, float *out);
tca_status tca_gen_uniform_f32(uint64_t seed, uint64_t stream, size_t n, double lo, double hi, float *out);
tca_status tca_f32_to_f16_host(const float *in, size_t n, uint16_t *out);   /* round to nearest even */
tca_status tca_f32_to_bf16_host(const float *in, size_t n, uint16_t *out);  /* round to nearest even */

/* ------------------------------------------------------------------------ */
/* Plan: the gfx950 panel tiling.  Replaces workload_manager                 */
/* (src/our_spmm.cu:347-471), the reference's row->warp packing.  Rows are   */
/* grouped `rows_per_wave` (16, 32 or 64, one wave per group; or 128 / 256,  */
/* one workgroup of rows/32 waves sharing each gathered B chunk) at a time;   */
/* each group's nonzero columns are condensed and cut into 32-wide k-chunks;  */
/* each (chunk, 16-row panel) lists its entries.  Values live outside the     */
/* plan, in "tiled order" */
/* (entry order); tca_pack_values converts from CSR order.  Built on the      */
/* host from host CSR arrays; device tables go to the current HIP device.    */
/* ------------------------------------------------------------------------ */
typedef struct tca_plan tca_plan;
typedef struct tca_plan_info {
    int n_rows, n_cols, nnz;
    int rows_per_wave;
    int n_groups;
    int n_chunks;
    int max_group_chunks;
    int transposed;            /* 1 for a plan made by tca_plan_create_transpose */
    size_t device_bytes;
    double build_ms;
    size_t dense_elems;        /* elements of the plan's panel-dense tiles (tca_pack_dense) */
} tca_plan_info;

tca_status tca_plan_create(int n_rows, int n_cols, int nnz, const int32_t *row_offset,
                           const int32_t *col_idx, int rows_per_wave, tca_plan **plan);
/* Same tables without device copies: for inspection (tca_plan_get_info,
 * tca_plan_export) on machines without a GPU.  Launch entry points reject it. */
tca_status tca_plan_create_host(int n_rows, int n_cols, int nnz, const int32_t *row_offset,
                                const int32_t *col_idx, int rows_per_wave, tca_plan **plan);
/* Plan of the transpose (n_cols x n_rows).  Its tiled values are read straight
 * from the forward plan's tiled values (indirect), so one value array serves
 * both directions. */
tca_status tca_plan_create_transpose(const tca_plan *fwd, int rows_per_wave, tca_plan **plan_t);
tca_status tca_plan_destroy(tca_plan *plan);
/* Transpose of a CSR pattern held on the device (its CSC), built on the device:
 * d_col_offset[n_cols+1], d_row_idx[nnz] (rows ascending within each column) and,
 * unless NULL, d_perm[nnz] = the CSR index of each CSC entry (the value
 * permutation).  Replaces the reference's host transpose of dense matrices
 * (src/miscalleneous.cu:146-157) for patterns that live on the GPU; allocates
 * stream-ordered scratch. */
tca_status tca_csr_transpose_device(int n_rows, int n_cols, int nnz, const int32_t *d_row_offset,
                                    const int32_t *d_col_idx, int32_t *d_col_offset, int32_t *d_row_idx,
                                    int32_t *d_perm, void *stream);

/* tca_csr_transpose_device without any allocation: the scratch (entry rows,
 * sort keys and values, radix-sort temp; about 16 bytes per entry plus the
 * sort's temp) lives in one fixed array of the library per GPU, of
 * tca_csr_transpose_noalloc_max_scratch() bytes (64 MiB, held until the process
 * ends); patterns that need more are rejected with TCA_ERR_UNSUPPORTED.  Same
 * arguments and results as tca_csr_transpose_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
tca_status tca_csr_transpose_device_noalloc(int n_rows, int n_cols, int nnz, const int32_t *d_row_offset,
                                            const int32_t *d_col_idx, int32_t *d_col_offset, int32_t *d_row_idx,
                                            int32_t *d_perm, void *stream);
size_t tca_csr_transpose_noalloc_max_scratch(void);

tca_status tca_plan_get_info(const tca_plan *plan, tca_plan_info *info);
/* Host copies of the tables (for the host-rebuild check).  Sizes:
 * grp_chunk_ptr[n_groups+1], chunk_cols[32*n_chunks], seg_ptr[n_chunks*rpw/16+1],
 * entry_slot[nnz], entry_src[nnz], work_order[n_groups].  NULL skips a table. */
tca_status tca_plan_export(const tca_plan *plan, int32_t *grp_chunk_ptr, int32_t *chunk_cols, int32_t *seg_ptr,
                           uint16_t *entry_slot, int32_t *entry_src, int32_t *work_order);

/* tiled[e] = csr[entry_src[e]] (and the inverse), elementwise on the device.
 * dtype selects the element size (F16/BF16: 2 B, E4M3: 1 B, F32: 4 B).  For a
 * transposed plan these use the forward plan's entry order. */
tca_status tca_pack_values(const tca_plan *plan, tca_dtype dtype, const void *d_csr, void *d_tiled, void *stream);
tca_status tca_unpack_values(const tca_plan *plan, tca_dtype dtype, const void *d_tiled, void *d_csr, void *stream);

/* Panel-dense tiles: every (chunk c, 16-row panel i) of the plan as a full
 * 16 x 32 tile t = c*(rows_per_wave/16) + i of info.dense_elems / 512 tiles,
 * element ((k/8)*16 + r)*8 + k%8 holding A[panel row r][chunk column k] (the
 * MFMA operand order: lane r + 16*(k/8) reads 8 consecutive k), zero where
 * the pattern has no entry.  Built from tiled values in the forward plan's
 * order (for a transposed plan: the forward plan's tiled values, placed
 * through its entry map).  F16/BF16 (1 KiB per tile) or E4M3 (512 B); d_dense
 * 16-byte aligned.  No counterpart in the reference, whose kernel reads CSR
 * values per nonzero (src/our_spmm.cu:290-346): the dense tiles trade a few
 * more bytes (a tile is 1/16 of the B rows a chunk gathers at N = 256) for a
 * chunk loop without per-entry work. */
tca_status tca_pack_dense(const tca_plan *plan, tca_dtype dtype, const void *d_tiled, void *d_dense, void *stream);

/* ------------------------------------------------------------------------ */
/* SpMM: C[M x N] (f32, row-major, ldc) = A (plan + tiled values) * B.       */
/* Replaces kernel_our_spmm_balanced (src/our_spmm.cu:290-346).              */
/*   a_dtype in {F16, BF16, E4M3}; b_dtype == a_dtype, or F32 with a_dtype in */
/*   {F16, BF16} (converted on load, round to nearest even).                 */
/*   B row-major K x N with ldb; N % 8 == 0; B, C 16-byte aligned, ldb and    */
/*   ldc multiples of 8 (elements).  Every row of C in [0, M) is written.    */
/*   row_scale[M] / col_scale[N] (may be NULL) multiply the FP32 result:     */
/*   C[r,n] = row_scale[r] * col_scale[n] * sum_k A[r,k] B[k,n] (E4M3).      */
/*   tile_n in {16, 32, 64, 128}, or 0 to pick from the built-in table.      */
/*   Plans of 128 / 256 rows per group: tile_n 64 or 128, F16/BF16 A only   */
/*   (tca_spmm_dense also takes E4M3 there).                                 */
/* ------------------------------------------------------------------------ */
tca_status tca_spmm(const tca_plan *plan, tca_dtype a_dtype, const void *d_a_tiled,
                    tca_dtype b_dtype, const void *d_B, int ldb, int N,
                    float *d_C, int ldc, const float *d_row_scale, const float *d_col_scale,
                    int tile_n, void *stream);
/* tca_spmm with A as panel-dense tiles (tca_pack_dense), forward or
 * transposed plan; same arguments.  Plans of 16 / 32 / 64 rows give the
 * results of tca_spmm bit for bit (the MFMAs see the same operands in the same
 * order); plans of 128 / 256 rows (F16 / BF16 A with an F16 / BF16 / F32 B, or
 * E4M3 A with an E4M3 B, tile_n 64 or 128) run the block kernels, whose K
 * groups (two per workgroup when the launch has fewer than 512 group x N-tile
 * units, else one) add their partial sums in a fixed order: the same results
 * to FP32 summation order, the same bits from launch to launch. */
tca_status tca_spmm_dense(const tca_plan *plan, tca_dtype a_dtype, const void *d_a_dense,
                          tca_dtype b_dtype, const void *d_B, int ldb, int N,
                          float *d_C, int ldc, const float *d_row_scale, const float *d_col_scale,
                          int tile_n, void *stream);

/* ------------------------------------------------------------------------ */
/* Blocked-ELL.  Replaces perform_cuSPARSE_BlockSpMM (src/spmm_cusparse.cu:   */
/* 181-445, decl src/spmm_cusparse.cuh:14): its host conversion (:254-331)   */
/* and the cusparseCreateBlockedEll + cusparseSpMM pair (:337-343, :394-398). */
/* Format: ceil(M/block) block rows of slots = ell_cols/block block slots;    */
/* col_ind[br*slots + j] = block column of slot j, or -1 (empty: its values  */
/* are not read); values (ceil(M/block)*block) x ell_cols row-major, row r,   */
/* slot j, column c of the block at values[r*ell_cols + j*block + c].  Values */
/* at columns >= n_cols of the last block column must be finite (zeros).      */
/* ------------------------------------------------------------------------ */
/* Host conversion from CSR (values 16-bit, F16 or BF16 bits; NULL = pattern
 * only, 1.0 at every entry like the .smtx reader).  A block is kept when it
 * holds a CSR entry; kept block columns ascending, -1 after the last.  Call
 * with col_ind / ell_values NULL to get ell_cols, then with buffers of
 * ceil(M/block)*(ell_cols/block) and ceil(M/block)*block*ell_cols elements.
 * block in {16, 32, 64, 128, 256}. */
tca_status tca_bell_from_csr(int n_rows, int n_cols, const int32_t *row_offset, const int32_t *col_idx,
                             const uint16_t *values, int block, int *ell_cols, int32_t *col_ind,
                             uint16_t *ell_values);
/* C[M x N] (f32, every row written) = A (Blocked-ELL, device arrays) * B (K x N
 * row-major, F16/BF16 like the values, or F32 rounded on load).  N % 8 == 0;
 * values, B, C 16-byte aligned, ldb % 8 == 0, ldc % 4 == 0.  tile_n (0 = auto,
 * 16/32/64) and rows_per_wave (0 = auto, 16/32/64, at most block) fix the
 * kernel's work unit. */
tca_status tca_bell_spmm(int n_rows, int n_cols, int block, int ell_cols, const int32_t *d_col_ind,
                         tca_dtype a_dtype, const void *d_values, tca_dtype b_dtype, const void *d_B, int ldb,
                         int N, float *d_C, int ldc, int tile_n, int rows_per_wave, void *stream);
/* The (rows_per_wave, tile_n) tca_bell_spmm (and tca_spmm_24, block 32) picks
 * for an n_rows x n_cols matrix times N columns when given 0s. */
tca_status tca_bell_shape(int n_rows, int n_cols, int N, int block, int tile_n, int *rows_per_wave, int *tile);

/* ------------------------------------------------------------------------ */
/* 2:4 structured sparsity on gfx950's sparse MFMA.  Replaces                 */
/* perform_cuSPARSELT_24sparsity (src/spmm_cusparse.cu:2-179, decl            */
/* src/spmm_cusparse.cuh:13): prune + compress (cusparseLtSpMMAPrune STRIP,   */
/* cusparseLtSpMMACompress, :95-136) and the product (cusparseLtMatmul,       */
/* :145-148; FP32 C here instead of the reference's FP16).  A different       */
/* product from tca_spmm's: A loses entries.                                 */
/* ------------------------------------------------------------------------ */
/* Prune a dense row-major M x K (ld) F16/BF16 matrix with the STRIP rule (per
 * 1x4 strip of a row the pair of largest L1 norm survives; ties to the first
 * pair of (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)) and compress it: values M x K/2
 * (the survivors in column order), meta M x K/16 uint16 (value j of a
 * 16-column block has its column-in-strip at bits 2j+1:2j).  K % 64 == 0. */
tca_status tca_compress_24(int n_rows, int n_cols, tca_dtype dtype, const void *d_dense, int ld,
                           void *d_values, uint16_t *d_meta, void *stream);
/* C[M x N] (f32, every row written) = compressed 2:4 A (M x K) * B (K x N
 * row-major, the values' dtype or F32 rounded on load).  K % 64 == 0, N % 8 == 0;
 * values, B, C 16-byte aligned.  tile_n / rows_per_wave: 0 = auto, 16/32/64. */
tca_status tca_spmm_24(int n_rows, int n_cols, int N, tca_dtype a_dtype, const void *d_values,
                       const uint16_t *d_meta, tca_dtype b_dtype, const void *d_B, int ldb, float *d_C,
                       int ldc, int tile_n, int rows_per_wave, void *stream);

/* SDDMM on the plan's pattern: out_tiled[e] = sum_n G[row(e), n] * X[col(e), n]
 * (the dW of Y = W X with dY = G).  G: M x N row-major (ldg), X: K x N (ldx),
 * dtypes F16/BF16, or F32 (G) rounded to X's type on load.  N % 32 == 0.
 * Runs on a 64-row tiling of the same pattern, built on the first call and
 * kept with the plan; out_tiled is in the plan's own tiled order. */
tca_status tca_sddmm(const tca_plan *plan, tca_dtype g_dtype, const void *d_G, int ldg,
                     tca_dtype x_dtype, const void *d_X, int ldx, int N, float *d_out_tiled, void *stream);

/* ------------------------------------------------------------------------ */
/* FP8 E4M3 quantisation with FP32 scales (scale = amax / 448).               */
/* ------------------------------------------------------------------------ */
/* Per-row scales of a CSR matrix (values in CSR order, fp16 bits). */
tca_status tca_quantize_rows_e4m3(int n_rows, const int32_t *d_row_offset, const uint16_t *d_values,
                                  uint8_t *d_q, float *d_row_scale, void *stream);
/* Per-column scales of a dense K x N fp16 matrix. */
tca_status tca_quantize_cols_e4m3(int K, int N, const uint16_t *d_B, int ldb, uint8_t *d_q, int ldq,
                                  float *d_col_scale, void *stream);

/* Plain SGD step on weight values held in any order (tiled or CSR):
 * w[i] = round_to_nearest_even(w[i] - lr * g[i]) for i < n, w F16 or BF16, g FP32
 * in the same order (e.g. the dW a batch's SDDMM item writes).  w and g 16-byte
 * aligned.  No counterpart in the reference (it has no training); a
 * per-call training loop updates all layers' weights (one buffer) in one
 * launch with it; a batch can instead fuse the update into its SDDMM items
 * (tca_batch_item.sgd_w). */
tca_status tca_sgd_update(tca_dtype w_dtype, void *d_w, const float *d_g, size_t n, float lr, void *stream);

/* ------------------------------------------------------------------------ */
/* Batched persistent launch over many layers (one kernel, capturable).      */
/* SPMM items: F16 values; forward with an F16 B on 16-row plans (tile_n     */
/* 16/32/64) or 32-row plans (tile_n 32); transposed with an F32 or F16 B on  */
/* 16-row plans (tile_n 16/32/64); panel-dense F16 A (a_dense), forward or  */
/* transposed, with an F16 B on 16-row (tile_n 16/32/64), 32-row (32/64),    */
/* 64-row (32/64) or 128-row (64) plans, or an F32 B on 16-row (16/32/64) or  */
/* 32-row (32) plans.                                                         */
/* SDDMM items: F16 X, F32 or F16 G.                                          */
/* BF16: the panel-dense SPMM items above with BF16 A (and a BF16 or F32 B),  */
/* SDDMM items with BF16 X (and a BF16 or F32 G); a batch takes F16 or BF16   */
/* items, not both.  Other combinations: UNSUPPORTED.                         */
/* An F32 operand (SpMM B, SDDMM G) gets a batch-owned 16-bit copy (2 bytes  */
/* per element, in the type its readers round it to: FP16 or BF16) that every */
/* run refreshes first -- same bits as rounding on load -- and the items read */
/* the copy: half the gather bytes.  The copy is per operand: one that the    */
/* items read with different widths, leading dimensions or 16-bit types, and */
/* any beyond the 256th, is read as F32 and rounded on load (same results).  */
/* a_dense = 2 (SPMM, F16/BF16): `a` holds the weight's VALUES, in CSR order */
/* of the forward matrix (values_order NULL) or in the tiled order of          */
/* values_order (a forward plan of the same pattern, e.g. the SDDMM item's    */
/* plan, so that dW and the weight share one order); the batch owns the        */
/* item's panel-dense tiles and rebuilds them from those values at the start */
/* of every run.  A training step that updates the weight in place then needs */
/* no separate repack: every run = one prologue launch (FP16 copies + dense   */
/* tiles of all such items) + the persistent launch.                          */
/* ------------------------------------------------------------------------ */
typedef enum tca_op { TCA_OP_SPMM = 0, TCA_OP_SDDMM = 1 } tca_op;
typedef struct tca_batch_item {
    tca_op op;
    const tca_plan *plan;
    int tile_n;            /* SPMM: 16/32/64/128 (0 = table); SDDMM: ignored */
    tca_dtype a_dtype;     /* SPMM: tiled values dtype; SDDMM: dtype of G    */
    const void *a;         /* SPMM: tiled values;       SDDMM: G             */
    int lda;               /* SDDMM: ldg                                     */
    tca_dtype b_dtype;     /* dense operand dtype                            */
    const void *b;         /* SPMM: B (K x N);          SDDMM: X (K x N)     */
    int ldb;
    int N;
    float *out;            /* SPMM: C (M x N, ldc);     SDDMM: out_tiled     */
    int ldc;
    int a_dense;           /* SPMM: 1 = `a` holds panel-dense tiles (tca_pack_dense);
                              2 = `a` holds values the batch packs every run (above) */
    const tca_plan *values_order;   /* a_dense = 2: order of `a` (NULL = CSR order) */
    /* SDDMM items: optional SGD step fused into the sampling (no counterpart in
       the reference; for training loops that would otherwise launch
       tca_sgd_update after the batch).  When sgd_w is not NULL, every entry
       the item writes to out[i] also updates sgd_w[i] <- round(fma(-sgd_lr,
       out[i], sgd_w[i])) (sgd_w_dtype F16 or BF16, same order and length as
       out), the same arithmetic as tca_sgd_update.  The item's own reads (G,
       X) do not involve sgd_w, so the update may target weights whose dense
       tiles the same batch packed in its prologue (a_dense = 2); an SPMM item
       of the batch that reads overlapping values directly (a_dense = 0), or a
       second SDDMM item updating overlapping weights, is rejected with
       TCA_ERR_INVALID_ARG. */
    void *sgd_w;
    tca_dtype sgd_w_dtype;
    float sgd_lr;
} tca_batch_item;
typedef struct tca_batch tca_batch;
tca_status tca_batch_create(int n_items, const tca_batch_item *items, tca_batch **batch);
/* sizeof(tca_batch_item) as this library was built: bindings that restate the
 * struct (ctypes) compare it at load time. */
size_t tca_batch_item_size(void);
tca_status tca_batch_run(tca_batch *batch, void *stream);
/* One stage of tca_batch_run: 1 = the prologue (FP16 copies of F32 operands and
 * dense tiles of a_dense = 2 items; nothing is launched when the batch has
 * neither), 2 = the persistent launch over the work units, 3 = both.  Lets a
 * caller time the persistent kernel alone. */
tca_status tca_batch_run_stage(tca_batch *batch, int stage, void *stream);
tca_status tca_batch_destroy(tca_batch *batch);
/* Work units of a batch and the number of workgroups its launch uses. */
tca_status tca_batch_get_info(const tca_batch *batch, int *n_units, int *n_workgroups);

/* ------------------------------------------------------------------------ */
/* Library baselines on device pointers (the reference times cuSPARSE and     */
/* cuBLAS GemmEx: src/spmm_cusparse.cu:446-579, src/gemm_cublas.cu:3-111).    */
/* ------------------------------------------------------------------------ */
/* rocSPARSE CSR SpMM, f16 A/B -> f32 C.  ms_per_iter = mean over n_iter after
 * warm-up (pass n_iter = 0 for a single untimed call). */
tca_status tca_rocsparse_spmm(int M, int N, int K, int nnz, const int32_t *d_row_offset, const int32_t *d_col_idx,
                              const uint16_t *d_values, const uint16_t *d_B, float *d_C, int n_iter,
                              double *ms_per_iter);
/* The transposed product C (K x N) = A^T B, B M x N: the library counterpart of
 * the dX launch on a transposed plan (no counterpart in the reference). */
tca_status tca_rocsparse_spmm_t(int M, int N, int K, int nnz, const int32_t *d_row_offset, const int32_t *d_col_idx,
                                const uint16_t *d_values, const uint16_t *d_B, float *d_C, int n_iter,
                                double *ms_per_iter);
/* hipBLASLt dense GEMM C = A_dense * B, f16 in, f32 out (row-major). */
tca_status tca_hipblaslt_gemm(int M, int N, int K, const uint16_t *d_A, const uint16_t *d_B, float *d_C,
                              int n_iter, double *ms_per_iter);
/* rocSPARSE SDDMM (the library counterpart of tca_sddmm): out (CSR order, f32)
 * = (G X^T) on the pattern, G M x N and X K x N row-major f16, FP32 compute.
 * No counterpart in the reference (it has no backward); ms_per_iter as above. */
tca_status tca_rocsparse_sddmm(int M, int K, int N, int nnz, const int32_t *d_row_offset, const int32_t *d_col_idx,
                               const uint16_t *d_G, const uint16_t *d_X, float *d_out, int n_iter,
                               double *ms_per_iter);
/* 2:4 structured sparsity on hipSPARSELt.  Replaces perform_cuSPARSELT_24sparsity
 * (src/spmm_cusparse.cu:2-179, decl src/spmm_cusparse.cuh:13): A (dense row-major
 * M x K, F16) pruned with the STRIP rule (per 1x4 strip of a row the pair of
 * largest L1 norm survives) into d_A_pruned (M x K, may be NULL), compressed,
 * and C (row-major M x N, F16, FP32 accumulation) = pruned A * B (K x N, F16).
 * A different product from tca_spmm's (A loses entries): a speed comparison.
 * M, N, K multiples of 8.  TCA_ERR_UNSUPPORTED when the process has bound a
 * hipSPARSELt other than the one the library was built against (PyTorch bundles
 * an older copy; use a process without it, e.g. tca_test --sparse24). */
tca_status tca_hipsparselt_spmm_24(int M, int N, int K, const uint16_t *d_A, uint16_t *d_A_pruned,
                                   const uint16_t *d_B, uint16_t *d_C, int n_iter, double *ms_per_iter);

#ifdef __cplusplus
}
#endif
#endif /* TCA_SPMM_H */
