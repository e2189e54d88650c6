/*
 * oracle.h -- CPU restatement of the TCA-SpMM row-balanced CSR x dense SpMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, calls or
 * loads this code; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it, and only as the checker.  Each function cites the
 * reference file:line whose behaviour it restates (paths relative to the
 * reference repository root).
 *
 * Pinning (see tests/test_oracle.py):
 *   orc_smtx_to_dense     -- bundled .smtx invariants (header 2048,512,1005; 28
 *                            non-empty rows; sorted distinct columns in [2,511]),
 *                            writer->reader round trips, malformed-file errors.
 *   orc_dense_to_csr /
 *   orc_csr_to_dense      -- round-trip identity, brute force on tiny inputs,
 *                            +-0 dropping, NaN kept (x != 0.0f is true for NaN).
 *   orc_compare_mat       -- crafted vectors for each branch of the rule.
 *   orc_spmm_f32 / _f64   -- closed forms: single-nnz rows (exact), identity
 *                            pattern (C == B), all-ones values (row sums), dense
 *                            numpy matmul on small shapes, f32 vs f64 bound.
 *   orc_sddmm_f64         -- dense numpy G @ X.T sampled on the pattern.
 *   orc_sddmm_f32         -- exact on small-integer operands (== numpy), a
 *                            dropped or transposed term fails; FP32 vs FP64 bound.
 *   orc_workload_manager  -- known answer on the bundled file at N=256:
 *                            24 blocks, 2 idle warp slots, 0 duplicates.
 *   orc_tiling_*          -- CSR reconstruction identity + brute-force union.
 *   orc_*_e4m3 / _bf16    -- exhaustive decode tables vs the format definitions;
 *                            encode == nearest-even over all decoded values.
 *   orc_dense_to_bell     -- hand-worked known answer, BELL -> dense identity,
 *                            slot count == brute-force maximum, sorted columns.
 *   orc_bell_spmm_f32     -- bit-equal to orc_spmm_f32 on the same matrix (same
 *                            ascending-column FP32 order); -1 slots anywhere.
 *   orc_prune_24_strip    -- hand-worked strips, brute force over all pairs on
 *                            tie-free rows, idempotence, <= 2 survivors per strip.
 */
#ifndef TCA_ORACLE_H
#define TCA_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- scalar formats ---------------------------------------------------- */
float    orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_f16(float f);          /* round to nearest even, IEEE binary16 */
float    orc_bf16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float f);         /* round to nearest even */
float    orc_e4m3_to_f32(uint8_t v);       /* OCP FP8 E4M3FN (bias 7, no inf, NaN=0x7f/0xff) */
uint8_t  orc_f32_to_e4m3(float f);         /* nearest even, saturate to +-448, NaN -> 0x7f */

/* ---- .smtx reader: src/file_input.cu:135-193 ----------------------------- */
/* Returns 0 on success, <0 on malformed input.  *dense is malloc'ed M*K fp16 with
 * 1.0 (0x3C00) at every pattern position (src/file_input.cu:189). */
int orc_smtx_to_dense(const char *path, int *M, int *K, int *nnz, uint16_t **dense);

/* ---- MatrixMarket reader: mtx_input_to_Dense, src/file_input.cu:8-61 ------ */
/* '%' lines skipped, header "rows cols nnz", entries "r c [v]" (v defaults to
 * 1.0), minus one when one_based; later duplicates overwrite.  Returns 0, or
 * <0 on a malformed file (the reference asserts on the count). */
int orc_mtx_to_dense(const char *path, int one_based, int *M, int *K, int *nnz, uint16_t **dense);

/* ---- dense <-> CSR: src/miscalleneous.cu:29-65, :66-82 -------------------- */
/* Counts entries with (float)x != 0.0f; if colidx/vals are non-NULL fills them
 * (row-major scan, sorted columns).  Returns nnz. */
int  orc_dense_to_csr(const uint16_t *dense, int M, int K, int *rowptr, int *colidx, uint16_t *vals);
void orc_csr_to_dense(int M, int K, const int *rowptr, const int *colidx, const uint16_t *vals, uint16_t *dense);

/* ---- compare_mat<float>: src/general.cuh:71-123 --------------------------- */
/* Returns 1 (success) / 0 (failure); *n_fail = failures seen before the
 * reference's early return after the 10th. */
int orc_compare_mat(const float *gold, const float *cmp, int n, float tol, int *n_fail);

/* ---- CPU SpMM, FP32 accumulate: naive_gemm semantics (src/miscalleneous.cu:114-134)
 * restricted to the nonzeros, ascending column order.  A, B are fp16 bits. */
void orc_spmm_f32(int M, int N, const int *rowptr, const int *colidx, const uint16_t *vals,
                  const uint16_t *B, int ldb, float *C, int ldc, int nthreads);
/* FP64 oracle on already-decoded values (exact decodes of fp16/bf16/fp8). */
void orc_spmm_f64(int M, int N, const int *rowptr, const int *colidx, const double *vals,
                  const double *B, int ldb, double *C, int ldc, int nthreads);
/* Sampled dense-dense product on the CSR pattern: out[e] = sum_n G[i,n] X[j,n]. */
void orc_sddmm_f64(int M, int N, const int *rowptr, const int *colidx,
                   const double *G, int ldg, const double *X, int ldx, double *out, int nthreads);
/* The same on fp16 bits with FP32 accumulation in ascending n (CPU baseline of dW). */
void orc_sddmm_f32(int M, int N, const int *rowptr, const int *colidx, const uint16_t *G, int ldg,
                   const uint16_t *X, int ldx, float *out, int nthreads);

/* ---- reference row->warp packing: src/our_spmm.cu:347-471 ---------------- */
/* warps_per_row[M]; row_idx/warpid capacity `cap` entries.  Returns the number
 * of warp slots written (num_blocks * warps_per_tb) or -1 (bad warps/TB or cap). */
int orc_workload_manager(int M, const int *rowptr, int N, int warps_per_tb, int *num_blocks,
                         int *warps_per_row, int *row_idx, int *warpid, int cap);

/* ---- independent rebuild of the gfx950 panel tiling ---------------------- */
/* Groups of `rpw` consecutive rows (rpw in {16,32,64}); per group the sorted
 * union of nonzero columns, cut into 32-wide chunks (pad -1); per (chunk,
 * 16-row panel) the entries ordered by (row, column), slot = r16*32 + k. */
int orc_tiling_sizes(int M, int K, const int *rowptr, const int *colidx, int rpw,
                     int *n_groups, int *n_chunks);
int orc_tiling_fill(int M, int K, const int *rowptr, const int *colidx, int rpw,
                    int *grp_chunk_ptr, int *chunk_cols, int *seg_ptr,
                    uint16_t *entry_slot, int *entry_src, int *work_order);

/* CSR transpose (CSC): rowptrT[K+1], colidxT[nnz], perm[nnz] = source index. */
void orc_transpose_csr(int M, int K, const int *rowptr, const int *colidx,
                       int *rowptrT, int *colidxT, int *perm);

/* ---- Blocked-ELL: the host conversion of perform_cuSPARSE_BlockSpMM,
 * src/spmm_cusparse.cu:254-331, and the product cusparseSpMM computes on it
 * (:394-398).  Block rows = ceil(M/b); block (br, bc) is kept when one of its
 * entries has (float)x != 0 (:258-266); slots per block row = the largest kept
 * count (:281-291); col_ind[br*slots + j] = the j-th kept block column of block
 * row br, ascending, -1 after the last (:293-311); values: (ceil(M/b)*b) x
 * (slots*b) row-major, values[r*ell_cols + j*b + c] = A[r][col_ind*b + c], zero
 * for -1 slots and past row M / column K (:313-331 read past K instead when K
 * is not a multiple of b; the math is what is restated here).  Returns
 * ell_cols = slots*b; col_ind / values may be NULL to size them. */
int orc_dense_to_bell(const uint16_t *dense, int M, int K, int b, int32_t *col_ind, uint16_t *values);
/* C[M x N] (f32) = BELL(A) * B: FP32 sums in ascending (slot, column-in-block)
 * order, -1 slots skipped, B rows at or past K read as zero. */
void orc_bell_spmm_f32(int M, int K, int N, int b, int ell_cols, const int32_t *col_ind,
                       const uint16_t *values, const uint16_t *B, float *C);

/* ---- 2:4 pruning: cusparseLtSpMMAPrune(..., CUSPARSELT_PRUNE_SPMMA_STRIP) as
 * perform_cuSPARSELT_24sparsity calls it (src/spmm_cusparse.cu:95-96), restated
 * from the library's published rule (cuSPARSELt is not in the reference; the
 * hipSPARSELt header words it the same way): in every 1x4 strip of a row the
 * pair of largest L1 norm survives, the other two become +0.  The order in
 * which the library breaks ties is not published: here the first pair in
 * (0,1) (0,2) (0,3) (1,2) (1,3) (2,3) order wins (parity unpinned on ties).
 * K % 4 == 0; returns -1 otherwise. */
int orc_prune_24_strip(const uint16_t *A, int M, int K, uint16_t *out);

void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
