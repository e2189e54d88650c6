/*
 * tca_spmm.hpp -- C++ drop-in API with the reference's names, argument order
 * and error behaviour (print + exit(EXIT_FAILURE) on a device/library error,
 * like CHECK_CUDA in src/general.cuh:8-16).  Implemented on the C ABI in
 * tca_spmm.h; `half` is HIP's __half.
 */
#ifndef TCA_SPMM_HPP
#define TCA_SPMM_HPP
#include <hip/hip_fp16.h>
#include <fstream>
#include "tca_spmm.h"

#define DEFAULT_TOL 1e-3 /* src/general.cuh:24 */

/* src/general.cuh:125-134 */
typedef struct CSR {
    int nRow;
    int nCol;
    int nnz;
    int *row_offset;
    int *col_idx;
    half *value;
} CSR;

/* src/file_input.cu:135 -- dense fp16 M x K with 1.0 at every pattern entry */
void smtx_input_to_Dense(std::ifstream &fin, half **mat, char *filename, int *nRow, int *nCol, int *nnz);
/* src/miscalleneous.cu:29 / :66 / :83 */
void convert_dense_to_csr(half *mat, CSR **descr, int nRow, int nCol);
void convert_csr_to_dense(CSR *descr, half **mat);
void destroy_csr(CSR *descr);
/* src/general.cuh:71 -- instantiated for float and half */
template <typename T> bool compare_mat(T *gold, T *comparing, int num_elem, float tolerance);
/* src/our_spmm.cuh:4 -- C (host, M x N, f32) is in/out: rows of A without
 * nonzeros keep the caller's values; *elapsed = mean ms per launch over n_iter
 * launches after 5 warm-ups. */
void our_spmm_balanced(CSR *A_csr, half *B, float *C, const int M, const int N, const int K, int n_iter, double *elapsed);
/* src/spmm_cusparse.cuh:15 and src/gemm_cublas.cuh:12, on rocSPARSE / hipBLASLt */
void perform_rocSPARSE_SpMM(half *A, half *B, float *C, int M, int N, int K, int num_iter, double *elapsed);
void perform_hipBLASLt_GemmEx(half *A, half *B, float *C, int M, int N, int K, int n_iter, double *elapsed);
/* src/spmm_cusparse.cuh:14 (perform_cuSPARSE_BlockSpMM): dense A -> Blocked-ELL
 * with blocks of ell_block_size (16/32/64/128/256), then the gfx950 Blocked-ELL
 * kernel (tca_bell_spmm); C (M x N) is overwritten. */
void perform_BlockSpMM(half *A, half *B, float *C, int M, int N, int K, int num_iter, int ell_block_size, double *elapsed);
/* src/spmm_cusparse.cuh:13 (perform_cuSPARSELT_24sparsity) on hipSPARSELt: A pruned to
 * 2:4 (STRIP), compressed, multiplied; the F16 result is widened into C. */
void perform_hipSPARSELt_24sparsity(half *A, half *B, float *C, int M, int N, int K, int num_iter, double *elapsed);
/* src/miscalleneous.cu:158-163 (warm_up_gpu_kernel<<<1,32>>>) */
void warm_up_gpu();

#endif
