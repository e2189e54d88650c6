// Library baselines on device pointers: rocSPARSE CSR SpMM and hipBLASLt dense
// GEMM, f16 inputs -> f32 output.  They take the place of the reference's
// cuSPARSE SpMM (src/spmm_cusparse.cu:446-579) and cuBLAS GemmEx
// (src/gemm_cublas.cu:3-111) comparisons: same warm-up / timed-loop shape.
#include "tca_internal.hpp"

#include <hipblaslt/hipblaslt.h>
#include <rocsparse/rocsparse.h>

using tca::fail;

namespace {

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

#define TCA_SPARSE_TRY(expr)                                                                          \
    do {                                                                                              \
        rocsparse_status _s = (expr);                                                                 \
        if (_s != rocsparse_status_success) {                                                         \
            st = fail(TCA_ERR_LIBRARY, std::string(#expr) + " -> rocsparse status " + std::to_string((int)_s)); \
            goto done;                                                                                \
        }                                                                                             \
    } while (0)

static tca_status rocsparse_spmm_op(rocsparse_operation op, int M, int N, int K, int nnz, const int32_t *d_row_offset,
                                    const int32_t *d_col_idx, const uint16_t *d_values, const uint16_t *d_B, float *d_C,
                                    int n_iter, double *ms_per_iter)
{
    if (M <= 0 || N <= 0 || K <= 0 || nnz < 0 || !d_row_offset || !d_B || !d_C)
        return fail(TCA_ERR_INVALID_ARG, "tca_rocsparse_spmm: bad argument");
    tca_status st = TCA_OK;
    rocsparse_handle h = nullptr;
    rocsparse_spmat_descr A = nullptr;
    rocsparse_dnmat_descr B = nullptr, C = nullptr;
    void *buf = nullptr;
    size_t bs = 0;
    const float alpha = 1.0f, beta = 0.0f;
    EventPair ev;
    TCA_SPARSE_TRY(rocsparse_create_handle(&h));
    TCA_SPARSE_TRY(rocsparse_create_csr_descr(&A, M, K, nnz, const_cast<int32_t *>(d_row_offset),
                                              const_cast<int32_t *>(d_col_idx), const_cast<uint16_t *>(d_values),
                                              rocsparse_indextype_i32, rocsparse_indextype_i32,
                                              rocsparse_index_base_zero, rocsparse_datatype_f16_r));
    {
        const int rb = op == rocsparse_operation_none ? K : M, rc = op == rocsparse_operation_none ? M : K;
        TCA_SPARSE_TRY(rocsparse_create_dnmat_descr(&B, rb, N, N, const_cast<uint16_t *>(d_B), rocsparse_datatype_f16_r,
                                                    rocsparse_order_row));
        TCA_SPARSE_TRY(rocsparse_create_dnmat_descr(&C, rc, N, N, d_C, rocsparse_datatype_f32_r, rocsparse_order_row));
    }
    TCA_SPARSE_TRY(rocsparse_spmm(h, op, rocsparse_operation_none, &alpha, A, B, &beta, C,
                                  rocsparse_datatype_f32_r, rocsparse_spmm_alg_default,
                                  rocsparse_spmm_stage_buffer_size, &bs, nullptr));
    if (bs && hipMalloc(&buf, bs) != hipSuccess) { st = fail(TCA_ERR_ALLOC, "rocsparse buffer"); goto done; }
    TCA_SPARSE_TRY(rocsparse_spmm(h, op, rocsparse_operation_none, &alpha, A, B, &beta, C,
                                  rocsparse_datatype_f32_r, rocsparse_spmm_alg_default,
                                  rocsparse_spmm_stage_preprocess, &bs, buf));
    {
        const int warm = n_iter > 0 ? 10 : 1;  // the reference warms cuSPARSE up 10x
        for (int i = 0; i < warm; i++)
            TCA_SPARSE_TRY(rocsparse_spmm(h, op, rocsparse_operation_none, &alpha, A, B, &beta,
                                          C, rocsparse_datatype_f32_r, rocsparse_spmm_alg_default,
                                          rocsparse_spmm_stage_compute, &bs, buf));
        if (n_iter > 0) {
            if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) { st = fail(TCA_ERR_HIP, "event"); goto done; }
            (void)hipEventRecord(ev.a, nullptr);
            for (int i = 0; i < n_iter; i++)
                TCA_SPARSE_TRY(rocsparse_spmm(h, op, rocsparse_operation_none, &alpha, A, B,
                                              &beta, C, rocsparse_datatype_f32_r, rocsparse_spmm_alg_default,
                                              rocsparse_spmm_stage_compute, &bs, buf));
            (void)hipEventRecord(ev.b, nullptr);
            (void)hipEventSynchronize(ev.b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev.a, ev.b);
            if (ms_per_iter) *ms_per_iter = ms / n_iter;
        }
        if (hipDeviceSynchronize() != hipSuccess) st = fail(TCA_ERR_HIP, "rocsparse: synchronize");
    }
done:
    if (buf) (void)hipFree(buf);
    if (A) rocsparse_destroy_spmat_descr(A);
    if (B) rocsparse_destroy_dnmat_descr(B);
    if (C) rocsparse_destroy_dnmat_descr(C);
    if (h) rocsparse_destroy_handle(h);
    return st;
}

extern "C" tca_status tca_rocsparse_spmm(int M, int N, int K, int nnz, const int32_t *d_row_offset,
                                         const int32_t *d_col_idx, const uint16_t *d_values, const uint16_t *d_B,
                                         float *d_C, int n_iter, double *ms_per_iter)
{
    return rocsparse_spmm_op(rocsparse_operation_none, M, N, K, nnz, d_row_offset, d_col_idx, d_values, d_B, d_C,
                             n_iter, ms_per_iter);
}

// The transposed product C (K x N) = A^T B (B: M x N), A (M x K) in CSR: the
// library counterpart of the dX launch on a transposed plan.
extern "C" tca_status tca_rocsparse_spmm_t(int M, int N, int K, int nnz, const int32_t *d_row_offset,
                                           const int32_t *d_col_idx, const uint16_t *d_values, const uint16_t *d_B,
                                           float *d_C, int n_iter, double *ms_per_iter)
{
    return rocsparse_spmm_op(rocsparse_operation_transpose, M, N, K, nnz, d_row_offset, d_col_idx, d_values, d_B, d_C,
                             n_iter, ms_per_iter);
}

#define TCA_LT_TRY(expr)                                                                              \
    do {                                                                                              \
        hipblasStatus_t _s = (expr);                                                                  \
        if (_s != HIPBLAS_STATUS_SUCCESS) {                                                           \
            st = fail(TCA_ERR_LIBRARY, std::string(#expr) + " -> hipblas status " + std::to_string((int)_s)); \
            goto done;                                                                                \
        }                                                                                             \
    } while (0)

// Row-major C[MxN] = A[MxK] B[KxN] computed as column-major C^T = B^T A^T
// (the same transpose trick as src/gemm_cublas.cu:76-83).
extern "C" tca_status tca_hipblaslt_gemm(int M, int N, int K, const uint16_t *d_A, const uint16_t *d_B, float *d_C,
                                         int n_iter, double *ms_per_iter)
{
    if (M <= 0 || N <= 0 || K <= 0 || !d_A || !d_B || !d_C) return fail(TCA_ERR_INVALID_ARG, "tca_hipblaslt_gemm: bad argument");
    tca_status st = TCA_OK;
    hipblasLtHandle_t h = nullptr;
    hipblasLtMatmulDesc_t desc = nullptr;
    hipblasLtMatrixLayout_t la = nullptr, lb = nullptr, lc = nullptr;
    hipblasLtMatmulPreference_t pref = nullptr;
    hipblasLtMatmulHeuristicResult_t heur[1];
    int found = 0;
    void *ws = nullptr;
    uint64_t ws_max = 64ull << 20;
    const float alpha = 1.0f, beta = 0.0f;
    EventPair ev;
    TCA_LT_TRY(hipblasLtCreate(&h));
    TCA_LT_TRY(hipblasLtMatmulDescCreate(&desc, HIPBLAS_COMPUTE_32F, HIP_R_32F));
    TCA_LT_TRY(hipblasLtMatrixLayoutCreate(&la, HIP_R_16F, N, K, N));   // B^T: N x K col-major
    TCA_LT_TRY(hipblasLtMatrixLayoutCreate(&lb, HIP_R_16F, K, M, K));   // A^T: K x M col-major
    TCA_LT_TRY(hipblasLtMatrixLayoutCreate(&lc, HIP_R_32F, N, M, N));   // C^T: N x M col-major
    TCA_LT_TRY(hipblasLtMatmulPreferenceCreate(&pref));
    TCA_LT_TRY(hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &ws_max, sizeof(ws_max)));
    TCA_LT_TRY(hipblasLtMatmulAlgoGetHeuristic(h, desc, la, lb, lc, lc, pref, 1, heur, &found));
    if (found < 1) { st = fail(TCA_ERR_LIBRARY, "hipBLASLt: no algorithm for f16 x f16 -> f32"); goto done; }
    if (heur[0].workspaceSize && hipMalloc(&ws, heur[0].workspaceSize) != hipSuccess) { st = fail(TCA_ERR_ALLOC, "hipBLASLt workspace"); goto done; }
    {
        const int warm = n_iter > 0 ? 5 : 1;  // the reference warms GemmEx up 5x
        for (int i = 0; i < warm; i++)
            TCA_LT_TRY(hipblasLtMatmul(h, desc, &alpha, d_B, la, d_A, lb, &beta, d_C, lc, d_C, lc, &heur[0].algo, ws,
                                       heur[0].workspaceSize, nullptr));
        if (n_iter > 0) {
            if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) { st = fail(TCA_ERR_HIP, "event"); goto done; }
            (void)hipEventRecord(ev.a, nullptr);
            for (int i = 0; i < n_iter; i++)
                TCA_LT_TRY(hipblasLtMatmul(h, desc, &alpha, d_B, la, d_A, lb, &beta, d_C, lc, d_C, lc, &heur[0].algo, ws,
                                           heur[0].workspaceSize, nullptr));
            (void)hipEventRecord(ev.b, nullptr);
            (void)hipEventSynchronize(ev.b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev.a, ev.b);
            if (ms_per_iter) *ms_per_iter = ms / n_iter;
        }
        if (hipDeviceSynchronize() != hipSuccess) st = fail(TCA_ERR_HIP, "hipBLASLt: synchronize");
    }
done:
    if (ws) (void)hipFree(ws);
    if (pref) hipblasLtMatmulPreferenceDestroy(pref);
    if (la) hipblasLtMatrixLayoutDestroy(la);
    if (lb) hipblasLtMatrixLayoutDestroy(lb);
    if (lc) hipblasLtMatrixLayoutDestroy(lc);
    if (desc) hipblasLtMatmulDescDestroy(desc);
    if (h) hipblasLtDestroy(h);
    return st;
}

// rocSPARSE SDDMM as the library counterpart of tca_sddmm (dW): out (CSR order,
// f32) = (G Xt) sampled on the pattern, G M x N and X K x N row-major f16 (X read
// as the column-major N x K matrix Xt), FP32 compute.  The reference has no
// backward; this is the baseline for the dW of its SpMM.
extern "C" tca_status tca_rocsparse_sddmm(int M, int K, int N, int nnz, const int32_t *d_row_offset,
                                          const int32_t *d_col_idx, const uint16_t *d_G, const uint16_t *d_X,
                                          float *d_out, int n_iter, double *ms_per_iter)
{
    if (M <= 0 || N <= 0 || K <= 0 || nnz < 0 || !d_row_offset || !d_G || !d_X || (nnz > 0 && (!d_col_idx || !d_out)))
        return fail(TCA_ERR_INVALID_ARG, "tca_rocsparse_sddmm: bad argument");
    tca_status st = TCA_OK;
    rocsparse_handle h = nullptr;
    rocsparse_spmat_descr C = nullptr;
    rocsparse_dnmat_descr A = nullptr, B = nullptr;
    void *buf = nullptr;
    size_t bs = 0;
    const float alpha = 1.0f, beta = 0.0f;
    EventPair ev;
    TCA_SPARSE_TRY(rocsparse_create_handle(&h));
    TCA_SPARSE_TRY(rocsparse_create_dnmat_descr(&A, M, N, N, const_cast<uint16_t *>(d_G), rocsparse_datatype_f16_r,
                                                rocsparse_order_row));
    TCA_SPARSE_TRY(rocsparse_create_dnmat_descr(&B, N, K, N, const_cast<uint16_t *>(d_X), rocsparse_datatype_f16_r,
                                                rocsparse_order_column));
    TCA_SPARSE_TRY(rocsparse_create_csr_descr(&C, M, K, nnz, const_cast<int32_t *>(d_row_offset),
                                              const_cast<int32_t *>(d_col_idx), d_out, rocsparse_indextype_i32,
                                              rocsparse_indextype_i32, rocsparse_index_base_zero,
                                              rocsparse_datatype_f32_r));
    TCA_SPARSE_TRY(rocsparse_sddmm_buffer_size(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, B,
                                               &beta, C, rocsparse_datatype_f32_r, rocsparse_sddmm_alg_default, &bs));
    if (bs && hipMalloc(&buf, bs) != hipSuccess) { st = fail(TCA_ERR_ALLOC, "rocsparse buffer"); goto done; }
    TCA_SPARSE_TRY(rocsparse_sddmm_preprocess(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, B,
                                              &beta, C, rocsparse_datatype_f32_r, rocsparse_sddmm_alg_default, buf));
    {
        const int warm = n_iter > 0 ? 10 : 1;
        for (int i = 0; i < warm; i++)
            TCA_SPARSE_TRY(rocsparse_sddmm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, B, &beta,
                                           C, rocsparse_datatype_f32_r, rocsparse_sddmm_alg_default, buf));
        if (n_iter > 0) {
            if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess) { st = fail(TCA_ERR_HIP, "event"); goto done; }
            (void)hipEventRecord(ev.a, nullptr);
            for (int i = 0; i < n_iter; i++)
                TCA_SPARSE_TRY(rocsparse_sddmm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, B,
                                               &beta, C, rocsparse_datatype_f32_r, rocsparse_sddmm_alg_default, buf));
            (void)hipEventRecord(ev.b, nullptr);
            (void)hipEventSynchronize(ev.b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev.a, ev.b);
            if (ms_per_iter) *ms_per_iter = ms / n_iter;
        }
        if (hipDeviceSynchronize() != hipSuccess) st = fail(TCA_ERR_HIP, "rocsparse: synchronize");
    }
done:
    if (buf) (void)hipFree(buf);
    if (C) rocsparse_destroy_spmat_descr(C);
    if (A) rocsparse_destroy_dnmat_descr(A);
    if (B) rocsparse_destroy_dnmat_descr(B);
    if (h) rocsparse_destroy_handle(h);
    return st;
}
