// Device-side transpose of a CSR pattern (its CSC): for a pattern that lives on
// the GPU, the column-major index arrays are built there -- a stable radix sort
// of the entries by column (entries arrive row-major, so rows stay ascending
// within each column), then the row of every sorted entry and the column
// offsets by binary search.  The reference transposes only dense matrices, on
// the host (src/miscalleneous.cu:146-157).
#include "tca_internal.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

using tca::fail;

namespace {

// row of every CSR entry: the last r with row_offset[r] <= e
__global__ void entry_rows_kernel(int n_rows, int nnz, const int32_t *__restrict__ row_offset, int32_t *__restrict__ rows)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += gridDim.x * blockDim.x) {
        int lo = 0, hi = n_rows;   // row_offset[lo] <= e < row_offset[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (row_offset[mid] <= e) lo = mid;
            else hi = mid;
        }
        rows[e] = lo;
    }
}

__global__ void iota_kernel(int n, int32_t *__restrict__ out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = i;
}

// col_offset[c] = number of sorted keys below c, c in [0, n_cols]
__global__ void col_offsets_kernel(int n_cols, int nnz, const int32_t *__restrict__ sorted_cols,
                                   int32_t *__restrict__ col_offset)
{
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c <= n_cols; c += gridDim.x * blockDim.x) {
        int lo = 0, hi = nnz;      // first index with key >= c
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (sorted_cols[mid] < c) lo = mid + 1;
            else hi = mid;
        }
        col_offset[c] = lo;
    }
}

__global__ void gather_rows_kernel(int nnz, const int32_t *__restrict__ perm, const int32_t *__restrict__ rows,
                                   int32_t *__restrict__ row_idx)
{
    for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < nnz; d += gridDim.x * blockDim.x) row_idx[d] = rows[perm[d]];
}

int grid_for(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

}  // namespace

extern "C" tca_status tca_csr_transpose_device(int n_rows, int n_cols, int nnz, const int32_t *d_row_offset,
                                               const int32_t *d_col_idx, int32_t *d_col_offset, int32_t *d_row_idx,
                                               int32_t *d_perm, void *stream)
{
    if (n_rows <= 0 || n_cols <= 0 || nnz < 0 || !d_row_offset || !d_col_offset || (nnz > 0 && (!d_col_idx || !d_row_idx)))
        return fail(TCA_ERR_INVALID_ARG, "tca_csr_transpose_device: bad argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nnz == 0) {
        hipLaunchKernelGGL(col_offsets_kernel, dim3(grid_for(n_cols + 1)), dim3(256), 0, s, n_cols, 0, d_col_idx, d_col_offset);
        TCA_HIP_TRY(hipGetLastError());
        return TCA_OK;
    }
    int bits = 1;
    while (bits < 31 && (1L << bits) < n_cols) bits++;
    // scratch: rows[nnz] | iota[nnz] | sorted keys[nnz] | perm (when the caller gives none)[nnz] | sort temp
    size_t temp_bytes = 0;
    TCA_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, d_col_idx, (int32_t *)nullptr, (int32_t *)nullptr,
                                                   (int32_t *)nullptr, nnz, 0, bits, s));
    const size_t arr = ((size_t)nnz * sizeof(int32_t) + 255) / 256 * 256;
    char *scratch = nullptr;
    bool pooled = hipMallocAsync(reinterpret_cast<void **>(&scratch), 4 * arr + temp_bytes, s) == hipSuccess;
    if (!pooled) {   // no stream-ordered pool: plain allocation, freed after the stream drains
        (void)hipGetLastError();
        if (hipMalloc(reinterpret_cast<void **>(&scratch), 4 * arr + temp_bytes) != hipSuccess)
            return fail(TCA_ERR_ALLOC, "tca_csr_transpose_device: scratch allocation");
    }
    int32_t *rows = reinterpret_cast<int32_t *>(scratch);
    int32_t *iota = reinterpret_cast<int32_t *>(scratch + arr);
    int32_t *keys = reinterpret_cast<int32_t *>(scratch + 2 * arr);
    int32_t *perm = d_perm ? d_perm : reinterpret_cast<int32_t *>(scratch + 3 * arr);
    void *temp = scratch + 4 * arr;
    tca_status st = TCA_OK;
    hipLaunchKernelGGL(entry_rows_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, n_rows, nnz, d_row_offset, rows);
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, nnz, iota);
    if (hipGetLastError() != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, d_col_idx, keys, iota, perm, nnz, 0, bits, s) != hipSuccess)
        st = fail(TCA_ERR_HIP, "tca_csr_transpose_device: sort");
    if (st == TCA_OK) {
        hipLaunchKernelGGL(col_offsets_kernel, dim3(grid_for(n_cols + 1)), dim3(256), 0, s, n_cols, nnz, keys, d_col_offset);
        hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, nnz, perm, rows, d_row_idx);
        if (hipGetLastError() != hipSuccess) st = fail(TCA_ERR_HIP, "tca_csr_transpose_device: launch");
    }
    if (pooled) {
        if (hipFreeAsync(scratch, s) != hipSuccess && st == TCA_OK) st = fail(TCA_ERR_HIP, "tca_csr_transpose_device: free");
    } else {
        if ((hipStreamSynchronize(s) != hipSuccess || hipFree(scratch) != hipSuccess) && st == TCA_OK)
            st = fail(TCA_ERR_HIP, "tca_csr_transpose_device: free");
    }
    return st;
}
