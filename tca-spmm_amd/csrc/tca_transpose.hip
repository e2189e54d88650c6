// Device-side transpose of a CSR pattern (its CSC): for a pattern that lives on
// the GPU, the column-major index arrays are built there -- a stable radix sort
// of the entries by column (entries arrive row-major, so rows stay ascending
// within each column), then the row of every sorted entry and the column
// offsets by binary search.  The reference transposes only dense matrices, on
// the host (src/miscalleneous.cu:146-157).
#include "tca_internal.hpp"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <mutex>
#include <string>

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

// Scratch layout of one transpose: rows[nnz] | iota[nnz] | sorted keys[nnz] |
// perm (when the caller gives none)[nnz] | radix-sort temp, every array on a
// 256-byte boundary.
size_t arr_bytes(int nnz) { return ((size_t)nnz * sizeof(int32_t) + 255) / 256 * 256; }

int key_bits(int n_cols)
{
    int bits = 1;
    while (bits < 31 && (1L << bits) < n_cols) bits++;
    return bits;
}

tca_status check_args(int n_rows, int n_cols, int nnz, const int32_t *d_row_offset, const int32_t *d_col_idx,
                      const int32_t *d_col_offset, const int32_t *d_row_idx, const char *who)
{
    if (n_rows <= 0 || n_cols <= 0 || nnz < 0 || !d_row_offset || !d_col_offset || (nnz > 0 && (!d_col_idx || !d_row_idx)))
        return fail(TCA_ERR_INVALID_ARG, std::string(who) + ": bad argument");
    return TCA_OK;
}

// Bytes of scratch one transpose of nnz entries over n_cols columns needs.
tca_status scratch_bytes(int n_cols, int nnz, hipStream_t s, size_t *temp_bytes, size_t *total)
{
    *temp_bytes = 0;
    TCA_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, *temp_bytes, (const int32_t *)nullptr, (int32_t *)nullptr,
                                                   (int32_t *)nullptr, (int32_t *)nullptr, nnz, 0, key_bits(n_cols), s));
    *total = 4 * arr_bytes(nnz) + *temp_bytes;
    return TCA_OK;
}

// The transpose on caller-provided scratch (at least scratch_bytes' total).
tca_status transpose_on(int n_rows, int n_cols, int nnz, const int32_t *d_row_offset, const int32_t *d_col_idx,
                        int32_t *d_col_offset, int32_t *d_row_idx, int32_t *d_perm, char *work, size_t temp_bytes,
                        hipStream_t s)
{
    const size_t arr = arr_bytes(nnz);
    int32_t *rows = reinterpret_cast<int32_t *>(work);
    int32_t *iota = reinterpret_cast<int32_t *>(work + arr);
    int32_t *keys = reinterpret_cast<int32_t *>(work + 2 * arr);
    int32_t *perm = d_perm ? d_perm : reinterpret_cast<int32_t *>(work + 3 * arr);
    void *temp = work + 4 * arr;
    hipLaunchKernelGGL(entry_rows_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, n_rows, nnz, d_row_offset, rows);
    hipLaunchKernelGGL(iota_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, nnz, iota);
    if (hipGetLastError() != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, d_col_idx, keys, iota, perm, nnz, 0, key_bits(n_cols), s) !=
            hipSuccess)
        return fail(TCA_ERR_HIP, "tca_csr_transpose_device: sort");
    hipLaunchKernelGGL(col_offsets_kernel, dim3(grid_for(n_cols + 1)), dim3(256), 0, s, n_cols, nnz, keys, d_col_offset);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, nnz, perm, rows, d_row_idx);
    if (hipGetLastError() != hipSuccess) return fail(TCA_ERR_HIP, "tca_csr_transpose_device: launch");
    return TCA_OK;
}

tca_status empty_transpose(int n_cols, int32_t *d_col_offset, hipStream_t s)
{
    hipLaunchKernelGGL(col_offsets_kernel, dim3(grid_for(n_cols + 1)), dim3(256), 0, s, n_cols, 0, (const int32_t *)nullptr,
                       d_col_offset);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

// Scratch of tca_csr_transpose_device_noalloc: one fixed array per GPU, in the
// library's device image (each GPU the library runs on holds its own copy until
// the process ends).  Host code never passes the array's name to a kernel: its
// device address on the current GPU comes from hipGetSymbolAddress, taken once
// per device on the first call there and kept.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
__device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
char *scratch_ptr[kMaxDevices] = {};

tca_status device_scratch(char **out)
{
    int dev = 0;
    TCA_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TCA_ERR_UNSUPPORTED, "tca_csr_transpose_device_noalloc: device index");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(TCA_ERR_HIP, std::string("tca_csr_transpose_device_noalloc: hipGetSymbolAddress: ") +
                                         hipGetErrorString(e));
        scratch_ptr[dev] = static_cast<char *>(p);
    }
    *out = scratch_ptr[dev];
    return TCA_OK;
}

}  // namespace

extern "C" tca_status tca_csr_transpose_device(int n_rows, int n_cols, int nnz, const int32_t *d_row_offset,
                                               const int32_t *d_col_idx, int32_t *d_col_offset, int32_t *d_row_idx,
                                               int32_t *d_perm, void *stream)
{
    tca_status st = check_args(n_rows, n_cols, nnz, d_row_offset, d_col_idx, d_col_offset, d_row_idx,
                               "tca_csr_transpose_device");
    if (st != TCA_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nnz == 0) return empty_transpose(n_cols, d_col_offset, s);
    size_t temp_bytes = 0, total = 0;
    st = scratch_bytes(n_cols, nnz, s, &temp_bytes, &total);
    if (st != TCA_OK) return st;
    char *scratch_mem = nullptr;
    bool pooled = hipMallocAsync(reinterpret_cast<void **>(&scratch_mem), total, s) == hipSuccess;
    if (!pooled) {   // no stream-ordered pool: plain allocation, freed after the stream drains
        (void)hipGetLastError();
        if (hipMalloc(reinterpret_cast<void **>(&scratch_mem), total) != hipSuccess)
            return fail(TCA_ERR_ALLOC, "tca_csr_transpose_device: scratch allocation");
    }
    st = transpose_on(n_rows, n_cols, nnz, d_row_offset, d_col_idx, d_col_offset, d_row_idx, d_perm, scratch_mem,
                      temp_bytes, s);
    if (pooled) {
        if (hipFreeAsync(scratch_mem, s) != hipSuccess && st == TCA_OK) st = fail(TCA_ERR_HIP, "tca_csr_transpose_device: free");
    } else {
        if ((hipStreamSynchronize(s) != hipSuccess || hipFree(scratch_mem) != hipSuccess) && st == TCA_OK)
            st = fail(TCA_ERR_HIP, "tca_csr_transpose_device: free");
    }
    return st;
}

extern "C" tca_status tca_csr_transpose_device_noalloc(int n_rows, int n_cols, int nnz, const int32_t *d_row_offset,
                                                       const int32_t *d_col_idx, int32_t *d_col_offset,
                                                       int32_t *d_row_idx, int32_t *d_perm, void *stream)
{
    tca_status st = check_args(n_rows, n_cols, nnz, d_row_offset, d_col_idx, d_col_offset, d_row_idx,
                               "tca_csr_transpose_device_noalloc");
    if (st != TCA_OK) return st;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (nnz == 0) return empty_transpose(n_cols, d_col_offset, s);
    size_t temp_bytes = 0, total = 0;
    st = scratch_bytes(n_cols, nnz, s, &temp_bytes, &total);
    if (st != TCA_OK) return st;
    if (total > kNoAllocScratchBytes)
        return fail(TCA_ERR_UNSUPPORTED, "tca_csr_transpose_device_noalloc: " + std::to_string(nnz) + " entries need " +
                                             std::to_string(total) + " bytes of scratch, above the fixed " +
                                             std::to_string(kNoAllocScratchBytes) + " (use tca_csr_transpose_device)");
    char *scratch_mem = nullptr;
    st = device_scratch(&scratch_mem);
    if (st != TCA_OK) return st;
    return transpose_on(n_rows, n_cols, nnz, d_row_offset, d_col_idx, d_col_offset, d_row_idx, d_perm, scratch_mem,
                        temp_bytes, s);
}

extern "C" size_t tca_csr_transpose_noalloc_max_scratch(void) { return kNoAllocScratchBytes; }
