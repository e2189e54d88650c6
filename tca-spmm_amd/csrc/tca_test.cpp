// tca_test -- the drop-in counterpart of the reference driver (src/test.cu:29-88):
// loads a .smtx pattern as a dense fp16 A (values 1.0), builds B (K x 256,
// nonzero with probability 0.9, U[0,1) values -- here from a seeded stream),
// computes the hipBLASLt gold, then checks rocSPARSE and our_spmm_balanced
// against it with compare_mat and prints the three timings.
#include "../../include/tca_spmm.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

// --sparse24 M N K n_iter dir: the hipSPARSELt 2:4 baseline in a process of its own
// (no other hipSPARSELt loaded).  Reads dir/A.bin (M x K) and dir/B.bin (K x N), F16
// row-major; writes dir/P.bin (pruned A) and dir/C.bin (M x N, F16); prints the mean
// ms per call over n_iter calls.
static int run_sparse24(int argc, char *argv[])
{
    if (argc < 7) {
        std::puts("Usage : ./tca_test --sparse24 M N K n_iter dir");
        return 2;
    }
    const int M = std::atoi(argv[2]), N = std::atoi(argv[3]), K = std::atoi(argv[4]), n_iter = std::atoi(argv[5]);
    const std::string dir = argv[6];
    if (M <= 0 || N <= 0 || K <= 0) return 2;
    std::vector<uint16_t> A((size_t)M * K), B((size_t)K * N), P(A.size()), C((size_t)M * N);
    auto rd = [&](const char *name, std::vector<uint16_t> &v) {
        FILE *f = std::fopen((dir + "/" + name).c_str(), "rb");
        if (!f) return false;
        const size_t n = std::fread(v.data(), sizeof(uint16_t), v.size(), f);
        std::fclose(f);
        return n == v.size();
    };
    auto wr = [&](const char *name, const std::vector<uint16_t> &v) {
        FILE *f = std::fopen((dir + "/" + name).c_str(), "wb");
        if (!f) return false;
        const size_t n = std::fwrite(v.data(), sizeof(uint16_t), v.size(), f);
        std::fclose(f);
        return n == v.size();
    };
    if (!rd("A.bin", A) || !rd("B.bin", B)) {
        std::fprintf(stderr, "cannot read %s/A.bin or B.bin\n", dir.c_str());
        return 2;
    }
    uint16_t *dA, *dB, *dP, *dC;
    if (hipMalloc(&dA, A.size() * 2) != hipSuccess || hipMalloc(&dB, B.size() * 2) != hipSuccess ||
        hipMalloc(&dP, P.size() * 2) != hipSuccess || hipMalloc(&dC, C.size() * 2) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 1;
    }
    (void)hipMemcpy(dA, A.data(), A.size() * 2, hipMemcpyHostToDevice);
    (void)hipMemcpy(dB, B.data(), B.size() * 2, hipMemcpyHostToDevice);
    double ms = 0.0;
    const tca_status s = tca_hipsparselt_spmm_24(M, N, K, dA, dP, dB, dC, n_iter, &ms);
    if (s != TCA_OK) {
        std::fprintf(stderr, "tca_hipsparselt_spmm_24: %s (%s)\n", tca_status_string(s), tca_last_error());
        return 1;
    }
    (void)hipMemcpy(P.data(), dP, P.size() * 2, hipMemcpyDeviceToHost);
    (void)hipMemcpy(C.data(), dC, C.size() * 2, hipMemcpyDeviceToHost);
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dP);
    (void)hipFree(dC);
    if (!wr("P.bin", P) || !wr("C.bin", C)) return 1;
    std::printf("hipSPARSELt 2:4 ms %f\n", ms);
    return 0;
}

int main(int argc, char *argv[])
{
    if (argc > 1 && std::strcmp(argv[1], "--sparse24") == 0) return run_sparse24(argc, argv);
    if (argc <= 1) {
        std::puts("Usage : ./tca_test {matrix.smtx} [N=256] [seed=0] [ell_block=0: no Blocked-ELL]");
        std::exit(EXIT_FAILURE);
    }
    std::ifstream fin(argv[1], std::ifstream::in);
    if (!fin.is_open()) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        std::exit(EXIT_FAILURE);
    }
    half *A;
    int M, K, nnz;
    smtx_input_to_Dense(fin, &A, argv[1], &M, &K, &nnz);
    const int N = argc > 2 ? std::atoi(argv[2]) : 256;
    const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 0;
    const int ell_block = argc > 4 ? std::atoi(argv[4]) : 0;
    double e_gemm = -1.0, e_sparse = -1.0, e_our = -1.0;
    for (int i = 0; i < 10; i++) warm_up_gpu();
    std::printf("Input file - %s\n", argv[1]);

    // B: nonzero with probability 0.9, value U[0,1)  (src/test.cu:56, general.cuh:140-174)
    std::vector<float> keep((size_t)K * N), val((size_t)K * N);
    tca_gen_uniform_f32(seed, 11, keep.size(), 0.0, 1.0, keep.data());
    tca_gen_uniform_f32(seed, 12, val.size(), 0.0, 1.0, val.data());
    for (size_t i = 0; i < val.size(); i++)
        if (keep[i] <= 0.1f) val[i] = 0.0f;
    half *B = static_cast<half *>(std::malloc(sizeof(half) * (size_t)K * N));
    tca_f32_to_f16_host(val.data(), val.size(), reinterpret_cast<uint16_t *>(B));
    float *C_gold = static_cast<float *>(std::calloc((size_t)M * N, sizeof(float)));
    float *C_cmp = static_cast<float *>(std::calloc((size_t)M * N, sizeof(float)));

    perform_hipBLASLt_GemmEx(A, B, C_gold, M, N, K, 10, &e_gemm);

    std::memset(C_cmp, 0, sizeof(float) * (size_t)M * N);
    perform_rocSPARSE_SpMM(A, B, C_cmp, M, N, K, 10, &e_sparse);
    if (!compare_mat(C_gold, C_cmp, M * N, DEFAULT_TOL)) std::printf("\033[35mvalidation failed (rocSPARSE)\033[0m\n");

    CSR *csrA;
    convert_dense_to_csr(A, &csrA, M, K);
    std::memset(C_cmp, 0, sizeof(float) * (size_t)M * N);
    our_spmm_balanced(csrA, B, C_cmp, M, N, K, 10, &e_our);
    bool ok = compare_mat(C_gold, C_cmp, M * N, DEFAULT_TOL);
    if (!ok) std::printf("\033[35mvalidation failed (Our, balanced)\033[0m\n");

    double e_bell = -1.0;
    if (ell_block > 0) {   // the reference's block-SpMM baseline (not run by its driver)
        std::memset(C_cmp, 0, sizeof(float) * (size_t)M * N);
        perform_BlockSpMM(A, B, C_cmp, M, N, K, 10, ell_block, &e_bell);
        if (!compare_mat(C_gold, C_cmp, M * N, DEFAULT_TOL)) {
            std::printf("\033[35mvalidation failed (Blocked-ELL)\033[0m\n");
            ok = false;
        }
    }

    std::free(A);
    destroy_csr(csrA);
    std::free(B);
    std::free(C_gold);
    std::free(C_cmp);
    std::fprintf(stdout, "GemmEx : \033[36m[%lf] ms\033[0m\n", e_gemm);
    std::fprintf(stdout, "rocSPARSE: \033[36m[%lf] ms\033[0m \n", e_sparse);
    std::fprintf(stdout, "our: \033[36m[%lf] ms\033[0m \n", e_our);
    if (ell_block > 0) std::fprintf(stdout, "Blocked-ELL(%d): \033[36m[%lf] ms\033[0m \n", ell_block, e_bell);
    return ok ? 0 : 1;
}
