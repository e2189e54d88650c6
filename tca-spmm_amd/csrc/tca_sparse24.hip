// 2:4 structured sparsity on gfx950's sparse MFMA (v_smfmac_f32_16x16x64_{f16,bf16}).
//
// The reference's 2:4 comparison (perform_cuSPARSELT_24sparsity,
// src/spmm_cusparse.cu:2-179) prunes A with the STRIP rule -- in every 1x4
// strip of a row the pair of largest L1 norm survives -- compresses it and
// multiplies.  Here both steps are kernels of our own:
//
//   compress24_kernel   dense A (M x K row-major) -> values (M x K/2: the two
//                       survivors of every strip, in column order) and meta
//                       (M x K/16 uint16: 2-bit column-in-strip of each value,
//                       value j of a 16-column block at bits 2j+1:2j).  Ties
//                       go to the first pair of (0,1) (0,2) (0,3) (1,2) (1,3)
//                       (2,3).
//   spmm24_kernel       C = A_2:4 * B.  The sparse MFMA covers 64 columns of
//                       A per instruction: lane (row r = l%16, q = l/16) holds
//                       the 8 values of A columns 16q..16q+15 (one 16-byte
//                       load) and their 16-bit meta word; its B operand is B
//                       rows 8q..8q+7 and 32+8q..32+8q+7 of the 64-row step
//                       at column l%16 (lane layout measured on the MI355X).
//                       One wave owns (RPW rows, NT columns); per step it
//                       stages the 64 B rows in its LDS tile and reads the
//                       operands with transposing reads (ds_read_b64_tr_b16),
//                       two steps of loads in flight.
#include "tca_device.hpp"

#include <algorithm>

namespace tca {
namespace {

typedef _Float16 f16x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x16 __attribute__((ext_vector_type(16)));

template <tca_dtype AT> struct Smfmac;
template <> struct Smfmac<TCA_F16> {
    __device__ static f32x4 mma(uint4 a, const uint4 (&b)[2], int idx, f32x4 c)
    {
        const f16x16 bv = __builtin_bit_cast(f16x16, b);
        return __builtin_amdgcn_smfmac_f32_16x16x64_f16(__builtin_bit_cast(f16x8, a), bv, c, idx, 0, 0);
    }
    __device__ static float mag(uint32_t h) { return (float)__builtin_bit_cast(_Float16, (uint16_t)(h & 0x7fffu)); }
};
template <> struct Smfmac<TCA_BF16> {
    __device__ static f32x4 mma(uint4 a, const uint4 (&b)[2], int idx, f32x4 c)
    {
        const bf16x16 bv = __builtin_bit_cast(bf16x16, b);
        return __builtin_amdgcn_smfmac_f32_16x16x64_bf16(__builtin_bit_cast(bf16x8, a), bv, c, idx, 0, 0);
    }
    __device__ static float mag(uint32_t h) { return __builtin_bit_cast(float, (h & 0x7fffu) << 16); }
};

// meta word of a 16-column block whose four strips all keep columns (0, 1):
// what rows past M read, so that their (zero) values pair with valid positions
constexpr uint32_t kMeta01 = 0x4444u;

// One thread per (row, 16-column block): 32 bytes in, 16 bytes + one meta word out.
template <tca_dtype AT>
__global__ __launch_bounds__(256) void compress24_kernel(int M, int K, const uint16_t *dense, int ld, uint16_t *values,
                                                         uint16_t *meta)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const int nb = K >> 4;
    if (t >= (long)M * nb) return;
    const int r = (int)(t / nb), c = (int)(t - (long)r * nb);
    const u32x4 *src = reinterpret_cast<const u32x4 *>(dense + (long)r * ld + 16 * c);
    const u32x4 v0 = gbl(src)[0], v1 = gbl(src)[1];
    const uint32_t w[8] = {v0[0], v0[1], v0[2], v0[3], v1[0], v1[1], v1[2], v1[3]};
    constexpr int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    uint32_t out[4];
    uint32_t m = 0;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        uint32_t e[4];
#pragma unroll
        for (int i = 0; i < 4; i++) e[i] = (w[2 * g + i / 2] >> (16 * (i & 1))) & 0xffffu;
        float mg[4];
#pragma unroll
        for (int i = 0; i < 4; i++) mg[i] = Smfmac<AT>::mag(e[i]);
        int best = 0;
        float bv = mg[pa[0]] + mg[pb[0]];
#pragma unroll
        for (int p = 1; p < 6; p++) {
            const float s = mg[pa[p]] + mg[pb[p]];
            if (s > bv) {
                bv = s;
                best = p;
            }
        }
        int i0 = 0, i1 = 1;
#pragma unroll
        for (int p = 0; p < 6; p++)
            if (p == best) {
                i0 = pa[p];
                i1 = pb[p];
            }
        const uint32_t x0 = i0 == 0 ? e[0] : i0 == 1 ? e[1] : e[2];
        const uint32_t x1 = i1 == 1 ? e[1] : i1 == 2 ? e[2] : e[3];
        out[g] = x0 | (x1 << 16);
        m |= (uint32_t)(i0 | (i1 << 2)) << (4 * g);
    }
    gbl(reinterpret_cast<u32x4 *>(values + (long)r * (K >> 1) + 8 * c))[0] = u32x4{out[0], out[1], out[2], out[3]};
    gbl(meta)[(long)r * nb + c] = (uint16_t)m;
}

struct S24Params {
    const void *values;
    const uint16_t *meta;
    const void *B;
    float *C;
    int M, K, N, ldb, ldc, steps, n_tiles, n_units;
    uint32_t val_bytes, meta_bytes, b_bytes;
};

constexpr int kS24Depth = 2;

template <int RPW, int NT, tca_dtype AT, tca_dtype BT>
__global__ __launch_bounds__(64) void spmm24_kernel(S24Params p)
{
    constexpr int PPG = RPW / 16;
    constexpr int NSUB = NT / 16;
    constexpr int SEGS = NT / 8;       // 8-column segments of a B row
    constexpr int LPS = 64 / SEGS;     // lanes sharing one segment
    constexpr int KPL = 64 / LPS;      // consecutive B rows per lane and step
    constexpr int LDB = tr_pitch<NT>();
    constexpr int D = kS24Depth;
    using Oct = uint4;
    using RL = RowLoad<AT, BT>;
    using RB = RowBuf<AT, BT>;
    struct Stage {
        typename RB::Raw rows[KPL];
        Oct aop[PPG];
        uint32_t idx[PPG];
    };
    __shared__ __attribute__((aligned(16))) uint8_t lds_b[64 * LDB * 2];   // rows by tr_row

    const int unit = blockIdx.x;
    if (unit >= p.n_units) return;
    const int rp = unit / p.n_tiles, tile = unit - rp * p.n_tiles;
    const int row0 = rp * RPW;
    const int lane = lane_id();
    const int n0 = tile * NT;
    const int seg = lane % SEGS, kq = lane / SEGS, k0 = kq * KPL;
    const int ncol = n0 + 8 * seg;
    const bool col_ok = ncol < p.N;
    const Rsrc r_b = make_rsrc(p.B, p.b_bytes);
    const Rsrc r_val = make_rsrc(p.values, p.val_bytes);
    const Rsrc r_meta = make_rsrc(p.meta, p.meta_bytes);
    const uint32_t ldb_bytes = (uint32_t)p.ldb * RB::BES;
    const uint32_t col_bytes = (uint32_t)ncol * RB::BES;
    const int half_k = p.K >> 1, n16 = p.K >> 4;
    bool row_ok[PPG];
    uint32_t a_off[PPG], m_off[PPG];
#pragma unroll
    for (int i = 0; i < PPG; i++) {
        const int r = row0 + 16 * i + (lane & 15);
        row_ok[i] = r < p.M;
        a_off[i] = ((uint32_t)r * (uint32_t)half_k + 8u * (uint32_t)(lane >> 4)) * 2u;
        m_off[i] = ((uint32_t)r * (uint32_t)n16 + (uint32_t)(lane >> 4)) * 2u;
    }

    f32x4 acc[PPG][NSUB];
#pragma unroll
    for (int i = 0; i < PPG; i++)
#pragma unroll
        for (int s = 0; s < NSUB; s++) acc[i][s] = f32x4{0.f, 0.f, 0.f, 0.f};

    auto load = [&](int s, Stage &st) {
#pragma unroll
        for (int j = 0; j < KPL; j++)
            st.rows[j] = RB::load(r_b, col_ok ? (uint32_t)(64 * s + k0 + j) * ldb_bytes + col_bytes : kOOB);
#pragma unroll
        for (int i = 0; i < PPG; i++) {
            st.aop[i] = buf_b128(r_val, row_ok[i] ? a_off[i] + 64u * (uint32_t)s : kOOB);
            st.idx[i] = buf_b16(r_meta, row_ok[i] ? m_off[i] + 8u * (uint32_t)s : kOOB);
        }
    };
    auto compute = [&](const Stage &st) {
#pragma unroll
        for (int j = 0; j < KPL; j++)
            *reinterpret_cast<Oct *>(lds_b + tr_row(k0 + j, LDB * 2) + 16 * seg) = RL::cvt(st.rows[j]);
        wave_lds_sync();
        int idx[PPG];
#pragma unroll
        for (int i = 0; i < PPG; i++) idx[i] = (int)(row_ok[i] ? st.idx[i] : kMeta01);
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            uint4 b[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int row = 32 * h + 8 * (lane >> 4) + ((lane & 15) >> 2), col = (16 * s + 4 * (lane & 3)) * 2;
                const uint2 lo = lds_read_tr16(lds_b, tr_row(row, LDB * 2) + col);
                const uint2 hi = lds_read_tr16(lds_b, tr_row(row + 4, LDB * 2) + col);
                b[h] = make_uint4(lo.x, lo.y, hi.x, hi.y);
            }
#pragma unroll
            for (int i = 0; i < PPG; i++) acc[i][s] = Smfmac<AT>::mma(st.aop[i], b, idx[i], acc[i][s]);
        }
        wave_lds_sync();
    };

    const int steps = p.steps;
    if (steps > 0) {
        const int last = steps - 1;
        Stage st[D];
#pragma unroll
        for (int s = 0; s < D; s++) load(min(s, last), st[s]);
        vm_drain();
        for (int t0 = 0; t0 < steps; t0 += D) {
#pragma unroll
            for (int s = 0; s < D; s++) {
                const int t = t0 + s;
                if (t < steps) compute(st[s]);
                load(min(t + D, last), st[s]);
            }
        }
    }

    const int rbase = row0 + 4 * (lane >> 4);
#pragma unroll
    for (int i = 0; i < PPG; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = rbase + 16 * i + j;
            if (row >= p.M) continue;
            float *crow = p.C + (long)row * p.ldc;
#pragma unroll
            for (int s = 0; s < NSUB; s++) {
                const int n = n0 + 16 * s + (lane & 15);
                if (n < p.N) gbl(crow)[n] = acc[i][s][j];
            }
        }
    }
}

using S24Fn = void (*)(S24Params);

template <int RPW, tca_dtype AT, tca_dtype BT>
S24Fn pick_s24_nt(int nt)
{
    switch (nt) {
    case 16: return spmm24_kernel<RPW, 16, AT, BT>;
    case 32: return spmm24_kernel<RPW, 32, AT, BT>;
    case 64: return spmm24_kernel<RPW, 64, AT, BT>;
    default: return nullptr;
    }
}

template <tca_dtype AT, tca_dtype BT>
S24Fn pick_s24_rpw(int rpw, int nt)
{
    switch (rpw) {
    case 16: return pick_s24_nt<16, AT, BT>(nt);
    case 32: return pick_s24_nt<32, AT, BT>(nt);
    case 64: return pick_s24_nt<64, AT, BT>(nt);
    default: return nullptr;
    }
}

S24Fn pick_s24(tca_dtype at, tca_dtype bt, int rpw, int nt)
{
    if (at == TCA_F16 && bt == TCA_F16) return pick_s24_rpw<TCA_F16, TCA_F16>(rpw, nt);
    if (at == TCA_BF16 && bt == TCA_BF16) return pick_s24_rpw<TCA_BF16, TCA_BF16>(rpw, nt);
    if (at == TCA_F16 && bt == TCA_F32) return pick_s24_rpw<TCA_F16, TCA_F32>(rpw, nt);
    if (at == TCA_BF16 && bt == TCA_F32) return pick_s24_rpw<TCA_BF16, TCA_F32>(rpw, nt);
    return nullptr;
}

}  // namespace

tca_status launch_compress24(int M, int K, tca_dtype dtype, const void *dense, int ld, void *values, uint16_t *meta,
                             hipStream_t stream)
{
    const long threads = (long)M * (K / 16);
    if (threads == 0) return TCA_OK;
    const unsigned blocks = (unsigned)((threads + 255) / 256);
    const uint16_t *d = static_cast<const uint16_t *>(dense);
    uint16_t *v = static_cast<uint16_t *>(values);
    if (dtype == TCA_F16)
        hipLaunchKernelGGL(compress24_kernel<TCA_F16>, dim3(blocks), dim3(256), 0, stream, M, K, d, ld, v, meta);
    else
        hipLaunchKernelGGL(compress24_kernel<TCA_BF16>, dim3(blocks), dim3(256), 0, stream, M, K, d, ld, v, meta);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

tca_status launch_spmm24(int M, int K, int N, tca_dtype a_dtype, const void *values, const uint16_t *meta,
                         tca_dtype b_dtype, const void *B, int ldb, float *C, int ldc, int tile_n, int rows_per_wave,
                         hipStream_t stream)
{
    int rpw = 0, nt = 0;
    bell_shape(M, K, N, 32, tile_n, &rpw, &nt);   // same unit-count rule as the Blocked-ELL kernel
    if (rows_per_wave) rpw = rows_per_wave;
    S24Fn fn = pick_s24(a_dtype, b_dtype, rpw, nt);
    if (!fn) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm_24: unsupported dtype pair or tile");
    if (M <= 0 || N <= 0) return TCA_OK;
    S24Params prm;
    prm.values = values;
    prm.meta = meta;
    prm.B = B;
    prm.C = C;
    prm.M = M;
    prm.K = K;
    prm.N = N;
    prm.ldb = ldb;
    prm.ldc = ldc;
    prm.steps = K / 64;
    prm.n_tiles = (N + nt - 1) / nt;
    const long units = (long)((M + rpw - 1) / rpw) * prm.n_tiles;
    const long vb = (long)M * (K / 2) * 2, mb = (long)M * (K / 16) * 2;
    prm.val_bytes = vb > 0 && vb < (1L << 31) - 64 ? (uint32_t)vb : (vb == 0 ? 16u : 0u);
    prm.meta_bytes = mb > 0 && mb < (1L << 31) - 64 ? (uint32_t)mb : (mb == 0 ? 16u : 0u);
    prm.b_bytes = operand_bytes(K, N, ldb, b_dtype == TCA_F32 ? 4 : 2);
    if (prm.val_bytes == 0 || prm.meta_bytes == 0 || prm.b_bytes == 0 || units >= (1L << 31))
        return fail(TCA_ERR_UNSUPPORTED, "tca_spmm_24: operands reach 2 GiB (32-bit offsets)");
    prm.n_units = (int)units;
    hipLaunchKernelGGL(fn, dim3((unsigned)units), dim3(64), 0, stream, prm);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

}  // namespace tca
