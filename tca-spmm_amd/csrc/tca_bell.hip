// Blocked-ELL SpMM on gfx950: C[M x N] (f32) = A (Blocked-ELL, block b) * B.
//
// The format is cuSPARSE's Blocked-ELL, as perform_cuSPARSE_BlockSpMM builds
// it (src/spmm_cusparse.cu:254-343) and as rocSPARSE declares it: ceil(M/b)
// block rows of `slots = ell_cols / b` block slots each; col_ind[br*slots + j]
// is the block column of slot j (-1: an empty slot, its values are ignored);
// values is (ceil(M/b)*b) x ell_cols row-major, row r / slot j / column c of
// the block at values[r*ell_cols + j*b + c].
//
// One wave (one workgroup) owns RPW rows of one block row and an N tile of NT
// columns.  A block row is a dense band: per 32-wide k step the wave's 32 B
// rows are b-contiguous runs (one slot, or two when b = 16), so the step needs
// one block-column index per lane and no gather table.  The wave stages the
// B rows row-major in its LDS tile and reads MFMA B operands with transposing
// reads (ds_read_b64_tr_b16); A operands (8 consecutive k of one row) are
// 16-byte loads straight from `values`, so A never passes through LDS.  Two
// steps are in flight (indices two steps ahead of their data); steps whose
// slots are all empty skip the MFMAs.
#include "tca_device.hpp"

#include <algorithm>

namespace tca {
namespace {

struct BellParams {
    const int32_t *col_ind;
    const void *values;
    const void *B;
    float *C;
    int M, K, N, ldb, ldc;
    int b_shift, slots, ell_cols, steps;
    int n_tiles, n_units;
    uint32_t val_bytes, b_bytes, ind_bytes;
};

constexpr int kBellDepth = 2;

template <int RPW, int NT, tca_dtype AT, tca_dtype BT>
__global__ __launch_bounds__(64) void bell_spmm_kernel(BellParams p)
{
    static_assert(Elt<AT>::ES == 2, "Blocked-ELL values are 16-bit");
    constexpr int PPG = RPW / 16;
    constexpr int NSUB = NT / 16;
    constexpr int SEGS = NT / 8;       // 8-column segments of a B row tile
    constexpr int LPS = 64 / SEGS;     // lanes sharing one segment
    constexpr int KPL = 32 / LPS;      // consecutive k rows per lane
    constexpr int LDB = tr_pitch<NT>();
    constexpr int D = kBellDepth;
    using Oct = typename Elt<AT>::Oct;
    using RL = RowLoad<AT, BT>;
    using RB = RowBuf<AT, BT>;
    struct Stage {
        typename RB::Raw rows[KPL];
        Oct aop[PPG];
        int blk;                       // block column of this lane's k rows (-1: none)
    };
    __shared__ __attribute__((aligned(16))) uint8_t lds_b[32 * LDB * 2];   // rows by tr_row

    const int unit = blockIdx.x;
    if (unit >= p.n_units) return;
    const int rp = unit / p.n_tiles, tile = unit - rp * p.n_tiles;
    const int row0 = rp * RPW;                   // first row of this wave (RPW divides b)
    const int br = row0 >> p.b_shift;
    const int bmask = (1 << p.b_shift) - 1;
    const int lane = lane_id();
    const int n0 = tile * NT;
    const int seg = lane % SEGS, kq = lane / SEGS, k0 = kq * KPL;
    const int ncol = n0 + 8 * seg;
    const bool col_ok = ncol < p.N;
    const Rsrc r_b = make_rsrc(p.B, p.b_bytes);
    const Rsrc r_val = make_rsrc(p.values, p.val_bytes);
    const Rsrc r_ind = make_rsrc(p.col_ind, p.ind_bytes);
    const uint32_t ldb_bytes = (uint32_t)p.ldb * RB::BES;
    const uint32_t col_bytes = (uint32_t)ncol * RB::BES;
    const uint32_t ind_base = (uint32_t)br * (uint32_t)p.slots;
    // A operand of panel i: row row0 + 16i + lane%16, k octet lane/16
    const uint32_t a_row = (uint32_t)(row0 + (lane & 15)) * (uint32_t)p.ell_cols + 8u * (uint32_t)(lane >> 4);

    f32x4 acc[PPG][NSUB];
#pragma unroll
    for (int i = 0; i < PPG; i++)
#pragma unroll
        for (int s = 0; s < NSUB; s++) acc[i][s] = f32x4{0.f, 0.f, 0.f, 0.f};

    // slot of this lane's k rows at step s (KPL divides 16, so they share one)
    auto load_index = [&](int s) -> uint32_t {
        const int slot = (32 * s + k0) >> p.b_shift;
        return buf_b32(r_ind, slot < p.slots ? (ind_base + (uint32_t)slot) * 4u : kOOB);
    };
    auto load_data = [&](uint32_t ix, int s, Stage &st) {
        const int e = 32 * s + k0;
        const bool in = (e >> p.b_shift) < p.slots;
        const int blk = in ? (int)ix : -1;
        st.blk = blk;
        const int krow = (blk << p.b_shift) + (e & bmask);
#pragma unroll
        for (int j = 0; j < KPL; j++) {
            const bool ok = blk >= 0 && krow + j < p.K && col_ok;
            st.rows[j] = RB::load(r_b, ok ? (uint32_t)(krow + j) * ldb_bytes + col_bytes : kOOB);
        }
        // A octet q (k = 8q..8q+7) lies in the slot of the lanes whose k rows
        // start at 8q: an empty slot's values are not read (they may hold anything)
        const int blk_a = __shfl(blk, ((8 * (lane >> 4)) / KPL) * SEGS);
        const bool a_ok = blk_a >= 0;
#pragma unroll
        for (int i = 0; i < PPG; i++)
            st.aop[i] = buf_b128(r_val, a_ok ? (a_row + (uint32_t)(16 * i) * (uint32_t)p.ell_cols + 32u * (uint32_t)s) * 2u
                                             : kOOB);
    };
    auto compute = [&](const Stage &st) {
        if (__builtin_amdgcn_ballot_w64(st.blk >= 0) == 0) return;   // every slot of the step is empty
        Oct rows[KPL];
#pragma unroll
        for (int j = 0; j < KPL; j++) rows[j] = RL::cvt(st.rows[j]);
#pragma unroll
        for (int j = 0; j < KPL; j++) *reinterpret_cast<Oct *>(lds_b + tr_row(k0 + j, LDB * 2) + 16 * seg) = rows[j];
        wave_lds_sync();
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            const int row = 8 * (lane >> 4) + ((lane & 15) >> 2), col = (16 * s + 4 * (lane & 3)) * 2;
            const uint2 lo = lds_read_tr16(lds_b, tr_row(row, LDB * 2) + col);
            const uint2 hi = lds_read_tr16(lds_b, tr_row(row + 4, LDB * 2) + col);
            const Oct b = make_uint4(lo.x, lo.y, hi.x, hi.y);
#pragma unroll
            for (int i = 0; i < PPG; i++) acc[i][s] = Elt<AT>::mma(st.aop[i], b, acc[i][s]);
        }
        wave_lds_sync();
    };

    // every load is unconditional (steps past the end re-read the last one);
    // only the compute is guarded
    const int steps = p.steps;
    if (steps > 0) {
        const int last = steps - 1;
        uint32_t ix[D];
        Stage st[D];
#pragma unroll
        for (int s = 0; s < D; s++) ix[s] = load_index(min(s, last));
#pragma unroll
        for (int s = 0; s < D; s++) {
            load_data(ix[s], min(s, last), st[s]);
            ix[s] = load_index(min(s + D, last));
        }
        vm_drain();
        for (int t0 = 0; t0 < steps; t0 += D) {
#pragma unroll
            for (int s = 0; s < D; s++) {
                const int t = t0 + s;
                if (t < steps) compute(st[s]);
                load_data(ix[s], min(t + D, last), st[s]);
                ix[s] = load_index(min(t + 2 * D, last));
            }
        }
    }

    // lane holds D[4*(lane/16)+j][lane%16] of each 16x16 tile
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

using BellFn = void (*)(BellParams);

template <int RPW, tca_dtype AT, tca_dtype BT>
BellFn pick_bell_nt(int nt)
{
    switch (nt) {
    case 16: return bell_spmm_kernel<RPW, 16, AT, BT>;
    case 32: return bell_spmm_kernel<RPW, 32, AT, BT>;
    case 64: return bell_spmm_kernel<RPW, 64, AT, BT>;
    default: return nullptr;
    }
}

template <tca_dtype AT, tca_dtype BT>
BellFn pick_bell_rpw(int rpw, int nt)
{
    switch (rpw) {
    case 16: return pick_bell_nt<16, AT, BT>(nt);
    case 32: return pick_bell_nt<32, AT, BT>(nt);
    case 64: return pick_bell_nt<64, AT, BT>(nt);
    default: return nullptr;
    }
}

BellFn pick_bell(tca_dtype at, tca_dtype bt, int rpw, int nt)
{
    if (at == TCA_F16 && bt == TCA_F16) return pick_bell_rpw<TCA_F16, TCA_F16>(rpw, nt);
    if (at == TCA_BF16 && bt == TCA_BF16) return pick_bell_rpw<TCA_BF16, TCA_BF16>(rpw, nt);
    if (at == TCA_F16 && bt == TCA_F32) return pick_bell_rpw<TCA_F16, TCA_F32>(rpw, nt);
    if (at == TCA_BF16 && bt == TCA_F32) return pick_bell_rpw<TCA_BF16, TCA_F32>(rpw, nt);
    return nullptr;
}

}  // namespace

// Rows per wave (16/32/64, at most b) and N tile (16/32/64): the largest that
// still leave min(max(K, 512), 2048) waves, N tile narrowed first.  A wave's
// dependent chain grows with K (its k steps run in order), so long
// contractions want more, shorter waves; short ones fewer, wider tiles
// (measured on the ten-product suite: 2048 x 512 best at 32 x 32, 512 x 2048
// at 16 x 16 for Blocked-ELL and 32 x 16 / 16 x 16 for 2:4).
void bell_shape(int M, int K, int N, int block, int tile_n, int *rpw, int *nt)
{
    const long min_units = std::min(std::max(K, 512), 2048);
    int r = std::min(block, 32), t = tile_n ? tile_n : 64;
    auto units = [&](int rr, int tt) { return (long)((M + rr - 1) / rr) * ((N + tt - 1) / tt); };
    if (!tile_n)
        while (t > 16 && units(r, t) < min_units) t >>= 1;
    while (r > 16 && units(r, t) < min_units) r >>= 1;
    *rpw = r;
    *nt = t;
}

tca_status launch_bell_spmm(int M, int K, int N, int block, int ell_cols, const int32_t *col_ind, tca_dtype a_dtype,
                            const void *values, tca_dtype b_dtype, const void *B, int ldb, float *C, int ldc,
                            int tile_n, int rows_per_wave, hipStream_t stream)
{
    int rpw = 0, nt = 0;
    bell_shape(M, K, N, block, tile_n, &rpw, &nt);
    if (rows_per_wave) rpw = rows_per_wave;
    if (rpw > block) return fail(TCA_ERR_INVALID_ARG, "tca_bell_spmm: rows_per_wave above the block size");
    BellFn fn = pick_bell(a_dtype, b_dtype, rpw, nt);
    if (!fn) return fail(TCA_ERR_UNSUPPORTED, "tca_bell_spmm: unsupported dtype pair or tile");
    if (M <= 0 || N <= 0) return TCA_OK;
    const int slots = ell_cols / block;
    const long ell_rows = (M + block - 1) / block;
    const long val_bytes = ell_rows * block * (long)ell_cols * 2;
    BellParams prm;
    prm.col_ind = col_ind;
    prm.values = values;
    prm.B = B;
    prm.C = C;
    prm.M = M;
    prm.K = K;
    prm.N = N;
    prm.ldb = ldb;
    prm.ldc = ldc;
    prm.b_shift = __builtin_ctz((unsigned)block);
    prm.slots = slots;
    prm.ell_cols = ell_cols;
    prm.steps = (ell_cols + 31) / 32;
    prm.n_tiles = (N + nt - 1) / nt;
    const long units = (long)((M + rpw - 1) / rpw) * prm.n_tiles;
    prm.b_bytes = operand_bytes(K, N, ldb, b_dtype == TCA_F32 ? 4 : 2);
    prm.val_bytes = val_bytes > 0 && val_bytes < (1L << 31) - 64 ? (uint32_t)val_bytes : (val_bytes == 0 ? 16u : 0u);
    prm.ind_bytes = ell_rows * slots > 0 ? (uint32_t)(ell_rows * slots * 4) : 16u;
    if (prm.b_bytes == 0 || prm.val_bytes == 0 || ell_rows * slots * 4 >= (1L << 31) || units >= (1L << 31))
        return fail(TCA_ERR_UNSUPPORTED, "tca_bell_spmm: B, values or col_ind reach 2 GiB (32-bit offsets)");
    prm.n_units = (int)units;
    hipLaunchKernelGGL(fn, dim3((unsigned)units), dim3(64), 0, stream, prm);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

}  // namespace tca
