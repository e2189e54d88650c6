// gfx950 kernels: panel-tiled SpMM (MFMA 16x16x32, FP32 accumulate), SDDMM,
// value gather/scatter and FP8 E4M3 quantisation.
//
// Data layout (built by build_tables):
//   rows are grouped RPW (16/32/64) at a time; group g's nonzero columns are
//   condensed and cut into 32-wide chunks (chunk_cols, -1 padded); each
//   (chunk, 16-row panel) segment lists entries as slot = (r16 << 5) | k with
//   the value at the same index of the tiled value array.
//
// One wave (64 lanes, one workgroup) owns one (group, N-tile of NT columns).
// Per chunk it gathers the 32 condensed B rows x NT columns into LDS in
// k-contiguous order (the MFMA B operand wants 8 consecutive k per lane),
// scatters the panel entries into a zeroed 16x32 LDS A tile, and issues
// PPG x NT/16 MFMAs.  Reference hot loop: src/our_spmm.cu:84-155 (gather,
// transpose into shared memory, mma.sync m16n8k8) -- here the operands are
// laid out for v_mfma_f32_16x16x32_{f16,bf16,fp8_fp8} on 64-wide waves.
#include "tca_device.hpp"

#include <algorithm>

namespace tca {
namespace {

struct SpmmParams {
    const int32_t *grp_chunk_ptr;
    const int32_t *chunk_cols;
    const int32_t *seg_ptr;
    const uint16_t *entry_slot;
    const uint16_t *entry_lds;   // entry_slot as element index into the LDS A tile (a_oct layout)
    const int32_t *entry_src;
    const int32_t *work_order;
    const void *a;
    const void *B;
    float *C;
    const float *row_scale;
    const float *col_scale;
    const int4 *units;       // standalone launch: (group, piece, c_lo, c_hi) per work unit
    int n_rows, N, ldb, ldc, n_tiles, n_groups, rpw, split, nnz;
    int n_units;             // standalone launch: entries of `units`
    int xcd_share;           // standalone launch: XCDs sharing one N-tile (see xcd_work)
    uint32_t b_bytes;        // bytes of B from its first to its last used element
    int n_chunks;            // chunks of the plan (chunk_cols holds 32 per chunk)
};

// Workgroup -> (unit, N-tile) for a 1-D standalone grid.  Workgroups are
// dispatched round-robin over the 8 XCDs (each with its own L2), so
// workgroup L runs on XCD L % 8.  Every workgroup of one N-tile reads the
// same gathered B columns, so the tiles are dealt to XCDs:
//   xcd_share == 0: n_tiles % 8 == 0, XCD x runs tiles x, x+8, ... (all units);
//   xcd_share == k (1..8, k * n_tiles == 8): k XCDs share a tile, taking its
//     units round-robin so that heavy and light units stay spread;
//   xcd_share == -2: units are dealt to XCDs instead (XCD x runs units x,
//     x+8, ..., each with all its N-tiles), for A operands that outweigh B:
//     a unit's A is then fetched by one L2, not by one per N-tile;
//   xcd_share == -1: plain order (tile = L / n_units).
// Returns false for the padding workgroups of a grid rounded up to 8.
__device__ __forceinline__ bool xcd_work(const SpmmParams &p, int L, int &unit, int &tile)
{
    if (p.xcd_share == -1) {
        tile = L / p.n_units;
        unit = L - tile * p.n_units;
        return true;
    }
    const int x = L & 7, l = L >> 3;
    if (p.xcd_share == -2) {
        const int r = l / p.n_tiles;
        tile = l - r * p.n_tiles;
        unit = x + 8 * r;
        return unit < p.n_units;
    }
    if (p.xcd_share == 0) {
        const int r = l / p.n_units;
        unit = l - r * p.n_units;
        tile = x + 8 * r;
    } else {
        tile = x % p.n_tiles;
        unit = l * p.xcd_share + x / p.n_tiles;
    }
    return unit < p.n_units && tile < p.n_tiles;
}

template <int ES> struct Store;
template <> struct Store<1> { using T = uint8_t; };
template <> struct Store<2> { using T = uint16_t; };
template <> struct Store<4> { using T = uint32_t; };
template <> struct Store<8> { using T = uint2; };
template <> struct Store<16> { using T = uint4; };

// Pack element e of KPL octets (k-consecutive) into KPL*ES bytes.
template <int KPL, int ES, class Oct>
__device__ inline typename Store<KPL * ES>::T pack_k(const Oct (&rows)[KPL], int e)
{
    constexpr int W = (KPL * ES + 3) / 4;
    uint32_t w[W > 0 ? W : 1] = {};
#pragma unroll
    for (int j = 0; j < KPL; j++) {
        const uint32_t v = oct_elem(rows[j], e);
        const int bit = j * ES * 8;
        w[bit / 32] |= v << (bit % 32);
    }
    typename Store<KPL * ES>::T out;
    if constexpr (KPL * ES == 1) out = (uint8_t)w[0];
    else if constexpr (KPL * ES == 2) out = (uint16_t)w[0];
    else if constexpr (KPL * ES == 4) out = w[0];
    else if constexpr (KPL * ES == 8) out = make_uint2(w[0], w[1]);
    else out = make_uint4(w[0], w[1], w[2], w[3]);
    return out;
}

constexpr int cmax(int a, int b) { return a > b ? a : b; }

// Entries prefetched per lane per panel and chunk (covers 128 entries of a
// 16x32 panel tile; denser tiles take an extra, unprefetched pass).
constexpr int kEP = 2;
// Chunks whose gathered data is in flight while one is computed.
constexpr int kDepth = 2;

// One wave = one (group g, N-tile `tile`, chunks [c_lo, c_hi) of g).  RPW rows
// (PPG = RPW/16 panels) x NT columns; lds_b holds the 32 x NT B tile, lds_a
// PPG*512 elements.  Software pipeline over a ring of kDepth register stages
// (indexed by compile-time slot, so nothing is copied between iterations):
//   index loads (column ids, segment bounds) run 2*kDepth chunks ahead,
//   data loads (gathered B rows, panel entries) run kDepth chunks ahead,
//   the current chunk is transposed / scattered into LDS and fed to the MFMAs.
// atomic: several waves share the output tile (chunk ranges split a group);
// the caller has zeroed C and the wave adds its partial sums.
// DENSE: p.a holds panel-dense tiles (every (chunk, panel) as a full 16 x 32
// tile in MFMA operand order, see pack_dense_kernel): the A operand octet of a
// lane is one vector load issued with the chunk's B rows, and neither the
// entry lists nor lds_a are touched.
template <int RPW, int NT, tca_dtype AT, tca_dtype BT, bool INDIRECT, bool DENSE = false>
__device__ __forceinline__ void spmm_unit(const SpmmParams &p, int g, int tile, int c_lo, int c_hi, uint8_t *lds_b,
                                          uint8_t *lds_a, f32x4 (&acc)[RPW / 16][NT / 16])
{
    static_assert(!(DENSE && INDIRECT), "panel-dense tiles are packed in the plan's own order");
    constexpr int PPG = RPW / 16;
    constexpr int ES = Elt<AT>::ES;
    constexpr int NSUB = NT / 16;
    constexpr int SEGS = NT / 8;       // 8-column segments of a B row tile
    constexpr int LPS = 64 / SEGS;     // lanes sharing one segment
    constexpr int KPL = 32 / LPS;      // consecutive k rows per lane (= NT/16)
    constexpr int OCT = 8 * ES;        // octet bytes
    constexpr int D = kDepth;
    // 16-bit types: B tile row-major [32 k][LDB], operands by transposing reads;
    // FP8: B tile [n][32 k] octets written k-packed (pack_k), plain operand reads
    constexpr bool TR = ES == 2;
    constexpr int LDB = tr_pitch<NT>();
    using Oct = typename Elt<AT>::Oct;
    using E = typename Store<ES>::T;

    struct Index {
        int cols[KPL];
        int segb;                      // lane j <= PPG: seg_ptr[c*PPG + j]
        int c;                         // the chunk (DENSE: its tiles' index)
    };
    using RL = RowLoad<AT, BT>;
    using RB = RowBuf<AT, BT>;
    struct Stage {
        typename RB::Raw rows[KPL];    // padded k rows and tail columns read as zeros
        int slot[PPG][kEP];            // raw entry slot (valid where eok has the bit)
        uint32_t eok;                  // bit i*kEP+k: slot[i][k] is a real entry
        uint32_t val[PPG][kEP];        // value bits, or entry_src when INDIRECT
        uint32_t vv[PPG][kEP];         // INDIRECT: value bits, loaded one chunk ahead
        int e0[PPG], e1[PPG];
        Oct aop[PPG];                  // DENSE: this lane's A operand octet of each panel
    };

    const int lane = lane_id();
    const int n0 = tile * NT;
    (void)g;

#pragma unroll
    for (int i = 0; i < PPG; i++)
#pragma unroll
        for (int s = 0; s < NSUB; s++) acc[i][s] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int seg = lane % SEGS, kq = lane / SEGS;
    const int k0 = kq * KPL;
    const int ncol = n0 + 8 * seg;
    const bool col_ok = ncol < p.N;
    const E *avals = static_cast<const E *>(p.a);
    const int seg_lane = lane <= PPG ? lane : PPG;
    const uint32_t nnz = (uint32_t)p.nnz;
    const Rsrc r_b = make_rsrc(p.B, p.b_bytes);
    const Rsrc r_cols = make_rsrc(p.chunk_cols, (uint32_t)p.n_chunks * 128u);
    const Rsrc r_lds = make_rsrc(p.entry_lds, nnz * 2u);
    const Rsrc r_val = make_rsrc(p.a, nnz * (uint32_t)ES);
    const Rsrc r_src = make_rsrc(p.entry_src, nnz * 4u);
    const Rsrc r_dense = make_rsrc(p.a, (uint32_t)p.n_chunks * (uint32_t)(PPG * 64 * OCT));
    const uint32_t ldb_bytes = (uint32_t)p.ldb * RB::BES;
    const uint32_t col_bytes = (uint32_t)ncol * RB::BES;

    if constexpr (!DENSE) {
#pragma unroll
        for (int i = 0; i < PPG; i++) reinterpret_cast<Oct *>(lds_a)[i * 64 + lane] = zero_oct<Oct>();
    }

    auto load_index = [&](int c, Index &ix) {
        const uint32_t cb = ((uint32_t)c * 32u + (uint32_t)k0) * 4u;
        if constexpr (KPL == 4) {
            const uint4 v = buf_b128(r_cols, cb);
            ix.cols[0] = (int)v.x; ix.cols[1] = (int)v.y; ix.cols[2] = (int)v.z; ix.cols[3] = (int)v.w;
        } else if constexpr (KPL == 2) {
            const uint2 v = buf_b64(r_cols, cb);
            ix.cols[0] = (int)v.x; ix.cols[1] = (int)v.y;
        } else {
#pragma unroll
            for (int j = 0; j < KPL; j++) ix.cols[j] = (int)buf_b32(r_cols, cb + 4u * j);
        }
        ix.c = c;
        if constexpr (!DENSE) ix.segb = gbl(p.seg_ptr)[c * PPG + seg_lane];
    };
    // Loads are issued unconditionally, results are only used at compute time:
    // masking a loaded value here would wait for it (and, the counter being in
    // order, for every load issued before it).
    auto load_data = [&](const Index &ix, Stage &st) {
        st.eok = 0;
#pragma unroll
        for (int j = 0; j < KPL; j++) {
            const bool ok = ix.cols[j] >= 0 && col_ok;
            st.rows[j] = RB::load(r_b, ok ? (uint32_t)ix.cols[j] * ldb_bytes + col_bytes : kOOB);
        }
        if constexpr (DENSE) {
#pragma unroll
            for (int i = 0; i < PPG; i++) {
                const uint32_t off = ((uint32_t)(ix.c * PPG + i) * 64u + (uint32_t)lane) * (uint32_t)OCT;
                if constexpr (OCT == 16) st.aop[i] = buf_b128(r_dense, off);
                else st.aop[i] = buf_b64(r_dense, off);
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < PPG; i++) {
            st.e0[i] = __shfl(ix.segb, i);
            st.e1[i] = __shfl(ix.segb, i + 1);
#pragma unroll
            for (int k = 0; k < kEP; k++) {
                const int e = st.e0[i] + k * 64 + lane;
                const bool ok = e < st.e1[i];
                const uint32_t eo = ok ? (uint32_t)e : (kOOB >> 2);
                st.slot[i][k] = (int)buf_b16(r_lds, eo * 2u);
                st.eok |= ok ? 1u << (i * kEP + k) : 0u;
                st.val[i][k] = INDIRECT ? buf_b32(r_src, eo * 4u) : load_value<AT>(r_val, eo);
            }
        }
    };
    // INDIRECT: the values behind entry_src, issued one chunk before use (the
    // entry_src words came a chunk earlier still, so the wait is partial)
    auto load_vals = [&](Stage &st) {
        if constexpr (INDIRECT) {
#pragma unroll
            for (int i = 0; i < PPG; i++)
#pragma unroll
                for (int k = 0; k < kEP; k++) st.vv[i][k] = load_value<AT>(r_val, st.val[i][k]);
        }
    };
    auto a_off = [&](int i, int x) { return (i * 512 + x) * ES; };   // x: entry_lds index
    auto compute_dense = [&](const Stage &st) {
        Oct rows[KPL];
#pragma unroll
        for (int j = 0; j < KPL; j++) rows[j] = RL::cvt(st.rows[j]);
        if constexpr (TR) {
#pragma unroll
            for (int j = 0; j < KPL; j++) *reinterpret_cast<Oct *>(lds_b + tr_row(k0 + j, LDB * 2) + 16 * seg) = rows[j];
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int n = 8 * seg + e;
                const int off = b_oct(n, k0 >> 3) * OCT + (k0 & 7) * ES;
                *reinterpret_cast<typename Store<KPL * ES>::T *>(lds_b + off) = pack_k<KPL, ES>(rows, e);
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            Oct b;
            if constexpr (TR) {
                const int row = 8 * (lane >> 4) + ((lane & 15) >> 2), col = (16 * s + 4 * (lane & 3)) * 2;
                const uint2 lo = lds_read_tr16(lds_b, tr_row(row, LDB * 2) + col);
                const uint2 hi = lds_read_tr16(lds_b, tr_row(row + 4, LDB * 2) + col);
                b = make_uint4(lo.x, lo.y, hi.x, hi.y);
            } else {
                b = reinterpret_cast<const Oct *>(lds_b)[b_oct(16 * s + (lane & 15), lane >> 4)];
            }
#pragma unroll
            for (int i = 0; i < PPG; i++) acc[i][s] = Elt<AT>::mma(st.aop[i], b, acc[i][s]);
        }
        wave_lds_sync();
    };
    auto compute = [&](const Stage &st) {
        if constexpr (DENSE) {
            compute_dense(st);
            return;
        }
        E v[PPG][kEP];
#pragma unroll
        for (int i = 0; i < PPG; i++)
#pragma unroll
            for (int k = 0; k < kEP; k++) v[i][k] = INDIRECT ? (E)st.vv[i][k] : (E)st.val[i][k];
        Oct rows[KPL];
#pragma unroll
        for (int j = 0; j < KPL; j++) rows[j] = RL::cvt(st.rows[j]);
        if constexpr (TR) {
#pragma unroll
            for (int j = 0; j < KPL; j++) *reinterpret_cast<Oct *>(lds_b + tr_row(k0 + j, LDB * 2) + 16 * seg) = rows[j];
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int n = 8 * seg + e;
                const int off = b_oct(n, k0 >> 3) * OCT + (k0 & 7) * ES;
                *reinterpret_cast<typename Store<KPL * ES>::T *>(lds_b + off) = pack_k<KPL, ES>(rows, e);
            }
        }
        bool overflow[PPG];
#pragma unroll
        for (int i = 0; i < PPG; i++) {
#pragma unroll
            for (int k = 0; k < kEP; k++)
                if ((st.eok >> (i * kEP + k)) & 1u) *reinterpret_cast<E *>(lds_a + a_off(i, st.slot[i][k])) = v[i][k];
            overflow[i] = st.e1[i] - st.e0[i] > kEP * 64;
            if (overflow[i])
                for (int e = st.e0[i] + kEP * 64 + lane; e < st.e1[i]; e += 64) {
                    const E vv = INDIRECT ? gbl(avals)[gbl(p.entry_src)[e]] : gbl(avals)[e];
                    *reinterpret_cast<E *>(lds_a + a_off(i, gbl(p.entry_lds)[e])) = vv;
                }
        }
        wave_lds_sync();
        Oct a[PPG];
#pragma unroll
        for (int i = 0; i < PPG; i++) a[i] = reinterpret_cast<const Oct *>(lds_a)[i * 64 + a_oct(lane & 15, lane >> 4)];
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            Oct b;
            if constexpr (TR) {
                const int row = 8 * (lane >> 4) + ((lane & 15) >> 2), col = (16 * s + 4 * (lane & 3)) * 2;
                const uint2 lo = lds_read_tr16(lds_b, tr_row(row, LDB * 2) + col);
                const uint2 hi = lds_read_tr16(lds_b, tr_row(row + 4, LDB * 2) + col);
                b = make_uint4(lo.x, lo.y, hi.x, hi.y);
            } else {
                b = reinterpret_cast<const Oct *>(lds_b)[b_oct(16 * s + (lane & 15), lane >> 4)];
            }
#pragma unroll
            for (int i = 0; i < PPG; i++) acc[i][s] = Elt<AT>::mma(a[i], b, acc[i][s]);
        }
        wave_lds_sync();
        // un-scatter: put the zeros back (a whole tile when it overflowed)
#pragma unroll
        for (int i = 0; i < PPG; i++) {
            if (overflow[i]) {
                reinterpret_cast<Oct *>(lds_a)[i * 64 + lane] = zero_oct<Oct>();
            } else {
#pragma unroll
                for (int k = 0; k < kEP; k++)
                    if ((st.eok >> (i * kEP + k)) & 1u) *reinterpret_cast<E *>(lds_a + a_off(i, st.slot[i][k])) = (E)0;
            }
        }
    };

    // Prefetches past the last chunk re-read the last chunk (clamped index),
    // so every load is unconditional; only the compute step is guarded.
    if (c_lo < c_hi) {
        const int c_last = c_hi - 1;
        Index ix[D];
        Stage st[D];
#pragma unroll
        for (int s = 0; s < D; s++) load_index(min(c_lo + s, c_last), ix[s]);
#pragma unroll
        for (int s = 0; s < D; s++) {
            load_data(ix[s], st[s]);
            load_index(min(c_lo + s + D, c_last), ix[s]);
        }
        load_vals(st[0]);
        vm_drain();
        for (int c0 = c_lo; c0 < c_hi; c0 += D) {
#pragma unroll
            for (int s = 0; s < D; s++) {
                const int c = c0 + s;
                if (c < c_hi) compute(st[s]);
                load_data(ix[s], st[s]);
                load_index(min(c + 2 * D, c_last), ix[s]);
                load_vals(st[(s + 1) % D]);
            }
        }
    }
    wave_lds_sync();  // this wave's LDS tiles are clean for its next use
}

// FP32 stores (or atomic adds) of one wave's accumulator tiles: acc[i][s] is
// the 16x16 tile of rows row0 + 16i .. and columns n0 + 16s .., its lane
// holding D[4*(lane/16)+j][lane%16].  The uniform flags pick one of three loops
// (plain stores; scaled stores; scaled atomic adds) so that no element pays a
// branch, or a scale load, that it does not need.
template <int NI, int NS>
__device__ __forceinline__ void store_tiles(const SpmmParams &p, int row0, int n0, const f32x4 (&acc)[NI][NS],
                                            bool atomic)
{
    const int lane = lane_id();
    const int rbase = row0 + 4 * (lane >> 4);
    auto run = [&](auto scaled, auto add) {
#pragma unroll
        for (int i = 0; i < NI; i++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = rbase + 16 * i + j;
                if (row >= p.n_rows) continue;
                float rs = 1.0f;
                if constexpr (decltype(scaled)::value) rs = p.row_scale ? gbl(p.row_scale)[row] : 1.0f;
                float *crow = p.C + (long)row * p.ldc;
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    const int n = n0 + 16 * s + (lane & 15);
                    if (n < p.N) {
                        float v = acc[i][s][j];
                        if constexpr (decltype(scaled)::value) {
                            v *= rs;
                            if (p.col_scale) v *= gbl(p.col_scale)[n];
                        }
                        if constexpr (decltype(add)::value) atomic_add_f32(crow + n, v);
                        else gbl(crow)[n] = v;
                    }
                }
            }
        }
    };
    if (atomic) run(std::true_type{}, std::true_type{});
    else if (p.row_scale || p.col_scale) run(std::true_type{}, std::false_type{});
    else run(std::false_type{}, std::false_type{});
}

template <int RPW, int NT>
__device__ __forceinline__ void spmm_store_direct(const SpmmParams &p, int g, int tile,
                                                  const f32x4 (&acc)[RPW / 16][NT / 16], bool atomic)
{
    store_tiles(p, g * RPW, tile * NT, acc, atomic);
}

// LDS bytes one wave of spmm_kernel needs: its B and A tiles during the chunk
// loop, then its partial RPW x NT tile (rows padded by 4 floats).
template <int RPW, int NT, tca_dtype AT>
struct SpmmLds {
    static constexpr int B = Elt<AT>::ES == 2 ? 32 * tr_pitch<NT>() * 2 : NT * 32 * Elt<AT>::ES;
    static constexpr int A = (RPW / 16) * 512 * Elt<AT>::ES;
    static constexpr int RED_LD = NT + 4;
    static constexpr int RED = RPW * RED_LD * 4;
    static constexpr int SLOT = (B + A) > RED ? (B + A) : RED;
    static constexpr int WMAX = 65536 / SLOT >= 8 ? 8 : (65536 / SLOT >= 1 ? 65536 / SLOT : 1);
};

// One workgroup of W = blockDim.x/64 waves on one (group g, N-tile, chunk
// range [c_lo, c_hi)): the waves split the range evenly, each into its own LDS
// slot of `slot` bytes; the W partial tiles are then summed in wave order
// (deterministic) and stored as float4 rows, or added atomically (`atomic`:
// other workgroups hold other chunk ranges of the group; C was zeroed).
template <int RPW, int NT, tca_dtype AT, tca_dtype BT, bool INDIRECT, bool DENSE = false>
__device__ __forceinline__ void spmm_wg_unit(const SpmmParams &p, int g, int tile, int c_lo, int c_hi, bool atomic,
                                             uint8_t *smem, int slot)
{
    using L = SpmmLds<RPW, NT, AT>;
    const int W = blockDim.x >> 6, w = threadIdx.x >> 6, lane = lane_id();
    uint8_t *mine = smem + w * slot;
    const int nch = c_hi - c_lo;
    const int lo = c_lo + (nch * w) / W, hi = c_lo + (nch * (w + 1)) / W;
    f32x4 acc[RPW / 16][NT / 16];
    spmm_unit<RPW, NT, AT, BT, INDIRECT, DENSE>(p, g, tile, lo, hi, mine, mine + L::B, acc);

    float *red = reinterpret_cast<float *>(mine);
#pragma unroll
    for (int i = 0; i < RPW / 16; i++)
#pragma unroll
        for (int s = 0; s < NT / 16; s++)
#pragma unroll
            for (int j = 0; j < 4; j++) red[(16 * i + 4 * (lane >> 4) + j) * L::RED_LD + 16 * s + (lane & 15)] = acc[i][s][j];
    __syncthreads();
    const int n0 = tile * NT;
    for (int idx = threadIdx.x * 4; idx < RPW * NT; idx += blockDim.x * 4) {
        const int r = idx / NT, c = idx - r * NT;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int q = 0; q < W; q++) {
            const float4 t = *reinterpret_cast<const float4 *>(reinterpret_cast<const float *>(smem + q * slot) +
                                                               r * L::RED_LD + c);
            v.x += t.x;
            v.y += t.y;
            v.z += t.z;
            v.w += t.w;
        }
        const int row = g * RPW + r, n = n0 + c;
        if (row >= p.n_rows || n >= p.N) continue;
        const float rs = p.row_scale ? gbl(p.row_scale)[row] : 1.0f;
        v.x *= rs;
        v.y *= rs;
        v.z *= rs;
        v.w *= rs;
        if (p.col_scale) {
            const auto cs = gbl(p.col_scale);
            v.x *= cs[n];
            v.y *= cs[n + 1];
            v.z *= cs[n + 2];
            v.w *= cs[n + 3];
        }
        float *dst = p.C + (long)row * p.ldc + n;
        if (atomic) {
            atomic_add_f32(dst, v.x);
            atomic_add_f32(dst + 1, v.y);
            atomic_add_f32(dst + 2, v.z);
            atomic_add_f32(dst + 3, v.w);
        } else {
            *gbl(reinterpret_cast<f32x4 *>(dst)) = f32x4{v.x, v.y, v.z, v.w};
        }
    }
}

// Standalone launch: workgroup -> (unit of the plan's unit table, N-tile) by
// xcd_work, run by spmm_wg_unit.  Units of a group split across workgroups
// (p.split) accumulate atomically into a zeroed C.
template <int RPW, int NT, tca_dtype AT, tca_dtype BT, bool INDIRECT, bool DENSE = false>
__global__ __launch_bounds__((64 * SpmmLds<RPW, NT, AT>::WMAX)) void spmm_kernel(SpmmParams p)
{
    using L = SpmmLds<RPW, NT, AT>;
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    int ui, tile;
    if (!xcd_work(p, blockIdx.x, ui, tile)) return;
    const int4 u = p.units[ui];   // (group, pieces, c_lo, c_hi)
    spmm_wg_unit<RPW, NT, AT, BT, INDIRECT, DENSE>(p, u.x, tile, u.z, u.w, p.split != 0, smem, L::SLOT);
}

// ---------------------------------------------------------------------------
// Wide groups (rows_per_wave 128 / 256): one workgroup of WM waves owns a
// (unit, N-tile) of a group of WM*PPW*16 rows, wave w owning panels
// [w*PPW, (w+1)*PPW).  The waves share every gathered B chunk: the workgroup
// loads the 32 x NT tile once (16-byte row pieces, one per thread and pass)
// into LDS in row-major order with a pitch of NT+16 elements -- four
// consecutive rows then start in distinct quarters of the 32 banks -- and
// each wave reads its MFMA B operands with ds_read_b64_tr_b16: lane 4r+j of
// a 16-lane group addresses row r, columns 4j..4j+3, and lane i receives
// column i of the four rows, i.e. 4 consecutive k of one n.  Two reads (rows
// 8q..8q+3 and 8q+4..8q+7 for lane group q) give the 8-k operand octet.
// B tiles are double buffered with one workgroup barrier per chunk; B row
// pieces, index words and panel entries are prefetched kWideDepth chunks
// ahead in a register ring.  A tiles are per wave, as in spmm_unit.  Per
// group and chunk the B rows are read once for all WM*PPW*16 rows instead of
// once per 16..64-row wave, which is what the wide plans exist for.
// ---------------------------------------------------------------------------
constexpr int kWideDepth = 4;
constexpr int kWideEP = 2;   // prefetched entries per lane per panel (128 of a 16x32 tile)

template <int NT, int WM, int PPW>
struct WideLds {
    static constexpr int LDB = NT + 16;                          // B row pitch (elements)
    static constexpr int BTILE = 32 * LDB * 2;                   // bytes per chunk tile (rows by tr_row)
    static constexpr int ATILE = PPW * 512 * 2;                  // bytes of one wave's A tiles
    static constexpr int BYTES = 2 * BTILE + WM * ATILE;
    static constexpr int SEGS = NT / 8;                          // 16-byte pieces per B row
    static constexpr int PIECES = 32 * SEGS;
    static constexpr int PPT = (PIECES + WM * 64 - 1) / (WM * 64);  // pieces per thread
};

template <int PPW, int NT, int WM, tca_dtype AT, tca_dtype BT, bool INDIRECT, bool DENSE = false>
__global__ __launch_bounds__(64 * WM) void spmm_wide_kernel(SpmmParams p)
{
    static_assert(!(DENSE && INDIRECT), "panel-dense tiles are packed in the plan's own order");
    using L = WideLds<NT, WM, PPW>;
    using RL = RowLoad<AT, BT>;
    using RB = RowBuf<AT, BT>;
    using Oct = uint4;
    using E = uint16_t;
    constexpr int PPG = WM * PPW;       // panels of the plan's group
    constexpr int RPG = PPG * 16;
    constexpr int NSUB = NT / 16;
    constexpr int D = kWideDepth;
    constexpr int EP = kWideEP;
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];

    const int w = threadIdx.x >> 6, lane = lane_id();
    const int tid = w * 64 + lane;
    int ui, tile;
    if (!xcd_work(p, blockIdx.x, ui, tile)) return;
    const int4 u = p.units[ui];   // (group, pieces, c_lo, c_hi)
    const int g = u.x, c_lo = u.z, c_hi = u.w;
    const int n0 = tile * NT;
    uint8_t *lds_a = smem + 2 * L::BTILE + w * L::ATILE;
    const E *avals = static_cast<const E *>(p.a);
    const uint32_t nnz = (uint32_t)p.nnz;
    const Rsrc r_b = make_rsrc(p.B, p.b_bytes);
    const Rsrc r_lds = make_rsrc(p.entry_lds, nnz * 2u);
    const Rsrc r_val = make_rsrc(p.a, nnz * 2u);
    const Rsrc r_src = make_rsrc(p.entry_src, nnz * 4u);
    const Rsrc r_dense = make_rsrc(p.a, (uint32_t)p.n_chunks * (uint32_t)(PPG * 64 * 16));
    const uint32_t ldb_bytes = (uint32_t)p.ldb * RB::BES;

    // this thread's B pieces: row k, 8-column segment
    int pk[L::PPT], pseg[L::PPT];
    bool pvalid[L::PPT];
#pragma unroll
    for (int j = 0; j < L::PPT; j++) {
        const int pc = tid + j * WM * 64;
        pvalid[j] = pc < L::PIECES;
        pk[j] = pvalid[j] ? piece_row<L::SEGS>(pc / L::SEGS) : 0;
        pseg[j] = pvalid[j] ? pc % L::SEGS : 0;
    }
    const int seg_lane = lane <= PPW ? lane : PPW;

    struct Index {
        int col[L::PPT];
        int segb;
        int c;                         // the chunk (DENSE: its tiles' index)
    };
    struct Stage {
        typename RB::Raw rows[L::PPT];  // absent rows / tail columns read as zeros
        int slot[PPW][EP];             // raw entry slot (valid where eok has the bit)
        uint32_t eok;
        uint32_t val[PPW][EP];         // value bits, or entry_src when INDIRECT
        uint32_t vv[PPW][EP];          // INDIRECT: value bits, loaded one chunk ahead
        int e0[PPW], e1[PPW];
        Oct aop[PPW];                  // DENSE: this lane's A operand octet of each of its panels
    };

    f32x4 acc[PPW][NSUB];
#pragma unroll
    for (int i = 0; i < PPW; i++)
#pragma unroll
        for (int s = 0; s < NSUB; s++) acc[i][s] = f32x4{0.f, 0.f, 0.f, 0.f};
    if constexpr (!DENSE) {
#pragma unroll
        for (int i = 0; i < PPW; i++) reinterpret_cast<Oct *>(lds_a)[i * 64 + lane] = zero_oct<Oct>();
    }

    auto load_index = [&](int c, Index &ix) {
#pragma unroll
        for (int j = 0; j < L::PPT; j++) ix.col[j] = p.chunk_cols[(long)c * 32 + pk[j]];
        ix.c = c;
        if constexpr (!DENSE) ix.segb = gbl(p.seg_ptr)[(long)c * PPG + w * PPW + seg_lane];
    };
    auto load_data = [&](const Index &ix, Stage &st) {
        st.eok = 0;
#pragma unroll
        for (int j = 0; j < L::PPT; j++) {
            const int ncol = n0 + 8 * pseg[j];
            const bool ok = pvalid[j] && ix.col[j] >= 0 && ncol < p.N;
            st.rows[j] = RB::load(r_b, ok ? (uint32_t)ix.col[j] * ldb_bytes + (uint32_t)ncol * RB::BES : kOOB);
        }
        if constexpr (DENSE) {
#pragma unroll
            for (int i = 0; i < PPW; i++)
                st.aop[i] = buf_b128(r_dense, ((uint32_t)(ix.c * PPG + w * PPW + i) * 64u + (uint32_t)lane) * 16u);
            return;
        }
#pragma unroll
        for (int i = 0; i < PPW; i++) {
            st.e0[i] = __shfl(ix.segb, i);
            st.e1[i] = __shfl(ix.segb, i + 1);
#pragma unroll
            for (int k = 0; k < EP; k++) {
                const int e = st.e0[i] + k * 64 + lane;
                const bool ok = e < st.e1[i];
                const uint32_t eo = ok ? (uint32_t)e : (kOOB >> 2);
                st.slot[i][k] = (int)buf_b16(r_lds, eo * 2u);
                st.eok |= ok ? 1u << (i * EP + k) : 0u;
                st.val[i][k] = INDIRECT ? buf_b32(r_src, eo * 4u) : buf_b16(r_val, eo * 2u);
            }
        }
    };
    auto load_vals = [&](Stage &st) {
        if constexpr (INDIRECT) {
#pragma unroll
            for (int i = 0; i < PPW; i++)
#pragma unroll
                for (int k = 0; k < EP; k++) st.vv[i][k] = buf_b16(r_val, st.val[i][k] * 2u);
        }
    };
    auto a_off = [&](int i, int x) { return (i * 512 + x) * 2; };   // x: entry_lds index
    // B operand addressing inside a tile (bytes): lane group q = lane/16, lane i = lane%16
    const int brow = 8 * (lane >> 4) + ((lane & 15) >> 2);
    const int bcol = 4 * (lane & 3);
    auto put_b = [&](const Stage &st, uint8_t *buf) {
#pragma unroll
        for (int j = 0; j < L::PPT; j++)
            if (pvalid[j]) {
                const Oct v = RL::cvt(st.rows[j]);
                *reinterpret_cast<Oct *>(buf + tr_row(pk[j], L::LDB * 2) + 16 * pseg[j]) = v;
            }
    };
    auto compute_dense = [&](const Stage &st, const uint8_t *buf) {
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            const int col = (16 * s + bcol) * 2;
            const uint2 lo = lds_read_tr16(buf, tr_row(brow, L::LDB * 2) + col);
            const uint2 hi = lds_read_tr16(buf, tr_row(brow + 4, L::LDB * 2) + col);
            const Oct b = make_uint4(lo.x, lo.y, hi.x, hi.y);
#pragma unroll
            for (int i = 0; i < PPW; i++) acc[i][s] = Elt<AT>::mma(st.aop[i], b, acc[i][s]);
        }
    };
    auto compute = [&](const Stage &st, const uint8_t *buf) {
        if constexpr (DENSE) {
            compute_dense(st, buf);
            return;
        }
        E v[PPW][EP];
#pragma unroll
        for (int i = 0; i < PPW; i++)
#pragma unroll
            for (int k = 0; k < EP; k++) v[i][k] = INDIRECT ? (E)st.vv[i][k] : (E)st.val[i][k];
        bool overflow[PPW];
#pragma unroll
        for (int i = 0; i < PPW; i++) {
#pragma unroll
            for (int k = 0; k < EP; k++)
                if ((st.eok >> (i * EP + k)) & 1u) *reinterpret_cast<E *>(lds_a + a_off(i, st.slot[i][k])) = v[i][k];
            overflow[i] = st.e1[i] - st.e0[i] > EP * 64;
            if (overflow[i])
                for (int e = st.e0[i] + EP * 64 + lane; e < st.e1[i]; e += 64) {
                    const E vv = INDIRECT ? gbl(avals)[gbl(p.entry_src)[e]] : gbl(avals)[e];
                    *reinterpret_cast<E *>(lds_a + a_off(i, gbl(p.entry_lds)[e])) = vv;
                }
        }
        wave_lds_sync();
        Oct a[PPW];
#pragma unroll
        for (int i = 0; i < PPW; i++) a[i] = reinterpret_cast<const Oct *>(lds_a)[i * 64 + a_oct(lane & 15, lane >> 4)];
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            const int col = (16 * s + bcol) * 2;
            const uint2 lo = lds_read_tr16(buf, tr_row(brow, L::LDB * 2) + col);
            const uint2 hi = lds_read_tr16(buf, tr_row(brow + 4, L::LDB * 2) + col);
            const Oct b = make_uint4(lo.x, lo.y, hi.x, hi.y);
#pragma unroll
            for (int i = 0; i < PPW; i++) acc[i][s] = Elt<AT>::mma(a[i], b, acc[i][s]);
        }
        wave_lds_sync();
#pragma unroll
        for (int i = 0; i < PPW; i++) {
            if (overflow[i]) {
                reinterpret_cast<Oct *>(lds_a)[i * 64 + lane] = zero_oct<Oct>();
            } else {
#pragma unroll
                for (int k = 0; k < EP; k++)
                    if ((st.eok >> (i * EP + k)) & 1u) *reinterpret_cast<E *>(lds_a + a_off(i, st.slot[i][k])) = (E)0;
            }
        }
    };

    if (c_lo < c_hi) {
        const int c_last = c_hi - 1;
        Index ix[D];
        Stage st[D];
#pragma unroll
        for (int s = 0; s < D; s++) load_index(min(c_lo + s, c_last), ix[s]);
#pragma unroll
        for (int s = 0; s < D; s++) {
            load_data(ix[s], st[s]);
            load_index(min(c_lo + s + D, c_last), ix[s]);
        }
        load_vals(st[0]);
        vm_drain();
        for (int c0 = c_lo; c0 < c_hi; c0 += D) {
#pragma unroll
            for (int s = 0; s < D; s++) {
                const int c = c0 + s;
                if (c < c_hi) {
                    uint8_t *buf = smem + (c & 1) * L::BTILE;
                    put_b(st[s], buf);
                    __syncthreads();
                    compute(st[s], buf);
                }
                load_data(ix[s], st[s]);
                load_index(min(c + 2 * D, c_last), ix[s]);
                load_vals(st[(s + 1) % D]);
            }
        }
    }

    // -> C (atomic when the group is split)
    store_tiles(p, g * RPG + w * PPW * 16, n0, acc, p.split != 0);
}

// ---------------------------------------------------------------------------
// Block units on panel-dense tiles (plans of 64 / 128 / 256 rows per group).
// One workgroup of KS x WM waves owns a (group, N-tile of NT columns, chunk
// range).  The WM waves of a K group share every 32 x NT B chunk tile: it is
// loaded once per K group (one or two 16-byte pieces per thread), written to
// LDS row-major with a pitch of NT+16 elements and read back as MFMA B operands
// with ds_read_b64_tr_b16 (as in the wide kernel), double buffered with one
// workgroup barrier per chunk.  Wave w of a K group owns panels
// [w*PPW, (w+1)*PPW) and all NT columns, its A operand octets come straight
// from the dense tiles.  The KS K groups split the chunk range (each with its
// own LDS buffers) and hand their partial tiles to group 0 through LDS, panel
// by panel, so long groups keep several waves per SIMD busy without atomics.
// Per chunk a K group reads 64*NT bytes of B and PPG KiB of A for PPG*NT/16
// MFMAs, against 64*NT + 1 KiB per panel for one wave per 16..64 rows.
// ---------------------------------------------------------------------------
template <int NT, int WM, int ES = 2>
struct BlockLds {
    // B row pitch (bytes): 16-bit rows NT+16 elements (rows by tr_row); 8-bit
    // rows NT+16 bytes, so that the eight 16-byte row windows a transposing
    // 8-bit read takes per lane group, and the next group's eight, fall on
    // distinct banks
    static constexpr int PITCH = (NT + 16) * ES;
    static constexpr int BTILE = 32 * PITCH;             // one chunk's B tile (bytes)
    static constexpr int SEGS = NT / 8;                  // 8-element pieces per B row
    static constexpr int PPT = 32 * SEGS / (WM * 64);    // pieces per thread
    static_assert(PPT >= 1 && (32 * SEGS) % (WM * 64) == 0, "B pieces must divide over the waves");
};
template <int PPW, int NT, int WM, int KS, int ES = 2>
constexpr int block_lds_bytes()
{
    return cmax(KS * 2 * BlockLds<NT, WM, ES>::BTILE, (KS - 1) * WM * (NT / 16) * 1024);
}

// Row of the m-th row slot when a workgroup writes an 8-bit B tile (pitch
// 8*SEGS + 16 bytes) in 8-byte pieces: the SEGS pieces of a row are
// consecutive lanes, so a 256-byte write cycle covers 32 / SEGS rows, which
// sit s rows apart (s = 4 for 64-byte rows, 8 for 128-byte rows) so that their
// row spans tile the 64 banks exactly.  A permutation of the 32 rows.
template <int SEGS> __device__ __forceinline__ int piece_row8(int m)
{
    constexpr int HW = 32 / SEGS;            // rows per write cycle
    constexpr int S = SEGS == 8 ? 4 : 8;     // their spacing
    const int h = m / HW, i = m % HW;
    return h % S + S * HW * (h / S) + S * i;
}

// Transposing 8-bit LDS read (ds_read_b64_tr_b8, gfx950): in each 16-lane
// group, lane 2r+j addresses row r, bytes 8j..8j+7 of an 8-row x 16-byte
// block, and lane i receives byte i of the eight rows -- 8 consecutive k of
// one n, an FP8 MFMA B operand (lane layout checked on the MI355X).
typedef int i32x2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint2 lds_read_tr8(const uint8_t *lds, int byte_off)
{
    const i32x2v v = __builtin_amdgcn_ds_read_tr8_b64_v2i32((__attribute__((address_space(3))) i32x2v *)(lds + byte_off));
    return __builtin_bit_cast(uint2, v);
}

template <int PPW, int NT, int WM, int KS, int D, tca_dtype AT, tca_dtype BT>
__device__ __forceinline__ void spmm_block_unit(const SpmmParams &p, int g, int tile, int u_lo, int u_hi, bool atomic,
                                                uint8_t *smem_all)
{
    constexpr int ES = Elt<AT>::ES;                     // 2: F16 / BF16, 1: E4M3 (A and B)
    using L = BlockLds<NT, WM, ES>;
    using RL = RowLoad<AT, BT>;
    using RB = RowBuf<AT, BT>;
    using Oct = typename Elt<AT>::Oct;
    constexpr int NSUB = NT / 16;
    constexpr int PPT = L::PPT;
    constexpr int PPG = WM * PPW;                       // panels of the plan's group
    constexpr int RED = WM * NSUB * 1024;               // one panel of a K group's waves (bytes)
    constexpr int ATILE = 512 * ES;                     // one 16 x 32 A tile (bytes)
    static_assert(ES == 2 || BT == AT, "8-bit block units take an 8-bit B");
    static_assert(D % 2 == 0, "the LDS buffer of a step is its register stage's parity");
    const int wv = threadIdx.x >> 6, lane = lane_id();
    const int h = wv / WM, w = wv - h * WM;
    const int tid = w * 64 + lane;
    uint8_t *smem = smem_all + h * 2 * L::BTILE;
    const int nch = u_hi - u_lo;
    const int c_lo = u_lo + (nch * h) / KS, c_hi = u_lo + (nch * (h + 1)) / KS;
    const int iters = (nch + KS - 1) / KS;              // barrier steps of every K group
    const int n0 = tile * NT;
    const Rsrc r_b = make_rsrc(p.B, p.b_bytes);
    const Rsrc r_cols = make_rsrc(p.chunk_cols, (uint32_t)p.n_chunks * 128u);
    const Rsrc r_a = make_rsrc(p.a, (uint32_t)p.n_chunks * (uint32_t)(PPG * ATILE));
    const uint32_t ldb_bytes = (uint32_t)p.ldb * RB::BES;
    // this thread's B pieces: row pk, 8-column segment, and where they go in LDS
    int pk[PPT];
    uint32_t pcol[PPT], plds[PPT];
    bool pok[PPT];
#pragma unroll
    for (int j = 0; j < PPT; j++) {
        const int pc = tid + j * WM * 64;
        pk[j] = ES == 2 ? piece_row<L::SEGS>(pc / L::SEGS) : piece_row8<L::SEGS>(pc / L::SEGS);
        const int seg = pc % L::SEGS;
        const int ncol = n0 + 8 * seg;
        pok[j] = ncol < p.N;
        pcol[j] = (uint32_t)ncol * RB::BES;
        plds[j] = (uint32_t)(ES == 2 ? tr_row(pk[j], L::PITCH) + 16 * seg : pk[j] * L::PITCH + 8 * seg);
    }
    // MFMA B operand reads: 16-bit, two 4-row transposing reads per 16 columns;
    // 8-bit, one 8-row read per 16 columns (lane group q: rows 8q..8q+7)
    const int brow = 8 * (lane >> 4) + ((lane & 15) >> 2);
    const int boff_lo = tr_row(brow, L::PITCH) + 8 * (lane & 3), boff_hi = tr_row(brow + 4, L::PITCH) + 8 * (lane & 3);
    const int boff8 = (8 * (lane >> 4) + ((lane & 15) >> 1)) * L::PITCH + 8 * (lane & 1);
    const uint32_t a_lane = (uint32_t)(w * PPW) * (uint32_t)ATILE + (uint32_t)lane * (uint32_t)(8 * ES);

    f32x4 acc[PPW][NSUB];
#pragma unroll
    for (int i = 0; i < PPW; i++)
#pragma unroll
        for (int s = 0; s < NSUB; s++) acc[i][s] = f32x4{0.f, 0.f, 0.f, 0.f};

    // Software pipeline over the K group's chunks, one barrier per step.  At
    // step k (chunk c_lo + k): the barrier; the B pieces of chunk k+2 go to LDS
    // buffer k&1 (its chunk k was read out at step k-1); the MFMA operands of
    // chunk k+1 are read from buffer (k+1)&1 (written at step k-1) into
    // registers; the MFMAs of chunk k run on operands read at step k-1, so they
    // overlap this step's LDS traffic and the next barrier.  Global loads run D
    // steps ahead: B pieces of chunk k+2+D, A tiles of chunk k+D, column
    // indices of chunk k+2+2D.
    struct BFrag {
        Oct b[NSUB];
    };
    typename RB::Raw pb[D][PPT];   // B pieces, chunk k+2 at step k (k % D)
    Oct pa[D][PPW];                // A tiles, chunk k at step k
    int ix[D][PPT];                // column indices, chunk k+2+D at step k
    BFrag fr[2];                   // MFMA B operands, chunk k at step k (k & 1)
    auto load_index = [&](int c, int (&x)[PPT]) {
#pragma unroll
        for (int j = 0; j < PPT; j++) x[j] = (int)buf_b32(r_cols, ((uint32_t)c * 32u + (uint32_t)pk[j]) * 4u);
    };
    auto load_b = [&](const int (&x)[PPT], typename RB::Raw (&rows)[PPT]) {
#pragma unroll
        for (int j = 0; j < PPT; j++) {
            const bool ok = x[j] >= 0 && pok[j];
            rows[j] = RB::load(r_b, ok ? (uint32_t)x[j] * ldb_bytes + pcol[j] : kOOB);
        }
    };
    auto load_a = [&](int c, Oct (&aop)[PPW]) {
#pragma unroll
        for (int i = 0; i < PPW; i++) {
            const uint32_t off = (uint32_t)c * (uint32_t)(PPG * ATILE) + a_lane + (uint32_t)(i * ATILE);
            if constexpr (ES == 2) aop[i] = buf_b128(r_a, off);
            else aop[i] = buf_b64(r_a, off);
        }
    };
    auto put_b = [&](uint8_t *buf, const typename RB::Raw (&rows)[PPT]) {
#pragma unroll
        for (int j = 0; j < PPT; j++) *reinterpret_cast<Oct *>(buf + plds[j]) = RL::cvt(rows[j]);
    };
    auto read_frag = [&](const uint8_t *buf, BFrag &f) {
#pragma unroll
        for (int sb = 0; sb < NSUB; sb++) {
            if constexpr (ES == 2) {
                const uint2 lo = lds_read_tr16(buf, boff_lo + 32 * sb);
                const uint2 hi = lds_read_tr16(buf, boff_hi + 32 * sb);
                f.b[sb] = make_uint4(lo.x, lo.y, hi.x, hi.y);
            } else {
                f.b[sb] = lds_read_tr8(buf, boff8 + 16 * sb);
            }
        }
    };
    // Every K group runs `iters` steps (the shorter ones idle in their last);
    // loads past a range re-read its last chunk (or, for an empty range, its
    // first, always a chunk of the plan or out of range for the buffer
    // descriptors, which read zeros), so every load is unconditional.
    const int c_last = max(c_hi - 1, c_lo);
    auto cc = [&](int k) { return min(c_lo + k, c_last); };
    {
        int x0[PPT], x1[PPT];
        typename RB::Raw b0[PPT], b1[PPT];
        load_index(cc(0), x0);
        load_index(cc(1), x1);
#pragma unroll
        for (int s = 0; s < D; s++) load_index(cc(s + 2), ix[s]);
        load_b(x0, b0);
        load_b(x1, b1);
        // the stages' loads in the loop's own order (B, A, index per stage), so
        // that the waits the compiler derives in the loop from its entry and
        // back-edge states together stay those of the steady state
#pragma unroll
        for (int s = 0; s < D; s++) {
            __builtin_amdgcn_sched_barrier(0);
            load_b(ix[s], pb[s]);
            __builtin_amdgcn_sched_barrier(0);
            load_a(cc(s), pa[s]);
            __builtin_amdgcn_sched_barrier(0);
            load_index(cc(s + 2 + D), ix[s]);
        }
        __builtin_amdgcn_sched_barrier(0);
        put_b(smem, b0);
        put_b(smem + L::BTILE, b1);
        __syncthreads();
        read_frag(smem, fr[0]);
    }
    auto step = [&](int k, int s) {   // s = k % D, a compile-time register stage
        const int fb = s & 1;             // D is even: k & 1 == s & 1
        __syncthreads();
        put_b(smem + fb * L::BTILE, pb[s]);
        read_frag(smem + (1 - fb) * L::BTILE, fr[1 - fb]);
        if (c_lo + k < c_hi) {
#pragma unroll
            for (int sb = 0; sb < NSUB; sb++)
#pragma unroll
                for (int i = 0; i < PPW; i++) acc[i][sb] = Elt<AT>::mma(pa[s][i], fr[fb].b[sb], acc[i][sb]);
        }
        load_b(ix[s], pb[s]);
        load_a(cc(k + D), pa[s]);
        load_index(cc(k + 2 + 2 * D), ix[s]);
    };
    // whole rounds of D steps with no guard inside (a path that skipped a
    // step's loads would reach the loop head with fewer loads in flight, and
    // the compiler would size every iteration's waits for that path), then the
    // last partial round
    int k0 = 0;
    for (; k0 + D <= iters; k0 += D) {
#pragma unroll
        for (int s = 0; s < D; s++) step(k0 + s, s);
    }
#pragma unroll
    for (int s = 0; s < D; s++)
        if (k0 + s < iters) step(k0 + s, s);
    if constexpr (KS > 1) {
        // K groups 1..KS-1 hand their partial tiles to group 0, one panel at a
        // time (the sum runs in K-group order: deterministic)
#pragma unroll
        for (int i = 0; i < PPW; i++) {
            __syncthreads();
            if (h > 0) {
#pragma unroll
                for (int sb = 0; sb < NSUB; sb++)
                    reinterpret_cast<f32x4 *>(smem_all + (h - 1) * RED + (w * NSUB + sb) * 1024)[lane] = acc[i][sb];
            }
            __syncthreads();
            if (h == 0) {
#pragma unroll
                for (int q = 1; q < KS; q++)
#pragma unroll
                    for (int sb = 0; sb < NSUB; sb++)
                        acc[i][sb] += reinterpret_cast<const f32x4 *>(smem_all + (q - 1) * RED + (w * NSUB + sb) * 1024)[lane];
            }
        }
    }
    if (h == 0) store_tiles(p, g * (PPG * 16) + w * PPW * 16, n0, acc, atomic);
    __syncthreads();   // LDS is clean for the workgroup's next unit
}

// Standalone launch of the block unit: workgroup -> (unit, N-tile) by xcd_work.
template <int PPW, int NT, int WM, int KS, tca_dtype AT, tca_dtype BT>
__global__ __launch_bounds__(64 * WM * KS) void spmm_block_kernel(SpmmParams p)
{
    __shared__ __attribute__((aligned(16))) uint8_t smem[block_lds_bytes<PPW, NT, WM, KS, Elt<AT>::ES>()];
    int ui, tile;
    if (!xcd_work(p, blockIdx.x, ui, tile)) return;
    const int4 u = p.units[ui];   // (group, pieces, c_lo, c_hi)
    spmm_block_unit<PPW, NT, WM, KS, 2, AT, BT>(p, u.x, tile, u.z, u.w, p.split != 0, smem);
}

using SpmmFn = void (*)(SpmmParams);

template <int RPW, tca_dtype AT, tca_dtype BT, bool IND, bool DN = false>
SpmmFn pick_nt(int nt)
{
    switch (nt) {
    case 16: return spmm_kernel<RPW, 16, AT, BT, IND, DN>;
    case 32: return spmm_kernel<RPW, 32, AT, BT, IND, DN>;
    case 64: return spmm_kernel<RPW, 64, AT, BT, IND, DN>;
    case 128: return spmm_kernel<RPW, 128, AT, BT, IND, DN>;
    }
    return nullptr;
}

template <tca_dtype AT, tca_dtype BT, bool IND, bool DN = false>
SpmmFn pick_rpw(int rpw, int nt)
{
    switch (rpw) {
    case 16: return pick_nt<16, AT, BT, IND, DN>(nt);
    case 32: return pick_nt<32, AT, BT, IND, DN>(nt);
    case 64: return pick_nt<64, AT, BT, IND, DN>(nt);
    }
    return nullptr;
}

template <int PPW, int WM, tca_dtype AT, tca_dtype BT, bool IND, bool DN = false>
SpmmFn pick_wide_nt(int nt)
{
    switch (nt) {
    case 64: return spmm_wide_kernel<PPW, 64, WM, AT, BT, IND, DN>;
    case 128: return spmm_wide_kernel<PPW, 128, WM, AT, BT, IND, DN>;
    }
    return nullptr;
}

template <tca_dtype AT, tca_dtype BT, bool IND, bool DN = false>
SpmmFn pick_wide_rpw(int rpw, int nt)
{
    switch (rpw) {
    case 128: return pick_wide_nt<2, 4, AT, BT, IND, DN>(nt);
    case 256: return pick_wide_nt<2, 8, AT, BT, IND, DN>(nt);
    }
    return nullptr;
}

template <bool IND>
SpmmFn pick_spmm(tca_dtype at, tca_dtype bt, int rpw, int nt)
{
    if (rpw >= 128) {
        if (at == TCA_F16 && bt == TCA_F16) return pick_wide_rpw<TCA_F16, TCA_F16, IND>(rpw, nt);
        if (at == TCA_BF16 && bt == TCA_BF16) return pick_wide_rpw<TCA_BF16, TCA_BF16, IND>(rpw, nt);
        if (at == TCA_F16 && bt == TCA_F32) return pick_wide_rpw<TCA_F16, TCA_F32, IND>(rpw, nt);
        if (at == TCA_BF16 && bt == TCA_F32) return pick_wide_rpw<TCA_BF16, TCA_F32, IND>(rpw, nt);
        return nullptr;
    }
    if (at == TCA_F16 && bt == TCA_F16) return pick_rpw<TCA_F16, TCA_F16, IND>(rpw, nt);
    if (at == TCA_BF16 && bt == TCA_BF16) return pick_rpw<TCA_BF16, TCA_BF16, IND>(rpw, nt);
    if (at == TCA_E4M3 && bt == TCA_E4M3) return pick_rpw<TCA_E4M3, TCA_E4M3, IND>(rpw, nt);
    if (at == TCA_F16 && bt == TCA_F32) return pick_rpw<TCA_F16, TCA_F32, IND>(rpw, nt);
    if (at == TCA_BF16 && bt == TCA_F32) return pick_rpw<TCA_BF16, TCA_F32, IND>(rpw, nt);
    return nullptr;
}

// Block kernels on panel-dense tiles of 128 / 256-row plans: 4 waves per K
// group (2 / 4 panels each), ks = 1 or 2 K groups per workgroup: two when the
// launch has fewer than kBlockKS2Units units (group x N-tile), so that each
// CU still runs eight waves (4096^2 x 512, 256 units: 26.6 against 34.0 us
// with one K group), one when there are enough units to put two or more
// workgroups on every CU (4096x1024 x 2048, 512 units: 24.3 against 33.8 us
// with two).
constexpr int kBlockWaves = 4;
constexpr long kBlockKS2Units = 512;   // two units per CU of the MI355X's 256

template <tca_dtype AT, tca_dtype BT, int KS>
SpmmFn pick_block_ks(int rpw, int nt)
{
    if (rpw == 128 && nt == 64) return spmm_block_kernel<2, 64, kBlockWaves, KS, AT, BT>;
    if (rpw == 128 && nt == 128) return spmm_block_kernel<2, 128, kBlockWaves, KS, AT, BT>;
    if (rpw == 256 && nt == 64) return spmm_block_kernel<4, 64, kBlockWaves, KS, AT, BT>;
    if (rpw == 256 && nt == 128) return spmm_block_kernel<4, 128, kBlockWaves, KS, AT, BT>;
    return nullptr;
}
template <tca_dtype AT, tca_dtype BT>
SpmmFn pick_block(int rpw, int nt, int ks)
{
    return ks == 2 ? pick_block_ks<AT, BT, 2>(rpw, nt) : pick_block_ks<AT, BT, 1>(rpw, nt);
}

// Panel-dense A (forward or transposed plans; 128 / 256 rows: F16 / BF16 A, or
// E4M3 A and B, on the block kernels).
SpmmFn pick_spmm_dense(tca_dtype at, tca_dtype bt, int rpw, int nt, int ks)
{
    if (rpw >= 128) {
        if (at == TCA_F16 && bt == TCA_F16) return pick_block<TCA_F16, TCA_F16>(rpw, nt, ks);
        if (at == TCA_E4M3 && bt == TCA_E4M3) return pick_block<TCA_E4M3, TCA_E4M3>(rpw, nt, ks);
        if (at == TCA_BF16 && bt == TCA_BF16) return pick_block<TCA_BF16, TCA_BF16>(rpw, nt, ks);
        if (at == TCA_F16 && bt == TCA_F32) return pick_block<TCA_F16, TCA_F32>(rpw, nt, ks);
        if (at == TCA_BF16 && bt == TCA_F32) return pick_block<TCA_BF16, TCA_F32>(rpw, nt, ks);
        return nullptr;
    }
    if (at == TCA_F16 && bt == TCA_F16) return pick_rpw<TCA_F16, TCA_F16, false, true>(rpw, nt);
    if (at == TCA_BF16 && bt == TCA_BF16) return pick_rpw<TCA_BF16, TCA_BF16, false, true>(rpw, nt);
    if (at == TCA_E4M3 && bt == TCA_E4M3) return pick_rpw<TCA_E4M3, TCA_E4M3, false, true>(rpw, nt);
    if (at == TCA_F16 && bt == TCA_F32) return pick_rpw<TCA_F16, TCA_F32, false, true>(rpw, nt);
    if (at == TCA_BF16 && bt == TCA_F32) return pick_rpw<TCA_BF16, TCA_F32, false, true>(rpw, nt);
    return nullptr;
}

// ---------------------------------------------------------------------------
// SDDMM: out[e] = sum_n G[row(e), n] X[col(e), n] on the plan's pattern.
// One wave = (group, block of CB chunks).  MFMA A operand = 8 consecutive n of
// a G row, B operand = 8 consecutive n of a gathered X row: both are
// contiguous in memory, so no LDS transpose is needed.  The 16x32 result tile
// of each (chunk, panel) is staged in LDS and sampled at the entry slots.
// ---------------------------------------------------------------------------
struct SddmmParams {
    const int32_t *grp_chunk_ptr;
    const int32_t *chunk_cols;
    const int32_t *seg_ptr;
    const uint16_t *entry_slot;
    const void *G;
    const void *X;
    float *out;
    int n_rows, N, ldg, ldx, rpw, n_blocks_per_group_max;
    const int32_t *unit_group;   // per workgroup: group id
    const int32_t *unit_chunk;   // per workgroup: first chunk
    uint32_t g_bytes, x_bytes;   // bytes of G and X from the first to the last used element
    const int32_t *dst;          // out index of each entry (null: the entry's own index)
    uint32_t nnz;                // entries of the tiling (entry_slot, dst)
    uint16_t *sgd_w;             // batch items: weights updated from each written entry (null: none)
    float sgd_lr;
    int sgd_bf16;
};

// One SGD step on a 16-bit weight: round(fma(-lr, g, w)) -- the arithmetic of
// sgd_kernel, so a fused update gives the bits of a separate tca_sgd_update.
__device__ __forceinline__ uint32_t sgd_step(uint32_t bits, float g, float lr, bool bf16)
{
    const float x = bf16 ? __builtin_bit_cast(float, bits << 16) : (float)__builtin_bit_cast(_Float16, (uint16_t)bits);
    const float y = fmaf(-lr, g, x);
    return bf16 ? f32_to_bf16_bits(y) : f32_to_f16_bits(y);
}

// out[d] = v, and the fused weight update when the item has one
__device__ __forceinline__ void sddmm_emit(const SddmmParams &p, uint32_t d, float v)
{
    gbl(p.out)[d] = v;
    if (p.sgd_w) {
        auto w = gbl(p.sgd_w);
        w[d] = (uint16_t)sgd_step(w[d], v, p.sgd_lr, p.sgd_bf16 != 0);
    }
}

// SDDMM with shared X rows.  The plan here has 64-row groups: a workgroup of 4
// waves owns (group g, chunks [c0, c0+CB)), wave w the group's panel w.  Per
// contraction step (32 n) the workgroup loads the gathered X rows of its
// chunks once -- CB*32 rows x 32 n, 16-byte pieces, double-buffered in LDS with
// one barrier per step -- and every wave reads its MFMA B operands (8
// consecutive n of one X row) from there; each wave's A operands are 8
// consecutive n of its own G rows, straight from memory, one step ahead.  X is
// fetched once per 64 rows instead of once per 16.  Each wave then stages its
// 16x32 result tiles in its own LDS area and samples them at the entry slots,
// writing out[dst[e]] (dW in the caller's plan's tiled order).
//
// X tile layout: 16-byte pieces, a 1 KB block per 16 tile rows; octet q (n =
// 8q .. 8q+7) of row 16*rb + r sits in slot 64*rb + 16q + (r ^ 2q).  Measured
// on gfx950's LDS (bank-conflict counters over lane patterns): the MFMA
// operand reads (lane 16q + r) and the piece writes (four lanes per X row, so
// that the global loads stay 64-byte coalesced) are both conflict-free with
// this XOR, while padded row pitches, rotations and XORs of the higher slot
// bits conflict on one side or the other.
__device__ __forceinline__ int sh_slot(int r, int q) { return 16 * q + (r ^ (2 * q)); }

constexpr int kShRows = 64;           // rows per group of the SDDMM tiling
// Row pitch (floats) of a wave's staged 16 x 32 result tile: the four lane
// groups' rows land 16 banks apart on a 64-bank LDS (and 16 apart on 32 banks)
constexpr int kShStage = 36;

template <int CB>
struct ShLds {
    static constexpr int XT = CB * 32 * 64;                     // one X tile (bytes)
    static constexpr int STAGE = 16 * kShStage * 4;             // one wave's result tile
    static constexpr int BYTES = 2 * XT + 4 * STAGE;
    static constexpr int PIECES = CB * 32 * 4;                  // 16-byte X pieces per step
    static constexpr int XP = (PIECES + 255) / 256;             // X pieces per thread and step
};

// Entries per lane and chunk whose sampling indices are fetched before the
// contraction (covers 128 entries of a 16 x 32 panel tile; denser tiles
// fetch the rest while sampling).
constexpr int kShEP = 2;

template <int CB, tca_dtype AT, tca_dtype GT>
__device__ __forceinline__ void sddmm_shared_unit(const SddmmParams &p, int g, int c0, uint8_t *smem)
{
    using L = ShLds<CB>;
    using Oct = typename Elt<AT>::Oct;
    using GR = RowLoad<AT, GT>;
    using GB = RowBuf<AT, GT>;
    using XB = RowBuf<AT, AT>;
    const int w = threadIdx.x >> 6, lane = lane_id(), tid = w * 64 + lane;
    const int nc = min(CB, gbl(p.grp_chunk_ptr)[g + 1] - c0);
    const Rsrc r_g = make_rsrc(p.G, p.g_bytes);
    const Rsrc r_x = make_rsrc(p.X, p.x_bytes);
    const Rsrc r_slot = make_rsrc(p.entry_slot, p.nnz * 2u);
    const Rsrc r_dst = make_rsrc(p.dst, p.nnz * 4u);
    const int plan_ppg = p.rpw >> 4;

    // this thread's X pieces: tile row (chunk b = row/32), 8-n octet q
    uint32_t xsrc[L::XP];
    int xdst[L::XP];
    bool xval[L::XP];
#pragma unroll
    for (int j = 0; j < L::XP; j++) {
        const int pc = tid + 256 * j, q = pc & 3, r = (pc >> 2) & 15;
        const int row = 16 * (pc >> 6) + r, b = row >> 5;
        xval[j] = pc < L::PIECES;
        const int col = xval[j] && b < nc ? gbl(p.chunk_cols)[(long)(c0 + b) * 32 + (row & 31)] : -1;
        xsrc[j] = col >= 0 ? ((uint32_t)col * (uint32_t)p.ldx + 8u * q) * 2u : kOOB;
        xdst[j] = 16 * (64 * (pc >> 6) + sh_slot(r, q));
    }
    const int grow = g * kShRows + 16 * w + (lane & 15);
    const uint32_t gsrc = grow < p.n_rows ? ((uint32_t)grow * (uint32_t)p.ldg + 8u * (lane >> 4)) * GB::BES : kOOB;
    // B operand of (chunk b, half h): X tile row 32b + 16h + lane%16, octet
    // lane/16 (block 2b + h)
    const int xread = 16 * sh_slot(lane & 15, lane >> 4);

    // sampling indices of this wave's panel in each chunk, issued now so that
    // they arrive during the contraction (out-of-range entries read as zero
    // and are skipped when sampling)
    // (all bounds first: interleaving them with the entry loads would make
    // each chunk's bounds wait for the previous chunk's entries)
    int e0[CB], e1[CB];
    uint32_t slot[CB][kShEP], dsti[CB][kShEP];
#pragma unroll
    for (int b = 0; b < CB; b++) {
        const int c = c0 + min(b, nc - 1);
        e0[b] = gbl(p.seg_ptr)[c * plan_ppg + w];
        e1[b] = gbl(p.seg_ptr)[c * plan_ppg + w + 1];
    }
#pragma unroll
    for (int b = 0; b < CB; b++) {
#pragma unroll
        for (int k = 0; k < kShEP; k++) {
            const uint32_t e = (uint32_t)(e0[b] + k * 64 + lane);
            slot[b][k] = buf_b16(r_slot, e * 2u);
            dsti[b][k] = p.dst ? buf_b32(r_dst, e * 4u) : e;
        }
    }

    f32x4 acc[CB][2];
#pragma unroll
    for (int b = 0; b < CB; b++)
#pragma unroll
        for (int h = 0; h < 2; h++) acc[b][h] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int steps = p.N >> 5;
    // X pieces in a two-step register ring (xr[t & 1] holds step t's pieces),
    // G rows two steps ahead
    typename XB::Raw xr[2][L::XP];
    typename GB::Raw gr[2];
    auto load_x = [&](int slot_, int t) {
#pragma unroll
        for (int j = 0; j < L::XP; j++) xr[slot_][j] = XB::load(r_x, xsrc[j], (uint32_t)(32 * t) * 2u);
    };
    auto put_x = [&](int slot_, int t) {
        uint8_t *buf = smem + (t & 1) * L::XT;
#pragma unroll
        for (int j = 0; j < L::XP; j++)
            if (xval[j]) *reinterpret_cast<Oct *>(buf + xdst[j]) = xr[slot_][j];
    };
    // prologue: X of step 0 in LDS; G of steps 0, 1 and X of steps 1, 2 in
    // registers, issued in the loop's own order (G(s), X(s+1), G(s+1), X(s+2)
    // in flight at the top of step s) so that the compiler's waits in the
    // loop, derived from its entry and back-edge states together, stay those
    // of the steady state
    // (scheduling barriers keep the compiler from reordering these loads)
    load_x(0, 0);
    put_x(0, 0);
    gr[0] = GB::load(r_g, gsrc, 0);
    __builtin_amdgcn_sched_barrier(0);
    load_x(1, min(1, steps - 1));
    __builtin_amdgcn_sched_barrier(0);
    gr[1] = GB::load(r_g, gsrc, (uint32_t)(32 * min(1, steps - 1)) * GB::BES);
    __builtin_amdgcn_sched_barrier(0);
    load_x(0, min(2, steps - 1));
    __syncthreads();
    auto step = [&](int k, int s) {   // k = s & 1, a compile-time register slot
        const uint8_t *buf = smem + (s & 1) * L::XT;
        const Oct ga = GR::cvt(gr[k]);
#pragma unroll
        for (int b = 0; b < CB; b++) {
            if (b < nc) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const Oct xb = *reinterpret_cast<const Oct *>(buf + xread + (2 * b + h) * 1024);
                    acc[b][h] = Elt<AT>::mma(ga, xb, acc[b][h]);
                }
            }
        }
        gr[k] = GB::load(r_g, gsrc, (uint32_t)(32 * min(s + 2, steps - 1)) * GB::BES);
        put_x(1 - k, s + 1);                      // X of step s+1, loaded two steps ago
        load_x(1 - k, min(s + 3, steps - 1));     // (clamped re-reads past the end)
        __syncthreads();
    };
    // Pairs of steps with no guard inside (a path that skipped the second
    // step's loads would reach the loop head with fewer loads in flight, and
    // the compiler would size every iteration's waits for that path), then
    // the odd last step.
    int s0 = 0;
    for (; s0 + 1 < steps; s0 += 2) {
        step(0, s0);
        step(1, s0 + 1);
    }
    if (s0 < steps) step(0, s0);
    // fused weight update: the weights of the prefetched entries, all issued
    // before the sampling so that their latency overlaps it
    uint32_t wv[CB][kShEP];
    if (p.sgd_w) {
        const Rsrc r_w = make_rsrc(p.sgd_w, p.nnz * 2u);
#pragma unroll
        for (int b = 0; b < CB; b++)
#pragma unroll
            for (int k = 0; k < kShEP; k++) {
                const bool ok = b < nc && e0[b] + k * 64 + lane < e1[b];
                wv[b][k] = buf_b16(r_w, ok ? dsti[b][k] * 2u : kOOB);
            }
    }
    // sample this wave's panel: stage each chunk's 16x32 tile in the wave's area
    float *tile = reinterpret_cast<float *>(smem + 2 * L::XT + w * L::STAGE);
#pragma unroll
    for (int b = 0; b < CB; b++) {
        if (b >= nc) break;
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int j = 0; j < 4; j++) tile[(4 * (lane >> 4) + j) * kShStage + 16 * h + (lane & 15)] = acc[b][h][j];
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < kShEP; k++) {
            if (e0[b] + k * 64 + lane < e1[b]) {
                const uint32_t sl = slot[b][k];
                const float v = tile[(sl >> 5) * kShStage + (sl & 31)];
                gbl(p.out)[dsti[b][k]] = v;
                if (p.sgd_w) gbl(p.sgd_w)[dsti[b][k]] = (uint16_t)sgd_step(wv[b][k], v, p.sgd_lr, p.sgd_bf16 != 0);
            }
        }
        for (int e = e0[b] + kShEP * 64 + lane; e < e1[b]; e += 64) {
            const int sl = gbl(p.entry_slot)[e];
            sddmm_emit(p, p.dst ? (uint32_t)gbl(p.dst)[e] : (uint32_t)e, tile[(sl >> 5) * kShStage + (sl & 31)]);
        }
        wave_lds_sync();
    }
}

template <int CB, tca_dtype AT, tca_dtype GT>
__global__ __launch_bounds__(256) void sddmm_shared_kernel(SddmmParams p)
{
    __shared__ __attribute__((aligned(16))) uint8_t smem[ShLds<CB>::BYTES];
    sddmm_shared_unit<CB, AT, GT>(p, p.unit_group[blockIdx.x], p.unit_chunk[blockIdx.x], smem);
}

// ---------------------------------------------------------------------------
// small elementwise kernels
// ---------------------------------------------------------------------------
// Panel-dense tiles: one wave per (chunk, panel) segment s = c*PPG + i.  The
// segment's entries are scattered into a zeroed 16 x 32 tile in LDS at
// element ((k/8)*16 + r)*8 + k%8 (entry slot = r << 5 | k), i.e. the octet of
// MFMA operand lane r + 16*(k/8); the tile is then written out whole, one
// octet per lane.  src holds tiled values; src_idx (a transposed plan's entry
// map into the forward plan's tiled order) may be null.
template <class T>
__global__ __launch_bounds__(256) void pack_dense_kernel(long n_segs, const int32_t *__restrict__ seg_ptr,
                                                         const uint16_t *__restrict__ entry_slot,
                                                         const int32_t *__restrict__ src_idx, const T *__restrict__ src,
                                                         T *__restrict__ dst)
{
    constexpr int OCT = 8 * (int)sizeof(T);
    using O = typename Store<OCT>::T;
    __shared__ __attribute__((aligned(16))) T tile[4][512];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long seg = (long)blockIdx.x * 4 + w;
    if (seg >= n_segs) return;   // whole waves; no workgroup barrier below
    T *t = tile[w];
    reinterpret_cast<O *>(t)[lane] = O{};
    wave_lds_sync();
    const int e1 = seg_ptr[seg + 1];
    for (int e = seg_ptr[seg] + lane; e < e1; e += 64) {
        const int slot = entry_slot[e], r = slot >> 5, k = slot & 31;
        t[((k >> 3) * 16 + r) * 8 + (k & 7)] = src[src_idx ? src_idx[e] : e];
    }
    wave_lds_sync();
    reinterpret_cast<O *>(dst + seg * 512)[lane] = reinterpret_cast<const O *>(t)[lane];
}

template <class T>
__global__ void gather_kernel(const int32_t *__restrict__ idx, int n, const T *__restrict__ src, T *__restrict__ dst,
                              bool scatter)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (scatter) dst[idx[i]] = src[i];
        else dst[i] = src[idx[i]];
    }
}

// w -= lr * g, 8 elements per thread (one 16-byte load of w, two of g, one
// 16-byte store); the last n % 8 elements one per thread of the first block.
template <bool BF>
__global__ __launch_bounds__(256) void sgd_kernel(uint16_t *__restrict__ w, const float *__restrict__ g, long n8,
                                                  long n, float lr)
{
    auto upd = [&](uint32_t bits, float gv) -> uint32_t { return sgd_step(bits, gv, lr, BF); };
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n8; i += (long)gridDim.x * blockDim.x) {
        const u32x4 wv = reinterpret_cast<const u32x4 *>(w)[i];
        const float4 g0 = reinterpret_cast<const float4 *>(g)[2 * i], g1 = reinterpret_cast<const float4 *>(g)[2 * i + 1];
        u32x4 o;
        o[0] = upd(wv[0] & 0xffffu, g0.x) | upd(wv[0] >> 16, g0.y) << 16;
        o[1] = upd(wv[1] & 0xffffu, g0.z) | upd(wv[1] >> 16, g0.w) << 16;
        o[2] = upd(wv[2] & 0xffffu, g1.x) | upd(wv[2] >> 16, g1.y) << 16;
        o[3] = upd(wv[3] & 0xffffu, g1.z) | upd(wv[3] >> 16, g1.w) << 16;
        reinterpret_cast<u32x4 *>(w)[i] = o;
    }
    if (blockIdx.x == 0) {
        const long t = 8 * n8 + threadIdx.x;
        if (t < n) w[t] = (uint16_t)upd(w[t], g[t]);
    }
}

// C[0:rows, 0:N] = 0 (before split-chunk launches accumulate into it)
__global__ void zero_rows_kernel(float *C, int rows, int N, int ldc)
{
    const long total = (long)rows * N;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
        C[(i / N) * ldc + (i % N)] = 0.0f;
}

// f32 -> OCP E4M3FN, round to nearest even, saturating to +-448; the sign is
// always kept (tiny negatives give 0x80).
__device__ inline uint8_t f32_to_e4m3(float x)
{
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7f;
    const uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    const float a = fabsf(x);
    if (a >= 448.0f) return sign | 0x7e;
    if (a < 0.015625f) {  // below 2^-6: multiples of 2^-9 (code 8 = 2^-6 on round-up)
        const int m = (int)__builtin_rintf(a * 512.0f);
        return (uint8_t)(sign | m);
    }
    const int e = (int)((u >> 23) & 0xff) - 127;
    const uint32_t mant = u & 0x7fffffu;
    const uint32_t keep = mant >> 20, rest = mant & 0xfffffu;
    uint32_t code = ((uint32_t)(e + 7) << 3) | keep;
    if (rest > 0x80000u || (rest == 0x80000u && (keep & 1u))) code++;
    if (code > 0x7eu) code = 0x7eu;
    return (uint8_t)(sign | code);
}

__device__ inline float f16_bits_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

__device__ inline float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// one wave per CSR row: scale = amax/448, q = e4m3(v / scale)
__global__ __launch_bounds__(256) void quantize_rows_kernel(int n_rows, const int32_t *row_offset, const uint16_t *vals,
                                                            uint8_t *q, float *scale)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n_rows) return;
    const int e0 = row_offset[row], e1 = row_offset[row + 1];
    float amax = 0.0f;
    for (int e = e0 + lane; e < e1; e += 64) amax = fmaxf(amax, fabsf(f16_bits_to_f32(vals[e])));
    amax = wave_max(amax);
    const float s = amax > 0.0f ? amax / 448.0f : 1.0f;
    if (lane == 0) scale[row] = s;
    for (int e = e0 + lane; e < e1; e += 64) q[e] = f32_to_e4m3(f16_bits_to_f32(vals[e]) / s);
}

// Column scales in three steps so that the K rows are spread over many
// workgroups: (1) partial column amax of a row slice, merged with an atomic
// max on the float bits (amax >= 0, so unsigned order is float order) into
// `scale`, which a fill kernel has zeroed; (2) q = e4m3(v / (amax / 448));
// (3) scale = amax / 448 (1 for an all-zero column).
constexpr int kQuantRowsPerSlice = 64;

__global__ __launch_bounds__(256) void quantize_cols_amax_kernel(int K, int N, const uint16_t *B, int ldb, float *amax_out)
{
    __shared__ float part[4][64];
    const int cl = threadIdx.x & 63, ph = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + cl;
    const int k0 = blockIdx.y * kQuantRowsPerSlice, k1 = min(K, k0 + kQuantRowsPerSlice);
    float amax = 0.0f;
    if (n < N)
        for (int k = k0 + ph; k < k1; k += 4) amax = fmaxf(amax, fabsf(f16_bits_to_f32(B[(long)k * ldb + n])));
    part[ph][cl] = amax;
    __syncthreads();
    if (ph == 0 && n < N) {
        const float m = fmaxf(fmaxf(part[0][cl], part[1][cl]), fmaxf(part[2][cl], part[3][cl]));
        atomicMax(reinterpret_cast<unsigned int *>(amax_out + n), __builtin_bit_cast(unsigned int, m));
    }
}

__device__ inline float col_scale_of(float amax) { return amax > 0.0f ? amax / 448.0f : 1.0f; }

__global__ __launch_bounds__(256) void quantize_cols_apply_kernel(int K, int N, const uint16_t *B, int ldb, uint8_t *q,
                                                                  int ldq, const float *amax)
{
    const long total = (long)K * N;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i / N), n = (int)(i - (long)k * N);
        q[(long)k * ldq + n] = f32_to_e4m3(f16_bits_to_f32(B[(long)k * ldb + n]) / col_scale_of(amax[n]));
    }
}

__global__ __launch_bounds__(256) void quantize_cols_finalize_kernel(int N, float *scale)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) scale[n] = col_scale_of(scale[n]);
}

__global__ void warm_up_kernel() {}

// ---------------------------------------------------------------------------
// Batched persistent kernel: every workgroup (kBatchWaves waves) takes work
// units until the list is exhausted, so one launch covers the forward SpMM,
// the transposed SpMM and the SDDMM of many layers.  Units are sorted heaviest
// first on the host and cut into partitions, each with its own counter (see
// the loop below); a last counter counts finished workgroups, and
// the last workgroup resets them all, so the next launch (stream-ordered)
// starts from zero without a memset.
// Variants (few and small, so that the kernel's register budget -- the maximum
// over its cases -- keeps three waves on each SIMD, four in the dense-only
// instantiation): tiled values, forward f16 x f16 on 16-row plans with tile_n
// 16/32/64 (0..2) and on 32-row plans with tile_n 32 (8); transposed f16 x f32
// on 16-row plans with tile_n 16/32/64 (3, 4, 7) and f16 x f16 (9..11);
// panel-dense f16 A (forward or transposed) with an f16 B on 16-row plans with
// tile_n 16/32/64 (12..14), 32-row plans with tile_n 32/64 (18, 19) and 64-row
// plans with tile_n 32 (20), with an f32 B on 16-row plans with tile_n 16/32/64
// (15..17) and 32-row plans with tile_n 32 (21); block units (the workgroup's
// four waves share each B chunk tile) on 64-row (22) and 128-row (23) plans
// with tile_n 64 and an f16 B; SDDMM with an f32 (5) or f16 (6) G on the plan's
// 64-row SDDMM tiling.  BF16: the panel-dense cases with BF16 operands (24..35,
// each its F16 counterpart + 12) and the SDDMM with a bf16 X and an f32 (36) or
// bf16 (37) G.
// ---------------------------------------------------------------------------
__device__ __forceinline__ SpmmParams to_spmm(const BatchItemDev &d)
{
    SpmmParams p;
    p.grp_chunk_ptr = d.grp_chunk_ptr;
    p.chunk_cols = d.chunk_cols;
    p.seg_ptr = d.seg_ptr;
    p.entry_slot = d.entry_slot;
    p.entry_lds = d.entry_lds;
    p.entry_src = d.entry_src;
    p.work_order = nullptr;
    p.a = d.a;
    p.B = d.b;
    p.C = d.out;
    p.row_scale = nullptr;
    p.col_scale = nullptr;
    p.units = nullptr;
    p.split = 0;
    p.nnz = d.nnz;
    p.n_rows = d.n_rows;
    p.N = d.N;
    p.ldb = d.ldb;
    p.ldc = d.ldc;
    p.n_tiles = d.n_tiles;
    p.n_groups = 0;
    p.rpw = d.rpw;
    p.n_units = 0;
    p.xcd_share = -1;
    p.b_bytes = d.b_bytes;
    p.n_chunks = d.n_chunks;
    return p;
}

__device__ __forceinline__ SddmmParams to_sddmm(const BatchItemDev &d)
{
    SddmmParams p;
    p.grp_chunk_ptr = d.grp_chunk_ptr;
    p.chunk_cols = d.chunk_cols;
    p.seg_ptr = d.seg_ptr;
    p.entry_slot = d.entry_slot;
    p.G = d.a;
    p.X = d.b;
    p.out = d.out;
    p.n_rows = d.n_rows;
    p.N = d.N;
    p.ldg = d.ldc;
    p.ldx = d.ldb;
    p.rpw = d.rpw;
    p.n_blocks_per_group_max = 0;
    p.unit_group = nullptr;
    p.unit_chunk = nullptr;
    p.g_bytes = d.a_bytes;
    p.x_bytes = d.b_bytes;
    p.nnz = (uint32_t)d.nnz;
    p.dst = d.dst;
    p.sgd_w = d.sgd_w;
    p.sgd_lr = d.sgd_lr;
    p.sgd_bf16 = d.sgd_bf16;
    return p;
}

// The batch kernel's workgroup: kBatchWaves waves share every unit (SpMM:
// chunk range split + LDS reduction; SDDMM: contraction split over N).
constexpr int kBatchWaves = 4;
constexpr int kBatchSlot = cmax(cmax(SpmmLds<16, 64, TCA_F16>::SLOT, SpmmLds<32, 32, TCA_F16>::SLOT),   // largest SpMM
                                cmax(SpmmLds<32, 64, TCA_F16>::SLOT, SpmmLds<64, 32, TCA_F16>::SLOT));  // variant's per-wave LDS
constexpr int kBatchLds = cmax(cmax(kBatchWaves * kBatchSlot, ShLds<kBatchSddmmCB>::BYTES),
                                cmax(block_lds_bytes<1, 64, kBatchWaves, 1>(), block_lds_bytes<2, 64, kBatchWaves, 1>()));

// MODE 0: every F16 case; 1 (DENSE_ONLY): the F16 SpMM cases of panel-dense
// items only (and the F16 SDDMM) -- the register budget is then that of the
// lean dense loop, which keeps four workgroups (16 waves) on each CU instead
// of three; 2: the same cases for BF16 operands (a batch of BF16 items, all
// panel-dense), an instantiation of its own so that neither type's cases
// weigh on the other's register budget.
template <int MODE>
__global__ __launch_bounds__(64 * kBatchWaves) __attribute__((amdgpu_waves_per_eu(MODE ? 4 : 3))) void batch_kernel(const BatchItemDev *__restrict__ items,
                                                                 const BatchUnitDev *__restrict__ units, int n_units,
                                                                 int *counters, const int *__restrict__ part_off,
                                                                 int nparts)
{
    __shared__ __attribute__((aligned(16))) uint8_t smem[kBatchLds];
    __shared__ int next_unit;
    // Partition x of P holds units [part_off[x], part_off[x+1]) (heaviest
    // first) and is served by the workgroups b with b % P == x.  A workgroup's
    // first unit is static (its rank r in the partition); later ones come from
    // the partition's counter, which hands out ranks gx + 0, 1, ....  One
    // atomic address per partition keeps the claims from queueing on a single
    // line.  A workgroup whose partition is exhausted moves on to the next
    // ones in turn (x+1, x+2, ... mod P) and claims from their counters, so
    // that no XCD idles while another still has units; it ends when every
    // partition is exhausted.  (P <= gridDim.x, so every partition has
    // workgroups.)
    (void)n_units;
    const int P = nparts;
    const int x = (int)blockIdx.x % P, r = (int)blockIdx.x / P;
    int xs = x, hops = 0;   // thread 0: the partition it claims from, partitions left
    for (bool first = true;; first = false) {
        int u;
        if (first && r < part_off[x + 1] - part_off[x]) {
            u = part_off[x] + r;
        } else {
            if (threadIdx.x == 0) {
                int found = -1;
                while (hops < P) {
                    const int gxs = ((int)gridDim.x - xs + P - 1) / P;
                    const int k = gxs + atomicAdd(counters + xs * kCounterStride, 1);
                    if (k < part_off[xs + 1] - part_off[xs]) {
                        found = part_off[xs] + k;
                        break;
                    }
                    xs = xs + 1 == P ? 0 : xs + 1;
                    hops++;
                }
                next_unit = found;
            }
            __syncthreads();
            u = __builtin_amdgcn_readfirstlane(next_unit);   // uniform: descriptors load into SGPRs
            __syncthreads();                                  // all have read next_unit
        }
        if (u < 0) break;
        const BatchUnitDev un = units[u];
        const BatchItemDev d = items[un.item];
        switch (d.variant) {
        case 0: if constexpr (MODE == 0) spmm_wg_unit<16, 16, TCA_F16, TCA_F16, false>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 1: if constexpr (MODE == 0) spmm_wg_unit<16, 32, TCA_F16, TCA_F16, false>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 2: if constexpr (MODE == 0) spmm_wg_unit<16, 64, TCA_F16, TCA_F16, false>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 3: if constexpr (MODE == 0) spmm_wg_unit<16, 16, TCA_F16, TCA_F32, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 4: if constexpr (MODE == 0) spmm_wg_unit<16, 32, TCA_F16, TCA_F32, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 7: if constexpr (MODE == 0) spmm_wg_unit<16, 64, TCA_F16, TCA_F32, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 8: if constexpr (MODE == 0) spmm_wg_unit<32, 32, TCA_F16, TCA_F16, false>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 9: if constexpr (MODE == 0) spmm_wg_unit<16, 16, TCA_F16, TCA_F16, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 10: if constexpr (MODE == 0) spmm_wg_unit<16, 32, TCA_F16, TCA_F16, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 11: if constexpr (MODE == 0) spmm_wg_unit<16, 64, TCA_F16, TCA_F16, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 12: if constexpr (MODE != 2) spmm_wg_unit<16, 16, TCA_F16, TCA_F16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 13: if constexpr (MODE != 2) spmm_wg_unit<16, 32, TCA_F16, TCA_F16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 14: if constexpr (MODE != 2) spmm_wg_unit<16, 64, TCA_F16, TCA_F16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 15: if constexpr (MODE != 2) spmm_wg_unit<16, 16, TCA_F16, TCA_F32, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 16: if constexpr (MODE != 2) spmm_wg_unit<16, 32, TCA_F16, TCA_F32, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 17: if constexpr (MODE != 2) spmm_wg_unit<16, 64, TCA_F16, TCA_F32, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 18: if constexpr (MODE != 2) spmm_wg_unit<32, 32, TCA_F16, TCA_F16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 19: if constexpr (MODE != 2) spmm_wg_unit<32, 64, TCA_F16, TCA_F16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 20: if constexpr (MODE != 2) spmm_wg_unit<64, 32, TCA_F16, TCA_F16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 21: if constexpr (MODE != 2) spmm_wg_unit<32, 32, TCA_F16, TCA_F32, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 22: if constexpr (MODE != 2) spmm_block_unit<1, 64, kBatchWaves, 1, 2, TCA_F16, TCA_F16>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem); break;
        case 23: if constexpr (MODE != 2) spmm_block_unit<2, 64, kBatchWaves, 1, 2, TCA_F16, TCA_F16>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem); break;
        case 5: if constexpr (MODE != 2) sddmm_shared_unit<kBatchSddmmCB, TCA_F16, TCA_F32>(to_sddmm(d), un.g, un.t, smem); break;
        case 6: if constexpr (MODE != 2) sddmm_shared_unit<kBatchSddmmCB, TCA_F16, TCA_F16>(to_sddmm(d), un.g, un.t, smem); break;
        // BF16: the panel-dense and SDDMM cases above with BF16 operands
        case 24: if constexpr (MODE == 2) spmm_wg_unit<16, 16, TCA_BF16, TCA_BF16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 25: if constexpr (MODE == 2) spmm_wg_unit<16, 32, TCA_BF16, TCA_BF16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 26: if constexpr (MODE == 2) spmm_wg_unit<16, 64, TCA_BF16, TCA_BF16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 27: if constexpr (MODE == 2) spmm_wg_unit<16, 16, TCA_BF16, TCA_F32, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 28: if constexpr (MODE == 2) spmm_wg_unit<16, 32, TCA_BF16, TCA_F32, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 29: if constexpr (MODE == 2) spmm_wg_unit<16, 64, TCA_BF16, TCA_F32, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 30: if constexpr (MODE == 2) spmm_wg_unit<32, 32, TCA_BF16, TCA_BF16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 31: if constexpr (MODE == 2) spmm_wg_unit<32, 64, TCA_BF16, TCA_BF16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 32: if constexpr (MODE == 2) spmm_wg_unit<64, 32, TCA_BF16, TCA_BF16, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 33: if constexpr (MODE == 2) spmm_wg_unit<32, 32, TCA_BF16, TCA_F32, false, true>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem, kBatchSlot); break;
        case 34: if constexpr (MODE == 2) spmm_block_unit<1, 64, kBatchWaves, 1, 2, TCA_BF16, TCA_BF16>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem); break;
        case 35: if constexpr (MODE == 2) spmm_block_unit<2, 64, kBatchWaves, 1, 2, TCA_BF16, TCA_BF16>(to_spmm(d), un.g, un.t, un.c_lo, un.c_hi, un.atomic != 0, smem); break;
        case 36: if constexpr (MODE == 2) sddmm_shared_unit<kBatchSddmmCB, TCA_BF16, TCA_F32>(to_sddmm(d), un.g, un.t, smem); break;
        case 37: if constexpr (MODE == 2) sddmm_shared_unit<kBatchSddmmCB, TCA_BF16, TCA_BF16>(to_sddmm(d), un.g, un.t, smem); break;
        default: break;
        }
        __syncthreads();  // LDS is reused by the next unit
    }
    // No fence before the count: every claim of this workgroup has returned
    // (its value was used), so the last workgroup's reset cannot overtake one;
    // the outputs need no ordering against the counters (the next launch is
    // stream-ordered).  A device-scope fence here would write back the XCD's
    // L2 once per exiting workgroup and stall the units still running.
    if (threadIdx.x == 0) {
        if (atomicAdd(&counters[kBatchParts * kCounterStride], 1) == (int)gridDim.x - 1) {
            for (int q = 0; q < kBatchParts; q++) atomicExch(&counters[q * kCounterStride], 0);
            atomicExch(&counters[kBatchParts * kCounterStride], 0);
        }
    }
}

// zero the outputs of split batch items (blockIdx.y = descriptor)
__global__ void zero_list_kernel(const ZeroDesc *__restrict__ list)
{
    const ZeroDesc z = list[blockIdx.y];
    const long total = (long)z.rows * z.N;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
        z.C[(i / z.N) * z.ldc + (i % z.N)] = 0.0f;
}

}  // namespace

int batch_variant(tca_op op, int rpw, int tile_n, tca_dtype a_dtype, tca_dtype b_dtype, bool transposed, bool dense)
{
    if (dense) {   // panel-dense tiles, forward or transposed plan alike
        if (op != TCA_OP_SPMM || (a_dtype != TCA_F16 && a_dtype != TCA_BF16) || (b_dtype != a_dtype && b_dtype != TCA_F32))
            return -1;
        const int f32 = b_dtype == TCA_F32;
        const int bf = a_dtype == TCA_BF16 ? 12 : 0;   // BF16: the F16 case + 12 (24..35)
        if (rpw == 16) {
            const int t = tile_n == 16 ? 0 : tile_n == 32 ? 1 : tile_n == 64 ? 2 : -1;
            return t < 0 ? -1 : (f32 ? 15 : 12) + t + bf;
        }
        // (F32 B on wider panels only at rows_per_wave 32 x tile_n 32: the others
        // would push the dense kernel past four workgroups per CU)
        if (rpw == 32 && tile_n == 32) return (f32 ? 21 : 18) + bf;
        if (!f32 && rpw == 32 && tile_n == 64) return 19 + bf;
        if (!f32 && rpw == 64 && tile_n == 32) return 20 + bf;
        // block units: the workgroup's four waves share each B chunk tile
        if (!f32 && rpw == 64 && tile_n == 64) return 22 + bf;
        if (!f32 && rpw == 128 && tile_n == 64) return 23 + bf;
        return -1;
    }
    if (op == TCA_OP_SPMM) {
        if (a_dtype != TCA_F16) return -1;
        if (!transposed && b_dtype == TCA_F16) {
            if (rpw == 16) return tile_n == 16 ? 0 : tile_n == 32 ? 1 : tile_n == 64 ? 2 : -1;
            if (rpw == 32) return tile_n == 32 ? 8 : -1;
        }
        if (transposed && b_dtype == TCA_F32 && rpw == 16) return tile_n == 16 ? 3 : tile_n == 32 ? 4 : tile_n == 64 ? 7 : -1;
        if (transposed && b_dtype == TCA_F16 && rpw == 16) return tile_n == 16 ? 9 : tile_n == 32 ? 10 : tile_n == 64 ? 11 : -1;
        return -1;
    }
    if (op == TCA_OP_SDDMM && b_dtype == TCA_F16 && !transposed) {
        if (a_dtype == TCA_F32) return 5;
        if (a_dtype == TCA_F16) return 6;
    }
    if (op == TCA_OP_SDDMM && b_dtype == TCA_BF16 && !transposed) {
        if (a_dtype == TCA_F32) return 36;
        if (a_dtype == TCA_BF16) return 37;
    }
    return -1;
}

int batch_blocks_per_cu(int mode)
{
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mode == 2 ? batch_kernel<2> : mode == 1 ? batch_kernel<1> : batch_kernel<0>,
                                                     64 * kBatchWaves, 0) != hipSuccess || n < 1)
        n = 1;
    return n;
}

tca_status launch_batch(const BatchItemDev *items, const BatchUnitDev *units, int n_units, int *counters,
                        const int *part_off, int nparts, int grid, const ZeroDesc *zero, int n_zero, int mode,
                        hipStream_t stream)
{
    if (n_zero > 0) {
        hipLaunchKernelGGL(zero_list_kernel, dim3(256, n_zero), dim3(256), 0, stream, zero);
        TCA_HIP_TRY(hipGetLastError());
    }
    if (n_units <= 0) return TCA_OK;
    if (nparts < 1 || nparts > kBatchParts || nparts > grid) return fail(TCA_ERR_INVALID_ARG, "batch: partitions");
    if (mode == 2)
        hipLaunchKernelGGL(batch_kernel<2>, dim3(grid), dim3(64 * kBatchWaves), 0, stream, items, units, n_units,
                           counters, part_off, nparts);
    else if (mode == 1)
        hipLaunchKernelGGL(batch_kernel<1>, dim3(grid), dim3(64 * kBatchWaves), 0, stream, items, units, n_units,
                           counters, part_off, nparts);
    else
        hipLaunchKernelGGL(batch_kernel<0>, dim3(grid), dim3(64 * kBatchWaves), 0, stream, items, units, n_units,
                           counters, part_off, nparts);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
tca_status launch_warm_up(hipStream_t stream)
{
    hipLaunchKernelGGL(warm_up_kernel, dim3(1), dim3(64), 0, stream);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

uint32_t operand_bytes(long rows, long N, long ld, int elem_bytes)
{
    if (rows <= 0 || N <= 0) return 16;   // never read; a non-empty range keeps the descriptor valid
    const long b = ((rows - 1) * ld + N) * (long)elem_bytes;
    return b < (1L << 31) - 64 ? (uint32_t)b : 0u;
}

int default_tile_n(const tca_plan *p, int N)
{
    if (p->host.rpw >= 128) return N >= 256 ? 128 : 64;
    if (N >= 128) return 64;
    if (N >= 32) return 32;
    return 16;
}

// Work decomposition of a standalone SpMM launch.  A workgroup runs one unit
// (group, chunk range) for one N-tile with W waves splitting the range.  W
// grows (up to what 64 KB of LDS allows, and while each wave keeps >= 4 chunks)
// until ~kTargetWaves waves exist; groups still longer than kMaxWaveChunks
// chunks per wave are split into S units that accumulate atomically.
static constexpr int kTargetWaves = 4096;
static constexpr int kMaxWaveChunks = 32;

static int spmm_slot_bytes(int rpw, int nt, int es)
{
    const int btile = es == 2 ? 32 * (nt == 16 ? 48 : nt + 16) * 2 : nt * 32 * es;   // = SpmmLds<...>::B
    const int tiles = btile + (rpw / 16) * 512 * es;
    const int red = rpw * (nt + 4) * 4;
    return tiles > red ? tiles : red;
}

// Split factor of the launches' unit tables: a group of maxc chunks under W
// waves per workgroup runs as S units (1 when the waves hold it whole).
static int split_for(int maxc, int W) { return maxc > W * kMaxWaveChunks ? (maxc + W * kMaxWaveChunks - 1) / (W * kMaxWaveChunks) : 1; }

// The unit table of split factor S: one (group, pieces, c_lo, c_hi) per unit,
// heavy groups first.  Built for every S a launch can choose when the plan is
// created (prepare_spmm_units), so that launches only look them up: tca_spmm
// allocates nothing, can be captured in a graph from its first call and is
// safe to call from several threads on one plan.
static tca_status spmm_units(const tca_plan *p, int S, const int4 **units, int *n_units, bool *split)
{
    for (const auto &u : p->unit_tables)
        if (u.pieces == S) {
            *units = static_cast<const int4 *>(u.dev);
            *n_units = u.count;
            *split = u.split;
            return TCA_OK;
        }
    return fail(TCA_ERR_INVALID_ARG, "tca_spmm: the plan has no unit table for split factor " + std::to_string(S));
}

tca_status prepare_spmm_units(tca_plan *p)
{
    const PlanTables &t = p->host;
    std::vector<int> splits = {1};
    for (int W = 1; W <= 8; W *= 2) {
        const int S = split_for(t.max_group_chunks, W);
        if (std::find(splits.begin(), splits.end(), S) == splits.end()) splits.push_back(S);
    }
    for (int S : splits) {
        std::vector<int4> h;
        bool any_split = false;
        for (int gi = 0; gi < t.n_groups; gi++) {
            const int g = t.work_order[gi];
            const int cb = t.grp_chunk_ptr[g], nch = t.grp_chunk_ptr[g + 1] - cb;
            const int P = std::max(1, std::min(S, nch / 4));
            if (P > 1) any_split = true;
            for (int k = 0; k < P; k++)
                h.push_back(make_int4(g, P, cb + (int)((long)nch * k / P), cb + (int)((long)nch * (k + 1) / P)));
        }
        tca_plan::UnitTable ut;
        ut.pieces = S;
        ut.count = (int)h.size();
        ut.split = any_split;
        if (!h.empty()) {
            if (hipMalloc(&ut.dev, sizeof(int4) * h.size()) != hipSuccess) return fail(TCA_ERR_ALLOC, "plan: unit table");
            p->unit_tables.push_back(ut);   // owned by the plan from here (freed with it)
            TCA_HIP_TRY(hipMemcpy(ut.dev, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice));
        } else {
            p->unit_tables.push_back(ut);
        }
    }
    return TCA_OK;
}

tca_status launch_spmm(const tca_plan *p, tca_dtype a_dtype, const void *a, tca_dtype b_dtype, const void *B, int ldb,
                       int N, float *C, int ldc, const float *row_scale, const float *col_scale, int tile_n,
                       bool dense, hipStream_t stream)
{
    const PlanTables &t = p->host;
    if (tile_n == 0) tile_n = default_tile_n(p, N);
    if (dense && dense_bytes(t, a_dtype) >= (1L << 31))
        return fail(TCA_ERR_UNSUPPORTED, "tca_spmm_dense: the dense tiles reach 2 GiB (32-bit offsets)");
    const int block_ks = (long)t.n_groups * ((N + tile_n - 1) / tile_n) < kBlockKS2Units ? 2 : 1;
    SpmmFn fn = dense           ? pick_spmm_dense(a_dtype, b_dtype, t.rpw, tile_n, block_ks)
                : p->transposed ? pick_spmm<true>(a_dtype, b_dtype, t.rpw, tile_n)
                                : pick_spmm<false>(a_dtype, b_dtype, t.rpw, tile_n);
    if (!fn) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm: unsupported dtype pair / rows_per_wave / tile_n");
    SpmmParams prm;
    prm.grp_chunk_ptr = p->d_grp_chunk_ptr;
    prm.chunk_cols = p->d_chunk_cols;
    prm.seg_ptr = p->d_seg_ptr;
    prm.entry_slot = p->d_entry_slot;
    prm.entry_lds = p->d_entry_lds;
    prm.entry_src = p->d_entry_src;
    prm.work_order = p->d_work_order;
    prm.a = a;
    prm.B = B;
    prm.C = C;
    prm.row_scale = row_scale;
    prm.col_scale = col_scale;
    prm.n_rows = t.n_rows;
    prm.N = N;
    prm.ldb = ldb;
    prm.ldc = ldc;
    prm.n_tiles = (N + tile_n - 1) / tile_n;
    prm.n_groups = t.n_groups;
    prm.rpw = t.rpw;
    prm.nnz = t.nnz;
    prm.n_chunks = t.n_chunks;
    prm.b_bytes = operand_bytes(t.n_cols, N, ldb, b_dtype == TCA_F32 ? 4 : b_dtype == TCA_E4M3 ? 1 : 2);
    if (prm.b_bytes == 0 || (long)t.nnz * 4 >= (1L << 31) || (long)t.n_chunks * 128 >= (1L << 31))
        return fail(TCA_ERR_UNSUPPORTED, "tca_spmm: B or the plan's tables reach 2 GiB (32-bit offsets)");

    const long base = (long)t.n_groups * prm.n_tiles;
    const int maxc = t.max_group_chunks;
    int W = 1, S = 1;
    size_t lds_bytes = 0;
    if (t.rpw >= 128 && dense) {
        // block kernel: one or two K groups of four waves share each unit (no
        // split: units that add atomically into C cost more than the
        // parallelism gains)
        W = kBlockWaves * block_ks;
        lds_bytes = 0;
    } else if (t.rpw >= 128) {
        // wide groups: W = rows / 32 waves share each B tile, one unit per group
        W = t.rpw / 32;
        lds_bytes = 2 * 32 * (tile_n + 16) * 2 + W * 2 * 512 * 2;   // = WideLds<...>::BYTES
    } else {
        const int es = a_dtype == TCA_E4M3 ? 1 : 2;
        const int slot = spmm_slot_bytes(t.rpw, tile_n, es);
        const int wcap = std::max(1, std::min(8, 65536 / slot));   // = SpmmLds<...>::WMAX
        while (W * 2 <= wcap && maxc / (W * 2) >= 4 && base * W < kTargetWaves) W *= 2;
        S = split_for(maxc, W);
        lds_bytes = (size_t)W * slot;
    }

    int n_units = 0;
    bool split = false;
    tca_status st = spmm_units(p, S, &prm.units, &n_units, &split);
    if (st != TCA_OK) return st;
    prm.split = split ? 1 : 0;
    if (split || n_units <= 0) {  // accumulate from zero (a kernel: no memset node under graph capture)
        const long total = (long)t.n_rows * N;
        const int blocks = (int)std::min<long>((total + 255) / 256, 4096);
        if (blocks > 0) {
            hipLaunchKernelGGL(zero_rows_kernel, dim3(blocks), dim3(256), 0, stream, C, t.n_rows, N, ldc);
            TCA_HIP_TRY(hipGetLastError());
        }
    }
    if (n_units <= 0) return TCA_OK;
    prm.n_units = n_units;
    long grid;
    // A read once per N-tile pass (tile-major) against B slabs read by every
    // XCD (unit-major): deal whichever is heavier once
    const int es_a = a_dtype == TCA_E4M3 ? 1 : 2;
    const double a_bytes = dense ? (double)dense_bytes(t, a_dtype) : (double)t.nnz * (es_a + 2 + (p->transposed ? 4 : 0));
    const double b_bytes = (double)t.n_cols * N * (b_dtype == TCA_F32 ? 4 : b_dtype == TCA_E4M3 ? 1 : 2);
    const int tiles_per_xcd_pass = std::min(prm.n_tiles, 8);
    if ((t.rpw < 128 || dense) && prm.n_tiles > 1 && a_bytes * (tiles_per_xcd_pass - 1) > b_bytes * 7) {
        prm.xcd_share = -2;
        grid = 8L * ((n_units + 7) / 8) * prm.n_tiles;
    } else if (prm.n_tiles % 8 == 0) {
        prm.xcd_share = 0;
        grid = (long)n_units * prm.n_tiles;
    } else if (8 % prm.n_tiles == 0) {
        prm.xcd_share = 8 / prm.n_tiles;
        grid = 8L * ((n_units + prm.xcd_share - 1) / prm.xcd_share);
    } else {
        prm.xcd_share = -1;
        grid = (long)n_units * prm.n_tiles;
    }
    if (grid > 0x7fffffffL) return fail(TCA_ERR_UNSUPPORTED, "tca_spmm: grid too large");
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(64 * W), lds_bytes, stream, prm);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

namespace {
using SddmmFn = void (*)(SddmmParams);

template <int CB>
SddmmFn pick_sddmm(tca_dtype g_dtype, tca_dtype x_dtype)
{
    if (x_dtype == TCA_F16 && g_dtype == TCA_F16) return sddmm_shared_kernel<CB, TCA_F16, TCA_F16>;
    if (x_dtype == TCA_F16 && g_dtype == TCA_F32) return sddmm_shared_kernel<CB, TCA_F16, TCA_F32>;
    if (x_dtype == TCA_BF16 && g_dtype == TCA_BF16) return sddmm_shared_kernel<CB, TCA_BF16, TCA_BF16>;
    if (x_dtype == TCA_BF16 && g_dtype == TCA_F32) return sddmm_shared_kernel<CB, TCA_BF16, TCA_F32>;
    return nullptr;
}
}  // namespace

// SDDMM on the plan's 64-row tiling (sddmm_tiling), dW written in the plan's
// tiled order.  Blocks of 4 chunks per workgroup, halved (to 2, then 1) while
// that leaves fewer than kSddmmMinGroups workgroups: every chunk block re-reads
// its group's G rows, so larger blocks win down to a little under one
// workgroup per CU (measured on the twelve-layer workload: 1024 -> 256 took the
// twelve dW calls from 148 to 117 us, 200 to 108 us -- the 1024 x 1024 layer at
// N = 1024 then runs 240 blocks of 2 instead of 480 of 1 --, 128 was slower
// again; blocks of 8 gained nothing).
static constexpr int kSddmmMinGroups = 200;

tca_status launch_sddmm(const tca_plan *p, tca_dtype g_dtype, const void *G, int ldg, tca_dtype x_dtype, const void *X,
                        int ldx, int N, float *out, hipStream_t stream)
{
    if (!pick_sddmm<4>(g_dtype, x_dtype)) return fail(TCA_ERR_UNSUPPORTED, "tca_sddmm: unsupported dtype pair");
    if (p->transposed) return fail(TCA_ERR_UNSUPPORTED, "tca_sddmm: needs a forward plan");
    const tca_plan *aux = nullptr;
    const int32_t *dst = nullptr;
    tca_status st = sddmm_tiling(p, &aux, &dst);
    if (st != TCA_OK) return st;
    const PlanTables &t = aux->host;
    const int32_t *units = nullptr;
    int n_units = 0, cb = 4;
    st = sddmm_units(aux, 4, &units, &n_units);
    while (st == TCA_OK && n_units < kSddmmMinGroups && cb > 1) {
        cb >>= 1;
        st = sddmm_units(aux, cb, &units, &n_units);
    }
    if (st != TCA_OK) return st;
    if (n_units == 0) return TCA_OK;
    SddmmParams prm;
    prm.sgd_w = nullptr;   // the fused weight update is a batch item option
    prm.sgd_lr = 0.0f;
    prm.sgd_bf16 = 0;
    prm.grp_chunk_ptr = aux->d_grp_chunk_ptr;
    prm.chunk_cols = aux->d_chunk_cols;
    prm.seg_ptr = aux->d_seg_ptr;
    prm.entry_slot = aux->d_entry_slot;
    prm.G = G;
    prm.X = X;
    prm.out = out;
    prm.n_rows = t.n_rows;
    prm.N = N;
    prm.ldg = ldg;
    prm.ldx = ldx;
    prm.rpw = t.rpw;
    prm.n_blocks_per_group_max = 0;
    prm.unit_group = units;
    prm.unit_chunk = units + n_units;
    prm.dst = dst;
    prm.g_bytes = operand_bytes(t.n_rows, N, ldg, g_dtype == TCA_F32 ? 4 : 2);
    prm.x_bytes = operand_bytes(t.n_cols, N, ldx, 2);
    prm.nnz = (uint32_t)t.nnz;
    if (prm.g_bytes == 0 || prm.x_bytes == 0 || (long)t.nnz * 4 >= (1L << 31))
        return fail(TCA_ERR_UNSUPPORTED, "tca_sddmm: G, X or the entry tables reach 2 GiB (32-bit offsets)");
    SddmmFn fn = cb == 4 ? pick_sddmm<4>(g_dtype, x_dtype) : cb == 2 ? pick_sddmm<2>(g_dtype, x_dtype)
                                                                   : pick_sddmm<1>(g_dtype, x_dtype);
    hipLaunchKernelGGL(fn, dim3(n_units), dim3(256), 0, stream, prm);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

namespace {

// Prologue of a batch run, one launch.  Blocks [0, conv_blocks): FP32 -> FP16
// copies, one thread per 8 consecutive elements of one row (two 16-byte loads,
// one 16-byte store); every operand starts on a 256-group boundary (first8),
// so a workgroup's operand is uniform.  Blocks from conv_blocks on: panel-dense
// tiles, one wave per tile as in pack_dense_kernel, the values read through the
// item's map (weight order -> the plan's entries).  Each block finds its
// descriptor by a binary search over the first groups / blocks staged in LDS.
__device__ __forceinline__ int find_first(const long *first, int n, long x)
{
    int lo = 0, hi = n - 1;   // last k with first[k] <= x (first[0] == 0)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void prologue_kernel(const ConvDesc *conv, int n_conv, long conv_blocks,
                                                       const PackDesc *pack, int n_pack)
{
    __shared__ long first[kMaxConv > kMaxPack ? kMaxConv : kMaxPack];
    __shared__ __attribute__((aligned(16))) uint16_t tile[kPackTilesPerBlock][512];
    const long blk = blockIdx.x;
    if (blk < conv_blocks) {
        for (int i = threadIdx.x; i < n_conv; i += 256) first[i] = conv[i].first8;
        __syncthreads();
        const long t0 = blk * 256;
        const int k = find_first(first, n_conv, t0);
        const ConvDesc d = conv[k];
        const long e = (t0 + threadIdx.x - d.first8) * 8;
        if (e >= (long)d.rows * d.N) return;   // this operand's padding
        const long r = e / d.N, col = e - r * d.N;
        const u32x4 *src = reinterpret_cast<const u32x4 *>(d.src + r * d.ld + col);
        const u32x4 a = gbl(src)[0], b = gbl(src)[1];
        const bool bf = d.bf16 != 0;
        auto h = [bf](uint32_t x) {
            const float f = __builtin_bit_cast(float, x);
            return bf ? f32_to_bf16_bits(f) : f32_to_f16_bits(f);
        };
        const u32x4 o = {h(a[0]) | h(a[1]) << 16, h(a[2]) | h(a[3]) << 16, h(b[0]) | h(b[1]) << 16, h(b[2]) | h(b[3]) << 16};
        gbl(reinterpret_cast<u32x4 *>(d.dst + r * d.N + col))[0] = o;
        return;
    }
    for (int i = threadIdx.x; i < n_pack; i += 256) first[i] = pack[i].first_block;
    __syncthreads();
    const long pb = blk - conv_blocks;
    const int k = find_first(first, n_pack, pb);
    const PackDesc d = pack[k];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long t = (pb - d.first_block) * kPackTilesPerBlock + w;
    if (t >= d.n_tiles) return;   // whole waves; no workgroup barrier below
    uint16_t *tl = tile[w];
    reinterpret_cast<uint4 *>(tl)[lane] = make_uint4(0, 0, 0, 0);
    wave_lds_sync();
    const int e1 = gbl(d.seg_ptr)[t + 1];
    for (int e = gbl(d.seg_ptr)[t] + lane; e < e1; e += 64) {
        const int slot = gbl(d.entry_slot)[e], r = slot >> 5, kk = slot & 31;
        tl[((kk >> 3) * 16 + r) * 8 + (kk & 7)] = gbl(d.src)[d.map ? gbl(d.map)[e] : e];
    }
    wave_lds_sync();
    gbl(reinterpret_cast<u32x4 *>(d.dst + t * 512))[lane] = reinterpret_cast<const u32x4 *>(tl)[lane];
}

}  // namespace

tca_status launch_prologue(const ConvDesc *conv, int n_conv, long total8, const PackDesc *pack, int n_pack,
                           long pack_blocks, hipStream_t stream)
{
    if (n_conv > kMaxConv || n_pack > kMaxPack) return fail(TCA_ERR_INVALID_ARG, "batch prologue: too many operands");
    const long conv_blocks = n_conv > 0 ? (total8 + 255) / 256 : 0;
    if (n_pack <= 0) pack_blocks = 0;
    const long blocks = conv_blocks + pack_blocks;
    if (blocks <= 0) return TCA_OK;
    if (blocks > 0x7fffffffL) return fail(TCA_ERR_UNSUPPORTED, "batch prologue: grid too large");
    hipLaunchKernelGGL(prologue_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, conv, n_conv, conv_blocks, pack,
                       n_pack);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

tca_status launch_gather(const int32_t *idx, int n, int elem_bytes, const void *src, void *dst, bool scatter,
                         hipStream_t stream)
{
    if (n <= 0) return TCA_OK;
    const int threads = 256;
    const int blocks = std::min((n + threads - 1) / threads, 65536);
    switch (elem_bytes) {
    case 1:
        hipLaunchKernelGGL(gather_kernel<uint8_t>, dim3(blocks), dim3(threads), 0, stream, idx, n,
                           static_cast<const uint8_t *>(src), static_cast<uint8_t *>(dst), scatter);
        break;
    case 2:
        hipLaunchKernelGGL(gather_kernel<uint16_t>, dim3(blocks), dim3(threads), 0, stream, idx, n,
                           static_cast<const uint16_t *>(src), static_cast<uint16_t *>(dst), scatter);
        break;
    case 4:
        hipLaunchKernelGGL(gather_kernel<uint32_t>, dim3(blocks), dim3(threads), 0, stream, idx, n,
                           static_cast<const uint32_t *>(src), static_cast<uint32_t *>(dst), scatter);
        break;
    default: return fail(TCA_ERR_INVALID_ARG, "gather: element size");
    }
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

tca_status launch_pack_dense(const tca_plan *p, int elem_bytes, const void *tiled, void *dense, hipStream_t stream)
{
    const long n_segs = (long)p->host.n_chunks * (p->host.rpw / 16);
    if (n_segs <= 0) return TCA_OK;
    const long blocks = (n_segs + 3) / 4;
    if (blocks > 0x7fffffffL) return fail(TCA_ERR_UNSUPPORTED, "tca_pack_dense: too many tiles");
    const int32_t *src_idx = p->transposed ? p->d_entry_src : nullptr;
    if (elem_bytes == 2)
        hipLaunchKernelGGL(pack_dense_kernel<uint16_t>, dim3((unsigned)blocks), dim3(256), 0, stream, n_segs, p->d_seg_ptr,
                           p->d_entry_slot, src_idx, static_cast<const uint16_t *>(tiled), static_cast<uint16_t *>(dense));
    else if (elem_bytes == 1)
        hipLaunchKernelGGL(pack_dense_kernel<uint8_t>, dim3((unsigned)blocks), dim3(256), 0, stream, n_segs, p->d_seg_ptr,
                           p->d_entry_slot, src_idx, static_cast<const uint8_t *>(tiled), static_cast<uint8_t *>(dense));
    else
        return fail(TCA_ERR_INVALID_ARG, "tca_pack_dense: element size");
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

tca_status launch_sgd(bool bf16, void *w, const float *g, long n, float lr, hipStream_t stream)
{
    if (n <= 0) return TCA_OK;
    const long n8 = n / 8;
    const long blocks = std::max<long>(1, std::min<long>((n8 + 255) / 256, 8192));
    if (bf16)
        hipLaunchKernelGGL(sgd_kernel<true>, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<uint16_t *>(w), g,
                           n8, n, lr);
    else
        hipLaunchKernelGGL(sgd_kernel<false>, dim3((unsigned)blocks), dim3(256), 0, stream, static_cast<uint16_t *>(w), g,
                           n8, n, lr);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

tca_status launch_quantize_rows(int n_rows, const int32_t *row_offset, const uint16_t *vals, uint8_t *q, float *scale,
                                hipStream_t stream)
{
    if (n_rows <= 0) return TCA_OK;
    hipLaunchKernelGGL(quantize_rows_kernel, dim3((n_rows + 3) / 4), dim3(256), 0, stream, n_rows, row_offset, vals, q,
                       scale);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}

tca_status launch_quantize_cols(int K, int N, const uint16_t *B, int ldb, uint8_t *q, int ldq, float *scale,
                                hipStream_t stream)
{
    if (K <= 0 || N <= 0) return TCA_OK;
    hipLaunchKernelGGL(zero_rows_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, scale, 1, N, N);
    TCA_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(quantize_cols_amax_kernel, dim3((N + 63) / 64, (K + kQuantRowsPerSlice - 1) / kQuantRowsPerSlice),
                       dim3(256), 0, stream, K, N, B, ldb, scale);
    TCA_HIP_TRY(hipGetLastError());
    const long total = (long)K * N;
    const int blocks = (int)std::min<long>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(quantize_cols_apply_kernel, dim3(blocks), dim3(256), 0, stream, K, N, B, ldb, q, ldq, scale);
    TCA_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(quantize_cols_finalize_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, N, scale);
    TCA_HIP_TRY(hipGetLastError());
    return TCA_OK;
}
}  // namespace tca
