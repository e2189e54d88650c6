// Device helpers shared by the gfx950 kernel files: MFMA element traits,
// buffer-resource loads, LDS transposing reads and wave-level ordering.
// Internal to each translation unit (anonymous namespace).
#pragma once
#include "tca_internal.hpp"

#include <hip/hip_runtime.h>
#include <type_traits>

namespace tca {
namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// MFMA element traits.  An "octet" is 8 consecutive-k elements of one operand
// lane: 16 bytes for f16/bf16, 8 bytes for fp8.
// ---------------------------------------------------------------------------
template <tca_dtype DT> struct Elt;
template <> struct Elt<TCA_F16> {
    static constexpr int ES = 2;
    using Oct = uint4;
    __device__ static f32x4 mma(Oct a, Oct b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
    }
};
template <> struct Elt<TCA_BF16> {
    static constexpr int ES = 2;
    using Oct = uint4;
    __device__ static f32x4 mma(Oct a, Oct b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    }
};
template <> struct Elt<TCA_E4M3> {
    static constexpr int ES = 1;
    using Oct = uint2;
    __device__ static f32x4 mma(Oct a, Oct b, f32x4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(__builtin_bit_cast(long, a), __builtin_bit_cast(long, b), c, 0, 0, 0);
    }
};

template <class Oct> __device__ inline Oct zero_oct();
template <> __device__ inline uint4 zero_oct<uint4>() { return make_uint4(0, 0, 0, 0); }
template <> __device__ inline uint2 zero_oct<uint2>() { return make_uint2(0, 0); }

// element e (0..7) of an octet, as raw bits
__device__ inline uint32_t oct_elem(const uint4 &o, int e)
{
    const uint32_t w = e < 2 ? o.x : e < 4 ? o.y : e < 6 ? o.z : o.w;
    return (w >> ((e & 1) * 16)) & 0xffffu;
}
__device__ inline uint32_t oct_elem(const uint2 &o, int e)
{
    const uint32_t w = e < 4 ? o.x : o.y;
    return (w >> ((e & 3) * 8)) & 0xffu;
}

__device__ inline uint32_t f32_to_f16_bits(float f) { return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)f); }
__device__ inline uint32_t f32_to_bf16_bits(float f) { return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)f); }

// 8 consecutive elements of a dense row, loaded raw in the storage type BT and
// converted to the MFMA element type AT only when used, so that a prefetched
// load is not waited for at the point where it is issued.
template <tca_dtype AT, tca_dtype BT> struct RowLoad;
template <tca_dtype AT> struct RowLoad<AT, AT> {
    using Oct = typename Elt<AT>::Oct;
    using Raw = Oct;
    __device__ static Raw load(const void *B, long off)
    {
        using E = typename std::conditional<Elt<AT>::ES == 2, uint16_t, uint8_t>::type;
        return *reinterpret_cast<const Oct *>(static_cast<const E *>(B) + off);
    }
    __device__ static Oct cvt(const Raw &r) { return r; }
};
struct F32x8 {
    float4 a, b;
};
template <> struct RowLoad<TCA_F16, TCA_F32> {
    using Raw = F32x8;
    __device__ static Raw load(const void *B, long off)
    {
        const float4 *q = reinterpret_cast<const float4 *>(static_cast<const float *>(B) + off);
        return F32x8{q[0], q[1]};
    }
    __device__ static uint4 cvt(const Raw &r)
    {
        return make_uint4(f32_to_f16_bits(r.a.x) | f32_to_f16_bits(r.a.y) << 16, f32_to_f16_bits(r.a.z) | f32_to_f16_bits(r.a.w) << 16,
                          f32_to_f16_bits(r.b.x) | f32_to_f16_bits(r.b.y) << 16, f32_to_f16_bits(r.b.z) | f32_to_f16_bits(r.b.w) << 16);
    }
};
template <> struct RowLoad<TCA_BF16, TCA_F32> {
    using Raw = F32x8;
    __device__ static Raw load(const void *B, long off)
    {
        const float4 *q = reinterpret_cast<const float4 *>(static_cast<const float *>(B) + off);
        return F32x8{q[0], q[1]};
    }
    __device__ static uint4 cvt(const Raw &r)
    {
        return make_uint4(f32_to_bf16_bits(r.a.x) | f32_to_bf16_bits(r.a.y) << 16, f32_to_bf16_bits(r.a.z) | f32_to_bf16_bits(r.a.w) << 16,
                          f32_to_bf16_bits(r.b.x) | f32_to_bf16_bits(r.b.y) << 16, f32_to_bf16_bits(r.b.z) | f32_to_bf16_bits(r.b.w) << 16);
    }
};

// ---------------------------------------------------------------------------
// Buffer-resource loads: 32-bit byte offsets from a uniform base, and reads at
// or past `bytes` return zeros.  Padding / masked lanes pass kOOB instead of a
// clamped address, so no select or 64-bit address arithmetic is spent on them.
// Every buffer addressed this way is below 2^31 bytes (checked on the host).
// ---------------------------------------------------------------------------
typedef __amdgpu_buffer_rsrc_t Rsrc;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kOOB = 0x80000000u;

__device__ __forceinline__ Rsrc make_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint4 buf_b128(Rsrc r, uint32_t off, uint32_t soff = 0)
{
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, (int)soff, 0));
}
__device__ __forceinline__ uint2 buf_b64(Rsrc r, uint32_t off, uint32_t soff = 0)
{
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0));
}
__device__ __forceinline__ uint32_t buf_b32(Rsrc r, uint32_t off, uint32_t soff = 0)
{
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0);
}
__device__ __forceinline__ uint32_t buf_b16(Rsrc r, uint32_t off) { return __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0); }
__device__ __forceinline__ uint32_t buf_b8(Rsrc r, uint32_t off) { return __builtin_amdgcn_raw_buffer_load_b8(r, (int)off, 0, 0); }

// 8 consecutive elements of a dense row by buffer load (byte offset; kOOB -> zeros)
template <tca_dtype AT, tca_dtype BT> struct RowBuf;
template <tca_dtype AT> struct RowBuf<AT, AT> {
    static constexpr int BES = Elt<AT>::ES;          // bytes per stored element
    using Raw = typename Elt<AT>::Oct;
    __device__ static Raw load(Rsrc r, uint32_t off, uint32_t soff = 0)
    {
        if constexpr (BES == 2) return buf_b128(r, off, soff);
        else return buf_b64(r, off, soff);
    }
};
template <tca_dtype AT> struct RowBuf<AT, TCA_F32> {
    static constexpr int BES = 4;
    using Raw = F32x8;
    __device__ static Raw load(Rsrc r, uint32_t off, uint32_t soff = 0)
    {
        return F32x8{__builtin_bit_cast(float4, buf_b128(r, off, soff)), __builtin_bit_cast(float4, buf_b128(r, off + 16, soff))};
    }
};
template <tca_dtype AT> __device__ __forceinline__ uint32_t load_value(Rsrc r, uint32_t idx)
{
    if constexpr (Elt<AT>::ES == 2) return buf_b16(r, idx * 2);
    else return buf_b8(r, idx);
}

// ---------------------------------------------------------------------------
// LDS layouts (offsets in octets).  B tile: NT rows (n) x 32 k; two n-rows
// share a 128-byte line of 8 octet slots, XOR-swizzled so that the 8 lanes of
// an MFMA operand read (n = 16s+0..7, fixed q) and the 16 lanes of a
// transposed write (8 n-segments x 2 k-halves) each hit distinct banks.
// A tile: 16 rows x 32 k, same pairing with a 2-bit swizzle.
// ---------------------------------------------------------------------------
__device__ inline int b_oct(int n, int q)
{
    const int m = n >> 1;
    const int sigma = ((m & 3) ^ ((m >> 2) & 7)) & 7;
    return m * 8 + ((((n & 1) << 2) | q) ^ sigma);
}
__device__ inline int a_oct(int r, int q)
{
    const int m = r >> 1;
    return m * 8 + ((((r & 1) << 2) | q) ^ (m & 3));
}

// Transposing LDS read (ds_read_b64_tr_b16): lane 4r+j of each 16-lane group
// addresses row r, columns 4j..4j+3 of a row-major 16-bit tile, and lane i of
// the group receives column i of the four rows -- 4 consecutive k of one n,
// half of an MFMA B operand octet.
typedef short s16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint2 lds_read_tr16(const uint8_t *lds, int byte_off)
{
    const s16x4 v = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) s16x4 *)(lds + byte_off));
    return __builtin_bit_cast(uint2, v);
}

// Row pitch (elements) of a row-major 32 x NT 16-bit B tile: 4 consecutive rows
// start in distinct quarters of the 32 banks (pitch = 32 or 96 bytes mod 128).
template <int NT> constexpr int tr_pitch() { return NT == 16 ? 48 : NT + 16; }

// Byte offset of row k in such a tile of pitch_bytes per row: within rows
// 8..15 (and 24..31, 40..47, ...) the two halves of four rows trade places.
// gfx950's LDS serves a transposing read 32 lanes (two lane groups q, rows
// 8q + 0..3, or 8q + 4..7) per cycle over 64 banks; eight rows of these
// pitches span a multiple of 256 bytes, so rows r and 8 + r fall on the same
// banks in plain order and on distinct ones in this order (every pitch here is
// an odd multiple of 32 bytes modulo 256: eight consecutive row slots take
// eight distinct 32-byte bank windows).  A permutation of the rows: the tile
// keeps its size.
__device__ __forceinline__ int tr_row(int k, int pitch_bytes) { return (k ^ ((k >> 1) & 4)) * pitch_bytes; }

// Row of the m-th row slot when a workgroup writes a B tile in 16-byte pieces
// of SEGS per row (piece p: row slot p / SEGS, segment p % SEGS).  A write
// cycle serves 16 lanes, i.e. 16 / SEGS rows; taken in order, consecutive rows
// of these pitches overlap on 64 banks, so the low three bits of m are rotated
// to spread each cycle's rows four (two rows per cycle) or two (four rows)
// apart, onto distinct banks.  A permutation of the rows.
template <int SEGS> __device__ __forceinline__ int piece_row(int m)
{
    constexpr int S = SEGS == 8 ? 1 : SEGS == 4 ? 2 : 0;
    if constexpr (S == 0) return m;
    const int low = m & 7;
    return (m & ~7) | (((low >> S) | (low << (3 - S))) & 7);
}

// Orders one wave's LDS accesses (write -> read -> write of its own tiles).
// LDS instructions of a wave execute in issue order, so only the compiler
// has to be kept from moving memory operations across this point.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wait for every vector-memory load issued so far.  Placed between a software
// pipeline's prologue and its loop: the compiler derives each wait in the loop
// from the state merged over the loop's entry and its back edge, and an entry
// with loads still in flight (issued in another order than in the steady
// state) makes it wait for nearly the whole ring at the top of every iteration.
// An entry with nothing in flight leaves the steady-state counts.  The first
// chunks' loads are needed at once anyway, so the drain costs nothing.
__device__ __forceinline__ void vm_drain()
{
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0), expcnt / lgkmcnt unconstrained
}

// Lane index that the compiler cannot treat as loop-invariant: inside the
// persistent batch loop, values derived from the lane would otherwise be
// hoisted out of the loop for every variant and stay live across all of them.
__device__ __forceinline__ int lane_id()
{
    int l = (int)threadIdx.x & 63;
    asm volatile("" : "+v"(l));
    return l;
}

// Global-address-space view of a pointer.  Pointers the batch kernel reads out
// of its item descriptors are generic to the compiler, which would then emit
// FLAT accesses; while a FLAT load is pending, every wait on the vector-memory
// counter becomes a full drain (vmcnt(0)), which serialises the chunk pipeline.
// Every device pointer here is global memory, so the cast is exact.
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T *gbl(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// Relaxed device-scope float add in global memory (hardware atomic, no CAS loop).
__device__ __forceinline__ void atomic_add_f32(float *p, float v)
{
    [[clang::atomic(ignore_denormal_mode)]] { __hip_atomic_fetch_add(gbl(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
}

}  // namespace
}  // namespace tca
