// Host-side pieces of the library: error state, .smtx I/O, dense<->CSR,
// compare_mat, seeded generators and the panel-tiling builder.
#include "tca_internal.hpp"

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

namespace tca {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

tca_status fail(tca_status s, const std::string &msg)
{
    g_last_error = msg;
    return s;
}

}  // namespace tca

using tca::fail;

extern "C" const char *tca_last_error(void) { return tca::g_last_error.c_str(); }
extern "C" int tca_version(void) { return TCA_VERSION; }
extern "C" void tca_host_free(void *p) { std::free(p); }

extern "C" const char *tca_status_string(tca_status s)
{
    switch (s) {
    case TCA_OK: return "ok";
    case TCA_ERR_INVALID_ARG: return "invalid argument";
    case TCA_ERR_UNSUPPORTED: return "unsupported";
    case TCA_ERR_HIP: return "HIP error";
    case TCA_ERR_ALLOC: return "allocation failed";
    case TCA_ERR_IO: return "I/O error";
    case TCA_ERR_FORMAT: return "malformed input";
    case TCA_ERR_LIBRARY: return "library error";
    }
    return "unknown status";
}

// ---------------------------------------------------------------------------
// .smtx: "rows, cols, nnz" / rows+1 offsets / nnz column ids
// (format of src/file_input.cu:135-193)
// ---------------------------------------------------------------------------
namespace {

struct Cursor {
    const char *p, *end;
    void skip_blank_inline()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) p++;
    }
    void skip_space()
    {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) p++;
    }
    bool read_long(long &v)
    {
        skip_blank_inline();
        if (p >= end) return false;
        char buf[32];
        size_t n = 0;
        const char *q = p;
        if (q < end && (*q == '-' || *q == '+')) buf[n++] = *q++;
        while (q < end && n < sizeof(buf) - 1 && *q >= '0' && *q <= '9') buf[n++] = *q++;
        if (n == 0 || (n == 1 && (buf[0] == '-' || buf[0] == '+'))) return false;
        buf[n] = 0;
        errno = 0;
        v = std::strtol(buf, nullptr, 10);
        if (errno) return false;
        p = q;
        return true;
    }
};

}  // namespace

namespace tca {

tca_status smtx_parse(const std::string &text, int *n_rows, int *n_cols, int *nnz, int32_t **row_offset,
                      int32_t **col_idx)
{
    Cursor c{text.data(), text.data() + text.size()};
    const char *eol = static_cast<const char *>(std::memchr(c.p, '\n', (size_t)(c.end - c.p)));
    if (!eol) return fail(TCA_ERR_FORMAT, "smtx: missing header line");
    Cursor h{c.p, eol};
    long hdr[3];
    for (int i = 0; i < 3; i++) {
        if (!h.read_long(hdr[i])) return fail(TCA_ERR_FORMAT, "smtx: header is not 'rows, cols, nnz'");
        h.skip_blank_inline();
        if (i < 2) {
            if (h.p >= h.end || *h.p != ',') return fail(TCA_ERR_FORMAT, "smtx: header is not 'rows, cols, nnz'");
            h.p++;
        }
    }
    if (hdr[0] <= 0 || hdr[1] <= 0 || hdr[2] < 0 || hdr[0] > (1L << 28) || hdr[1] > (1L << 28) || hdr[2] > (1L << 31) - 1)
        return fail(TCA_ERR_FORMAT, "smtx: header values out of range");
    const int M = (int)hdr[0], K = (int)hdr[1], NZ = (int)hdr[2];
    if ((long)NZ > (long)M * (long)K) return fail(TCA_ERR_FORMAT, "smtx: nnz exceeds rows*cols");

    c.p = eol + 1;
    const char *eol2 = static_cast<const char *>(std::memchr(c.p, '\n', (size_t)(c.end - c.p)));
    Cursor r{c.p, eol2 ? eol2 : c.end};
    int32_t *rp = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * ((size_t)M + 1)));
    if (!rp) return fail(TCA_ERR_ALLOC, "smtx: row offsets");
    int cnt = 0;
    for (;;) {
        r.skip_blank_inline();
        if (r.p >= r.end) break;
        long v;
        if (!r.read_long(v)) { std::free(rp); return fail(TCA_ERR_FORMAT, "smtx: bad row-offset token"); }
        if (cnt >= M + 1) { std::free(rp); return fail(TCA_ERR_FORMAT, "smtx: more than rows+1 row offsets"); }
        rp[cnt++] = (int32_t)v;
    }
    if (cnt != M + 1) { std::free(rp); return fail(TCA_ERR_FORMAT, "smtx: fewer than rows+1 row offsets"); }
    if (rp[0] != 0 || rp[M] != NZ) { std::free(rp); return fail(TCA_ERR_FORMAT, "smtx: row offsets do not span [0, nnz]"); }
    for (int i = 0; i < M; i++)
        if (rp[i + 1] < rp[i]) { std::free(rp); return fail(TCA_ERR_FORMAT, "smtx: row offsets decrease"); }

    c.p = eol2 ? eol2 + 1 : c.end;
    int32_t *ci = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * ((size_t)NZ + 1)));
    if (!ci) { std::free(rp); return fail(TCA_ERR_ALLOC, "smtx: column ids"); }
    int k = 0;
    for (;;) {
        c.skip_space();
        if (c.p >= c.end) break;
        long v;
        if (!c.read_long(v)) { std::free(rp); std::free(ci); return fail(TCA_ERR_FORMAT, "smtx: bad column token"); }
        if (k >= NZ) { std::free(rp); std::free(ci); return fail(TCA_ERR_FORMAT, "smtx: more than nnz column ids"); }
        if (v < 0 || v >= K) { std::free(rp); std::free(ci); return fail(TCA_ERR_FORMAT, "smtx: column id out of range"); }
        ci[k++] = (int32_t)v;
    }
    if (k != NZ) { std::free(rp); std::free(ci); return fail(TCA_ERR_FORMAT, "smtx: fewer than nnz column ids"); }
    *n_rows = M;
    *n_cols = K;
    *nnz = NZ;
    *row_offset = rp;
    *col_idx = ci;
    return TCA_OK;
}

}  // namespace tca

extern "C" tca_status tca_smtx_read(const char *path, int *n_rows, int *n_cols, int *nnz, int32_t **row_offset,
                                    int32_t **col_idx)
{
    if (!path || !n_rows || !n_cols || !nnz || !row_offset || !col_idx)
        return fail(TCA_ERR_INVALID_ARG, "tca_smtx_read: NULL argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(TCA_ERR_IO, std::string("tca_smtx_read: cannot open ") + path);
    std::string text;
    char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) text.append(chunk, got);
    std::fclose(f);
    return tca::smtx_parse(text, n_rows, n_cols, nnz, row_offset, col_idx);
}

extern "C" tca_status tca_smtx_write(const char *path, int n_rows, int n_cols, int nnz, const int32_t *row_offset,
                                     const int32_t *col_idx)
{
    if (!path || !row_offset || (nnz > 0 && !col_idx) || n_rows <= 0 || n_cols <= 0 || nnz < 0)
        return fail(TCA_ERR_INVALID_ARG, "tca_smtx_write: bad argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(TCA_ERR_IO, std::string("tca_smtx_write: cannot open ") + path);
    std::fprintf(f, "%d, %d, %d\n", n_rows, n_cols, nnz);
    for (int i = 0; i <= n_rows; i++) std::fprintf(f, i ? " %d" : "%d", row_offset[i]);
    std::fputc('\n', f);
    for (int i = 0; i < nnz; i++) std::fprintf(f, i ? " %d" : "%d", col_idx[i]);
    std::fputc('\n', f);
    if (std::fclose(f) != 0) return fail(TCA_ERR_IO, "tca_smtx_write: write failed");
    return TCA_OK;
}

// ---------------------------------------------------------------------------
// MatrixMarket coordinate reader (src/file_input.cu:8-134)
// ---------------------------------------------------------------------------
namespace {
uint16_t f32_to_f16_rne(float f);
}

extern "C" tca_status tca_mtx_read(const char *path, int one_based, int pattern_only, int *n_rows, int *n_cols,
                                   int *nnz, int32_t **row_offset, int32_t **col_idx, uint16_t **values)
{
    if (!path || !n_rows || !n_cols || !nnz || !row_offset || !col_idx || !values)
        return fail(TCA_ERR_INVALID_ARG, "tca_mtx_read: NULL argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(TCA_ERR_IO, std::string("tca_mtx_read: cannot open ") + path);
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    std::fclose(f);

    struct Ent { int32_t r, c; float v; int32_t seq; };
    std::vector<Ent> ents;
    long M = -1, K = -1, NZ = -1;
    size_t pos = 0;
    int seq = 0;
    while (pos < text.size()) {
        size_t eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        const std::string line = text.substr(pos, eol - pos);
        pos = eol + 1;
        size_t first = line.find_first_not_of(" \t\r");
        if (first == std::string::npos) continue;     // blank line
        if (line[first] == '%') continue;             // comment / banner
        const char *p = line.c_str() + first;
        char *end;
        if (M < 0) {
            M = std::strtol(p, &end, 10);
            if (end == p) return fail(TCA_ERR_FORMAT, "mtx: bad header");
            p = end;
            K = std::strtol(p, &end, 10);
            if (end == p) return fail(TCA_ERR_FORMAT, "mtx: bad header");
            p = end;
            NZ = std::strtol(p, &end, 10);
            if (end == p) return fail(TCA_ERR_FORMAT, "mtx: bad header");
            if (M <= 0 || K <= 0 || NZ < 0 || M > (1L << 28) || K > (1L << 28) || NZ > 0x7fffffffL)
                return fail(TCA_ERR_FORMAT, "mtx: header values out of range");
            continue;
        }
        long r = std::strtol(p, &end, 10);
        if (end == p) return fail(TCA_ERR_FORMAT, "mtx: bad entry line");
        p = end;
        long c = std::strtol(p, &end, 10);
        if (end == p) return fail(TCA_ERR_FORMAT, "mtx: bad entry line");
        p = end;
        float v = 1.0f;
        if (!pattern_only) {
            const float x = std::strtof(p, &end);
            if (end != p) v = x;
        }
        if (one_based) { r--; c--; }
        if (r < 0 || r >= M || c < 0 || c >= K) return fail(TCA_ERR_FORMAT, "mtx: entry index out of range");
        ents.push_back({(int32_t)r, (int32_t)c, v, seq++});
    }
    if (M < 0) return fail(TCA_ERR_FORMAT, "mtx: missing header");
    if ((long)ents.size() != NZ) return fail(TCA_ERR_FORMAT, "mtx: entry count differs from the header");
    std::stable_sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
        return a.r != b.r ? a.r < b.r : a.c < b.c;
    });
    // repeated (r, c): the last one in file order wins (stable sort keeps file order)
    std::vector<Ent> uniq;
    for (const Ent &e : ents) {
        if (!uniq.empty() && uniq.back().r == e.r && uniq.back().c == e.c) uniq.back() = e;
        else uniq.push_back(e);
    }
    int32_t *rp = static_cast<int32_t *>(std::calloc((size_t)M + 1, sizeof(int32_t)));
    int32_t *ci = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * (uniq.size() + 1)));
    uint16_t *vv = static_cast<uint16_t *>(std::malloc(sizeof(uint16_t) * (uniq.size() + 1)));
    if (!rp || !ci || !vv) {
        std::free(rp); std::free(ci); std::free(vv);
        return fail(TCA_ERR_ALLOC, "mtx: CSR arrays");
    }
    for (size_t k = 0; k < uniq.size(); k++) {
        rp[uniq[k].r + 1]++;
        ci[k] = uniq[k].c;
        vv[k] = f32_to_f16_rne(uniq[k].v);
    }
    for (long r = 0; r < M; r++) rp[r + 1] += rp[r];
    *n_rows = (int)M;
    *n_cols = (int)K;
    *nnz = (int)uniq.size();
    *row_offset = rp;
    *col_idx = ci;
    *values = vv;
    return TCA_OK;
}

// ---------------------------------------------------------------------------
// dense <-> CSR on fp16 bits (src/miscalleneous.cu:29-65, :66-82)
// A half is nonzero iff its float value != 0.0f: +-0 are dropped, NaN is kept.
// ---------------------------------------------------------------------------
static inline bool f16_nonzero(uint16_t h) { return (h & 0x7fffu) != 0; }

extern "C" tca_status tca_dense_to_csr_f16(const uint16_t *dense, int n_rows, int n_cols, int32_t *row_offset,
                                           int32_t *col_idx, uint16_t *values, int capacity, int *nnz)
{
    if (!dense || n_rows < 0 || n_cols < 0 || !nnz) return fail(TCA_ERR_INVALID_ARG, "tca_dense_to_csr_f16: bad argument");
    long count = 0;
    if (row_offset) row_offset[0] = 0;
    for (int r = 0; r < n_rows; r++) {
        const uint16_t *row = dense + (size_t)r * n_cols;
        for (int c = 0; c < n_cols; c++) {
            if (!f16_nonzero(row[c])) continue;
            if (col_idx) {
                if (count >= capacity) return fail(TCA_ERR_INVALID_ARG, "tca_dense_to_csr_f16: capacity too small");
                col_idx[count] = c;
                if (values) values[count] = row[c];
            }
            count++;
        }
        if (row_offset) row_offset[r + 1] = (int32_t)count;
    }
    if (count > 0x7fffffffL) return fail(TCA_ERR_UNSUPPORTED, "tca_dense_to_csr_f16: nnz exceeds int32");
    *nnz = (int)count;
    return TCA_OK;
}

extern "C" tca_status tca_csr_to_dense_f16(int n_rows, int n_cols, const int32_t *row_offset, const int32_t *col_idx,
                                           const uint16_t *values, uint16_t *dense)
{
    if (!row_offset || !dense || n_rows < 0 || n_cols < 0) return fail(TCA_ERR_INVALID_ARG, "tca_csr_to_dense_f16: bad argument");
    std::memset(dense, 0, sizeof(uint16_t) * (size_t)n_rows * n_cols);
    for (int r = 0; r < n_rows; r++)
        for (int e = row_offset[r]; e < row_offset[r + 1]; e++) {
            int c = col_idx[e];
            if (c < 0 || c >= n_cols) return fail(TCA_ERR_FORMAT, "tca_csr_to_dense_f16: column out of range");
            dense[(size_t)r * n_cols + c] = values[e];
        }
    return TCA_OK;
}

// ---------------------------------------------------------------------------
// CSR -> Blocked-ELL (the host conversion of perform_cuSPARSE_BlockSpMM,
// src/spmm_cusparse.cu:254-331, from CSR instead of the dense matrix): a block
// is kept when it holds an entry; each block row lists its kept block columns
// ascending, -1 padded to the longest block row; values are the padded
// (ceil(M/b)*b) x ell_cols row-major array, zeros off the pattern.
// ---------------------------------------------------------------------------
extern "C" tca_status tca_bell_from_csr(int n_rows, int n_cols, const int32_t *row_offset, const int32_t *col_idx,
                                        const uint16_t *values, int block, int *ell_cols, int32_t *col_ind,
                                        uint16_t *ell_values)
{
    if (n_rows < 0 || n_cols < 0 || !row_offset || !ell_cols || (row_offset[n_rows] > 0 && !col_idx))
        return fail(TCA_ERR_INVALID_ARG, "tca_bell_from_csr: bad argument");
    if (block < 16 || block > 256 || (block & (block - 1)))
        return fail(TCA_ERR_UNSUPPORTED, "tca_bell_from_csr: block must be 16, 32, 64, 128 or 256");
    const int BR = (n_rows + block - 1) / block, BC = (n_cols + block - 1) / block;
    std::vector<int> stamp((size_t)std::max(BC, 1), -1), pos((size_t)std::max(BC, 1), 0);
    std::vector<int> list;
    int slots = 0;
    for (int br = 0; br < BR; br++) {
        int n = 0;
        for (int r = br * block; r < std::min(n_rows, br * block + block); r++)
            for (int e = row_offset[r]; e < row_offset[r + 1]; e++) {
                const int c = col_idx[e];
                if (c < 0 || c >= n_cols) return fail(TCA_ERR_FORMAT, "tca_bell_from_csr: column out of range");
                if (stamp[c / block] != br) {
                    stamp[c / block] = br;
                    n++;
                }
            }
        slots = std::max(slots, n);
    }
    if ((long)slots * block > 0x7fffffffL) return fail(TCA_ERR_UNSUPPORTED, "tca_bell_from_csr: ell_cols exceeds int32");
    *ell_cols = slots * block;
    if (!col_ind || !ell_values) return TCA_OK;
    const size_t width = (size_t)slots * block;
    std::memset(ell_values, 0, sizeof(uint16_t) * (size_t)BR * block * width);
    std::fill(stamp.begin(), stamp.end(), -1);
    for (int br = 0; br < BR; br++) {
        list.clear();
        const int r1 = std::min(n_rows, br * block + block);
        for (int r = br * block; r < r1; r++)
            for (int e = row_offset[r]; e < row_offset[r + 1]; e++)
                if (stamp[col_idx[e] / block] != br) {
                    stamp[col_idx[e] / block] = br;
                    list.push_back(col_idx[e] / block);
                }
        std::sort(list.begin(), list.end());
        int32_t *ci = col_ind + (size_t)br * slots;
        for (int j = 0; j < slots; j++) ci[j] = j < (int)list.size() ? list[j] : -1;
        for (int j = 0; j < (int)list.size(); j++) pos[list[j]] = j;
        for (int r = br * block; r < r1; r++)
            for (int e = row_offset[r]; e < row_offset[r + 1]; e++) {
                const int c = col_idx[e];
                ell_values[(size_t)r * width + (size_t)pos[c / block] * block + c % block] = values ? values[e] : 0x3C00;
            }
    }
    return TCA_OK;
}

// compare_mat<float> (src/general.cuh:71-123)
extern "C" int tca_compare_mat_f32(const float *gold, const float *cmp, int n, float tol, int *n_fail)
{
    int count = 0, ok = 1;
    for (int i = 0; i < n; i++) {
        const float d = std::fabs(gold[i] - cmp[i]);
        const float rel = std::fabs(d / gold[i]);
        if (rel >= tol && d >= tol) {
            ok = 0;
            if (++count >= 10) break;
        }
    }
    if (n_fail) *n_fail = count;
    return ok;
}

// ---------------------------------------------------------------------------
// Counter-based generators.  Stream (seed, stream) is a SplitMix64 sequence
// started at key = mix(seed ^ mix(stream)); element i is its (i+1)-th output,
// so any element can be produced independently.
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

inline uint64_t finalize(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline uint64_t stream_key(uint64_t seed, uint64_t stream) { return finalize(seed ^ finalize(stream + kGolden)); }
inline uint64_t draw(uint64_t key, uint64_t i) { return finalize(key + (i + 1) * kGolden); }
inline double u01(uint64_t key, uint64_t i) { return (double)(draw(key, i) >> 11) * 0x1.0p-53; }
inline double normal01(uint64_t key, uint64_t i)
{
    const double u1 = u01(key, 2 * i), u2 = u01(key, 2 * i + 1);
    return std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(6.283185307179586476925 * u2);
}

uint16_t f32_to_f16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint16_t s = (uint16_t)((u >> 16) & 0x8000u);
    const uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(s | 0x7e00u);
    if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00u);
    if (a < 0x38800000u) {
        // subnormal range: the value in units of 2^-24, rounded half to even
        const uint32_t mant = (a & 0x7fffffu) | 0x800000u;
        const int exp = (int)(a >> 23);
        if (exp < 102) return s;  // below half of the smallest subnormal
        const int shift = 126 - exp;  // value / 2^-24 = mant >> shift (mant has 23 fraction bits)
        uint32_t q = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1))) q++;
        return (uint16_t)(s | q);
    }
    uint32_t out = ((((a >> 23) - 112u) << 10) | ((a >> 13) & 0x3ffu));
    const uint32_t rest = a & 0x1fffu;
    if (rest > 0x1000u || (rest == 0x1000u && (out & 1u))) out++;
    return (uint16_t)(s | out);
}

}  // namespace

extern "C" uint64_t tca_rng_u64(uint64_t seed, uint64_t stream, uint64_t index)
{
    return draw(stream_key(seed, stream), index);
}

extern "C" tca_status tca_gen_normal_f32(uint64_t seed, uint64_t stream, size_t n, double mean, double std_,
                                         float *out)
{
    if (!out && n) return fail(TCA_ERR_INVALID_ARG, "tca_gen_normal_f32: NULL out");
    const uint64_t key = stream_key(seed, stream);
    for (size_t i = 0; i < n; i++) out[i] = (float)(mean + std_ * normal01(key, i));
    return TCA_OK;
}

extern "C" tca_status tca_gen_uniform_f32(uint64_t seed, uint64_t stream, size_t n, double lo, double hi, float *out)
{
    if (!out && n) return fail(TCA_ERR_INVALID_ARG, "tca_gen_uniform_f32: NULL out");
    const uint64_t key = stream_key(seed, stream);
    for (size_t i = 0; i < n; i++) out[i] = (float)(lo + (hi - lo) * u01(key, i));
    return TCA_OK;
}

extern "C" tca_status tca_f32_to_f16_host(const float *in, size_t n, uint16_t *out)
{
    if ((!in || !out) && n) return fail(TCA_ERR_INVALID_ARG, "tca_f32_to_f16_host: NULL");
    for (size_t i = 0; i < n; i++) out[i] = f32_to_f16_rne(in[i]);
    return TCA_OK;
}

extern "C" tca_status tca_f32_to_bf16_host(const float *in, size_t n, uint16_t *out)
{
    if ((!in || !out) && n) return fail(TCA_ERR_INVALID_ARG, "tca_f32_to_bf16_host: NULL");
    for (size_t i = 0; i < n; i++) {
        uint32_t u;
        std::memcpy(&u, &in[i], 4);
        if ((u & 0x7fffffffu) > 0x7f800000u) out[i] = (uint16_t)((u >> 16) | 0x40u);
        else out[i] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
    return TCA_OK;
}

// Row lengths then distinct sorted columns (Floyd's sampling per row).
extern "C" tca_status tca_gen_pattern(tca_pattern_kind kind, int n_rows, int n_cols, double density, double param,
                                      uint64_t seed, uint64_t stream, int32_t *row_offset, int32_t **col_idx, int *nnz)
{
    if (n_rows <= 0 || n_cols <= 0 || !row_offset || !col_idx || !nnz || density < 0.0 || density > 1.0)
        return fail(TCA_ERR_INVALID_ARG, "tca_gen_pattern: bad argument");
    std::vector<int> len(n_rows, 0);
    const uint64_t klen = stream_key(seed, stream * 4 + 1);
    const uint64_t kcol = stream_key(seed, stream * 4 + 2);
    std::vector<int32_t> cols;
    if (kind == TCA_PATTERN_BERNOULLI) {
        row_offset[0] = 0;
        for (int r = 0; r < n_rows; r++) {
            for (int c = 0; c < n_cols; c++)
                if (u01(kcol, (uint64_t)r * (uint64_t)n_cols + (uint64_t)c) < density) cols.push_back(c);
            row_offset[r + 1] = (int32_t)cols.size();
        }
    } else {
        const double target = density * (double)n_rows * (double)n_cols;
        if (kind == TCA_PATTERN_LOGNORMAL) {
            std::vector<double> w(n_rows);
            double sum = 0.0;
            for (int r = 0; r < n_rows; r++) { w[r] = std::exp(param * normal01(klen, (uint64_t)r)); sum += w[r]; }
            for (int r = 0; r < n_rows; r++) {
                const double v = std::nearbyint(w[r] / sum * target);
                len[r] = (int)std::min<double>(std::max(v, 0.0), (double)n_cols);
            }
        } else if (kind == TCA_PATTERN_UNIFORM) {
            const double mean = density * (double)n_cols;
            for (int r = 0; r < n_rows; r++) {
                const double f = 1.0 - param + 2.0 * param * u01(klen, (uint64_t)r);
                len[r] = (int)std::min<double>(std::max(std::nearbyint(mean * f), 0.0), (double)n_cols);
            }
        } else {
            return fail(TCA_ERR_INVALID_ARG, "tca_gen_pattern: unknown kind");
        }
        std::vector<uint8_t> mark(n_cols, 0);
        row_offset[0] = 0;
        for (int r = 0; r < n_rows; r++) {
            const int L = len[r];
            // Floyd: for j = K-L .. K-1 pick t in [0, j]; take t unless taken, else j
            for (int step = 0, j = n_cols - L; j < n_cols; j++, step++) {
                const uint64_t idx = ((uint64_t)r << 32) | (uint64_t)step;
                int t = (int)(u01(kcol, idx) * (double)(j + 1));
                if (t > j) t = j;
                if (mark[t]) t = j;
                mark[t] = 1;
            }
            for (int c = 0; c < n_cols; c++)
                if (mark[c]) { cols.push_back(c); mark[c] = 0; }
            row_offset[r + 1] = (int32_t)cols.size();
        }
    }
    int32_t *out = static_cast<int32_t *>(std::malloc(sizeof(int32_t) * (cols.size() + 1)));
    if (!out) return fail(TCA_ERR_ALLOC, "tca_gen_pattern: columns");
    std::copy(cols.begin(), cols.end(), out);
    *col_idx = out;
    *nnz = (int)cols.size();
    return TCA_OK;
}

// ---------------------------------------------------------------------------
// Panel tiling builder (sort-merge per row group).
// ---------------------------------------------------------------------------
namespace tca {

tca_status build_tables(int n_rows, int n_cols, int nnz, const int32_t *row_offset, const int32_t *col_idx, int rpw,
                        PlanTables &t)
{
    if (rpw != 16 && rpw != 32 && rpw != 64 && rpw != 128 && rpw != 256)
        return fail(TCA_ERR_INVALID_ARG, "rows_per_wave must be 16, 32, 64, 128 or 256");
    if (n_rows <= 0 || n_cols <= 0 || nnz < 0 || !row_offset || (nnz > 0 && !col_idx))
        return fail(TCA_ERR_INVALID_ARG, "plan: bad CSR arguments");
    if (row_offset[0] != 0 || row_offset[n_rows] != nnz) return fail(TCA_ERR_FORMAT, "plan: row offsets do not span [0, nnz]");
    // offsets first (monotone, inside [0, nnz]), so that the column pass below
    // never reads past col_idx[nnz - 1]
    for (int r = 0; r < n_rows; r++)
        if (row_offset[r + 1] < row_offset[r] || row_offset[r + 1] > nnz)
            return fail(TCA_ERR_FORMAT, "plan: row offsets decrease or pass nnz");
    for (int r = 0; r < n_rows; r++) {
        for (int e = row_offset[r]; e < row_offset[r + 1]; e++) {
            const int c = col_idx[e];
            if (c < 0 || c >= n_cols) return fail(TCA_ERR_FORMAT, "plan: column out of range");
            if (e > row_offset[r] && col_idx[e - 1] >= c) return fail(TCA_ERR_FORMAT, "plan: columns not sorted/distinct");
        }
    }
    const int G = (n_rows + rpw - 1) / rpw, ppg = rpw / 16;
    t.n_rows = n_rows;
    t.n_cols = n_cols;
    t.nnz = nnz;
    t.rpw = rpw;
    t.n_groups = G;

    // pass 1: per-group union (sorted) -> chunk counts; per-entry (chunk, slot)
    std::vector<std::vector<int32_t>> unions(G);
    std::vector<int32_t> ent_chunk(nnz);  // chunk within group
    std::vector<uint16_t> ent_slot(nnz);
    t.grp_chunk_ptr.assign(G + 1, 0);
    for (int g = 0; g < G; g++) {
        const int r0 = g * rpw, r1 = std::min(n_rows, r0 + rpw);
        std::vector<int32_t> &u = unions[g];
        u.assign(col_idx + row_offset[r0], col_idx + row_offset[r1]);
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        for (int r = r0; r < r1; r++) {
            // merge walk: the row's columns are ascending, so its union positions are too
            size_t p = 0;
            for (int e = row_offset[r]; e < row_offset[r + 1]; e++) {
                while (u[p] != col_idx[e]) p++;
                ent_chunk[e] = (int32_t)(p >> 5);
                ent_slot[e] = (uint16_t)(((r - r0) & 15) << 5 | (p & 31));
            }
        }
        t.grp_chunk_ptr[g + 1] = t.grp_chunk_ptr[g] + (int32_t)((u.size() + 31) / 32);
    }
    const int C = t.grp_chunk_ptr[G];
    t.n_chunks = C;
    t.max_group_chunks = 0;
    for (int g = 0; g < G; g++) t.max_group_chunks = std::max(t.max_group_chunks, t.grp_chunk_ptr[g + 1] - t.grp_chunk_ptr[g]);
    t.chunk_cols.assign((size_t)C * 32, -1);
    for (int g = 0; g < G; g++)
        std::copy(unions[g].begin(), unions[g].end(), t.chunk_cols.begin() + (size_t)t.grp_chunk_ptr[g] * 32);

    // pass 2: segment sizes, prefix, then fill in (row, column) order
    std::vector<int32_t> seg_count((size_t)C * ppg, 0);
    for (int g = 0; g < G; g++) {
        const int r0 = g * rpw, r1 = std::min(n_rows, r0 + rpw);
        for (int r = r0; r < r1; r++)
            for (int e = row_offset[r]; e < row_offset[r + 1]; e++)
                seg_count[(size_t)(t.grp_chunk_ptr[g] + ent_chunk[e]) * ppg + ((r - r0) >> 4)]++;
    }
    t.seg_ptr.assign((size_t)C * ppg + 1, 0);
    for (size_t s = 0; s < seg_count.size(); s++) t.seg_ptr[s + 1] = t.seg_ptr[s] + seg_count[s];
    std::vector<int32_t> cursor(t.seg_ptr.begin(), t.seg_ptr.end() - 1);
    t.entry_slot.assign(nnz, 0);
    t.entry_src.assign(nnz, 0);
    for (int g = 0; g < G; g++) {
        const int r0 = g * rpw, r1 = std::min(n_rows, r0 + rpw);
        for (int r = r0; r < r1; r++)
            for (int e = row_offset[r]; e < row_offset[r + 1]; e++) {
                const int32_t d = cursor[(size_t)(t.grp_chunk_ptr[g] + ent_chunk[e]) * ppg + ((r - r0) >> 4)]++;
                t.entry_slot[d] = ent_slot[e];
                t.entry_src[d] = e;
            }
    }
    // heavy groups first (stable for equal chunk counts)
    t.work_order.resize(G);
    std::iota(t.work_order.begin(), t.work_order.end(), 0);
    std::stable_sort(t.work_order.begin(), t.work_order.end(), [&](int a, int b) {
        return (t.grp_chunk_ptr[a + 1] - t.grp_chunk_ptr[a]) > (t.grp_chunk_ptr[b + 1] - t.grp_chunk_ptr[b]);
    });
    return TCA_OK;
}

}  // namespace tca
