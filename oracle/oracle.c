/*
 * oracle.c -- CPU restatement of the TCA-SpMM SpMM path (test infrastructure only).
 * Built by oracle/Makefile with gcc -O2 -fopenmp -ffp-contract=off.
 * See oracle.h for what each function restates and how it is pinned.
 */
#include "oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <ctype.h>
#include <errno.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* scalar formats                                                           */
/* ------------------------------------------------------------------------ */
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

float orc_f16_to_f32(uint16_t h)
{
    uint32_t s = (uint32_t)(h >> 15) << 31;
    uint32_t e = (h >> 10) & 0x1f, m = h & 0x3ff;
    if (e == 0) {
        /* zero or subnormal: m * 2^-24, exact in float */
        float v = (float)m * 5.9604644775390625e-08f;
        return s ? -v : v;
    }
    if (e == 31) return u2f(s | 0x7f800000u | (m << 13));
    return u2f(s | ((e - 15 + 127) << 23) | (m << 13));
}

uint16_t orc_f32_to_f16(float f)
{
    uint32_t u = f2u(f);
    uint16_t s = (uint16_t)((u >> 16) & 0x8000);
    uint32_t a = u & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(s | 0x7e00); /* NaN */
    if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00); /* rounds to >= 65520 -> inf */
    if (a < 0x38800000u) {
        /* result is subnormal (or zero) in f16: value / 2^-24, round half even */
        float v = u2f(a) * 16777216.0f; /* exact scaling by 2^24 */
        float r = nearbyintf(v);          /* default rounding mode: nearest even */
        return (uint16_t)(s | (uint16_t)r);
    }
    uint32_t e = (a >> 23) - 127 + 15;
    uint32_t m = a & 0x7fffff;
    uint32_t keep = m >> 13, rest = m & 0x1fff;
    uint32_t out = (e << 10) | keep;
    if (rest > 0x1000 || (rest == 0x1000 && (keep & 1))) out += 1; /* may carry into exponent: correct */
    return (uint16_t)(s | out);
}

float orc_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

uint16_t orc_f32_to_bf16(float f)
{
    uint32_t u = f2u(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
    uint32_t lsb = (u >> 16) & 1;
    u += 0x7fffu + lsb;
    return (uint16_t)(u >> 16);
}

float orc_e4m3_to_f32(uint8_t v)
{
    int s = v >> 7, e = (v >> 3) & 15, m = v & 7;
    float r;
    if (e == 15 && m == 7) return NAN;
    if (e == 0) r = ldexpf((float)m, -9);          /* m/8 * 2^-6 */
    else r = ldexpf(1.0f + (float)m / 8.0f, e - 7);
    return s ? -r : r;
}

uint8_t orc_f32_to_e4m3(float f)
{
    if (isnan(f)) return 0x7f;
    uint8_t sign = signbit(f) ? 0x80 : 0x00;
    float a = fabsf(f);
    if (a >= 448.0f) return sign | 0x7e;            /* saturate */
    /* codes 0x00..0x7e are increasing in value: binary search the floor code */
    int lo = 0, hi = 0x7e;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (orc_e4m3_to_f32((uint8_t)mid) <= a) lo = mid; else hi = mid - 1;
    }
    int c = lo;
    if (c < 0x7e) {
        float v0 = orc_e4m3_to_f32((uint8_t)c), v1 = orc_e4m3_to_f32((uint8_t)(c + 1));
        float d0 = a - v0, d1 = v1 - a;
        if (d1 < d0 || (d1 == d0 && (c & 1))) c += 1;
    }
    return (uint8_t)(sign | c); /* the sign is kept, so -tiny -> 0x80 (-0) */
}

/* ------------------------------------------------------------------------ */
/* .smtx reader -- src/file_input.cu:135-193                                */
/* Line 1: "rows, cols, nnz" (comma separated, stoi tolerates blanks);      */
/* line 2: rows+1 row offsets (space separated); lines 3+: nnz column ids.  */
/* The reference trusts every count; this restatement rejects bad files.    */
/* ------------------------------------------------------------------------ */
static char *slurp(const char *path, size_t *len)
{
    FILE *f = fopen(path, "rb");
    if (!f) return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    if (n < 0) { fclose(f); return NULL; }
    char *buf = (char *)malloc((size_t)n + 1);
    if (!buf) { fclose(f); return NULL; }
    size_t got = fread(buf, 1, (size_t)n, f);
    fclose(f);
    buf[got] = 0;
    *len = got;
    return buf;
}

/* parse one integer token the way std::stoi does (leading blanks, sign, digits) */
static int parse_int(const char **p, const char *end, long *out)
{
    const char *s = *p;
    while (s < end && (*s == ' ' || *s == '\t' || *s == '\r')) s++;
    if (s >= end) return -1;
    char *e;
    errno = 0;
    long v = strtol(s, &e, 10);
    if (e == s || errno) return -1;
    *out = v;
    *p = e;
    return 0;
}

int orc_smtx_to_dense(const char *path, int *M, int *K, int *nnz, uint16_t **dense)
{
    size_t len = 0;
    char *buf = slurp(path, &len);
    if (!buf) return -1;
    const char *p = buf, *end = buf + len;
    const char *eol = memchr(p, '\n', (size_t)(end - p));
    if (!eol) { free(buf); return -2; }
    long hdr[3];
    for (int i = 0; i < 3; i++) {
        if (parse_int(&p, eol, &hdr[i])) { free(buf); return -2; }
        while (p < eol && (*p == ' ' || *p == '\t' || *p == '\r')) p++;
        if (i < 2) {
            if (p >= eol || *p != ',') { free(buf); return -2; }
            p++;
        }
    }
    if (hdr[0] <= 0 || hdr[1] <= 0 || hdr[2] < 0 || hdr[0] > 1L << 24 || hdr[1] > 1L << 24) { free(buf); return -2; }
    int rows = (int)hdr[0], cols = (int)hdr[1], nz = (int)hdr[2];
    p = eol + 1;
    eol = memchr(p, '\n', (size_t)(end - p));
    const char *line2_end = eol ? eol : end;
    int *rowptr = (int *)malloc(sizeof(int) * ((size_t)rows + 1));
    int cnt = 0;
    for (;;) {
        long v;
        const char *q = p;
        while (q < line2_end && (*q == ' ' || *q == '\t' || *q == '\r')) q++;
        if (q >= line2_end) break;
        if (parse_int(&p, line2_end, &v)) { free(rowptr); free(buf); return -3; }
        if (cnt >= rows + 1) { free(rowptr); free(buf); return -3; }
        rowptr[cnt++] = (int)v;
    }
    if (cnt != rows + 1 || rowptr[0] != 0 || rowptr[rows] != nz) { free(rowptr); free(buf); return -3; }
    for (int i = 0; i < rows; i++)
        if (rowptr[i + 1] < rowptr[i]) { free(rowptr); free(buf); return -3; }
    p = eol ? eol + 1 : end;
    int *colidx = (int *)malloc(sizeof(int) * ((size_t)nz + 1));
    int c = 0;
    for (;;) {
        while (p < end && isspace((unsigned char)*p)) p++;
        if (p >= end) break;
        long v;
        if (parse_int(&p, end, &v)) { free(colidx); free(rowptr); free(buf); return -4; }
        if (c >= nz) { free(colidx); free(rowptr); free(buf); return -4; }
        colidx[c++] = (int)v;
    }
    if (c != nz) { free(colidx); free(rowptr); free(buf); return -4; }
    for (int i = 0; i < nz; i++)
        if (colidx[i] < 0 || colidx[i] >= cols) { free(colidx); free(rowptr); free(buf); return -5; }
    uint16_t *d = (uint16_t *)calloc((size_t)rows * cols, sizeof(uint16_t));
    for (int i = 0; i < rows; i++)
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++)
            d[(size_t)i * cols + colidx[j]] = 0x3C00; /* __float2half(1.0f) */
    *M = rows; *K = cols; *nnz = nz; *dense = d;
    free(colidx); free(rowptr); free(buf);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* MatrixMarket -> dense -- mtx_input_to_Dense, src/file_input.cu:8-61       */
/* ------------------------------------------------------------------------ */
int orc_mtx_to_dense(const char *path, int one_based, int *M, int *K, int *nnz, uint16_t **dense)
{
    size_t len = 0;
    char *buf = slurp(path, &len);
    if (!buf) return -1;
    long rows = -1, cols = -1, meta = -1, count = 0;
    uint16_t *d = NULL;
    char *save = NULL;
    for (char *line = strtok_r(buf, "\n", &save); line; line = strtok_r(NULL, "\n", &save)) {
        char *s = line;
        while (*s == ' ' || *s == '\t' || *s == '\r') s++;
        if (*s == 0 || *s == '%') continue;
        if (rows < 0) {
            if (sscanf(s, "%ld %ld %ld", &rows, &cols, &meta) != 3 || rows <= 0 || cols <= 0 || meta < 0) {
                free(buf);
                return -2;
            }
            d = (uint16_t *)calloc((size_t)rows * cols, sizeof(uint16_t));
            continue;
        }
        long r, c;
        float v = 1.0f;
        int got = sscanf(s, "%ld %ld %f", &r, &c, &v);
        if (got < 2) { free(d); free(buf); return -3; }
        if (one_based) { r--; c--; }
        if (r < 0 || r >= rows || c < 0 || c >= cols) { free(d); free(buf); return -4; }
        d[(size_t)r * cols + c] = orc_f32_to_f16(v);
        count++;
    }
    free(buf);
    if (rows < 0) return -2;
    if (count != meta) { free(d); return -5; }
    *M = (int)rows; *K = (int)cols; *nnz = (int)meta; *dense = d;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* dense <-> CSR -- src/miscalleneous.cu:29-65 and :66-82                    */
/* ------------------------------------------------------------------------ */
int orc_dense_to_csr(const uint16_t *dense, int M, int K, int *rowptr, int *colidx, uint16_t *vals)
{
    int nz = 0;
    if (rowptr) rowptr[0] = 0;
    for (int r = 0; r < M; r++) {
        for (int c = 0; c < K; c++) {
            uint16_t h = dense[(size_t)r * K + c];
            if (orc_f16_to_f32(h) != 0.0f) {
                if (colidx) colidx[nz] = c;
                if (vals) vals[nz] = h;
                nz++;
            }
        }
        if (rowptr) rowptr[r + 1] = nz;
    }
    return nz;
}

void orc_csr_to_dense(int M, int K, const int *rowptr, const int *colidx, const uint16_t *vals, uint16_t *dense)
{
    memset(dense, 0, sizeof(uint16_t) * (size_t)M * K);
    for (int r = 0; r < M; r++)
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++)
            dense[(size_t)r * K + colidx[e]] = vals[e];
}

/* ------------------------------------------------------------------------ */
/* compare_mat<float> -- src/general.cuh:71-123                              */
/* fail iff |d|/|gold| >= tol AND |d| >= tol; return after the 10th failure  */
/* ------------------------------------------------------------------------ */
int orc_compare_mat(const float *gold, const float *cmp, int n, float tol, int *n_fail)
{
    int ok = 1, count = 0;
    for (int i = 0; i < n; i++) {
        float v1 = gold[i], v2 = cmp[i];
        float disc = fabsf(v1 - v2);
        float rel = fabsf(disc / v1);
        if (rel >= tol && disc >= tol) {
            ok = 0;
            count++;
            if (count >= 10) break;
        }
    }
    if (n_fail) *n_fail = count;
    return ok;
}

/* ------------------------------------------------------------------------ */
/* CPU SpMM -- naive_gemm (src/miscalleneous.cu:114-134) on nonzeros only     */
/* ------------------------------------------------------------------------ */
void orc_spmm_f32(int M, int N, const int *rowptr, const int *colidx, const uint16_t *vals,
                  const uint16_t *B, int ldb, float *C, int ldc, int nthreads)
{
    /* B rows are decoded once (exact), then each output row accumulates
     * a * B[col] in ascending column order in FP32. */
    int K = 0;
    for (int e = 0; e < rowptr[M]; e++) if (colidx[e] + 1 > K) K = colidx[e] + 1;
    float *Bf = (float *)malloc(sizeof(float) * ((size_t)K * N + 1));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int k = 0; k < K; k++)
        for (int n = 0; n < N; n++) Bf[(size_t)k * N + n] = orc_f16_to_f32(B[(size_t)k * ldb + n]);
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
#endif
    for (int r = 0; r < M; r++) {
        float *c = C + (size_t)r * ldc;
        for (int n = 0; n < N; n++) c[n] = 0.0f;
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
            const float a = orc_f16_to_f32(vals[e]);
            const float *b = Bf + (size_t)colidx[e] * N;
            for (int n = 0; n < N; n++) c[n] += a * b[n];
        }
    }
    free(Bf);
}

void orc_spmm_f64(int M, int N, const int *rowptr, const int *colidx, const double *vals,
                  const double *B, int ldb, double *C, int ldc, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
#endif
    for (int r = 0; r < M; r++) {
        double *c = C + (size_t)r * ldc;
        for (int n = 0; n < N; n++) c[n] = 0.0;
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
            double a = vals[e];
            const double *b = B + (size_t)colidx[e] * ldb;
            for (int n = 0; n < N; n++) c[n] += a * b[n];
        }
    }
}

void orc_sddmm_f64(int M, int N, const int *rowptr, const int *colidx,
                   const double *G, int ldg, const double *X, int ldx, double *out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 16)
#endif
    for (int r = 0; r < M; r++) {
        const double *g = G + (size_t)r * ldg;
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
            const double *x = X + (size_t)colidx[e] * ldx;
            double s = 0.0;
            for (int n = 0; n < N; n++) s += g[n] * x[n];
            out[e] = s;
        }
    }
}

/* FP32 SDDMM on fp16 bits: out[e] = sum_n f32(G[r,n]) * f32(X[col(e),n]),
 * FP32 accumulation in ascending n (the CPU baseline of the dW product; the
 * reference has no backward, so this follows the same naive_gemm semantics,
 * src/miscalleneous.cu:114-134, restricted to the pattern). */
void orc_sddmm_f32(int M, int N, const int *rowptr, const int *colidx, const uint16_t *G, int ldg,
                   const uint16_t *X, int ldx, float *out, int nthreads)
{
    int K = 0;
    for (int e = 0; e < rowptr[M]; e++) if (colidx[e] + 1 > K) K = colidx[e] + 1;
    float *Xf = (float *)malloc(sizeof(float) * ((size_t)K * N + 1));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads) schedule(static)
#endif
    for (int k = 0; k < K; k++)
        for (int n = 0; n < N; n++) Xf[(size_t)k * N + n] = orc_f16_to_f32(X[(size_t)k * ldx + n]);
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
        float *g = (float *)malloc(sizeof(float) * ((size_t)N + 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int r = 0; r < M; r++) {
            for (int n = 0; n < N; n++) g[n] = orc_f16_to_f32(G[(size_t)r * ldg + n]);
            for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
                const float *x = Xf + (size_t)colidx[e] * N;
                float s = 0.0f;
                for (int n = 0; n < N; n++) s += g[n] * x[n];
                out[e] = s;
            }
        }
        free(g);
    }
    free(Xf);
}

/* ------------------------------------------------------------------------ */
/* workload_manager -- src/our_spmm.cu:347-471 (row -> warp packing)          */
/* ------------------------------------------------------------------------ */
int orc_workload_manager(int M, const int *rowptr, int N, int warps_per_tb, int *num_blocks,
                         int *warps_per_row, int *row_idx, int *warpid, int cap)
{
    if (warps_per_tb < 8 || warps_per_tb > 32) { *num_blocks = 0; return -1; }
    for (int i = 0; i < M; i++) {
        int nz = rowptr[i + 1] - rowptr[i], w;
        if (nz == 0) w = 0;
        else if (nz <= 8) w = 1;
        else if (nz <= 16) w = 2;
        else if (nz <= 32) w = 4;
        else if (nz <= 64) w = 8;
        else {
            int blocks = (int)ceilf((float)nz / 64.0f);
            int lim = N / 64;
            w = (blocks < lim ? blocks : lim) * warps_per_tb;
        }
        warps_per_row[i] = w;
    }
    int *tb_of_row = (int *)malloc(sizeof(int) * (size_t)(M > 0 ? M : 1));
    for (int i = 0; i < M; i++) tb_of_row[i] = -1;
    int *slot_w = (int *)malloc(sizeof(int) * (size_t)warps_per_tb);
    int *slot_r = (int *)malloc(sizeof(int) * (size_t)warps_per_tb);
    int tb = 0, written = 0, overflow = 0;
    for (int i = 0; i < M; i++) {
        int nw = warps_per_row[i];
        if (nw == 0 || tb_of_row[i] != -1) continue;
        int ntb;
        if (nw <= warps_per_tb) {
            ntb = 1;
            tb_of_row[i] = tb;
            for (int s = 0; s < warps_per_tb; s++) { slot_w[s] = -1; slot_r[s] = -1; }
            for (int s = 0; s < nw; s++) { slot_w[s] = s; slot_r[s] = i; }
            for (int j = i + 1; j < M; j++) {
                int nnw = warps_per_row[j];
                if (nnw == 0) continue;
                if (nw + nnw <= warps_per_tb) {
                    /* no check of tb_of_row[j]: the reference can pack a row twice */
                    tb_of_row[j] = tb;
                    for (int s = nw; s < nw + nnw; s++) { slot_w[s] = s - nw; slot_r[s] = j; }
                    nw += nnw;
                }
                if (nw == warps_per_tb) break;
            }
            if (written + warps_per_tb > cap) overflow = 1;
            else {
                memcpy(warpid + written, slot_w, sizeof(int) * (size_t)warps_per_tb);
                memcpy(row_idx + written, slot_r, sizeof(int) * (size_t)warps_per_tb);
                written += warps_per_tb;
            }
        } else {
            tb_of_row[i] = tb;
            ntb = nw / warps_per_tb;
            if (written + nw > cap) overflow = 1;
            else {
                for (int s = 0; s < nw; s++) { warpid[written + s] = s; row_idx[written + s] = i; }
                written += nw;
            }
        }
        tb += ntb;
    }
    free(slot_w); free(slot_r); free(tb_of_row);
    *num_blocks = tb;
    return overflow ? -1 : written;
}

/* ------------------------------------------------------------------------ */
/* Panel tiling, rebuilt with dense per-group column bitmaps (the product    */
/* builds it with a sort-merge; the two must agree byte for byte).           */
/* ------------------------------------------------------------------------ */
int orc_tiling_sizes(int M, int K, const int *rowptr, const int *colidx, int rpw,
                     int *n_groups, int *n_chunks)
{
    if (rpw != 16 && rpw != 32 && rpw != 64 && rpw != 128 && rpw != 256) return -1;
    int G = (M + rpw - 1) / rpw;
    unsigned char *present = (unsigned char *)malloc((size_t)K);
    long total = 0;
    for (int g = 0; g < G; g++) {
        memset(present, 0, (size_t)K);
        int r1 = (g + 1) * rpw < M ? (g + 1) * rpw : M;
        for (int r = g * rpw; r < r1; r++)
            for (int e = rowptr[r]; e < rowptr[r + 1]; e++) present[colidx[e]] = 1;
        int u = 0;
        for (int j = 0; j < K; j++) u += present[j];
        total += (u + 31) / 32;
    }
    free(present);
    *n_groups = G;
    *n_chunks = (int)total;
    return 0;
}

int orc_tiling_fill(int M, int K, const int *rowptr, const int *colidx, int rpw,
                    int *grp_chunk_ptr, int *chunk_cols, int *seg_ptr,
                    uint16_t *entry_slot, int *entry_src, int *work_order)
{
    if (rpw != 16 && rpw != 32 && rpw != 64 && rpw != 128 && rpw != 256) return -1;
    int G = (M + rpw - 1) / rpw, ppg = rpw / 16;
    unsigned char *present = (unsigned char *)malloc((size_t)K);
    int *pos = (int *)malloc(sizeof(int) * (size_t)K);
    int chunk = 0, ent = 0;
    grp_chunk_ptr[0] = 0;
    seg_ptr[0] = 0;
    for (int g = 0; g < G; g++) {
        memset(present, 0, (size_t)K);
        int r0 = g * rpw, r1 = (g + 1) * rpw < M ? (g + 1) * rpw : M;
        for (int r = r0; r < r1; r++)
            for (int e = rowptr[r]; e < rowptr[r + 1]; e++) present[colidx[e]] = 1;
        int u = 0;
        for (int j = 0; j < K; j++) if (present[j]) pos[j] = u++;
        int nch = (u + 31) / 32;
        for (int s = 0; s < nch * 32; s++) chunk_cols[(size_t)chunk * 32 + s] = -1;
        for (int j = 0; j < K; j++) if (present[j]) chunk_cols[(size_t)chunk * 32 + pos[j]] = j;
        for (int c = 0; c < nch; c++) {
            for (int i = 0; i < ppg; i++) {
                for (int r16 = 0; r16 < 16; r16++) {
                    int r = r0 + 16 * i + r16;
                    if (r >= r1) continue;
                    for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
                        int p = pos[colidx[e]];
                        if (p / 32 != c) continue;
                        entry_slot[ent] = (uint16_t)(r16 * 32 + (p % 32));
                        entry_src[ent] = e;
                        ent++;
                    }
                }
                seg_ptr[(size_t)(chunk + c) * ppg + i + 1] = ent;
            }
        }
        chunk += nch;
        grp_chunk_ptr[g + 1] = chunk;
    }
    /* work order: groups by chunk count, descending; ties by group id */
    for (int g = 0; g < G; g++) work_order[g] = g;
    for (int a = 1; a < G; a++) { /* insertion sort keeps ties stable */
        int w = work_order[a], cw = grp_chunk_ptr[w + 1] - grp_chunk_ptr[w];
        int b = a - 1;
        while (b >= 0) {
            int o = work_order[b], co = grp_chunk_ptr[o + 1] - grp_chunk_ptr[o];
            if (co >= cw) break;
            work_order[b + 1] = o;
            b--;
        }
        work_order[b + 1] = w;
    }
    free(present); free(pos);
    return ent;
}

void orc_transpose_csr(int M, int K, const int *rowptr, const int *colidx,
                       int *rowptrT, int *colidxT, int *perm)
{
    for (int j = 0; j <= K; j++) rowptrT[j] = 0;
    for (int r = 0; r < M; r++)
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) rowptrT[colidx[e] + 1]++;
    for (int j = 0; j < K; j++) rowptrT[j + 1] += rowptrT[j];
    int *fill = (int *)malloc(sizeof(int) * (size_t)(K > 0 ? K : 1));
    for (int j = 0; j < K; j++) fill[j] = rowptrT[j];
    for (int r = 0; r < M; r++)
        for (int e = rowptr[r]; e < rowptr[r + 1]; e++) {
            int j = colidx[e], d = fill[j]++;
            colidxT[d] = r;
            perm[d] = e;
        }
    free(fill);
}

/* src/spmm_cusparse.cu:254-331 (perform_cuSPARSE_BlockSpMM's host conversion) */
int orc_dense_to_bell(const uint16_t *dense, int M, int K, int b, int32_t *col_ind, uint16_t *values)
{
    const int BR = (M + b - 1) / b, BC = (K + b - 1) / b;
    unsigned char *kept = (unsigned char *)calloc((size_t)(BR > 0 ? BR : 1) * (size_t)(BC > 0 ? BC : 1), 1);
    for (int r = 0; r < M; r++)
        for (int c = 0; c < K; c++)
            if (orc_f16_to_f32(dense[(size_t)r * K + c]) != 0.0f) kept[(size_t)(r / b) * BC + c / b] = 1;
    int slots = 0;
    for (int br = 0; br < BR; br++) {
        int n = 0;
        for (int bc = 0; bc < BC; bc++) n += kept[(size_t)br * BC + bc];
        if (n > slots) slots = n;
    }
    const int ell_cols = slots * b;
    if (col_ind && values) {
        for (int br = 0; br < BR; br++) {
            int j = 0;
            for (int bc = 0; bc < BC; bc++)
                if (kept[(size_t)br * BC + bc]) col_ind[(size_t)br * slots + j++] = bc;
            for (; j < slots; j++) col_ind[(size_t)br * slots + j] = -1;
        }
        for (int r = 0; r < BR * b; r++)
            for (int e = 0; e < ell_cols; e++) {
                const int bc = col_ind[(size_t)(r / b) * slots + e / b];
                const int c = bc * b + e % b;
                values[(size_t)r * ell_cols + e] = (bc >= 0 && r < M && c < K) ? dense[(size_t)r * K + c] : 0;
            }
    }
    free(kept);
    return ell_cols;
}

/* cusparseSpMM on the Blocked-ELL descriptor, src/spmm_cusparse.cu:394-398 */
void orc_bell_spmm_f32(int M, int K, int N, int b, int ell_cols, const int32_t *col_ind,
                       const uint16_t *values, const uint16_t *B, float *C)
{
    const int slots = b > 0 ? ell_cols / b : 0;
    for (int r = 0; r < M; r++)
        for (int n = 0; n < N; n++) {
            float acc = 0.0f;
            for (int j = 0; j < slots; j++) {
                const int bc = col_ind[(size_t)(r / b) * slots + j];
                if (bc < 0) continue;
                for (int c = 0; c < b; c++) {
                    const int k = bc * b + c;
                    if (k >= K) break;
                    const float a = orc_f16_to_f32(values[(size_t)r * ell_cols + j * b + c]);
                    if (a != 0.0f) acc += a * orc_f16_to_f32(B[(size_t)k * N + n]);
                }
            }
            C[(size_t)r * N + n] = acc;
        }
}

/* cusparseLtSpMMAPrune, STRIP rule (src/spmm_cusparse.cu:95-96) */
int orc_prune_24_strip(const uint16_t *A, int M, int K, uint16_t *out)
{
    static const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {1, 2, 3, 2, 3, 3};
    if (K % 4) return -1;
    for (int r = 0; r < M; r++)
        for (int g = 0; g < K; g += 4) {
            const uint16_t *a = A + (size_t)r * K + g;
            uint16_t *o = out + (size_t)r * K + g;
            float mag[4];
            for (int i = 0; i < 4; i++) mag[i] = fabsf(orc_f16_to_f32(a[i]));
            int best = 0;
            float bestv = mag[pa[0]] + mag[pb[0]];
            for (int c = 1; c < 6; c++)
                if (mag[pa[c]] + mag[pb[c]] > bestv) {
                    bestv = mag[pa[c]] + mag[pb[c]];
                    best = c;
                }
            for (int i = 0; i < 4; i++) o[i] = (i == pa[best] || i == pb[best]) ? a[i] : 0;
        }
    return 0;
}
