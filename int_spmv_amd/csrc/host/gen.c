/*
 * gen.c -- deterministic synthetic inputs.
 *
 * The reference reads matrices from files and sets x[i] = i (initVec,
 * intSpmv_main.c:361-367).  This engine also generates its benchmark inputs:
 * every random draw comes from xoshiro256** seeded through splitmix64, with
 * an independent stream per row (or per block of edges), so results do not
 * depend on the number of OpenMP threads.
 */
#include "intspmv_host.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    uint64_t s[4];
} rng_t;

static inline uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static void rng_init(rng_t *r, uint64_t seed, uint64_t stream)
{
    uint64_t sm = seed ^ (stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull);
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&sm);
}

static inline uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}

/* uniform in [0, 1) with 53 random bits */
static inline double rng_u01(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

/* uniform integer in [0, m) (m > 0), Lemire-style rejection-free via 128-bit mul */
static inline uint64_t rng_below(rng_t *r, uint64_t m)
{
    return (uint64_t)(((unsigned __int128)rng_next(r) * m) >> 64);
}

static inline double rng_normal(rng_t *r)
{
    double u1 = 1.0 - rng_u01(r); /* (0, 1] */
    double u2 = rng_u01(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

void intspmv_fill_uniform(double *v, size_t n, double lo, double hi, uint64_t seed)
{
    rng_t r;
    rng_init(&r, seed, 0);
    for (size_t i = 0; i < n; i++) v[i] = lo + (hi - lo) * rng_u01(&r);
}

void intspmv_fill_normal(double *v, size_t n, double mean, double sigma, uint64_t seed)
{
    rng_t r;
    rng_init(&r, seed, 1);
    for (size_t i = 0; i < n; i++) v[i] = mean + sigma * rng_normal(&r);
}

int intspmv_gen_laplace2d(uint32_t g, intspmv_csr *out)
{
    if (!out || g == 0 || (uint64_t)g * g > 0xFFFFFFFFull / 5) return INTSPMV_HOST_EARG;
    uint32_t n = g * g;
    uint32_t nnz = 5 * n - 4 * g;
    int st = intspmv_csr_alloc(out, n, nnz);
    if (st) return st;
    uint32_t p = 0;
    for (uint32_t i = 0; i < g; i++)
        for (uint32_t j = 0; j < g; j++) {
            uint32_t r = i * g + j;
            if (i > 0) { out->col[p] = r - g; out->val[p++] = -1.0; }
            if (j > 0) { out->col[p] = r - 1; out->val[p++] = -1.0; }
            out->col[p] = r; out->val[p++] = 4.0;
            if (j + 1 < g) { out->col[p] = r + 1; out->val[p++] = -1.0; }
            if (i + 1 < g) { out->col[p] = r + g; out->val[p++] = -1.0; }
            out->row[r + 1] = p;
        }
    return intspmv_csr_stats(out, NULL);
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

/* Fills the columns of one row: `band_frac` of the draws uniform in
 * [lo, hi] around the diagonal, the rest uniform in [0, n); duplicates are
 * redrawn using the caller's scratch bitmap (all-zero on entry and exit). */
static void draw_row(rng_t *r, uint32_t n, uint32_t len, uint32_t lo, uint32_t hi,
                     double band_frac, uint8_t *mark, uint32_t *cols)
{
    uint64_t bandw = (uint64_t)hi - lo + 1;
    /* a row longer than the band cannot be drawn mostly inside it */
    int band_ok = bandw * 2 >= len;
    for (uint32_t k = 0; k < len; k++) {
        uint32_t c;
        do {
            if (band_ok && rng_u01(r) < band_frac)
                c = lo + (uint32_t)rng_below(r, bandw);
            else
                c = (uint32_t)rng_below(r, n);
        } while (mark[c]);
        mark[c] = 1;
        cols[k] = c;
    }
    for (uint32_t k = 0; k < len; k++) mark[cols[k]] = 0;
    qsort(cols, len, sizeof(uint32_t), cmp_u32);
}

/* Builds CSR from per-row lengths; columns and values per row from the row's
 * own stream. */
static int fill_rows(uint32_t n, const uint32_t *len, double band_frac, uint32_t band_half,
                     uint64_t seed, double vmean, double vsig, intspmv_csr *out)
{
    uint64_t tot = 0;
    for (uint32_t i = 0; i < n; i++) tot += len[i];
    if (tot > 0xFFFFFFFFull) return INTSPMV_HOST_EOVERFLOW;
    int st = intspmv_csr_alloc(out, n, (uint32_t)tot);
    if (st) return st;
    for (uint32_t i = 0; i < n; i++) out->row[i + 1] = out->row[i] + len[i];
    int fail = 0;
#pragma omp parallel
    {
        uint8_t *mark = (uint8_t *)calloc(n, 1);
        if (!mark) {
#pragma omp atomic write
            fail = 1;
        }
#pragma omp for schedule(dynamic, 512)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            if (!mark) continue;
            uint32_t i = (uint32_t)ii;
            rng_t r;
            rng_init(&r, seed, 1000003ull + i);
            uint32_t lo = i > band_half ? i - band_half : 0;
            uint32_t hi = (uint64_t)i + band_half < n ? i + band_half : n - 1;
            uint32_t *cols = out->col + out->row[i];
            draw_row(&r, n, len[i], lo, hi, band_frac, mark, cols);
            for (uint32_t k = 0; k < len[i]; k++)
                out->val[out->row[i] + k] = vmean + vsig * rng_normal(&r);
        }
        free(mark);
    }
    if (fail) {
        intspmv_csr_free(out);
        return INTSPMV_HOST_ENOMEM;
    }
    return intspmv_csr_stats(out, NULL);
}

int intspmv_gen_powerlaw(uint32_t n, uint64_t target_nnz, double s, uint32_t maxlen,
                         double band_frac, uint32_t band_half, uint64_t seed,
                         intspmv_csr *out)
{
    if (!out || n == 0 || maxlen == 0 || maxlen > n || s <= 0.0) return INTSPMV_HOST_EARG;
    double *cdf = (double *)malloc((size_t)maxlen * sizeof(double));
    uint32_t *base = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    uint32_t *len = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    if (!cdf || !base || !len) {
        free(cdf);
        free(base);
        free(len);
        return INTSPMV_HOST_ENOMEM;
    }
    double acc = 0.0;
    for (uint32_t k = 1; k <= maxlen; k++) cdf[k - 1] = (acc += pow((double)k, -s));
    rng_t r;
    rng_init(&r, seed, 7);
    for (uint32_t i = 0; i < n; i++) {
        double u = rng_u01(&r) * acc;
        uint32_t lo = 0, hi = maxlen - 1; /* first k with cdf >= u */
        while (lo < hi) {
            uint32_t mid = (lo + hi) / 2;
            if (cdf[mid] >= u) hi = mid; else lo = mid + 1;
        }
        base[i] = lo + 1;
    }
    /* bisection on the scale c so that sum clip(round(c*k)) ~ target */
    double clo = 1e-3, chi = 1e4;
    for (int it = 0; it < 80; it++) {
        double c = 0.5 * (clo + chi);
        uint64_t tot = 0;
        for (uint32_t i = 0; i < n; i++) {
            double v = floor(c * base[i] + 0.5);
            tot += v < 1.0 ? 1 : (v > maxlen ? maxlen : (uint64_t)v);
        }
        if (tot < target_nnz) clo = c; else chi = c;
    }
    for (uint32_t i = 0; i < n; i++) {
        double v = floor(chi * base[i] + 0.5);
        len[i] = v < 1.0 ? 1u : (v > maxlen ? maxlen : (uint32_t)v);
    }
    free(cdf);
    free(base);
    int st = fill_rows(n, len, band_frac, band_half, seed, 0.0, 1.0, out);
    free(len);
    return st;
}

int intspmv_gen_banded(uint32_t n, uint32_t lmin, uint32_t lmax, double band_frac,
                       uint32_t band_w, uint64_t seed, intspmv_csr *out)
{
    if (!out || n == 0 || lmin > lmax || lmax > n) return INTSPMV_HOST_EARG;
    uint32_t *len = (uint32_t *)malloc((size_t)n * sizeof(uint32_t));
    if (!len) return INTSPMV_HOST_ENOMEM;
    rng_t r;
    rng_init(&r, seed, 11);
    for (uint32_t i = 0; i < n; i++) len[i] = lmin + (uint32_t)rng_below(&r, lmax - lmin + 1);
    int st = fill_rows(n, len, band_frac, band_w / 2, seed, 0.0, 1.0, out);
    free(len);
    return st;
}

int intspmv_gen_stencil27(uint32_t g, double sigma, uint64_t seed, intspmv_csr *out)
{
    if (!out || g == 0 || (uint64_t)g * g * g * 27 > 0xFFFFFFFFull) return INTSPMV_HOST_EARG;
    uint32_t n = g * g * g;
    uint64_t side = 3ull * g - 2;
    uint32_t nnz = (uint32_t)(side * side * side);
    int st = intspmv_csr_alloc(out, n, nnz);
    if (st) return st;
    /* row pointers: count neighbours per row */
    for (uint32_t i = 0; i < g; i++)
        for (uint32_t j = 0; j < g; j++)
            for (uint32_t k = 0; k < g; k++) {
                uint32_t ci = 1 + (i > 0) + (i + 1 < g), cj = 1 + (j > 0) + (j + 1 < g),
                         ck = 1 + (k > 0) + (k + 1 < g);
                out->row[(i * g + j) * g + k + 1] = ci * cj * ck;
            }
    for (uint32_t r = 0; r < n; r++) out->row[r + 1] += out->row[r];
#pragma omp parallel for schedule(static)
    for (int64_t rr = 0; rr < (int64_t)n; rr++) {
        uint32_t r = (uint32_t)rr, i = r / (g * g), j = (r / g) % g, k = r % g;
        rng_t rg;
        rng_init(&rg, seed, 5000011ull + r);
        uint32_t p = out->row[r];
        for (int di = -1; di <= 1; di++)
            for (int dj = -1; dj <= 1; dj++)
                for (int dk = -1; dk <= 1; dk++) {
                    int64_t a = (int64_t)i + di, b = (int64_t)j + dj, c = (int64_t)k + dk;
                    if (a < 0 || b < 0 || c < 0 || a >= g || b >= g || c >= g) continue;
                    uint32_t cc = (uint32_t)((a * g + b) * g + c);
                    out->col[p] = cc;
                    out->val[p] = (cc == r) ? 26.0 : -1.0 + sigma * rng_normal(&rg);
                    p++;
                }
    }
    return intspmv_csr_stats(out, NULL);
}

int intspmv_gen_rmat(uint32_t scale, uint32_t edgefactor, double a, double b, double c,
                     uint64_t seed, intspmv_csr *out)
{
    if (!out || scale == 0 || scale > 30 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return INTSPMV_HOST_EARG;
    uint64_t n = 1ull << scale, m = (uint64_t)edgefactor * n;
    if (m > 0xFFFFFFFFull) return INTSPMV_HOST_EOVERFLOW;
    uint32_t *ri = (uint32_t *)malloc(m * sizeof(uint32_t));
    uint32_t *ci = (uint32_t *)malloc(m * sizeof(uint32_t));
    double *v = (double *)malloc(m * sizeof(double));
    if (!ri || !ci || !v) {
        free(ri);
        free(ci);
        free(v);
        return INTSPMV_HOST_ENOMEM;
    }
    const uint64_t blk = 1u << 16;
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t bb = 0; bb < (int64_t)((m + blk - 1) / blk); bb++) {
        rng_t r;
        rng_init(&r, seed, 9000007ull + (uint64_t)bb);
        uint64_t e1 = ((uint64_t)bb + 1) * blk < m ? ((uint64_t)bb + 1) * blk : m;
        for (uint64_t e = (uint64_t)bb * blk; e < e1; e++) {
            uint32_t row = 0, col = 0;
            for (uint32_t l = 0; l < scale; l++) {
                double u = rng_u01(&r);
                uint32_t rb = u >= ab, cb = (u >= a && u < ab) || u >= abc;
                row = (row << 1) | rb;
                col = (col << 1) | cb;
            }
            ri[e] = row;
            ci[e] = col;
            v[e] = 1.0 - rng_u01(&r); /* (0, 1] */
        }
    }
    int st = intspmv_csr_from_coo((uint32_t)n, (size_t)m, ri, ci, v, out);
    free(ri);
    free(ci);
    free(v);
    if (st) return st;
    return intspmv_csr_stats(out, NULL);
}
