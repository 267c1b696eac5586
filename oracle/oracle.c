/*
 * oracle.c -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * CPU restatement of the reference kernels of intSpmv_main.c, written for
 * clarity, one function per reference function, each citing the lines it
 * follows.  Compiled with -ffp-contract=off so that the only fused
 * multiply-adds are the explicit fma() calls that pin the reference's
 * contraction (oracle.h lists the conventions).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define HB 8u /* HBRICK_SIZE, intSpmv_main.c:101-103 */

/* _mm512_reduce_add_pd as the halving tree of the 512-bit vector:
 * (lo256 + hi256) -> (lo128 + hi128) -> (e0 + e1). */
double orc_reduce8(const double *p)
{
    double a0 = p[0] + p[4], a1 = p[1] + p[5], a2 = p[2] + p[6], a3 = p[3] + p[7];
    double b0 = a0 + a2, b1 = a1 + a3;
    return b0 + b1;
}

/* cf_CSR, intSpmv_main.c:382-396 */
void orc_cf_csr(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                const double *x, double *y)
{
    for (uint32_t i = 0; i < n; i++) {
        double aux = 0.0;
        for (uint32_t j = row[i]; j < row[i + 1]; j++) aux = fma(val[j], x[col[j]], aux);
        y[i] = aux;
    }
}

/* getErrors, intSpmv_main.c:404-425 (compares magnitudes, global scale) */
void orc_get_errors(uint32_t len, const double *ar, const double *ac, double *aErr,
                    double *rErr, uint32_t *pos)
{
    double maxDif = 0.0, maxVal = 0.0;
    uint32_t p = 0;
    for (uint32_t i = 0; i < len; i++) {
        double v = fabs(ar[i]);
        if (v > maxVal) maxVal = v;
        double d = fabs(fabs(ac[i]) - v);
        if (d > maxDif) {
            maxDif = d;
            p = i;
        }
    }
    *aErr = maxDif;
    *rErr = maxDif / maxVal;
    *pos = p;
}

/* Per-row scaled error max_i |y_i - yref_i| / sum_j |a_ij x_j| (the
 * correctness criterion of this engine; not in the reference).  A row whose
 * scale is zero counts as infinite error unless both values are equal. */
double orc_scaled_err(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                      const double *x, const double *yref, const double *y, uint32_t *argmax)
{
    double worst = 0.0;
    uint32_t arg = 0;
    for (uint32_t i = 0; i < n; i++) {
        double s = 0.0;
        for (uint32_t j = row[i]; j < row[i + 1]; j++) s += fabs(val[j] * x[col[j]]);
        double d = fabs(y[i] - yref[i]);
        double e = (s > 0.0) ? d / s : (d == 0.0 ? 0.0 : INFINITY);
        if (e != e) e = INFINITY; /* NaN counts as failure */
        if (e > worst) {
            worst = e;
            arg = i;
        }
    }
    if (argmax) *argmax = arg;
    return worst;
}

/* cf_CSR for k right-hand sides at once: Y[i*ldy + j] is column j's cf_CSR
 * chain (one fused multiply-add per entry in CSR order from +0.0), so every
 * column has the bits of orc_cf_csr on that column.  Rows are independent,
 * so they are spread over the host's threads (checker only; not timed). */
void orc_csr_mm(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                const double *X, uint32_t ldx, uint32_t k, double *Y, uint32_t ldy)
{
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const uint32_t i = (uint32_t)ii;
        double *y = Y + (size_t)i * ldy;
        for (uint32_t j = 0; j < k; j++) y[j] = 0.0;
        for (uint32_t e = row[i]; e < row[i + 1]; e++) {
            const double a = val[e], *x = X + (size_t)col[e] * ldx;
            for (uint32_t j = 0; j < k; j++) y[j] = fma(a, x[j], y[j]);
        }
    }
}

/* orc_scaled_err over k right-hand sides: the largest
 * |Y_ij - Yref_ij| / sum_e |a_ie X_col(e),j| over rows i and columns j, with
 * the same conventions as orc_scaled_err (zero scale: infinite unless equal;
 * NaN: infinite).  *argmax = i * k + j of the worst entry. */
double orc_scaled_err_mm(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                         const double *X, uint32_t ldx, uint32_t k, const double *Yref,
                         const double *Y, uint32_t ldy, uint64_t *argmax)
{
    double worst = 0.0;
    uint64_t arg = 0;
#pragma omp parallel
    {
        double w = 0.0;
        uint64_t a = 0;
        double *s = (double *)calloc(k ? k : 1, sizeof(double));
#pragma omp for schedule(dynamic, 1024) nowait
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const uint32_t i = (uint32_t)ii;
            for (uint32_t j = 0; j < k; j++) s[j] = 0.0;
            for (uint32_t e = row[i]; e < row[i + 1]; e++) {
                const double *x = X + (size_t)col[e] * ldx;
                for (uint32_t j = 0; j < k; j++) s[j] += fabs(val[e] * x[j]);
            }
            for (uint32_t j = 0; j < k; j++) {
                double d = fabs(Y[(size_t)i * ldy + j] - Yref[(size_t)i * ldy + j]);
                double e = (s[j] > 0.0) ? d / s[j] : (d == 0.0 ? 0.0 : INFINITY);
                if (e != e) e = INFINITY;
                if (e > w) {
                    w = e;
                    a = (uint64_t)i * k + j;
                }
            }
        }
        free(s);
#pragma omp critical
        if (w > worst || (w == worst && a < arg && w > 0.0)) {
            worst = w;
            arg = a;
        }
    }
    if (argmax) *argmax = arg;
    return worst;
}

/* x_bar = A^T y_bar by a sequential scatter in CSR order (the transpose
 * product the adjoint is checked against; not in the reference). */
void orc_csr_transpose_product(uint32_t n, const uint32_t *row, const uint32_t *col,
                               const double *val, const double *ybar, double *xbar)
{
    for (uint32_t j = 0; j < n; j++) xbar[j] = 0.0;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t k = row[i]; k < row[i + 1]; k++) xbar[col[k]] = fma(val[k], ybar[i], xbar[col[k]]);
}

/* int_axc, intSpmv_main.c:711-732: per brick reduce8(val*x), sum += red. */
void orc_int_axc(uint32_t nrows, const double *ax, const uint32_t *brp, double *y)
{
    for (uint32_t r = 0; r < nrows; r++) {
        double sum = 0.0;
        for (uint32_t p = brp[r]; p < brp[r + 1]; p += 2 * HB) {
            double pr[HB];
            for (uint32_t i = 0; i < HB; i++) pr[i] = ax[p + i] * ax[p + HB + i];
            sum = sum + orc_reduce8(pr);
        }
        y[r] = sum;
    }
}

/* int_k1, intSpmv_main.c:943-965: lane-wise fma chain per 8-row chunk. */
void orc_int_k1(uint32_t chunkNum, const double *val, const uint32_t *col, const uint32_t *chp,
                const uint32_t *nmc, const double *x, double *res)
{
    for (uint32_t c = 0; c < chunkNum; c++) {
        double s[HB] = {0};
        uint32_t off = chp[c];
        for (uint32_t k = 0; k < nmc[c]; k++, off += HB)
            for (uint32_t l = 0; l < HB; l++) s[l] = fma(val[off + l], x[col[off + l]], s[l]);
        for (uint32_t l = 0; l < HB; l++) res[c * HB + l] = s[l];
    }
}

/* orderArrayK1, intSpmv_main.c:969-976 */
void orc_order_k1(uint32_t len, const uint32_t *permi, const double *res, double *y)
{
    for (uint32_t i = 0; i < len; i++) y[permi[i]] = res[i];
}

/* int_k1p, intSpmv_main.c:1026-1066 with the tail mask of test_int_k1p
 * (intSpmv_main.c:1077-1086): the last chunk scatters only its
 * 8 - (chunkNum*8 - nrows) valid lanes. */
void orc_int_k1p(uint32_t chunkNum, uint32_t nrows, const double *val, const uint32_t *col,
                 const uint32_t *chp, const uint32_t *nmc, const uint32_t *permi,
                 const double *x, double *y)
{
    for (uint32_t c = 0; c < chunkNum; c++) {
        double s[HB] = {0};
        uint32_t off = chp[c];
        for (uint32_t k = 0; k < nmc[c]; k++, off += HB)
            for (uint32_t l = 0; l < HB; l++) s[l] = fma(val[off + l], x[col[off + l]], s[l]);
        uint32_t lanes = HB;
        if (c == chunkNum - 1) lanes = HB - (chunkNum * HB - nrows);
        for (uint32_t l = 0; l < lanes; l++) y[permi[c * HB + l]] = s[l];
    }
}

/* int_axt_noc_h1, intSpmv_main.c:1385-1404: y[sec[t]] += reduce8(val*x);
 * the caller zeroes y (intSpmv_main.c:1421). */
void orc_int_axt_noc_h1(uint32_t tileN, const double *ax, const uint32_t *sec, double *y)
{
    for (uint32_t t = 0; t < tileN; t++) {
        double pr[HB];
        const double *a = ax + (size_t)t * 2 * HB;
        for (uint32_t i = 0; i < HB; i++) pr[i] = a[i] * a[HB + i];
        y[sec[t]] = y[sec[t]] + orc_reduce8(pr);
    }
}

/* int_axt_noc, intSpmv_main.c:1443-1474: per tile, lane-wise fma chain over
 * the th brick-rows, then y[sec[8t+i]] += tmp[i]; the caller zeroes y. */
void orc_int_axt_noc(uint32_t tileN, uint32_t th, const double *ax, const uint32_t *sec,
                     double *y)
{
    for (uint32_t t = 0; t < tileN; t++) {
        double s[HB] = {0};
        const double *a = ax + (size_t)t * th * 2 * HB;
        for (uint32_t r = 0; r < th; r++)
            for (uint32_t i = 0; i < HB; i++) s[i] = fma(a[r * 2 * HB + i], a[r * 2 * HB + HB + i], s[i]);
        for (uint32_t i = 0; i < HB; i++) y[sec[t * HB + i]] = y[sec[t * HB + i]] + s[i];
    }
}

/* int_axt_com_h1, intSpmv_main.c:1517-1566, restated faithfully including
 * its defects: segments whose sec word is 0 (row 0 with offset 0) are
 * skipped, and segment sums are prefix differences.  Tiles past tileN, which
 * the reference reads beyond the end of ax, are taken as zero; they only
 * feed prefix entries that no segment reads. */
/* One COM block b: products of its tiles in blk1, in-tile prefix sums
 * continued across the block's tiles in blk2, then every segment start of the
 * block adds its segment sum to y (atomically when other blocks run at the
 * same time). */
static void com_block(uint32_t b, uint32_t bs, uint32_t log, uint32_t thw, uint32_t lenAX,
                      uint32_t lenSEC, const double *ax, const uint32_t *sec, double *blk1,
                      double *blk2, double *y, int atomic)
{
    const uint32_t tpb = bs / thw, lbs = tpb * 2 * thw;
    double acu = 0.0;
    for (uint32_t t = 0; t < tpb; t++) {
        size_t p = (size_t)b * lbs + (size_t)t * 2 * thw;
        double v[HB], r1[HB], r2[HB], r3[HB];
        for (uint32_t i = 0; i < HB; i++)
            v[i] = (p + HB + i < lenAX) ? ax[p + i] * ax[p + HB + i] : 0.0;
        for (uint32_t i = 0; i < HB; i++) r1[i] = i >= 1 ? v[i] + v[i - 1] : v[i];
        for (uint32_t i = 0; i < HB; i++) r2[i] = i >= 2 ? r1[i] + r1[i - 2] : r1[i];
        for (uint32_t i = 0; i < HB; i++) r3[i] = i >= 4 ? r2[i] + r2[i - 4] : r2[i];
        for (uint32_t i = 0; i < HB; i++) {
            blk1[t * HB + i] = v[i];
            blk2[t * HB + i] = r3[i] + acu;
        }
        acu = blk2[t * HB + HB - 1];
    }
    uint32_t sp = b * bs, ep = (b + 1) * bs;
    ep = ep > lenSEC ? lenSEC : ep;
    for (uint32_t q = sp; q < ep; q++) {
        uint32_t ro = sec[q];
        if (ro != 0) {
            uint32_t r = ro >> log, o = ro & (bs - 1), l = q - sp;
            double val = blk2[l + o] - blk2[l] + blk1[l];
            if (atomic) {
#pragma omp atomic
                y[r] += val;
            } else {
                y[r] = y[r] + val;
            }
        }
    }
}

/* int_axt_com_h1, intSpmv_main.c:1517-1566, restated faithfully including
 * its defects: segments whose sec word is 0 (row 0 with offset 0) are
 * skipped, and segment sums are prefix differences.  Tiles past tileN, which
 * the reference reads beyond the end of ax, are taken as zero; they only
 * feed prefix entries that no segment reads.  Blocks in order (serial). */
void orc_int_axt_com_h1(uint32_t bn, uint32_t bs, uint32_t log, uint32_t thw, uint32_t tn,
                        uint32_t lenAX, uint32_t lenSEC, const double *ax, const uint32_t *sec,
                        double *y)
{
    double *blk1 = (double *)calloc((size_t)bs, sizeof(double));
    double *blk2 = (double *)calloc((size_t)bs, sizeof(double));
    (void)tn;
    for (uint32_t b = 0; b < bn; b++) com_block(b, bs, log, thw, lenAX, lenSEC, ax, sec, blk1, blk2, y, 0);
    free(blk1);
    free(blk2);
}

/* ------------------------------------------------------------ baseline */

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void omp_axc(uint32_t nrows, const double *ax, const uint32_t *brp, double *y)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t rr = 0; rr < (int64_t)nrows; rr++) {
        uint32_t r = (uint32_t)rr;
        double sum = 0.0;
        for (uint32_t p = brp[r]; p < brp[r + 1]; p += 2 * HB) {
            double pr[HB];
            for (uint32_t i = 0; i < HB; i++) pr[i] = ax[p + i] * ax[p + HB + i];
            sum = sum + orc_reduce8(pr);
        }
        y[r] = sum;
    }
}

static void omp_k1(uint32_t chunkNum, const double *val, const uint32_t *col,
                   const uint32_t *chp, const uint32_t *nmc, const double *x, double *res)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t cc = 0; cc < (int64_t)chunkNum; cc++) {
        uint32_t c = (uint32_t)cc;
        double s[HB] = {0};
        uint32_t off = chp[c];
        for (uint32_t k = 0; k < nmc[c]; k++, off += HB)
            for (uint32_t l = 0; l < HB; l++) s[l] = fma(val[off + l], x[col[off + l]], s[l]);
        for (uint32_t l = 0; l < HB; l++) res[c * HB + l] = s[l];
    }
}

static void omp_k1p(uint32_t chunkNum, uint32_t nrows, const double *val, const uint32_t *col,
                    const uint32_t *chp, const uint32_t *nmc, const uint32_t *permi,
                    const double *x, double *y)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t cc = 0; cc < (int64_t)chunkNum; cc++) {
        uint32_t c = (uint32_t)cc;
        double s[HB] = {0};
        uint32_t off = chp[c];
        for (uint32_t k = 0; k < nmc[c]; k++, off += HB)
            for (uint32_t l = 0; l < HB; l++) s[l] = fma(val[off + l], x[col[off + l]], s[l]);
        uint32_t lanes = (c == chunkNum - 1) ? HB - (chunkNum * HB - nrows) : HB;
        for (uint32_t l = 0; l < lanes; l++) y[permi[c * HB + l]] = s[l];
    }
}

static void omp_axt_h1(uint32_t tileN, const double *ax, const uint32_t *sec, double *y)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t tt = 0; tt < (int64_t)tileN; tt++) {
        double pr[HB];
        const double *a = ax + (size_t)tt * 2 * HB;
        for (uint32_t i = 0; i < HB; i++) pr[i] = a[i] * a[HB + i];
        double red = orc_reduce8(pr);
        uint32_t r = sec[tt];
#pragma omp atomic
        y[r] += red;
    }
}

static void omp_axt(uint32_t tileN, uint32_t th, const double *ax, const uint32_t *sec,
                    double *y)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t tt = 0; tt < (int64_t)tileN; tt++) {
        double s[HB] = {0};
        const double *a = ax + (size_t)tt * th * 2 * HB;
        for (uint32_t r = 0; r < th; r++)
            for (uint32_t i = 0; i < HB; i++) s[i] = fma(a[r * 2 * HB + i], a[r * 2 * HB + HB + i], s[i]);
        for (uint32_t i = 0; i < HB; i++) {
            uint32_t r = sec[tt * HB + i];
#pragma omp atomic
            y[r] += s[i];
        }
    }
}

/* int_axt_com_h1 over the host's threads: blocks dealt dynamically, each
 * thread with its own block scratch, y updated atomically (as the
 * reference's parallel loop, intSpmv_main.c:1532-1564).
 * prm = {bs, log, thw, lenAX, lenSEC}. */
static void omp_axt_com(uint32_t bn, const uint32_t *prm, const double *ax, const uint32_t *sec,
                        double *y)
{
#pragma omp parallel
    {
        double *blk1 = (double *)calloc((size_t)prm[0], sizeof(double));
        double *blk2 = (double *)calloc((size_t)prm[0], sizeof(double));
#pragma omp for schedule(dynamic)
        for (int64_t b = 0; b < (int64_t)bn; b++)
            com_block((uint32_t)b, prm[0], prm[1], prm[2], prm[3], prm[4], ax, sec, blk1, blk2, y, 1);
        free(blk1);
        free(blk2);
    }
}

double orc_bench(int kind, int threads, int iters, uint32_t n, uint32_t len, uint32_t th,
                 const void *a0, const void *a1, const void *a2, const void *a3,
                 const void *a4, const void *a5, double *y)
{
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    double total = 0.0;
    double *res = NULL;
    if (kind == 2) res = (double *)calloc((size_t)len * HB + HB, sizeof(double));
    for (int it = 0; it < iters; it++) {
        for (uint32_t i = 0; i < n; i++) y[i] = 0.0; /* untimed, as fill_array */
        double t0 = now();
        switch (kind) {
        case 0:
            orc_cf_csr(n, (const uint32_t *)a0, (const uint32_t *)a1, (const double *)a2,
                       (const double *)a3, y);
            break;
        case 1:
            omp_axc(n, (const double *)a0, (const uint32_t *)a1, y);
            break;
        case 2: /* len = chunkNum */
            omp_k1(len, (const double *)a0, (const uint32_t *)a1, (const uint32_t *)a2,
                   (const uint32_t *)a3, (const double *)a4, res);
            break;
        case 3:
            omp_k1p(len, n, (const double *)a0, (const uint32_t *)a1, (const uint32_t *)a2,
                    (const uint32_t *)a3, (const uint32_t *)a5, (const double *)a4, y);
            break;
        case 4: /* len = tileN */
            omp_axt_h1(len, (const double *)a0, (const uint32_t *)a1, y);
            break;
        case 5:
            omp_axt(len, th, (const double *)a0, (const uint32_t *)a1, y);
            break;
        case 6: /* len = bn, a2 = {bs, log, thw, lenAX, lenSEC} */
            omp_axt_com(len, (const uint32_t *)a2, (const double *)a0, (const uint32_t *)a1, y);
            break;
        default:
            return -1.0;
        }
        total += now() - t0;
    }
    if (kind == 2) {
        orc_order_k1(n, (const uint32_t *)a5, res, y);
        free(res);
    }
    return iters > 0 ? total / iters : 0.0;
}
