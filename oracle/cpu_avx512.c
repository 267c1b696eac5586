/*
 * cpu_avx512.c -- TEST INFRASTRUCTURE ONLY: the CPU baseline.
 *
 * An AVX-512 restatement of the reference's OpenMP kernels, for timing on the
 * GPU host's cores beside the GPU (the reference itself cannot be built here:
 * it includes <mkl.h>).  Same parallel structure as the reference: OpenMP
 * `schedule(dynamic)` over rows / chunks / tiles / blocks, `omp atomic` into
 * y for the AXT kernels, 512-bit loads of the format arrays (the unaligned
 * forms, as fast as the aligned ones on aligned data), gathers of x for K1.
 * Each function names the reference function it follows.  Numerical
 * conventions are the oracle's (oracle.h): the lane-wise
 * chains are fused multiply-adds, reductions are _mm512_reduce_add_pd (the
 * halving tree), cf_CSR is one fused multiply-add per entry.
 *
 * Differences from the reference, all without effect on what is timed:
 *   - gcc has no _mm512_i32logather_pd / _mm512_i32loscatter_pd (icc-only
 *     names); their 256-bit-index forms _mm512_i32gather_pd /
 *     _mm512_i32scatter_pd take the same eight 32-bit indexes;
 *   - int_axt_noc's inner `omp parallel for num_threads(8)` over the tile's
 *     eight lanes (intSpmv_main.c:1465) is a nested region; with nested
 *     parallelism off (the OpenMP default) it runs on the calling thread, so
 *     it is written as the loop it executes;
 *   - int_axt_com_h1's per-thread block scratch is allocated once per call
 *     instead of leaking (D4), and tiles past tileN read as zero.
 *
 * Built only with -mavx512f (oracle/build.py); bench.py calls it only when
 * the host CPU reports avx512f, otherwise the baseline reads "n/a".
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <omp.h>

#define HB 8u

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* cf_CSR, intSpmv_main.c:382-396 (serial, as in the reference) */
static void v_cf_csr(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                     const double *x, double *y)
{
    for (uint32_t i = 0; i < n; i++) {
        double aux = 0.0;
        for (uint32_t j = row[i]; j < row[i + 1]; j++) aux = fma(val[j], x[col[j]], aux);
        y[i] = aux;
    }
}

/* int_axc, intSpmv_main.c:711-732 */
static void v_axc(uint32_t nrows, const double *ax, const uint32_t *brp, double *y)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t rr = 0; rr < (int64_t)nrows; rr++) {
        double sum = 0.0;
        for (uint32_t p = brp[rr]; p < brp[rr + 1]; p += 2 * HB) {
            __m512d v = _mm512_loadu_pd(ax + p), xv = _mm512_loadu_pd(ax + p + HB);
            sum = sum + _mm512_reduce_add_pd(_mm512_mul_pd(v, xv));
        }
        y[rr] = sum;
    }
}

/* int_k1, intSpmv_main.c:943-965 */
static void v_k1(uint32_t chunkNum, const double *val, const uint32_t *col, const uint32_t *chp,
                 const uint32_t *nmc, const double *x, double *res)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t cc = 0; cc < (int64_t)chunkNum; cc++) {
        __m512d s = _mm512_setzero_pd();
        uint32_t off = chp[cc];
        for (uint32_t k = 0; k < nmc[cc]; k++, off += HB) {
            __m256i ci = _mm256_loadu_si256((const __m256i *)(col + off));
            s = _mm512_fmadd_pd(_mm512_loadu_pd(val + off), _mm512_i32gather_pd(ci, x, 8), s);
        }
        _mm512_storeu_pd(res + cc * HB, s);
    }
}

/* orderArrayK1, intSpmv_main.c:969-976 (timed separately by the reference) */
static void v_order_k1(uint32_t len, const uint32_t *permi, const double *res, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)len; i++) y[permi[i]] = res[i];
}

/* int_k1p, intSpmv_main.c:1026-1066, with the tail mask of test_int_k1p
 * (intSpmv_main.c:1077-1086) on the last chunk's scatter. */
static void v_k1p(uint32_t chunkNum, uint32_t nrows, const double *val, const uint32_t *col,
                  const uint32_t *chp, const uint32_t *nmc, const uint32_t *permi,
                  const double *x, double *y)
{
    const __mmask8 last = (__mmask8)(0xFFu >> (chunkNum * HB - nrows));
#pragma omp parallel for schedule(dynamic)
    for (int64_t cc = 0; cc < (int64_t)chunkNum; cc++) {
        __m512d s = _mm512_setzero_pd();
        uint32_t off = chp[cc];
        for (uint32_t k = 0; k < nmc[cc]; k++, off += HB) {
            __m256i ci = _mm256_loadu_si256((const __m256i *)(col + off));
            s = _mm512_fmadd_pd(_mm512_loadu_pd(val + off), _mm512_i32gather_pd(ci, x, 8), s);
        }
        __m256i pi = _mm256_loadu_si256((const __m256i *)(permi + cc * HB));
        _mm512_mask_i32scatter_pd(y, cc == (int64_t)chunkNum - 1 ? last : (__mmask8)0xFF, pi, s, 8);
    }
}

/* int_axt_noc_h1, intSpmv_main.c:1385-1404 */
static void v_axt_h1(uint32_t tileN, const double *ax, const uint32_t *sec, double *y)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t t = 0; t < (int64_t)tileN; t++) {
        const double *a = ax + (size_t)t * 2 * HB;
        double red = _mm512_reduce_add_pd(_mm512_mul_pd(_mm512_loadu_pd(a), _mm512_loadu_pd(a + HB)));
        uint32_t r = sec[t];
#pragma omp atomic
        y[r] += red;
    }
}

/* int_axt_noc, intSpmv_main.c:1443-1474 */
static void v_axt(uint32_t tileN, uint32_t th, const double *ax, const uint32_t *sec, double *y)
{
#pragma omp parallel for schedule(dynamic)
    for (int64_t t = 0; t < (int64_t)tileN; t++) {
        const double *a = ax + (size_t)t * th * 2 * HB;
        __m512d s = _mm512_setzero_pd();
        for (uint32_t r = 0; r < th; r++)
            s = _mm512_fmadd_pd(_mm512_loadu_pd(a + r * 2 * HB), _mm512_loadu_pd(a + r * 2 * HB + HB), s);
        double tmp[HB] __attribute__((aligned(64)));
        _mm512_storeu_pd(tmp, s);
        for (uint32_t i = 0; i < HB; i++) {
            uint32_t r = sec[(size_t)t * HB + i];
#pragma omp atomic
            y[r] += tmp[i];
        }
    }
}

/* lanes shifted up by K (lane i takes lane i-K, zeros below K) */
#define SHIFT_UP(v, K)                                                                  \
    _mm512_castsi512_pd(_mm512_alignr_epi64(_mm512_castpd_si512(v), _mm512_setzero_si512(), \
                                            8 - (K)))

/* int_axt_com_h1, intSpmv_main.c:1517-1566, defects kept (segments whose head
 * word is 0 are skipped; segment sums are prefix differences).
 * prm = {bs, log, thw, lenAX, lenSEC}. */
static void v_axt_com(uint32_t bn, const uint32_t *prm, const double *ax, const uint32_t *sec,
                      double *y)
{
    const uint32_t bs = prm[0], lg = prm[1], thw = prm[2], lenAX = prm[3], lenSEC = prm[4];
    const uint32_t tpb = bs / thw;
#pragma omp parallel
    {
        double *blk1 = (double *)malloc((size_t)bs * sizeof(double));
        double *blk2 = (double *)malloc((size_t)bs * sizeof(double));
#pragma omp for schedule(dynamic)
        for (int64_t b = 0; b < (int64_t)bn; b++) {
            __m512d acu = _mm512_setzero_pd();
            for (uint32_t t = 0; t < tpb; t++) {
                size_t p = (size_t)b * tpb * 2 * thw + (size_t)t * 2 * thw;
                __m512d v = p + 2 * HB <= lenAX
                                ? _mm512_mul_pd(_mm512_loadu_pd(ax + p), _mm512_loadu_pd(ax + p + HB))
                                : _mm512_setzero_pd();
                __m512d r = _mm512_add_pd(v, SHIFT_UP(v, 1));
                r = _mm512_add_pd(r, SHIFT_UP(r, 2));
                r = _mm512_add_pd(r, SHIFT_UP(r, 4));
                r = _mm512_add_pd(r, acu);
                _mm512_storeu_pd(blk1 + t * HB, v);
                _mm512_storeu_pd(blk2 + t * HB, r);
                acu = _mm512_set1_pd(blk2[t * HB + HB - 1]);
            }
            uint32_t sp = (uint32_t)b * bs, ep = sp + bs > lenSEC ? lenSEC : sp + bs;
            for (uint32_t q = sp; q < ep; q++) {
                uint32_t ro = sec[q];
                if (ro != 0) {
                    uint32_t rw = ro >> lg, o = ro & (bs - 1), l = q - sp;
                    double val = blk2[l + o] - blk2[l] + blk1[l];
#pragma omp atomic
                    y[rw] += val;
                }
            }
        }
        free(blk1);
        free(blk2);
    }
}

/* The orc_bench protocol (oracle.h) on these kernels: iters calls, each timed
 * alone, y zeroed outside the timed region; the mean seconds per call. */
double orc_bench_avx512(int kind, int threads, int iters, uint32_t n, uint32_t len, uint32_t th,
                        const void *a0, const void *a1, const void *a2, const void *a3,
                        const void *a4, const void *a5, double *y)
{
    if (threads > 0) omp_set_num_threads(threads);
    double total = 0.0;
    double *res = NULL;
    if (kind == 2) {
        res = (double *)malloc(((size_t)len * HB + HB) * sizeof(double));
        if (!res) return -1.0;
    }
    for (int it = 0; it < iters; it++) {
        memset(y, 0, (size_t)n * sizeof(double));
        double t0 = now();
        switch (kind) {
        case 0:
            v_cf_csr(n, a0, a1, a2, a3, y);
            break;
        case 1:
            v_axc(n, a0, a1, y);
            break;
        case 2:
            v_k1(len, a0, a1, a2, a3, a4, res);
            break;
        case 3:
            v_k1p(len, n, a0, a1, a2, a3, a5, a4, y);
            break;
        case 4:
            v_axt_h1(len, a0, a1, y);
            break;
        case 5:
            v_axt(len, th, a0, a1, y);
            break;
        case 6:
            v_axt_com(len, a2, a0, a1, y);
            break;
        default:
            free(res);
            return -1.0;
        }
        total += now() - t0;
    }
    if (kind == 2) {
        v_order_k1(n, a5, res, y);
        free(res);
    }
    return iters > 0 ? total / iters : 0.0;
}
