/*
 * oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU kernels (intSpmv_main.c),
 * used as the checker of the GPU library and as the CPU baseline that
 * bench.py times beside it.  Nothing in the product (int_spmv_amd/, include/)
 * links or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.
 *
 * Rounding-order conventions, fixed here because the reference leaves them to
 * the compiler:
 *   - cf_CSR's `aux = aux + val*x` is one fused multiply-add per term, in CSR
 *     order, starting from +0.0 (gcc and clang contract the expression);
 *   - _mm512_reduce_add_pd is the halving tree
 *     ((p0+p4)+(p2+p6)) + ((p1+p5)+(p3+p7)) of rounded products;
 *   - the lane-wise `vtSum = vtSum + val*x` of int_k1 / int_k1p /
 *     int_axt_noc is one fused multiply-add per step (icc -O3, the
 *     reference's build, contracts it);
 *   - the reference's `#pragma omp atomic` accumulations into y are taken in
 *     serial order (tile order, then lane order) from a zeroed y.
 */
#ifndef INTSPMV_ORACLE_H
#define INTSPMV_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void orc_cf_csr(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                const double *x, double *y);
void orc_get_errors(uint32_t len, const double *ar, const double *ac, double *aErr,
                    double *rErr, uint32_t *pos);
double orc_scaled_err(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                      const double *x, const double *yref, const double *y, uint32_t *argmax);
void orc_csr_mm(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                const double *X, uint32_t ldx, uint32_t k, double *Y, uint32_t ldy);
double orc_scaled_err_mm(uint32_t n, const uint32_t *row, const uint32_t *col, const double *val,
                         const double *X, uint32_t ldx, uint32_t k, const double *Yref,
                         const double *Y, uint32_t ldy, uint64_t *argmax);
void orc_csr_transpose_product(uint32_t n, const uint32_t *row, const uint32_t *col,
                               const double *val, const double *ybar, double *xbar);
double orc_reduce8(const double *p);

void orc_int_axc(uint32_t nrows, const double *ax, const uint32_t *brp, double *y);
void orc_int_k1(uint32_t chunkNum, const double *val, const uint32_t *col, const uint32_t *chp,
                const uint32_t *nmc, const double *x, double *res);
void orc_order_k1(uint32_t len, const uint32_t *permi, const double *res, double *y);
void orc_int_k1p(uint32_t chunkNum, uint32_t nrows, const double *val, const uint32_t *col,
                 const uint32_t *chp, const uint32_t *nmc, const uint32_t *permi,
                 const double *x, double *y);
void orc_int_axt_noc_h1(uint32_t tileN, const double *ax, const uint32_t *sec, double *y);
void orc_int_axt_noc(uint32_t tileN, uint32_t th, const double *ax, const uint32_t *sec,
                     double *y);
void orc_int_axt_com_h1(uint32_t bn, uint32_t bs, uint32_t log, uint32_t thw, uint32_t tn,
                        uint32_t lenAX, uint32_t lenSEC, const double *ax, const uint32_t *sec,
                        double *y);

/* CPU baseline: the reference's OpenMP parallel structure (dynamic schedule,
 * atomics into y), timed per call with y zeroed outside the timed region as
 * in test_* (intSpmv_main.c:1419-1427).  kind: 0 cf_CSR (serial),
 * 1 int_axc, 2 int_k1 (+orderArrayK1 untimed), 3 int_k1p, 4 int_axt_noc_h1,
 * 5 int_axt_noc, 6 int_axt_com_h1 (a2 = {bs, log, thw, lenAX, lenSEC}, len =
 * the block count).  Returns the mean seconds per call; y receives the last
 * result.  The pointer arguments a0..a5 follow the kernel's own order. */
double orc_bench(int kind, int threads, int iters, uint32_t n, uint32_t len, uint32_t th,
                 const void *a0, const void *a1, const void *a2, const void *a3,
                 const void *a4, const void *a5, double *y);

#ifdef __cplusplus
}
#endif
#endif
