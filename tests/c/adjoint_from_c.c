/*
 * adjoint_from_c.c -- x_bar = A^T y_bar from a C host through the two C
 * libraries only (intspmv_host.h builds, intspmv.h runs on the GPU):
 *
 *   A's AXT NOC layout with its column map  -> intspmv_axt_build(want_acol)
 *   the transposed tile map, once           -> intspmv_axt_transpose_map
 *   A's values in A^T's slot order          -> spmv_permute_vals(permT)
 *   x_bar = A^T y_bar                       -> spmv_axt_gather on A^T's map
 *   the same, reading A's values directly   -> spmv_axt_gather_permuted
 *
 * The result is checked against a sequential transpose product on the host
 * (per column j: |x_bar_j - ref_j| <= 1e-12 * sum_i |a_ij y_bar_i|), and the
 * two device paths must agree bit for bit.  Exit status 0 when both hold.
 *
 * usage: adjoint_from_c <n> <target_nnz> <tileHeight>
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "intspmv.h"
#include "intspmv_host.h"

#define CK(call)                                                                    \
    do {                                                                            \
        int st_ = (call);                                                           \
        if (st_) {                                                                  \
            fprintf(stderr, "%s failed: %d (%s)\n", #call, st_, intspmv_strerror(st_)); \
            return 2;                                                               \
        }                                                                           \
    } while (0)

static intspmv_ctx *ctx;

static void *up(const void *h, size_t bytes)
{
    void *d = NULL;
    if (intspmv_malloc(ctx, &d, bytes ? bytes : 8) || intspmv_memcpy_htod(ctx, d, h, bytes))
        return NULL;
    return d;
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s <n> <target_nnz> <tileHeight>\n", argv[0]);
        return 2;
    }
    const uint32_t n = (uint32_t)atol(argv[1]), th = (uint32_t)atol(argv[3]);
    const uint64_t target = (uint64_t)atoll(argv[2]);
    intspmv_csr m;
    CK(intspmv_gen_powerlaw(n, target, 2.0, n < 2048 ? n : 2048, 0.8, n / 40, 7, &m));
    double *zero = calloc(n, sizeof(double));
    intspmv_axt a;
    CK(intspmv_axt_build(&m, zero, 0, INTSPMV_HBRICK_SIZE, th, "NOC", 1, &a, NULL));
    intspmv_axt_tmap t;
    CK(intspmv_axt_transpose_map(&m, &a, &t));
    const intspmv_axt *at = &t.at;

    /* A's values in A's slot order: what a forward product (spmv_axt_gather
     * on A's own map) reads, and what an optimiser would update */
    const size_t ns = a.lenAX / 2;
    double *vals = malloc((ns ? ns : 1) * sizeof(double));
    CK(intspmv_axt_value_plane(&a, vals));

    CK(intspmv_create(&ctx, 0));
    intspmv_axt_desc d;
    memset(&d, 0, sizeof d);
    d.nrows = at->nrows;
    d.nnz = at->nnz;
    d.tileHW = at->tileHW;
    d.tileH = at->tileH;
    d.tileN = at->tileN;
    d.lenAX = at->lenAX;
    d.lenSEC = at->lenSEC;
    d.colN = at->colN;
    d.mode = INTSPMV_NOC;
    d.val_type = INTSPMV_F64;
    d.x_type = INTSPMV_F64;
    d.layout = INTSPMV_AXT_COLMAP;
    intspmv_plan *p;
    CK(intspmv_axt_plan(ctx, &d, at->con, &p));

    double *ybar = malloc(n * sizeof(double));
    intspmv_fill_normal(ybar, n, 0.0, 1.0, 11);
    void *d_vals = up(vals, ns * sizeof(double));
    void *d_perm = up(t.permT, (size_t)t.lenPERM * sizeof(uint32_t));
    void *d_acolT = up(at->acol, (size_t)t.lenPERM * sizeof(uint32_t));
    void *d_secT = up(at->sec, (size_t)at->lenSEC * sizeof(uint32_t));
    void *d_ybar = up(ybar, n * sizeof(double));
    void *d_valsT = NULL, *d_x1 = NULL, *d_x2 = NULL;
    CK(intspmv_malloc(ctx, &d_valsT, ((size_t)t.lenPERM + 1) * sizeof(double)));
    CK(intspmv_malloc(ctx, &d_x1, n * sizeof(double)));
    CK(intspmv_malloc(ctx, &d_x2, n * sizeof(double)));
    if (!d_vals || !d_perm || !d_acolT || !d_secT || !d_ybar) return 2;

    /* path 1: refresh A^T's copy of the values, then the plain product */
    CK(spmv_permute_vals(ctx, t.lenPERM, INTSPMV_F64, d_perm, d_vals, d_valsT));
    CK(spmv_axt_gather(ctx, p, d_valsT, d_acolT, d_secT, d_ybar, 1, d_x1, 1, 1, 0));
    /* path 2: A's values read through the slot map inside the product */
    CK(spmv_axt_gather_permuted(ctx, p, d_vals, d_perm, d_acolT, d_secT, d_ybar, 1, d_x2, 1, 1,
                                0));
    double *x1 = malloc(n * sizeof(double)), *x2 = malloc(n * sizeof(double));
    CK(intspmv_memcpy_dtoh(ctx, x1, d_x1, n * sizeof(double)));
    CK(intspmv_memcpy_dtoh(ctx, x2, d_x2, n * sizeof(double)));

    /* sequential transpose product and its per-column scale */
    double *ref = calloc(n, sizeof(double)), *scale = calloc(n, sizeof(double));
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t k = m.row[i]; k < m.row[i + 1]; k++) {
            ref[m.col[k]] = fma(m.val[k], ybar[i], ref[m.col[k]]);
            scale[m.col[k]] += fabs(m.val[k] * ybar[i]);
        }
    double worst = 0.0;
    size_t differ = 0;
    for (uint32_t j = 0; j < n; j++) {
        double e = fabs(x1[j] - ref[j]);
        e = scale[j] > 0 ? e / scale[j] : (e == 0 ? 0 : INFINITY);
        if (!(e <= worst)) worst = e;
        differ += memcmp(&x1[j], &x2[j], sizeof(double)) != 0;
    }
    printf("n %u nnz %u th %u slots_AT %u max_scaled_err %.3e paths_differ %zu\n", n, m.nnz, th,
           t.lenPERM, worst, differ);

    intspmv_plan_destroy(p);
    void *dev[] = {d_vals, d_perm, d_acolT, d_secT, d_ybar, d_valsT, d_x1, d_x2};
    for (size_t i = 0; i < sizeof dev / sizeof dev[0]; i++) intspmv_free(ctx, dev[i]);
    intspmv_destroy(ctx);
    intspmv_axt_tmap_free(&t);
    intspmv_axt_free(&a);
    intspmv_csr_free(&m);
    free(zero);
    free(vals);
    free(ybar);
    free(x1);
    free(x2);
    free(ref);
    free(scale);
    return (worst <= 1e-12 && differ == 0) ? 0 : 1;
}
