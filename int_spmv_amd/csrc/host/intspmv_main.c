/*
 * intspmv_main.c -- the C host driver: the reference's main sequence
 * (intSpmv_main.c:1617-1761) with every kernel running on the GPU through
 * the C ABI of libintspmv (include/intspmv.h).
 *
 *   read matrix -> row statistics -> reference vector yr -> CSR kernels ->
 *   AXC -> K1 / K1p -> AXT NOC (th 1, 4, 8, 16, 32, 48) -> AXT COM (bs 128,
 *   256, 512) -> "formats' data" and "kernels' results" tables.
 *
 * Differences from the reference's main, each deliberate:
 *   - each kernel is timed as NUM_ITE back-to-back launches between two
 *     stream synchronisations (per-call host timing would measure launch
 *     latency, not the kernel); y is filled with NaN once before a kernel's
 *     launches instead of zeroed before each (every entry point overwrites y,
 *     and a row a kernel leaves unwritten then shows as an infinite error);
 *   - besides getErrors' global figures the table prints the per-row scaled
 *     error max_i |y_i - yr_i| / sum_j |a_ij x_j|;
 *   - the COM rows are correct results (the reference's COM kernel drops
 *     row-0 segments and subtracts prefix sums);
 *   - matrices may also be Matrix Market files or generated (--gen).
 *
 * usage: intspmv_main <matrix.csr|.bin|.mtx> [--ite N] [--x ramp|uniform] [--device D]
 *        intspmv_main --gen laplace:G | powerlaw:N:NNZ | stencil27:G [options]
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "intspmv.h"
#include "intspmv_host.h"

#define HDL printf("---------------------------------------------------------------------------------------------------------\n")

typedef struct {
    char name[40];
    double et, ot, flops, aErr, rErr, sErr;
    unsigned pos;
    int ok;
} result_t;

static intspmv_ctx *ctx;
static int n_ite = 250;

static void die(const char *what, int st)
{
    fprintf(stderr, "intspmv_main: %s failed: %s (%s)\n", what, intspmv_strerror(st),
            ctx ? intspmv_last_hip_error(ctx) : "");
    exit(EXIT_FAILURE);
}

#define CK(call)                                                                          \
    do {                                                                                  \
        int st_ = (call);                                                                 \
        if (st_) die(#call, st_);                                                         \
    } while (0)

/* cf_CSR (intSpmv_main.c:382-396), the reference vector yr, on the host as in
 * the reference's main (1645-1648): per row one fused multiply-add per entry
 * in CSR order from +0.0 (the reference's compilers contract
 * aux = aux + val * x). */
static void cf_csr_host(const intspmv_csr *m, const double *x, double *y)
{
    for (uint32_t i = 0; i < m->nrows; i++) {
        double aux = 0.0;
        for (uint32_t k = m->row[i]; k < m->row[i + 1]; k++) aux = fma(m->val[k], x[m->col[k]], aux);
        y[i] = aux;
    }
}

static double *nan_host; /* NaNs, as many as the longest output vector */

/* Fills count doubles of a device vector with NaN before a kernel runs. */
static void poison(void *d, size_t count)
{
    CK(intspmv_memcpy_htod(ctx, d, nan_host, count * sizeof(double)));
}

static void *up(const void *h, size_t bytes)
{
    void *d = NULL;
    CK(intspmv_malloc(ctx, &d, bytes));
    CK(intspmv_memcpy_htod(ctx, d, h, bytes));
    return d;
}

/* getErrors (intSpmv_main.c:404-425) plus the per-row scaled error. */
static void errors(const intspmv_csr *m, const double *scale, const double *yr, const double *y,
                   result_t *r)
{
    double maxDif = 0.0, maxVal = 0.0, se = 0.0;
    unsigned pos = 0;
    for (uint32_t i = 0; i < m->nrows; i++) {
        double v = fabs(yr[i]);
        if (v > maxVal) maxVal = v;
        double d = fabs(fabs(y[i]) - v);
        if (d > maxDif) {
            maxDif = d;
            pos = i;
        }
        double e = fabs(y[i] - yr[i]);
        e = scale[i] > 0 ? e / scale[i] : (e == 0 ? 0 : INFINITY);
        if (!(e <= se)) se = e;
    }
    r->aErr = maxDif;
    r->rErr = maxVal > 0 ? maxDif / maxVal : 0;
    r->pos = pos;
    r->sErr = se;
}

typedef int (*launch_fn)(void *arg);

/* Times n_ite launches; returns seconds per launch. */
static double time_launches(launch_fn f, void *arg)
{
    CK(f(arg)); /* warm-up */
    CK(intspmv_sync(ctx));
    double t0 = intspmv_wtime();
    for (int i = 0; i < n_ite; i++) CK(f(arg));
    CK(intspmv_sync(ctx));
    return (intspmv_wtime() - t0) / n_ite;
}

/* ---- launch closures ----------------------------------------------------- */
typedef struct {
    const intspmv_plan *plan;
    uint32_t n;
    const void *a0, *a1, *a2, *a3, *a4, *a5;
    void *y;
} call_t;

static int l_csr_ordered(void *p)
{
    call_t *c = p;
    return spmv_csr_ordered(ctx, c->n, c->a0, c->a1, c->a2, c->a3, c->y);
}
static int l_csr(void *p)
{
    call_t *c = p;
    return spmv_csr(ctx, c->plan, c->a0, c->a1, c->a2, c->a3, c->y);
}
static int l_axc(void *p)
{
    call_t *c = p;
    return spmv_axc(ctx, c->plan, c->a0, c->a1, c->y);
}
static int l_k1(void *p)
{
    call_t *c = p;
    return spmv_k1(ctx, c->plan, c->a0, c->a1, c->a2, c->a3, c->a4, c->a5, c->y);
}
static int l_k1_order(void *p)
{
    call_t *c = p;
    return spmv_k1_order(ctx, c->n, c->a0, c->a1, c->y);
}
static int l_axt(void *p)
{
    call_t *c = p;
    return spmv_axt(ctx, c->plan, c->a0, NULL, c->a1, c->y);
}

static void finish(const char *name, double et, double ot, const intspmv_csr *m,
                   const double *scale, const double *yr, void *dy, result_t *r)
{
    double *y = malloc((size_t)m->nrows * sizeof(double) + 8);
    CK(intspmv_memcpy_dtoh(ctx, y, dy, (size_t)m->nrows * sizeof(double)));
    snprintf(r->name, sizeof r->name, "%s", name);
    r->et = et;
    r->ot = ot;
    r->flops = 2.0 * m->nnz / et;
    errors(m, scale, yr, y, r);
    r->ok = 1;
    free(y);
}

static int parse_gen(const char *spec, intspmv_csr *m)
{
    unsigned long long a = 0, b = 0;
    if (sscanf(spec, "laplace:%llu", &a) == 1) return intspmv_gen_laplace2d((uint32_t)a, m);
    if (sscanf(spec, "powerlaw:%llu:%llu", &a, &b) == 2)
        return intspmv_gen_powerlaw((uint32_t)a, b, 2.0, 8192 < a ? 8192 : (uint32_t)a, 0.8,
                                    (uint32_t)(0.025 * a), 0, m);
    if (sscanf(spec, "stencil27:%llu", &a) == 1) return intspmv_gen_stencil27((uint32_t)a, 0.01, 0, m);
    return INTSPMV_HOST_EARG;
}

int main(int argc, char **argv)
{
    const char *path = NULL, *gen = NULL, *xmode = "ramp";
    int device = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--ite") && i + 1 < argc) n_ite = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--x") && i + 1 < argc) xmode = argv[++i];
        else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--gen") && i + 1 < argc) gen = argv[++i];
        else path = argv[i];
    }
    if ((!path && !gen) || n_ite < 1) {
        printf("\n\tUsage:\n\n\t\t%s <matFileName.csr|.bin|.mtx> [--ite N] [--x ramp|uniform]"
               " [--device D]\n\t\t%s --gen laplace:G|powerlaw:N:NNZ|stencil27:G [...]\n\n",
               argv[0], argv[0]);
        return EXIT_FAILURE;
    }
    HDL;
    printf("run settings\n");
    HDL;
    printf("matFileName:        %s\n", path ? path : gen);
    printf("FPT:                double\n");
    printf("NUM_ITE:            %d\n", n_ite);
    printf("HBRICK_SIZE:        %u\n", INTSPMV_HBRICK_SIZE);
    printf("CHUNK_SIZE:         %u\n", INTSPMV_CHUNK_SIZE);
    printf("device:             %d\n", device);
    fflush(stdout);

    intspmv_csr m;
    int st = gen ? parse_gen(gen, &m) : intspmv_csr_read(path, &m);
    if (st) {
        fprintf(stderr, "cannot read / generate matrix (%d)\n", st);
        return EXIT_FAILURE;
    }
    if (intspmv_csr_stats(&m, path)) return EXIT_FAILURE; /* <matrix>.sta beside a read matrix */
    CK(intspmv_create(&ctx, device));

    const uint32_t n = m.nrows;
    double *x = malloc((size_t)n * sizeof(double) + 8);
    double *scale = calloc((size_t)n + 1, sizeof(double));
    double *yr = calloc((size_t)n + 1, sizeof(double));
    if (!strcmp(xmode, "uniform")) intspmv_fill_uniform(x, n, -1.0, 1.0, 0);
    else
        for (uint32_t i = 0; i < n; i++) x[i] = (double)i; /* initVec, intSpmv_main.c:361-367 */
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t k = m.row[i]; k < m.row[i + 1]; k++) scale[i] += fabs(m.val[k] * x[m.col[k]]);

    result_t res[32];
    int nres = 0;
    intspmv_format_data fds[16];
    int nfd = 0;
    void *dy;
    CK(intspmv_malloc(ctx, &dy, ((size_t)n + 8) * sizeof(double)));
    nan_host = malloc(((size_t)n + 16) * sizeof(double));
    for (size_t i = 0; i < (size_t)n + 16; i++) nan_host[i] = NAN;
    cf_csr_host(&m, x, yr);

    /* ---- CSR ------------------------------------------------------------- */
    intspmv_csr_format_data(&m, &fds[nfd++]);
    void *dval = up(m.val, (size_t)m.nnz * 8), *dcol = up(m.col, (size_t)m.nnz * 4);
    void *drow = up(m.row, ((size_t)n + 1) * 4), *dx = up(x, (size_t)n * 8);
    call_t c = {0};
    c.n = n;
    c.a0 = dval;
    c.a1 = dcol;
    c.a2 = drow;
    c.a3 = dx;
    c.y = dy;
    poison(dy, n);
    double et = time_launches(l_csr_ordered, &c);
    finish("spmv_csr_ordered", et, 0, &m, scale, yr, dy, &res[nres++]);
    intspmv_plan *pcsr;
    CK(intspmv_csr_plan(ctx, n, m.row, &pcsr));
    c.plan = pcsr;
    poison(dy, n);
    et = time_launches(l_csr, &c);
    finish("spmv_csr", et, 0, &m, scale, yr, dy, &res[nres++]);
    intspmv_plan_destroy(pcsr);

    /* ---- AXC ------------------------------------------------------------- */
    {
        intspmv_axc a;
        if (intspmv_axc_build(&m, x, &a, &fds[nfd]) == 0) {
            nfd++;
            intspmv_plan *p;
            CK(intspmv_axc_plan(ctx, n, a.brp, &p));
            call_t ca = {0};
            ca.plan = p;
            ca.a0 = up(a.ax, (size_t)a.lenAX * 8);
            ca.a1 = up(a.brp, (size_t)a.lenBRP * 4);
            ca.y = dy;
            poison(dy, n);
            et = time_launches(l_axc, &ca);
            finish("spmv_axc", et, 0, &m, scale, yr, dy, &res[nres++]);
            intspmv_free(ctx, (void *)ca.a0);
            intspmv_free(ctx, (void *)ca.a1);
            intspmv_plan_destroy(p);
            intspmv_axc_free(&a);
        }
    }

    /* ---- K1 / K1p -------------------------------------------------------- */
    {
        intspmv_k1 k;
        if (intspmv_k1_build(&m, &k, &fds[nfd]) == 0) {
            nfd++;
            intspmv_plan *p;
            CK(intspmv_k1_plan(ctx, n, k.chunkNum, k.nmc, &p));
            void *dres;
            CK(intspmv_malloc(ctx, &dres, ((size_t)k.chunkNum * 8 + 8) * sizeof(double)));
            call_t ck = {0};
            ck.plan = p;
            ck.a0 = up(k.val, (size_t)k.lenVC * 8);
            ck.a1 = up(k.col, (size_t)k.lenVC * 4);
            ck.a2 = up(k.chp, (size_t)k.chunkNum * 4 + 4);
            ck.a3 = up(k.nmc, (size_t)k.chunkNum * 4 + 4);
            void *dperm = up(k.permi, ((size_t)k.chunkNum + 1) * 8 * 4);
            ck.a4 = NULL;
            ck.a5 = dx;
            ck.y = dres;
            poison(dres, (size_t)k.chunkNum * 8);
            et = time_launches(l_k1, &ck);
            call_t co = {0};
            co.n = n;
            co.a0 = dperm;
            co.a1 = dres;
            co.y = dy;
            poison(dy, n);
            double ot = time_launches(l_k1_order, &co);
            finish("spmv_k1", et, ot, &m, scale, yr, dy, &res[nres++]);
            ck.a4 = dperm;
            ck.y = dy;
            poison(dy, n);
            et = time_launches(l_k1, &ck);
            finish("spmv_k1p", et, 0, &m, scale, yr, dy, &res[nres++]);
            intspmv_free(ctx, dres);
            intspmv_free(ctx, dperm);
            for (const void **q = &ck.a0; q <= &ck.a3; q++) intspmv_free(ctx, (void *)*q);
            intspmv_plan_destroy(p);
            intspmv_k1_free(&k);
        }
    }

    /* ---- AXT ------------------------------------------------------------- */
    static const struct {
        uint32_t th, bs;
        const char *mode;
    } axts[] = {{1, 128, "NOC"},  {4, 128, "NOC"},  {8, 128, "NOC"},
                {16, 128, "NOC"}, {32, 128, "NOC"}, {48, 128, "NOC"},
                {1, 128, "COM"},  {1, 256, "COM"},  {1, 512, "COM"}};
    for (size_t i = 0; i < sizeof axts / sizeof axts[0]; i++) {
        intspmv_axt a;
        if (intspmv_axt_build(&m, x, axts[i].bs, 8, axts[i].th, axts[i].mode, 0, &a, &fds[nfd]))
            continue;
        nfd++;
        intspmv_axt_desc d = {0};
        d.nrows = a.nrows;
        d.nnz = a.nnz;
        d.tileHW = a.tileHW;
        d.tileH = a.tileH;
        d.tileN = a.tileN;
        d.lenAX = a.lenAX;
        d.lenSEC = a.lenSEC;
        d.colN = a.colN;
        d.bs = a.bs;
        d.log = a.log;
        d.mode = strcmp(a.mode, "NOC") ? INTSPMV_COM : INTSPMV_NOC;
        d.val_type = INTSPMV_F64;
        d.x_type = INTSPMV_F64;
        d.layout = INTSPMV_AXT_INTERLEAVED;
        intspmv_plan *p;
        CK(intspmv_axt_plan(ctx, &d, a.con, &p));
        call_t ca = {0};
        ca.plan = p;
        ca.a0 = up(a.ax, (size_t)a.lenAX * 8);
        ca.a1 = up(a.sec, (size_t)a.lenSEC * 4);
        ca.y = dy;
        poison(dy, n);
        et = time_launches(l_axt, &ca);
        char nm[40];
        if (d.mode == INTSPMV_NOC) snprintf(nm, sizeof nm, "spmv_axt_noc_h%u_hw8", a.tileH);
        else snprintf(nm, sizeof nm, "spmv_axt_com_h1_hw8_bs%u", a.bs);
        finish(nm, et, 0, &m, scale, yr, dy, &res[nres++]);
        intspmv_free(ctx, (void *)ca.a0);
        intspmv_free(ctx, (void *)ca.a1);
        intspmv_plan_destroy(p);
        intspmv_axt_free(&a);
    }

    HDL;
    printf("formats' data \n");
    HDL;
    printf("%22s %20s %10s %20s\n", "format", "memory [Mbytes]", "occupancy", "convTime [s]");
    for (int i = 0; i < nfd; i++)
        printf("%22s %20.2lf %10.2lf %20.6lf\n", fds[i].name, fds[i].mfp * 1e-6, fds[i].beta,
               fds[i].ct);
    HDL;
    printf("kernels' results\n");
    HDL;
    printf("%25s %15s %8s %15s %13s %13s %10s\n", "kernel", "exeTime [s]", "Gflops",
           "ordTime [s]", "aErr||.||inf", "rErr||.||inf", "rowInd");
    for (int i = 0; i < nres; i++)
        printf("%25s %15.7lf %8.3lf %15.7lf %11.3le %13.3le %12u\n", res[i].name, res[i].et,
               res[i].flops * 1e-9, res[i].ot, res[i].aErr, res[i].rErr, res[i].pos);
    /* beside the reference's table (kept column for column): the per-row
     * error scaled by sum |a_ij x_j|, the criterion the results are held to */
    HDL;
    printf("kernels' scaled errors\n");
    HDL;
    printf("%25s %13s\n", "kernel", "rowScaledErr");
    for (int i = 0; i < nres; i++) printf("%25s %13.3le\n", res[i].name, res[i].sErr);
    fflush(stdout);
    intspmv_free(ctx, dval);
    intspmv_free(ctx, dcol);
    intspmv_free(ctx, drow);
    intspmv_free(ctx, dx);
    intspmv_free(ctx, dy);
    intspmv_destroy(ctx);
    free(x);
    free(nan_host);
    free(scale);
    free(yr);
    intspmv_csr_free(&m);
    return EXIT_SUCCESS;
}
