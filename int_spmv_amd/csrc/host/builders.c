/*
 * builders.c -- CSR -> AXC / K1 / AXT conversion on the host.
 *
 * Each builder reproduces the arrays of the reference builder bit for bit:
 *   AXC  getArrayBrpAXC / getArrayAxAXC            intSpmv_main.c:619-658
 *   K1   getArrayPermiK1 / getArraysNmcChpK1 /
 *        getArraysValColK1                         intSpmv_main.c:782-837
 *   AXT  getArraysLenAXT                           intSpmv_main.c:1118-1148
 *        getArraysAxSecAXT_NOC_H1                  intSpmv_main.c:1152-1184
 *        getArraysAxSecAXT_NOC                     intSpmv_main.c:1188-1221
 *        getArraysAxSecAXT_COM_H1                  intSpmv_main.c:1225-1260
 *        getFormatDataAXT                          intSpmv_main.c:1275-1381
 * Differences, all on inputs the reference leaves undefined:
 *   - AXC ax and K1 val/col padding is zero (the reference _mm_malloc's them
 *     without initialising, intSpmv_main.c:684, 877-878);
 *   - K1's row order uses an explicit stable sort (rows by length descending,
 *     ties by ascending row), which is what glibc's merge-sort qsort gives for
 *     orderFunction (intSpmv_main.c:775-791);
 *   - lengths that would overflow the 32-bit index type are rejected.
 * The reference times NUM_ITE repetitions of each builder and reports the
 * mean; here ct is the wall time of one build.
 */
#include "intspmv_host.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static void *aligned_zalloc(size_t bytes)
{
    void *p = NULL;
    size_t b = (bytes + 63) & ~(size_t)63;
    if (b == 0) b = 64;
    if (posix_memalign(&p, 64, b) != 0) return NULL;
    memset(p, 0, b);
    return p;
}

/* ------------------------------------------------------------------ AXC */

int intspmv_axc_build(const intspmv_csr *m, const double *x, intspmv_axc *out,
                      intspmv_format_data *fd)
{
    if (!m || !x || !out || !m->rl) return INTSPMV_HOST_EARG;
    memset(out, 0, sizeof(*out));
    double t0 = intspmv_wtime();
    const uint32_t n = m->nrows;
    out->nrows = n;
    out->nnz = m->nnz;
    out->lenBRP = n + 1;
    out->brp = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
    if (!out->brp) return INTSPMV_HOST_ENOMEM;
    uint64_t acc = 0;
    for (uint32_t r = 0; r < n; r++) {
        uint64_t bricks = (m->rl[r] + (uint64_t)INTSPMV_HBRICK_SIZE - 1) / INTSPMV_HBRICK_SIZE;
        acc += 2 * bricks * INTSPMV_HBRICK_SIZE;
        if (acc > 0xFFFFFFFFull) {
            intspmv_axc_free(out);
            return INTSPMV_HOST_EOVERFLOW;
        }
        out->brp[r + 1] = (uint32_t)acc;
    }
    out->lenAX = out->brp[n];
    out->ax = (double *)aligned_zalloc((size_t)out->lenAX * sizeof(double));
    if (!out->ax) {
        intspmv_axc_free(out);
        return INTSPMV_HOST_ENOMEM;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ri = 0; ri < (int64_t)n; ri++) {
        uint32_t r = (uint32_t)ri;
        uint32_t pos = out->brp[r], cnt = 0;
        for (uint32_t k = m->row[r]; k < m->row[r + 1]; k++) {
            out->ax[pos] = m->val[k];
            out->ax[pos + INTSPMV_HBRICK_SIZE] = x[m->col[k]];
            if (cnt == INTSPMV_HBRICK_SIZE - 1) {
                pos += 1 + INTSPMV_HBRICK_SIZE;
                cnt = 0;
            } else {
                pos++;
                cnt++;
            }
        }
    }
    if (fd) {
        memset(fd, 0, sizeof(*fd));
        strcpy(fd->name, "AXC");
        fd->mfp = (double)out->lenAX * sizeof(double) + (double)out->lenBRP * sizeof(uint32_t);
        fd->beta = out->lenAX ? (double)out->nnz / (double)(out->lenAX >> 1) : 0.0;
        fd->ct = intspmv_wtime() - t0;
    }
    return INTSPMV_HOST_OK;
}

void intspmv_axc_free(intspmv_axc *a)
{
    if (!a) return;
    free(a->ax);
    free(a->brp);
    memset(a, 0, sizeof(*a));
}

/* ------------------------------------------------------------------- K1 */

int intspmv_k1_build(const intspmv_csr *m, intspmv_k1 *out, intspmv_format_data *fd)
{
    if (!m || !out || !m->rl) return INTSPMV_HOST_EARG;
    memset(out, 0, sizeof(*out));
    double t0 = intspmv_wtime();
    const uint32_t n = m->nrows, CS = INTSPMV_CHUNK_SIZE;
    out->nrows = n;
    out->nnz = m->nnz;
    out->chunkNum = (n + CS - 1) / CS;
    out->permi = (uint32_t *)aligned_zalloc(((size_t)out->chunkNum + 1) * CS * sizeof(uint32_t));
    out->nmc = (uint32_t *)calloc(out->chunkNum ? out->chunkNum : 1, sizeof(uint32_t));
    out->chp = (uint32_t *)calloc(out->chunkNum ? out->chunkNum : 1, sizeof(uint32_t));
    if (!out->permi || !out->nmc || !out->chp) {
        intspmv_k1_free(out);
        return INTSPMV_HOST_ENOMEM;
    }
    /* stable counting sort by row length, descending */
    uint32_t rmax = 0;
    for (uint32_t r = 0; r < n; r++) rmax = m->rl[r] > rmax ? m->rl[r] : rmax;
    size_t *start = (size_t *)calloc((size_t)rmax + 2, sizeof(size_t));
    if (!start) {
        intspmv_k1_free(out);
        return INTSPMV_HOST_ENOMEM;
    }
    for (uint32_t r = 0; r < n; r++) start[rmax - m->rl[r] + 1]++;
    for (uint32_t b = 0; b <= rmax; b++) start[b + 1] += start[b];
    for (uint32_t r = 0; r < n; r++) out->permi[start[rmax - m->rl[r]]++] = r;
    free(start);
    /* nmc / chp (getArraysNmcChpK1) */
    uint64_t l = 0;
    for (uint32_t c = 0; c < out->chunkNum; c++) {
        out->nmc[c] = m->rl[out->permi[c * CS]];
        l += (uint64_t)CS * out->nmc[c];
    }
    if (l > 0xFFFFFFFFull) {
        intspmv_k1_free(out);
        return INTSPMV_HOST_EOVERFLOW;
    }
    for (uint32_t c = 1; c < out->chunkNum; c++) out->chp[c] = out->chp[c - 1] + out->nmc[c - 1] * CS;
    out->lenVC = (uint32_t)l;
    out->val = (double *)aligned_zalloc((size_t)out->lenVC * sizeof(double));
    out->col = (uint32_t *)aligned_zalloc((size_t)out->lenVC * sizeof(uint32_t));
    if (!out->val || !out->col) {
        intspmv_k1_free(out);
        return INTSPMV_HOST_ENOMEM;
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ci = 0; ci < (int64_t)out->chunkNum; ci++) {
        uint32_t c = (uint32_t)ci;
        for (uint32_t rid = 0; rid < CS; rid++) {
            uint32_t s = c * CS + rid;
            if (s >= n) break;
            uint32_t r = out->permi[s];
            uint32_t off = 0;
            for (uint32_t k = m->row[r]; k < m->row[r + 1]; k++, off++) {
                uint32_t p = out->chp[c] + off * CS + rid;
                out->val[p] = m->val[k];
                out->col[p] = m->col[k];
            }
        }
    }
    if (fd) {
        memset(fd, 0, sizeof(*fd));
        strcpy(fd->name, "K1");
        fd->mfp = 2.0 * out->chunkNum * sizeof(uint32_t) + (double)out->lenVC * sizeof(double) +
                  (double)out->lenVC * sizeof(uint32_t) + (double)n * sizeof(uint32_t) +
                  (double)n * sizeof(double);
        fd->beta = out->lenVC ? (double)out->nnz / (double)out->lenVC : 0.0;
        fd->ct = intspmv_wtime() - t0;
    }
    return INTSPMV_HOST_OK;
}

void intspmv_k1_free(intspmv_k1 *k)
{
    if (!k) return;
    free(k->permi);
    free(k->nmc);
    free(k->chp);
    free(k->val);
    free(k->col);
    memset(k, 0, sizeof(*k));
}

/* ------------------------------------------------------------------ AXT */

static int mode_is_noc(const char *mode) { return strcmp(mode, "NOC") == 0; }
static int mode_is_com(const char *mode) { return strcmp(mode, "COM") == 0; }

int intspmv_axt_lengths(const intspmv_csr *m, uint32_t thw, uint32_t th, const char *mode,
                        uint32_t *con, uint32_t *tileN, uint32_t *lenAX, uint32_t *lenSEC,
                        uint32_t *colN)
{
    if (!m || !m->rl || !mode || thw == 0 || th == 0 || !con) return INTSPMV_HOST_EARG;
    if (!mode_is_noc(mode) && !mode_is_com(mode)) return INTSPMV_HOST_EARG;
    const int noc = mode_is_noc(mode);
    const uint64_t ths = (uint64_t)thw * th;
    const uint64_t grp = (th == 1) ? thw : th;
    uint64_t rsp = 0, total_cols = 0;
    for (uint32_t r = 0; r < m->nrows; r++) {
        uint64_t T = rsp / ths;                 /* ((rsp + ths)/ths) - 1 */
        uint64_t R = rsp % th;
        uint64_t C = rsp / th - T * thw;        /* ((rsp + th)/th) - 1 - T*thw */
        uint64_t pos = T * (2 * ths) + R * (2 * (uint64_t)thw) + C;
        if (pos > 0xFFFFFFFFull) return INTSPMV_HOST_EOVERFLOW;
        con[r] = (uint32_t)pos;
        uint64_t rl = m->rl[r];
        uint64_t positions = noc ? ((rl + grp - 1) / grp) * grp : rl;
        total_cols += (positions + th - 1) / th;
        rsp += positions;
    }
    uint64_t tiles = (total_cols + thw - 1) / thw;
    uint64_t lax = tiles * 2 * ths;
    uint64_t lsec = (noc && th == 1) ? tiles : (noc ? tiles * thw : tiles * ths);
    if (lax > 0xFFFFFFFFull || lsec > 0xFFFFFFFFull) return INTSPMV_HOST_EOVERFLOW;
    *tileN = (uint32_t)tiles;
    *lenAX = (uint32_t)lax;
    *lenSEC = (uint32_t)lsec;
    if (colN) *colN = (uint32_t)total_cols;
    return INTSPMV_HOST_OK;
}

/* slot index of the value at ax position p: tile T, brick-row R, lane C */
static inline uint32_t axt_slot(uint32_t p, uint32_t thw, uint32_t ths)
{
    uint32_t T = p / (2 * ths), w = p % (2 * ths);
    return T * ths + (w / (2 * thw)) * thw + (w % (2 * thw));
}

static void axt_noc_h1(const intspmv_csr *m, const double *x, intspmv_axt *a)
{
    const uint32_t thw = a->tileHW, ths = thw * a->tileH;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ri = 0; ri < (int64_t)a->nrows; ri++) {
        uint32_t r = (uint32_t)ri;
        if (m->rl[r] == 0) continue;
        uint32_t pos = a->con[r], psec = pos / (2 * ths), cnt = 0;
        for (uint32_t k = m->row[r]; k < m->row[r + 1]; k++) {
            a->ax[pos] = m->val[k];
            a->ax[pos + thw] = x[m->col[k]];
            a->sec[psec] = r;
            if (a->acol) {
                a->acol[axt_slot(pos, thw, ths)] = m->col[k];
                a->nzslot[k] = axt_slot(pos, thw, ths);
            }
            pos++;
            cnt++;
            if (cnt % thw == 0) {
                pos += thw;
                psec++;
            }
        }
    }
}

static void axt_noc(const intspmv_csr *m, const double *x, intspmv_axt *a)
{
    const uint32_t thw = a->tileHW, th = a->tileH, ths = thw * th;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ri = 0; ri < (int64_t)a->nrows; ri++) {
        uint32_t r = (uint32_t)ri;
        if (m->rl[r] == 0) continue;
        uint32_t pos = a->con[r];
        uint32_t psec = (pos / (2 * ths)) * thw + pos % thw, cnt = 0;
        for (uint32_t k = m->row[r]; k < m->row[r + 1]; k++) {
            a->ax[pos] = m->val[k];
            a->ax[pos + thw] = x[m->col[k]];
            a->sec[psec] = r;
            if (a->acol) {
                a->acol[axt_slot(pos, thw, ths)] = m->col[k];
                a->nzslot[k] = axt_slot(pos, thw, ths);
            }
            pos += 2 * thw;
            cnt++;
            if (cnt % th == 0) {
                pos = pos + 1 - 2 * th * thw;
                psec++;
                if (pos % thw == 0) pos += (2 * th - 1) * thw;
            }
        }
    }
}

static void axt_com_h1(const intspmv_csr *m, const double *x, intspmv_axt *a)
{
    const uint32_t thw = a->tileHW, bs = a->bs, lg = a->log;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ri = 0; ri < (int64_t)a->nrows; ri++) {
        uint32_t r = (uint32_t)ri, rl = m->rl[r], e = 0;
        if (rl == 0) continue;
        for (uint32_t k = m->row[r]; k < m->row[r + 1]; k++, e++) {
            uint32_t bco = k % bs, tid = k / thw, tco = k % thw;
            uint32_t pos = tid * 2 * thw + tco;
            a->ax[pos] = m->val[k];
            a->ax[pos + thw] = x[m->col[k]];
            if (a->acol) {
                a->acol[k] = m->col[k];
                a->nzslot[k] = k;
            }
            if (e == 0 || bco == 0) {
                uint32_t q1 = rl - e - 1, q2 = bs - 1 - bco;
                uint32_t off = q1 > q2 ? q2 : q1;
                a->sec[k] = (r << lg) | off;
            }
        }
    }
}

int intspmv_axt_build(const intspmv_csr *m, const double *x, uint32_t bs, uint32_t thw,
                      uint32_t th, const char *mode, int want_acol, intspmv_axt *out,
                      intspmv_format_data *fd)
{
    if (!m || !x || !mode || !out || !m->rl) return INTSPMV_HOST_EARG;
    memset(out, 0, sizeof(*out));
    if (thw != INTSPMV_HBRICK_SIZE || th == 0 || th > 512) return INTSPMV_HOST_EARG;
    const int noc = mode_is_noc(mode), com = mode_is_com(mode);
    if (!noc && !com) return INTSPMV_HOST_EARG;
    uint32_t lg = 0;
    if (com) {
        if (th != 1 || bs < 2 || bs > 512 || (bs & (bs - 1)) != 0) return INTSPMV_HOST_EARG;
        for (uint32_t i = 1; i < 10; i++)
            if ((bs >> i) == 1) {
                lg = i;
                break;
            }
        if (m->nrows > 0 && ((uint64_t)(m->nrows - 1) << lg) > 0xFFFFFFFFull)
            return INTSPMV_HOST_EOVERFLOW;
    }
    double t0 = intspmv_wtime();
    out->nrows = m->nrows;
    out->nnz = m->nnz;
    strcpy(out->mode, noc ? "NOC" : "COM");
    out->tileHW = thw;
    out->tileH = th;
    for (uint32_t i = 0; i < 10; i++)
        if ((th >> i) == 1) out->logTH = i;
    out->lenCON = m->nrows;
    out->bs = com ? bs : 0;
    out->log = lg;
    out->con = (uint32_t *)calloc(m->nrows ? m->nrows : 1, sizeof(uint32_t));
    if (!out->con) return INTSPMV_HOST_ENOMEM;
    int st = intspmv_axt_lengths(m, thw, th, mode, out->con, &out->tileN, &out->lenAX,
                                 &out->lenSEC, &out->colN);
    if (st) {
        intspmv_axt_free(out);
        return st;
    }
    out->ax = (double *)aligned_zalloc((size_t)out->lenAX * sizeof(double));
    out->sec = (uint32_t *)aligned_zalloc((size_t)out->lenSEC * sizeof(uint32_t));
    if (want_acol) {
        size_t ns = (size_t)out->lenAX / 2;
        out->acol = (uint32_t *)malloc((ns ? ns : 1) * sizeof(uint32_t));
        out->nzslot = (uint32_t *)malloc((m->nnz ? m->nnz : 1) * sizeof(uint32_t));
        if (out->acol)
            for (size_t s = 0; s < ns; s++) out->acol[s] = INTSPMV_ACOL_PAD;
    }
    if (!out->ax || !out->sec || (want_acol && (!out->acol || !out->nzslot))) {
        intspmv_axt_free(out);
        return INTSPMV_HOST_ENOMEM;
    }
    char name[32];
    if (noc && th == 1) {
        axt_noc_h1(m, x, out);
        strcpy(name, "AXT_NOC_H1_HW8");
    } else if (noc) {
        axt_noc(m, x, out);
        snprintf(name, sizeof name, "AXT_NOC_H%u_HW8", th);
    } else {
        axt_com_h1(m, x, out);
        snprintf(name, sizeof name, "AXT_COM_H1_HW8_BS%u", bs);
    }
    if (fd) {
        memset(fd, 0, sizeof(*fd));
        strcpy(fd->name, name);
        fd->mfp = (double)out->lenAX * sizeof(double) + (double)out->lenSEC * sizeof(uint32_t);
        fd->beta = out->lenAX ? (double)out->nnz / (double)(out->lenAX >> 1) : 0.0;
        fd->ct = intspmv_wtime() - t0;
    }
    return INTSPMV_HOST_OK;
}

int intspmv_axt_refresh_x(intspmv_axt *a, const double *x)
{
    if (!a || !x || !a->acol || !a->ax) return INTSPMV_HOST_EARG;
    const uint32_t thw = a->tileHW, ths = thw * a->tileH;
    const int64_t ns = (int64_t)a->lenAX / 2;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < ns; s++) {
        uint32_t c = a->acol[s];
        if (c == INTSPMV_ACOL_PAD) continue;
        uint32_t T = (uint32_t)s / ths, w = (uint32_t)s % ths;
        size_t p = (size_t)T * 2 * ths + (size_t)(w / thw) * 2 * thw + thw + (w % thw);
        a->ax[p] = x[c];
    }
    return INTSPMV_HOST_OK;
}

void intspmv_axt_free(intspmv_axt *a)
{
    if (!a) return;
    free(a->ax);
    free(a->sec);
    free(a->con);
    free(a->acol);
    free(a->nzslot);
    memset(a, 0, sizeof(*a));
}

/* ------------------------------------------------------ transposed map */

int intspmv_axt_value_plane(const intspmv_axt *a, double *vals)
{
    if (!a || !vals || (a->lenAX && !a->ax)) return INTSPMV_HOST_EARG;
    const uint32_t thw = a->tileHW;
    const int64_t nb = (int64_t)a->lenAX / (2 * thw); /* brick-rows */
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; b++) memcpy(vals + b * thw, a->ax + b * 2 * thw, thw * sizeof(double));
    return INTSPMV_HOST_OK;
}

/* A^T's AXT NOC layout is built by the same builder as A's, from the CSR
 * transpose (entries of a row of A^T in ascending row order of A).  An entry
 * k' of A^T is entry perm[k'] of A, which sits in A's slot a->nzslot[perm[k']]
 * and in A^T's slot at.nzslot[k'], so
 *   permT[at.nzslot[k']] = a->nzslot[perm[k']],
 * and every slot of A^T that holds no entry is padding. */
int intspmv_axt_transpose_map(const intspmv_csr *m, const intspmv_axt *a,
                              intspmv_axt_tmap *out)
{
    if (!m || !a || !out) return INTSPMV_HOST_EARG;
    memset(out, 0, sizeof(*out));
    if (!mode_is_noc(a->mode) || !a->acol || !a->nzslot || a->nrows != m->nrows ||
        a->nnz != m->nnz || a->tileHW != INTSPMV_HBRICK_SIZE)
        return INTSPMV_HOST_EARG;
    double t0 = intspmv_wtime();
    intspmv_csr at;
    uint32_t *perm = (uint32_t *)malloc((m->nnz ? m->nnz : 1) * sizeof(uint32_t));
    double *zero = (double *)calloc(m->nrows ? m->nrows : 1, sizeof(double));
    if (!perm || !zero) {
        free(perm);
        free(zero);
        return INTSPMV_HOST_ENOMEM;
    }
    int st = intspmv_csr_transpose(m, &at, perm);
    if (st == 0) {
        st = intspmv_axt_build(&at, zero, 0, a->tileHW, a->tileH, "NOC", 1, &out->at, NULL);
        intspmv_csr_free(&at);
    }
    free(zero);
    if (st) {
        free(perm);
        memset(out, 0, sizeof(*out));
        return st;
    }
    const size_t ns = (size_t)out->at.lenAX / 2;
    out->lenPERM = (uint32_t)ns;
    out->permT = (uint32_t *)malloc((ns ? ns : 1) * sizeof(uint32_t));
    if (!out->permT) {
        free(perm);
        intspmv_axt_tmap_free(out);
        return INTSPMV_HOST_ENOMEM;
    }
    for (size_t s = 0; s < ns; s++) out->permT[s] = INTSPMV_ACOL_PAD;
    for (uint32_t k = 0; k < m->nnz; k++) out->permT[out->at.nzslot[k]] = a->nzslot[perm[k]];
    free(perm);
    out->ct = intspmv_wtime() - t0;
    return INTSPMV_HOST_OK;
}

void intspmv_axt_tmap_free(intspmv_axt_tmap *t)
{
    if (!t) return;
    intspmv_axt_free(&t->at);
    free(t->permT);
    memset(t, 0, sizeof(*t));
}
