/*
 * csr_io.c -- CSR matrix container, readers and writers.
 *
 * .csr text and .bin binary follow matrixReading (intSpmv_main.c:187-249):
 *   .csr: "nrows nnz", then nnz lines "val col", then nrows+1 row pointers;
 *   .bin: uint32 nrows, uint32 nnz, double val[nnz], uint32 col[nnz],
 *         uint32 row[nrows+1].
 * Indices are 0-based and the matrix is square (x has nrows entries,
 * intSpmv_main.c:1645).  Matrix Market (.mtx) is an addition: coordinate
 * format, 1-based, converted to CSR sorted by (row, col) with duplicates
 * summed.  Unlike the reference, errors are returned, not exit()ed.
 */
#define _POSIX_C_SOURCE 200809L
#include "intspmv_host.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

double intspmv_wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int intspmv_csr_alloc(intspmv_csr *m, uint32_t nrows, uint32_t nnz)
{
    if (!m) return INTSPMV_HOST_EARG;
    memset(m, 0, sizeof(*m));
    m->nrows = nrows;
    m->nnz = nnz;
    m->val = (double *)calloc(nnz ? nnz : 1, sizeof(double));
    m->col = (uint32_t *)calloc(nnz ? nnz : 1, sizeof(uint32_t));
    m->row = (uint32_t *)calloc((size_t)nrows + 1, sizeof(uint32_t));
    m->rl = (uint32_t *)calloc(nrows ? nrows : 1, sizeof(uint32_t));
    if (!m->val || !m->col || !m->row || !m->rl) {
        intspmv_csr_free(m);
        return INTSPMV_HOST_ENOMEM;
    }
    return INTSPMV_HOST_OK;
}

void intspmv_csr_free(intspmv_csr *m)
{
    if (!m) return;
    free(m->val);
    free(m->col);
    free(m->row);
    free(m->rl);
    memset(m, 0, sizeof(*m));
}

static int ends_with(const char *s, const char *suf)
{
    size_t a = strlen(s), b = strlen(suf);
    return a >= b && strcmp(s + a - b, suf) == 0;
}

/* Checks that row[] is a valid pointer array for nnz entries and that every
 * column index is inside [0, nrows). */
static int csr_validate(const intspmv_csr *m)
{
    if (m->row[0] != 0 || m->row[m->nrows] != m->nnz) return INTSPMV_HOST_EFORMAT;
    for (uint32_t i = 0; i < m->nrows; i++)
        if (m->row[i + 1] < m->row[i]) return INTSPMV_HOST_EFORMAT;
    for (uint32_t k = 0; k < m->nnz; k++)
        if (m->col[k] >= m->nrows) return INTSPMV_HOST_EFORMAT;
    return INTSPMV_HOST_OK;
}

static int read_csr_text(const char *path, intspmv_csr *out)
{
    FILE *fh = fopen(path, "r");
    if (!fh) return INTSPMV_HOST_EIO;
    unsigned nrows, nnz;
    int st = INTSPMV_HOST_OK;
    if (fscanf(fh, "%u %u", &nrows, &nnz) != 2) {
        fclose(fh);
        return INTSPMV_HOST_EFORMAT;
    }
    st = intspmv_csr_alloc(out, nrows, nnz);
    if (st) {
        fclose(fh);
        return st;
    }
    for (uint32_t i = 0; i < nnz && !st; i++) {
        unsigned c;
        if (fscanf(fh, "%lf %u", &out->val[i], &c) != 2) st = INTSPMV_HOST_EFORMAT;
        out->col[i] = c;
    }
    for (uint32_t i = 0; i <= nrows && !st; i++) {
        unsigned r;
        if (fscanf(fh, "%u", &r) != 1) st = INTSPMV_HOST_EFORMAT;
        out->row[i] = r;
    }
    fclose(fh);
    if (!st) st = csr_validate(out);
    if (st) intspmv_csr_free(out);
    return st;
}

static int read_csr_bin(const char *path, intspmv_csr *out)
{
    FILE *fh = fopen(path, "rb");
    if (!fh) return INTSPMV_HOST_EIO;
    uint32_t hdr[2];
    if (fread(hdr, sizeof(uint32_t), 2, fh) != 2) {
        fclose(fh);
        return INTSPMV_HOST_EFORMAT;
    }
    int st = intspmv_csr_alloc(out, hdr[0], hdr[1]);
    if (st) {
        fclose(fh);
        return st;
    }
    if (fread(out->val, sizeof(double), hdr[1], fh) != hdr[1] ||
        fread(out->col, sizeof(uint32_t), hdr[1], fh) != hdr[1] ||
        fread(out->row, sizeof(uint32_t), (size_t)hdr[0] + 1, fh) != (size_t)hdr[0] + 1)
        st = INTSPMV_HOST_EFORMAT;
    fclose(fh);
    if (!st) st = csr_validate(out);
    if (st) intspmv_csr_free(out);
    return st;
}

/* Stable (row, col) ordering of COO entries by two counting passes. */
int intspmv_csr_from_coo(uint32_t nrows, size_t n_entries, const uint32_t *ri,
                         const uint32_t *ci, const double *v, intspmv_csr *out)
{
    if (!out || (n_entries && (!ri || !ci || !v))) return INTSPMV_HOST_EARG;
    for (size_t k = 0; k < n_entries; k++)
        if (ri[k] >= nrows || ci[k] >= nrows) return INTSPMV_HOST_EARG;
    /* pass 1: sort by column (stable), pass 2: by row (stable) */
    size_t *cnt = (size_t *)calloc((size_t)nrows + 1, sizeof(size_t));
    size_t *ord1 = (size_t *)malloc((n_entries ? n_entries : 1) * sizeof(size_t));
    size_t *ord2 = (size_t *)malloc((n_entries ? n_entries : 1) * sizeof(size_t));
    if (!cnt || !ord1 || !ord2) {
        free(cnt);
        free(ord1);
        free(ord2);
        return INTSPMV_HOST_ENOMEM;
    }
    for (size_t k = 0; k < n_entries; k++) cnt[ci[k] + 1]++;
    for (uint32_t i = 0; i < nrows; i++) cnt[i + 1] += cnt[i];
    for (size_t k = 0; k < n_entries; k++) ord1[cnt[ci[k]]++] = k;
    memset(cnt, 0, ((size_t)nrows + 1) * sizeof(size_t));
    for (size_t k = 0; k < n_entries; k++) cnt[ri[k] + 1]++;
    for (uint32_t i = 0; i < nrows; i++) cnt[i + 1] += cnt[i];
    for (size_t t = 0; t < n_entries; t++) {
        size_t k = ord1[t];
        ord2[cnt[ri[k]]++] = k;
    }
    free(ord1);
    /* count unique (row, col) pairs */
    size_t nu = 0;
    for (size_t t = 0; t < n_entries; t++) {
        size_t k = ord2[t];
        if (t == 0 || ri[k] != ri[ord2[t - 1]] || ci[k] != ci[ord2[t - 1]]) nu++;
    }
    if (nu > 0xFFFFFFFFull) {
        free(cnt);
        free(ord2);
        return INTSPMV_HOST_EOVERFLOW;
    }
    int st = intspmv_csr_alloc(out, nrows, (uint32_t)nu);
    if (st) {
        free(cnt);
        free(ord2);
        return st;
    }
    size_t p = (size_t)-1;
    for (size_t t = 0; t < n_entries; t++) {
        size_t k = ord2[t];
        if (t == 0 || ri[k] != ri[ord2[t - 1]] || ci[k] != ci[ord2[t - 1]]) {
            p++;
            out->col[p] = ci[k];
            out->val[p] = v[k];
            out->row[ri[k] + 1]++;
        } else {
            out->val[p] += v[k];
        }
    }
    for (uint32_t i = 0; i < nrows; i++) out->row[i + 1] += out->row[i];
    free(cnt);
    free(ord2);
    return INTSPMV_HOST_OK;
}

int intspmv_csr_transpose(const intspmv_csr *a, intspmv_csr *at, uint32_t *perm)
{
    if (!a || !at) return INTSPMV_HOST_EARG;
    int st = intspmv_csr_alloc(at, a->nrows, a->nnz);
    if (st) return st;
    for (uint32_t k = 0; k < a->nnz; k++) at->row[a->col[k] + 1]++;
    for (uint32_t i = 0; i < a->nrows; i++) at->row[i + 1] += at->row[i];
    uint32_t *next = (uint32_t *)malloc(((size_t)a->nrows + 1) * sizeof(uint32_t));
    if (!next) {
        intspmv_csr_free(at);
        return INTSPMV_HOST_ENOMEM;
    }
    memcpy(next, at->row, ((size_t)a->nrows + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < a->nrows; i++)
        for (uint32_t k = a->row[i]; k < a->row[i + 1]; k++) {
            uint32_t q = next[a->col[k]]++;
            at->col[q] = i;
            at->val[q] = a->val[k];
            if (perm) perm[q] = k;
        }
    free(next);
    return intspmv_csr_stats(at, NULL);
}

static int read_mtx(const char *path, intspmv_csr *out)
{
    FILE *fh = fopen(path, "r");
    if (!fh) return INTSPMV_HOST_EIO;
    char line[1024];
    if (!fgets(line, sizeof line, fh)) {
        fclose(fh);
        return INTSPMV_HOST_EFORMAT;
    }
    for (char *q = line; *q; q++) *q = (char)tolower((unsigned char)*q);
    if (strncmp(line, "%%matrixmarket", 14) != 0 || !strstr(line, "coordinate")) {
        fclose(fh);
        return INTSPMV_HOST_EFORMAT;
    }
    int pattern = strstr(line, "pattern") != NULL;
    int symmetric = strstr(line, "symmetric") != NULL;
    int skew = strstr(line, "skew-symmetric") != NULL;
    if (strstr(line, "complex") || strstr(line, "hermitian")) {
        fclose(fh);
        return INTSPMV_HOST_EFORMAT;
    }
    do {
        if (!fgets(line, sizeof line, fh)) {
            fclose(fh);
            return INTSPMV_HOST_EFORMAT;
        }
    } while (line[0] == '%');
    unsigned long long M, N, L;
    if (sscanf(line, "%llu %llu %llu", &M, &N, &L) != 3 || M != N || M > 0xFFFFFFFFull) {
        fclose(fh);
        return INTSPMV_HOST_EFORMAT;
    }
    size_t cap = (size_t)L * ((symmetric || skew) ? 2 : 1);
    uint32_t *ri = (uint32_t *)malloc((cap ? cap : 1) * sizeof(uint32_t));
    uint32_t *ci = (uint32_t *)malloc((cap ? cap : 1) * sizeof(uint32_t));
    double *vv = (double *)malloc((cap ? cap : 1) * sizeof(double));
    int st = (ri && ci && vv) ? INTSPMV_HOST_OK : INTSPMV_HOST_ENOMEM;
    size_t ne = 0;
    for (unsigned long long e = 0; e < L && !st; e++) {
        unsigned long long i, j;
        double v = 1.0;
        if (!fgets(line, sizeof line, fh)) {
            st = INTSPMV_HOST_EFORMAT;
            break;
        }
        int got = pattern ? sscanf(line, "%llu %llu", &i, &j)
                          : sscanf(line, "%llu %llu %lf", &i, &j, &v);
        if (got != (pattern ? 2 : 3) || i < 1 || j < 1 || i > M || j > N) {
            st = INTSPMV_HOST_EFORMAT;
            break;
        }
        ri[ne] = (uint32_t)(i - 1);
        ci[ne] = (uint32_t)(j - 1);
        vv[ne] = v;
        ne++;
        if ((symmetric || skew) && i != j) {
            ri[ne] = (uint32_t)(j - 1);
            ci[ne] = (uint32_t)(i - 1);
            vv[ne] = skew ? -v : v;
            ne++;
        }
    }
    fclose(fh);
    if (!st) st = intspmv_csr_from_coo((uint32_t)M, ne, ri, ci, vv, out);
    free(ri);
    free(ci);
    free(vv);
    return st;
}

int intspmv_csr_read(const char *path, intspmv_csr *out)
{
    if (!path || !out) return INTSPMV_HOST_EARG;
    memset(out, 0, sizeof(*out));
    if (ends_with(path, ".csr")) return read_csr_text(path, out);
    if (ends_with(path, ".bin")) return read_csr_bin(path, out);
    if (ends_with(path, ".mtx")) return read_mtx(path, out);
    return INTSPMV_HOST_EFORMAT;
}

int intspmv_csr_write_bin(const char *path, const intspmv_csr *m)
{
    if (!path || !m) return INTSPMV_HOST_EARG;
    FILE *fh = fopen(path, "wb");
    if (!fh) return INTSPMV_HOST_EIO;
    uint32_t hdr[2] = {m->nrows, m->nnz};
    int ok = fwrite(hdr, sizeof(uint32_t), 2, fh) == 2 &&
             fwrite(m->val, sizeof(double), m->nnz, fh) == m->nnz &&
             fwrite(m->col, sizeof(uint32_t), m->nnz, fh) == m->nnz &&
             fwrite(m->row, sizeof(uint32_t), (size_t)m->nrows + 1, fh) == (size_t)m->nrows + 1;
    ok = (fclose(fh) == 0) && ok;
    return ok ? INTSPMV_HOST_OK : INTSPMV_HOST_EIO;
}

int intspmv_csr_write_csr(const char *path, const intspmv_csr *m)
{
    if (!path || !m) return INTSPMV_HOST_EARG;
    FILE *fh = fopen(path, "w");
    if (!fh) return INTSPMV_HOST_EIO;
    fprintf(fh, "%u %u\n", m->nrows, m->nnz);
    for (uint32_t k = 0; k < m->nnz; k++) fprintf(fh, "%.17g %u\n", m->val[k], m->col[k]);
    for (uint32_t i = 0; i <= m->nrows; i++) fprintf(fh, "%u\n", m->row[i]);
    return fclose(fh) == 0 ? INTSPMV_HOST_OK : INTSPMV_HOST_EIO;
}

int intspmv_csr_write_mtx(const char *path, const intspmv_csr *m)
{
    if (!path || !m) return INTSPMV_HOST_EARG;
    FILE *fh = fopen(path, "w");
    if (!fh) return INTSPMV_HOST_EIO;
    fprintf(fh, "%%%%MatrixMarket matrix coordinate real general\n");
    fprintf(fh, "%u %u %u\n", m->nrows, m->nrows, m->nnz);
    for (uint32_t i = 0; i < m->nrows; i++)
        for (uint32_t k = m->row[i]; k < m->row[i + 1]; k++)
            fprintf(fh, "%u %u %.17g\n", i + 1, m->col[k] + 1, m->val[k]);
    return fclose(fh) == 0 ? INTSPMV_HOST_OK : INTSPMV_HOST_EIO;
}
