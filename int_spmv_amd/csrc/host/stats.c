/*
 * stats.c -- row lengths and matrix statistics.
 *
 * Follows printMatrixStats (intSpmv_main.c:253-302): rl[i] = row[i+1]-row[i]
 * (the array every builder needs, intSpmv_main.c:261), rmin / rave / rmax /
 * rsd (population standard deviation) and bandwidth max |i - col|.  The .sta
 * file keeps the reference's layout; its path is given by the caller instead
 * of being derived with strtok on '.'.
 */
#include "intspmv_host.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int intspmv_csr_stats(intspmv_csr *m, const char *mat_file_name)
{
    if (!m || !m->row || !m->rl) return INTSPMV_HOST_EARG;
    uint32_t rmin = 1000000000u, rmax = 0, bw = 0;
    double rave1 = 0.0, rave2 = 0.0;
    for (uint32_t i = 0; i < m->nrows; i++) {
        uint32_t rl = m->row[i + 1] - m->row[i];
        m->rl[i] = rl;
        rave1 += rl;
        rave2 += (double)rl * (double)rl;
        rmin = rmin < rl ? rmin : rl;
        rmax = rmax > rl ? rmax : rl;
        for (uint32_t j = m->row[i]; j < m->row[i + 1]; j++) {
            uint32_t d = (i > m->col[j]) ? i - m->col[j] : m->col[j] - i;
            bw = d > bw ? d : bw;
        }
    }
    double n = m->nrows ? (double)m->nrows : 1.0;
    rave1 /= n;
    rave2 /= n;
    double var = rave2 - rave1 * rave1;
    m->rmin = rmin;
    m->rave = rave1;
    m->rmax = rmax;
    m->rsd = sqrt(var > 0.0 ? var : 0.0);
    m->bw = bw;
    if (mat_file_name) {
        /* <name>.sta beside the matrix: the extension of the file name (from
         * its first '.') becomes .sta, as the reference's strtok does for a
         * path without dots in its directories */
        size_t len = strlen(mat_file_name);
        char *sta = (char *)malloc(len + 5);
        if (!sta) return INTSPMV_HOST_ENOMEM;
        memcpy(sta, mat_file_name, len + 1);
        char *base = strrchr(sta, '/');
        char *dot = strchr(base ? base + 1 : sta, '.');
        strcpy(dot ? dot : sta + len, ".sta");
        FILE *fh = fopen(sta, "w+");
        free(sta);
        if (!fh) return INTSPMV_HOST_EIO;
        fprintf(fh, "------------------------------------\n");
        fprintf(fh, "matrix's statistics\n");
        fprintf(fh, "------------------------------------\n");
        fprintf(fh, "name:  %28s\n", mat_file_name);
        fprintf(fh, "nrows: %28u\n", m->nrows);
        fprintf(fh, "nnz:   %28u\n", m->nnz);
        fprintf(fh, "rmin:  %28u\n", m->rmin);
        fprintf(fh, "rave:  %28.2lf\n", m->rave);
        fprintf(fh, "rmax:  %28u\n", m->rmax);
        fprintf(fh, "rsd:   %28.2lf\n", m->rsd);
        fprintf(fh, "rsdp:  %28.2lf\n", rave1 > 0 ? (m->rsd / rave1) * 100 : 0.0);
        fprintf(fh, "bw:    %28u\n", m->bw);
        fclose(fh);
    }
    return INTSPMV_HOST_OK;
}

void intspmv_csr_format_data(const intspmv_csr *m, intspmv_format_data *fd)
{
    /* intSpmv_main.c:344-355: val + col + row + vec bytes, beta = 1. */
    memset(fd, 0, sizeof(*fd));
    strcpy(fd->name, "CSR");
    fd->mfp = (double)m->nnz * sizeof(double) + (double)m->nnz * sizeof(uint32_t) +
              ((double)m->nrows + 1) * sizeof(uint32_t) + (double)m->nrows * sizeof(double);
    fd->beta = m->nnz ? 1.0 : 0.0;
    fd->ct = 0.0;
}
