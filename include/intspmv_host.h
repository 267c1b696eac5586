/*
 * intspmv_host.h -- host-side C API of the int_spmv MI355X engine.
 *
 * This is the C host that feeds the GPU library (intspmv.h): matrix I/O,
 * row statistics, the CSR -> AXC / K1 / AXT builders and the synthetic
 * matrix generators.  Every builder produces arrays that are bit-identical to
 * the ones the reference builds from the same CSR matrix and vector
 * (intSpmv_main.c), with the padding the reference leaves uninitialised
 * defined as zero.  All functions return 0 on success and a negative
 * INTSPMV_HOST_E* status on failure; none of them calls exit().
 *
 * Index type: uint32_t, as the reference's UIN (intSpmv_main.c:77-79).
 * Value type: double, as the reference's FP64 build (intSpmv_main.c:70-73).
 */
#ifndef INTSPMV_HOST_H
#define INTSPMV_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layout constants (intSpmv_main.c:101-109). */
#define INTSPMV_HBRICK_SIZE 8u /* AXT/AXC tile width, K1 lanes */
#define INTSPMV_CHUNK_SIZE  8u /* K1 slice height */

enum {
    INTSPMV_HOST_OK = 0,
    INTSPMV_HOST_EARG = -1,      /* invalid argument */
    INTSPMV_HOST_ENOMEM = -3,    /* allocation failed */
    INTSPMV_HOST_EIO = -6,       /* file could not be opened / parsed */
    INTSPMV_HOST_EFORMAT = -7,   /* unsupported file extension / content */
    INTSPMV_HOST_EOVERFLOW = -5  /* a length would not fit in 32 bits */
};

/* CSR matrix; mirrors str_matCSR (intSpmv_main.c:183).  rowStart/rowEnd of
 * the reference are only used by its MKL calls and are not kept here. */
typedef struct {
    uint32_t nrows, nnz;
    uint32_t rmin;
    double rave;
    uint32_t rmax;
    double rsd;
    uint32_t bw;
    double *val;   /* [nnz] */
    uint32_t *row; /* [nrows+1] */
    uint32_t *col; /* [nnz] */
    uint32_t *rl;  /* [nrows], filled by intspmv_csr_stats */
} intspmv_csr;

/* Per-format record printed in the "formats' data" table; mirrors
 * str_formatData (intSpmv_main.c:306).  ct is the builder time in seconds. */
typedef struct {
    char name[32];
    double mfp;  /* memory footprint, bytes */
    double beta; /* occupancy nnz / stored slots */
    double ct;   /* conversion time, seconds */
} intspmv_format_data;

/* AXC (intSpmv_main.c:615): brick = [8 values | 8 gathered x]. */
typedef struct {
    uint32_t nrows, nnz, lenAX, lenBRP;
    double *ax;    /* [lenAX], 64-byte aligned, padding = 0 */
    uint32_t *brp; /* [lenBRP] */
} intspmv_axc;

/* K1 (intSpmv_main.c:771): rows sorted by length, 8-row column-major slices. */
typedef struct {
    uint32_t nrows, nnz, chunkNum, lenVC;
    uint32_t *permi; /* [(chunkNum+1)*8], tail zero */
    uint32_t *nmc;   /* [chunkNum] */
    uint32_t *chp;   /* [chunkNum] */
    double *val;     /* [lenVC], padding = 0 */
    uint32_t *col;   /* [lenVC], padding = 0 */
} intspmv_k1;

/* AXT (intSpmv_main.c:1114).  Fields after `con` are additions of this
 * engine: colN is the number of occupied tile columns (the reference's local
 * totalColumns), acol the column index of every slot in ax order (slot s of
 * tile T, brick-row R, lane C is s = T*tileH*8 + R*8 + C) with padding slots
 * set to INTSPMV_ACOL_PAD. */
#define INTSPMV_ACOL_PAD 0xFFFFFFFFu
typedef struct {
    uint32_t nrows, nnz;
    char mode[8]; /* "NOC" or "COM" */
    uint32_t tileHW, tileH, logTH, tileN, lenAX, lenSEC, lenCON, log;
    double *ax;    /* [lenAX], 64-byte aligned, zero-filled before building */
    uint32_t *sec; /* [lenSEC] */
    uint32_t *con; /* [lenCON = nrows] */
    uint32_t bs;   /* COM block size (0 for NOC) */
    uint32_t colN; /* occupied columns (th>1) / slots (th==1) */
    uint32_t *acol; /* [lenAX/2] or NULL when not requested */
    uint32_t *nzslot; /* [nnz]: slot of CSR entry k (with acol), else NULL */
} intspmv_axt;

/* ---- matrix I/O (intSpmv_main.c:187-249, plus Matrix Market) ---------- */
/* Reads by extension: ".csr" text, ".bin" binary (reference formats) or
 * ".mtx" Matrix Market (coordinate real/integer/pattern, general or
 * symmetric; entries sorted by (row,col), duplicates summed). */
int intspmv_csr_read(const char *path, intspmv_csr *out);
int intspmv_csr_write_bin(const char *path, const intspmv_csr *m);
int intspmv_csr_write_csr(const char *path, const intspmv_csr *m);
int intspmv_csr_write_mtx(const char *path, const intspmv_csr *m);
/* Allocates val/col/row/rl for a matrix of the given size (zeroed). */
int intspmv_csr_alloc(intspmv_csr *m, uint32_t nrows, uint32_t nnz);
void intspmv_csr_free(intspmv_csr *m);
/* Transpose: at = A^T in CSR form (rows of at = columns of A, entries of a
 * row in ascending column order); perm[k'] = index in A of entry k' of at
 * (perm may be NULL). */
int intspmv_csr_transpose(const intspmv_csr *a, intspmv_csr *at, uint32_t *perm);
/* Builds CSR from COO triplets: sorts by (row, col), sums duplicates. */
int intspmv_csr_from_coo(uint32_t nrows, size_t n_entries, const uint32_t *ri,
                         const uint32_t *ci, const double *v, intspmv_csr *out);

/* ---- statistics (intSpmv_main.c:253-302) ------------------------------- */
/* Fills rl[] and rmin/rave/rmax/rsd/bw.  When mat_file_name is not NULL the
 * statistics are also written, in the reference's layout, to the file of that
 * name with its extension (from the first '.' of the last path component)
 * replaced by .sta -- printMatrixStats(matFileName, ...). */
int intspmv_csr_stats(intspmv_csr *m, const char *mat_file_name);
/* Format data of plain CSR (intSpmv_main.c:324-357). */
void intspmv_csr_format_data(const intspmv_csr *m, intspmv_format_data *fd);

/* ---- builders ----------------------------------------------------------- */
int intspmv_axc_build(const intspmv_csr *m, const double *x, intspmv_axc *out,
                      intspmv_format_data *fd);
void intspmv_axc_free(intspmv_axc *a);

int intspmv_k1_build(const intspmv_csr *m, intspmv_k1 *out, intspmv_format_data *fd);
void intspmv_k1_free(intspmv_k1 *k);

/* mode "NOC" (th >= 1) or "COM" (th == 1, bs a power of two in [2, 512]);
 * thw must be 8.  want_acol != 0 also fills acol and nzslot. */
int intspmv_axt_build(const intspmv_csr *m, const double *x, uint32_t bs, uint32_t thw,
                      uint32_t th, const char *mode, int want_acol, intspmv_axt *out,
                      intspmv_format_data *fd);
void intspmv_axt_free(intspmv_axt *a);
/* Rewrites the embedded-x half of ax from acol and a new x (padding stays 0). */
int intspmv_axt_refresh_x(intspmv_axt *a, const double *x);
/* Stored values in slot order (slot s = T*tileH*8 + R*8 + C): the value
 * half of every brick-row of ax, lenAX/2 entries -- the aval array of the
 * column-map entry points (spmv_axt_gather and the adjoint). */
int intspmv_axt_value_plane(const intspmv_axt *a, double *vals);

/* The transposed tile map for the adjoint X_bar = A^T Y_bar, built once per
 * matrix, as getFormatDataAXT (intSpmv_main.c:1275-1381) builds the arrays
 * int_axt_noc (1443) consumes:
 *   at     A^T in AXT NOC form with A's tile height, built from the CSR
 *          transpose with acol/nzslot (values in its value half, x half 0);
 *   permT  [lenPERM = at.lenAX/2]: permT[s] = the slot of A's layout holding
 *          the entry that slot s of A^T's layout holds (INTSPMV_ACOL_PAD on
 *          padding).  spmv_permute_vals(permT) moves A's slot values into
 *          A^T's order; spmv_axt_gather_permuted reads them through it.
 * a must be A's NOC layout built with want_acol (for nzslot).  ct is the
 * build time in seconds. */
typedef struct {
    intspmv_axt at;
    uint32_t lenPERM;
    uint32_t *permT;
    double ct;
} intspmv_axt_tmap;
int intspmv_axt_transpose_map(const intspmv_csr *m, const intspmv_axt *a,
                              intspmv_axt_tmap *out);
void intspmv_axt_tmap_free(intspmv_axt_tmap *t);

/* Lengths only (getArraysLenAXT): fills con/tileN/lenAX/lenSEC/colN. */
int intspmv_axt_lengths(const intspmv_csr *m, uint32_t thw, uint32_t th, const char *mode,
                        uint32_t *con, uint32_t *tileN, uint32_t *lenAX, uint32_t *lenSEC,
                        uint32_t *colN);

/* ---- autotuner model ------------------------------------------------------
 * Ranks AXT NOC tile heights th in {1, 2, 4, 8, 16} and right-hand-side
 * blockings kc in {1, 2, 4, 8} (any th) and {16, 32} (th = 1, wide passes;
 * FP64 x also takes 4 and 8 as wide passes at th = 1) for a column-map product
 * with k right-hand sides, sv-byte values and sx-byte x (4 or 8), by modelled
 * bytes moved:
 *   passes * (slots * (sv + 4) + nnz * xline(th, kc) + 4 * lenSEC) + k * n * sx,
 * passes = ceil(k / kc), slots = lenAX / 2 from the exact length formula of
 * getArraysLenAXT (no arrays are built), xline = the 64-byte lines one
 * entry's x gathers touch per pass: one per 16-byte request of a lane (narrow)
 * or the kc * sx bytes the lanes of a wide group read together.  Fills up to
 * cap entries of out in ascending cost; *n receives the number of
 * candidates (at most 30). */
typedef struct {
    uint32_t tileH, kc;
    uint64_t slots, bytes;
} intspmv_tune_choice;
int intspmv_tune_axt(const intspmv_csr *m, uint32_t k, uint32_t sv, uint32_t sx,
                     intspmv_tune_choice *out, uint32_t cap, uint32_t *n);

/* ---- synthetic inputs --------------------------------------------------- */
/* Deterministic generator: splitmix64-seeded xoshiro256**. */
void intspmv_fill_uniform(double *v, size_t n, double lo, double hi, uint64_t seed);
void intspmv_fill_normal(double *v, size_t n, double mean, double sigma, uint64_t seed);
/* 2-D 5-point Laplacian on a g x g grid (diag 4, off-diagonal -1). */
int intspmv_gen_laplace2d(uint32_t g, intspmv_csr *out);
/* Power-law rows: lengths ~ Zipf(s) on [1, maxlen], rescaled so the total is
 * close to target_nnz; band_frac of each row's columns uniform within
 * +-band_half of the diagonal, the rest uniform in [0, n); no duplicates in a
 * row; values N(0,1). */
int intspmv_gen_powerlaw(uint32_t n, uint64_t target_nnz, double s, uint32_t maxlen,
                         double band_frac, uint32_t band_half, uint64_t seed,
                         intspmv_csr *out);
/* 3-D 27-point stencil on g^3: diag 26, off-diagonal -1 + N(0, sigma). */
int intspmv_gen_stencil27(uint32_t g, double sigma, uint64_t seed, intspmv_csr *out);
/* R-MAT graph, 2^scale vertices, edgefactor*2^scale edges, duplicates summed,
 * values U(0,1]. */
int intspmv_gen_rmat(uint32_t scale, uint32_t edgefactor, double a, double b, double c,
                     uint64_t seed, intspmv_csr *out);
/* Banded-random: row lengths uniform in [lmin, lmax]; band_frac of columns in
 * a band of width band_w around the diagonal, the rest uniform; N(0,1). */
int intspmv_gen_banded(uint32_t n, uint32_t lmin, uint32_t lmax, double band_frac,
                       uint32_t band_w, uint64_t seed, intspmv_csr *out);

/* Monotonic wall clock in seconds (replaces the reference's gettimeofday). */
double intspmv_wtime(void);

#ifdef __cplusplus
}
#endif
#endif /* INTSPMV_HOST_H */
