/*
 * intspmv.h -- C ABI of the MI355X (gfx950) SpMV library libintspmv.so.
 *
 * The reference's kernels are static C functions called with raw host
 * pointers (intSpmv_main.c:382-1566).  Each entry point below replaces one of
 * them with a hand-written gfx950 HIP kernel that takes DEVICE pointers to the
 * same arrays the reference builders produce (see intspmv_host.h).
 *
 * Conventions
 *   - Every function returns INTSPMV_OK (0) or a negative INTSPMV_ERR_* code;
 *     intspmv_strerror() names it.  Nothing calls exit().
 *   - Launches are asynchronous on the context's stream (intspmv_set_stream).
 *     One context per host thread; a plan must not be used by two launches
 *     that can overlap in time (its split-row scratch is shared).
 *   - y semantics: all spmv_* entry points OVERWRITE y (y = A*x), including
 *     rows with no entries.  The reference's AXT kernels instead accumulate
 *     into a y their caller zeroed (intSpmv_main.c:1421, 1492, 1591); a host
 *     harness that zeroes y first observes the same result.
 *   - A plan (intspmv_*_plan) is the per-matrix analysis done once, on the
 *     host, like the reference's getFormatData* conversion step: it cuts the
 *     rows into wave-sized work blocks and owns the small device scratch used
 *     to combine rows that are split across waves.
 */
#ifndef INTSPMV_H
#define INTSPMV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    INTSPMV_OK = 0,
    INTSPMV_ERR_ARG = -1,         /* invalid argument / inconsistent sizes */
    INTSPMV_ERR_HIP = -2,         /* a HIP runtime call failed */
    INTSPMV_ERR_NOMEM = -3,       /* device or host allocation failed */
    INTSPMV_ERR_UNSUPPORTED = -4, /* combination not implemented */
    INTSPMV_ERR_OVERFLOW = -5,    /* a size does not fit the 32-bit index type */
    INTSPMV_ERR_NODEVICE = -6,    /* no HIP device visible */
    INTSPMV_ERR_TOO_LARGE = -7    /* beyond spmv_axt_h1_noalloc's fixed scratch */
};

/* element types of stored values / vectors */
enum { INTSPMV_F64 = 0, INTSPMV_F32 = 1, INTSPMV_BF16 = 2 };
/* AXT modes (str_matAXT.mode, intSpmv_main.c:1114) */
enum { INTSPMV_NOC = 0, INTSPMV_COM = 1 };

typedef struct intspmv_ctx intspmv_ctx;
typedef struct intspmv_plan intspmv_plan;

/* ---- context, memory, errors ------------------------------------------ */
int intspmv_create(intspmv_ctx **ctx, int device);
int intspmv_destroy(intspmv_ctx *ctx);
int intspmv_set_stream(intspmv_ctx *ctx, void *hip_stream); /* NULL = null stream */
int intspmv_sync(intspmv_ctx *ctx);
const char *intspmv_strerror(int status);
const char *intspmv_last_hip_error(intspmv_ctx *ctx);
int intspmv_device_count(int *count);
int intspmv_malloc(intspmv_ctx *ctx, void **dptr, size_t bytes);
int intspmv_free(intspmv_ctx *ctx, void *dptr);
int intspmv_memcpy_htod(intspmv_ctx *ctx, void *dst, const void *src, size_t bytes);
int intspmv_memcpy_dtoh(intspmv_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- CSR ----------------------------------------------------------------
 * Replaces cf_CSR (intSpmv_main.c:382-396).
 * spmv_csr_ordered: one lane per row, fused multiply-add in CSR order from
 *   +0.0 -- the same bits as cf_CSR compiled with contraction.
 * spmv_csr: nnz-balanced wave blocks; rows that fit a block (all rows of up to
 *   `long_units` entries) get the same bits as spmv_csr_ordered, longer rows
 *   are reduced by a wave tree and split across waves. */
int intspmv_csr_plan(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *row_host,
                     intspmv_plan **plan);
/* A plan under which spmv_csr gives the bits of cf_CSR (intSpmv_main.c:382-396)
 * for every row: rows longer than a block are never split, and each is one fma
 * chain in CSR order from +0.0, staged through LDS by one wave.  Rows of at
 * least 2^31 entries are refused.  The plan-free spmv_csr_ordered gives the
 * same bits with one lane per row. */
int intspmv_csr_plan_exact(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *row_host,
                           intspmv_plan **plan);
/* The same plan with rows in groups pinned to XCDs, as intspmv_axt_plan_grouped
 * (group g on XCD g % 8; group_rows = NULL: ngroups groups of equal entries).
 * spmv_csr gives the same bits with either plan. */
int intspmv_csr_plan_grouped(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *row_host,
                             uint32_t ngroups, const uint32_t *group_rows, intspmv_plan **plan);
int spmv_csr_ordered(intspmv_ctx *ctx, uint32_t nrows, const double *val, const uint32_t *col,
                     const uint32_t *row, const double *x, double *y);
int spmv_csr(intspmv_ctx *ctx, const intspmv_plan *plan, const double *val,
             const uint32_t *col, const uint32_t *row, const double *x, double *y);

/* ---- AXC ----------------------------------------------------------------
 * Replaces int_axc (intSpmv_main.c:711-732): per brick the halving-tree sum
 * of val*x, accumulated per row in brick order. */
int intspmv_axc_plan(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *brp_host,
                     intspmv_plan **plan);
int spmv_axc(intspmv_ctx *ctx, const intspmv_plan *plan, const double *ax, const uint32_t *brp,
             double *y);

/* ---- K1 -----------------------------------------------------------------
 * Replaces int_k1 + orderArrayK1 (intSpmv_main.c:943-976) and int_k1p
 * (intSpmv_main.c:1026-1066).  With permi == NULL the permuted result
 * res[chunkNum*8] is written (int_k1); otherwise y[permi[i]] (int_k1p, which
 * masks the padded lanes of the last chunk).  val must be 16-byte and col
 * 8-byte aligned (as device allocations are; the reference's _mm_malloc gives
 * 64). */
int intspmv_k1_plan(intspmv_ctx *ctx, uint32_t nrows, uint32_t chunkNum, const uint32_t *nmc_host,
                    intspmv_plan **plan);
int spmv_k1(intspmv_ctx *ctx, const intspmv_plan *plan, const double *val, const uint32_t *col,
            const uint32_t *chp, const uint32_t *nmc, const uint32_t *permi, const double *x,
            double *y);
/* orderArrayK1 alone: y[permi[i]] = res[i] for i < nrows. */
int spmv_k1_order(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *permi, const double *res,
                  double *y);
/* The K1 conversion on the device from a device CSR (row[nrows+1], col, val):
 * getFormatDataK1 with orderFunction, getArrayPermiK1, getArraysNmcChpK1 and
 * getArraysValColK1 (intSpmv_main.c:775-905), the same arrays as the host
 * builder (rows sorted by length, longest first, ties in row order).  Writes
 * permi[(chunkNum+1)*8], nmc[chunkNum], chp[chunkNum] (chunkNum =
 * ceil(nrows/8)) and *lenVC; with val_out/col_out non-NULL also the lenVC
 * entries of val/col (INTSPMV_ERR_OVERFLOW if cap_vc < lenVC: call once with
 * NULL to learn lenVC).  Synchronous; allocates temporaries. */
int intspmv_k1_build_device(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *row,
                            const uint32_t *col, const double *val, uint32_t *permi,
                            uint32_t *nmc, uint32_t *chp, double *val_out, uint32_t *col_out,
                            uint32_t cap_vc, uint32_t *lenVC);

/* ---- AXT ----------------------------------------------------------------
 * Replaces int_axt_noc_h1 (intSpmv_main.c:1385-1404), int_axt_noc
 * (intSpmv_main.c:1443-1474) and int_axt_com_h1 (intSpmv_main.c:1517-1566).
 * The descriptor carries the lengths of str_matAXT plus colN, the number of
 * occupied tile columns (NOC th>1) or slots (th==1): the reference's local
 * totalColumns of getArraysLenAXT (intSpmv_main.c:1127-1141). */
typedef struct {
    uint32_t nrows, nnz;
    uint32_t tileHW, tileH, tileN;
    uint32_t lenAX, lenSEC, colN;
    uint32_t bs, log; /* COM only */
    int32_t mode;     /* INTSPMV_NOC or INTSPMV_COM */
    int32_t val_type; /* stored value type: F64, F32 or BF16 */
    int32_t x_type;   /* embedded / gathered x type: F64 or F32 */
    int32_t layout;   /* INTSPMV_AXT_INTERLEAVED, _PLANAR or _COLMAP */
} intspmv_axt_desc;

/* INTERLEAVED: the reference's ax: per brick-row 8 values then 8 x, both of
 *   the same type (F64/F64 is the drop-in; F32/F32 the FP32 storage variant).
 * PLANAR: values and embedded x in two separate slot-ordered arrays (slot s of
 *   tile T, brick-row R, lane C is T*tileH*8 + R*8 + C); used for mixed types
 *   such as BF16 values with FP32 x.
 * COLMAP: a plan for the column-map entry points below (spmv_axt_gather,
 *   spmv_axt_grad_vals); spmv_axt refuses it.  Plans of the embedded layouts
 *   may use larger blocks than the column-map kernels stage and are then
 *   refused by spmv_axt_gather. */
enum { INTSPMV_AXT_INTERLEAVED = 0, INTSPMV_AXT_PLANAR = 1, INTSPMV_AXT_COLMAP = 2 };

/* con_host: str_matAXT.con (lenCON = nrows) from the builder. */
int intspmv_axt_plan(intspmv_ctx *ctx, const intspmv_axt_desc *desc, const uint32_t *con_host,
                     intspmv_plan **plan);
/* A column-map plan (layout INTSPMV_AXT_COLMAP) whose rows come in ngroups
 * contiguous groups, rows [group_rows[g], group_rows[g+1]), group_rows[0] = 0,
 * group_rows[ngroups] = nrows; group_rows = NULL cuts ngroups groups of about
 * equal tiles.  Group g runs on XCD g % 8 of the MI355X (its work blocks, long
 * rows included, go only to the workgroups that XCD receives), so the x
 * values of a band around a group's rows stay in that XCD's L2.  Same
 * products and the same bits as intspmv_axt_plan's plan; which is faster
 * depends on the matrix (power-law rows with a diagonal band: grouped;
 * R-MAT with multi-RHS wide passes: ungrouped).  No reference counterpart
 * (the reference schedules rows with OpenMP, intSpmv_main.c:1391). */
int intspmv_axt_plan_grouped(intspmv_ctx *ctx, const intspmv_axt_desc *desc,
                             const uint32_t *con_host, uint32_t ngroups,
                             const uint32_t *group_rows, intspmv_plan **plan);
/* The autotuner's placement choice behind the C ABI: builds both column-map
 * plans of a matrix (intspmv_axt_plan, and intspmv_axt_plan_grouped with eight
 * groups of equal tiles), times spmv_axt_gather with k right-hand sides under
 * each (reps launches after one untimed, on the context's stream, on a zero X
 * it allocates and frees) and returns the faster plan in *plan.  *grouped = 1
 * when that is the grouped plan; the two times in microseconds per product go
 * to us_plain / us_grouped (each may be NULL).  The plans give the same bits,
 * so the choice changes only the time.  No reference counterpart. */
int intspmv_axt_plan_tuned(intspmv_ctx *ctx, const intspmv_axt_desc *desc,
                           const uint32_t *con_host, const void *aval, const uint32_t *acol,
                           const uint32_t *sec, uint32_t k, uint32_t reps, intspmv_plan **plan,
                           int *grouped, float *us_plain, float *us_grouped);
/* Embedded-x product.  INTERLEAVED: ax = the reference array, xplane = NULL.
 * PLANAR: ax = value plane, xplane = x plane.  y has the accumulation type
 * (double for F64 values, float otherwise). */
int spmv_axt(intspmv_ctx *ctx, const intspmv_plan *plan, const void *ax, const void *xplane,
             const uint32_t *sec, void *y);

/* ---- AXT with a column map (x not embedded) -----------------------------
 * aval[s] / acol[s]: value and column of slot s in the tile order above
 * (acol = 0xFFFFFFFF on padding), sec as built.  The plan is the AXT plan of
 * the same matrix, built with layout INTSPMV_AXT_COLMAP (its val_type /
 * x_type select the kernel).  X is n x k row-major with leading dimension ldx >= k, Y likewise;
 * 1 <= k <= 32 right-hand sides.  Y = A X.  No reference counterpart: the
 * reference embeds x (intSpmv_main.c:1171, 1207), which fixes one x per build. */
int spmv_axt_gather(intspmv_ctx *ctx, const intspmv_plan *plan, const void *aval,
                    const uint32_t *acol, const uint32_t *sec, const void *x, uint32_t ldx,
                    void *y, uint32_t ldy, uint32_t k, uint32_t kc);
/* kc: right-hand sides per pass over the matrix.  1, 2, 4, 8: narrow passes
 * (lanes along the slots, each gathering its slot's kc values of x); 16:
 * narrow for F32/BF16 values at any tile height; with tile height 1,
 * kc >= 2 * (16 / sizeof x) -- 8, 16, 32 for F32 x; 4, 8, 16, 32 for F64 x --
 * is a wide pass: lanes along the right-hand sides, a group of lanes reading
 * each x row of kc values together, the matrix streamed once per pass, long
 * rows included.  Both give the same bits for rows inside a block; long rows
 * (more than a block) agree within rounding.  0 = automatic: at tile height 1
 * one wide pass of the smallest width >= what is left of k (up to 32), else
 * the largest narrow width <= it. */
/* The same product with the slot values read through a permutation: value
 * of slot s = vals[vperm[s]] (0 where vperm[s] = 0xFFFFFFFF; vperm 8-byte
 * aligned).  On the transposed tile map, with its slot map (A^T slot -> A
 * slot) and A's own values, this is the adjoint X_bar = A^T Y_bar straight
 * after a value update, without first moving the values into A^T's slot
 * order (spmv_permute_vals).  F64 and F32.  No reference counterpart. */
int spmv_axt_gather_permuted(intspmv_ctx *ctx, const intspmv_plan *plan, const void *vals,
                             const uint32_t *vperm, const uint32_t *acol, const uint32_t *sec,
                             const void *x, uint32_t ldx, void *y, uint32_t ldy, uint32_t k,
                             uint32_t kc);
/* Value gradient on the pattern: vbar[s] = ybar[row(s)] * x[acol[s]] (0 on
 * padding), one entry per slot (lenAX/2).  F64 or F32. */
int spmv_axt_grad_vals(intspmv_ctx *ctx, const intspmv_plan *plan, const uint32_t *acol,
                       const uint32_t *sec, const void *x, const void *ybar, void *vbar);
/* On the transposed tile map of A (intspmv_axt_transpose_map) with A^T's own
 * copy of the values: x_bar = A^T y_bar (one right-hand side, the product of
 * spmv_axt_gather) and, in the same pass, A's value gradient in A^T's slot
 * order, vbar[s] = y_bar[acol[s]] * x[row(s)] (0 on padding) -- the value of
 * spmv_axt_grad_vals for the entry slot s holds, bit for bit.  For callers
 * that keep A's values in both slot orders (updated alike by an elementwise
 * optimiser), so that the adjoint never reads A's values through the slot
 * map.  F64 and F32; vbar 16-byte (F64) / 8-byte (F32) aligned.  No reference
 * counterpart. */
int spmv_axt_gather_grad(intspmv_ctx *ctx, const intspmv_plan *plan, const void *aval,
                         const uint32_t *acol, const uint32_t *sec, const void *ybar, void *xbar,
                         const void *x, void *vbar);
/* Same for the reference layout, whose tiles already hold x[col]:
 * vbar[s] = ybar[row(s)] * (x half of slot s). */
int spmv_axt_grad_vals_embedded(intspmv_ctx *ctx, const intspmv_plan *plan, const double *ax,
                                const uint32_t *sec, const double *ybar, double *vbar);
/* dst[s] = src[perm[s]] (0 where perm[s] == 0xFFFFFFFF) for s < n: moves slot
 * values between the slot orders of A and of A^T (the transposed map). */
int spmv_permute_vals(intspmv_ctx *ctx, uint32_t n, int val_type, const uint32_t *perm,
                      const void *src, void *dst);
/* Device-side conversion: rewrites the embedded x of every slot of the
 * reference layout from acol and a new x (the host builder's ax[..+8] =
 * vec[col] step, intSpmv_main.c:1171, 1207; 0 on padding, as the builder
 * leaves it), so x can change without a rebuild.  acol 8-byte and ax 16-byte
 * aligned. */
int spmv_axt_refresh_x(intspmv_ctx *ctx, const intspmv_plan *plan, const uint32_t *acol,
                       const double *x, double *ax);

/* ---- batches ------------------------------------------------------------
 * Many AXT NOC matrices in column-map form, each with its own tile height,
 * 1..32 right-hand sides and right-hand-side blocking kc (1, 2, 4, 8: narrow
 * passes; at tile height 1 also the wide passes of spmv_axt_gather, kc >=
 * 2 * (16 / sizeof x) up to 32; 0 = as spmv_axt_gather chooses), run by two
 * persistent kernels -- narrow and wide passes, on the caller's stream and a
 * second stream of the batch, joined back -- plus one fix-up kernel when some
 * row is longer than a wave streams alone.  A matrix's rows get the bits of
 * spmv_axt_gather with the same kc.  The work list is built once on the host.  intspmv_batch_graph_capture records the
 * launch in a hipGraph once; intspmv_batch_graph_launch replays it.  Pointers
 * in intspmv_batch_matrix are device pointers except con_host.  All matrices of
 * a batch share val_type / x_type (F64/F64, F32/F32 or BF16/F32).  No
 * reference counterpart (the reference times one matrix and one x). */
typedef struct {
    const void *aval;          /* slot-order values */
    const uint32_t *acol, *sec;
    const void *x;             /* n x k, leading dimension ldx */
    void *y;                   /* n x k, leading dimension ldy */
    uint32_t ldx, ldy, k, kc;
    uint32_t nrows, tileH, tileN, lenAX, colN;
    uint32_t umax;             /* units per work block (8..128; 0 = 128) */
    const uint32_t *con_host;  /* str_matAXT.con of this matrix */
} intspmv_batch_matrix;
typedef struct intspmv_batch intspmv_batch;
int intspmv_batch_create(intspmv_ctx *ctx, int val_type, int x_type, uint32_t nmat,
                         const intspmv_batch_matrix *mats, intspmv_batch **batch);
int intspmv_batch_launch(intspmv_ctx *ctx, const intspmv_batch *batch);
int intspmv_batch_graph_capture(intspmv_ctx *ctx, intspmv_batch *batch);
int intspmv_batch_graph_launch(intspmv_ctx *ctx, const intspmv_batch *batch);
int intspmv_batch_info(const intspmv_batch *batch, uint32_t *nitems, uint32_t *grid,
                       uint32_t *nsplit);
int intspmv_batch_destroy(intspmv_batch *batch);

/* ---- allocation-free product --------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc):
 * the reference-layout product of spmv_axt for NOC, tile height 1, FP64
 * values and x interleaved (int_axt_noc_h1, intSpmv_main.c:1385-1404), with
 * no plan: y = A x from the builder's ax[tileN*16] and sec[tileN] alone.  It
 * calls no hipMalloc / hipFree: rows that cross its fixed chunks of 64 tiles
 * leave partial sums in a fixed device scratch array of
 * INTSPMV_NOALLOC_SCRATCH_BYTES per device, set up by intspmv_create, which
 * limits tileN to INTSPMV_NOALLOC_MAX_TILES (INTSPMV_ERR_TOO_LARGE beyond,
 * nothing launched).  Every scratch entry a call reads is written earlier in
 * the same call; calls that overlap in time on one GPU (several streams or
 * contexts) share the array and race.  Rows inside one chunk get the bits of
 * the serial tile order; rows across chunks add the chunks' partial sums in
 * chunk order.  sec must be non-decreasing (as built). */
#define INTSPMV_NOALLOC_SCRATCH_BYTES (8u << 20)
#define INTSPMV_NOALLOC_MAX_TILES (INTSPMV_NOALLOC_SCRATCH_BYTES / 16u * 64u)
int spmv_axt_h1_noalloc(intspmv_ctx *ctx, uint32_t nrows, uint32_t tileN, const double *ax,
                        const uint32_t *sec, double *y);

/* Measurement utility, no reference counterpart: the rate (gathers per
 * second) of a kernel that streams `count` 4-byte indexes and gathers one
 * `elem_bytes`-wide element (4, 8 or 16) at each, uniformly over a buffer of
 * `span_bytes`, timed over `reps` launches on the context's stream.  It is
 * the ceiling of a product whose entries each gather x alone (the column
 * map, CSR, K1): a span inside one XCD's L2 gives the L2 request ceiling,
 * the span of x the ceiling for uniformly placed columns.  Allocates and
 * frees its own buffers; synchronises the stream. */
int intspmv_gather_rate(intspmv_ctx *ctx, size_t span_bytes, uint32_t elem_bytes, size_t count,
                        uint32_t reps, double *gathers_per_s);
/* Measurement utility, no reference counterpart: the rate (bytes per second)
 * of a plain read of a `bytes`-long device buffer (16-byte nontemporal loads,
 * whole lines per load instruction, each wave a contiguous 1, 2, 4 or 8 KB:
 * the best of the four), timed over `reps` back-to-back launches on the
 * context's stream.  It is the ceiling the streaming products (the
 * embedded-x AXT, AXC) are measured against, beside the 8 TB/s peak.
 * Allocates and frees its own buffer; synchronises the stream. */
int intspmv_stream_rate(intspmv_ctx *ctx, size_t bytes, uint32_t reps, double *bytes_per_s);
int intspmv_plan_destroy(intspmv_plan *plan);
/* Host-only inspection of the work decomposition a plan uses for a row ->
 * unit pointer uptr[nrows+1]: blocks (8 words each: ubeg, uend, rbeg, rend,
 * b0, k, piece, 0) and split rows (4 words: row, first piece, pieces, 0).
 * Arrays are filled when their capacity (in entries) suffices; the counts are
 * always returned.  No device is touched. */
int intspmv_plan_layout(uint32_t nrows, const uint32_t *uptr, uint32_t umax, uint32_t rmax,
                        uint32_t long_units, uint32_t long_max, uint32_t *blocks,
                        size_t cap_blocks, uint32_t *nblocks, uint32_t *splits,
                        size_t cap_splits, uint32_t *nsplits, uint32_t *nlong_wg);
/* The same for a plan with row groups pinned to XCDs (intspmv_axt_plan_grouped). */
int intspmv_plan_layout_grouped(uint32_t nrows, const uint32_t *uptr, uint32_t umax,
                                uint32_t rmax, uint32_t long_units, uint32_t long_max,
                                uint32_t ngroups, const uint32_t *group_rows, uint32_t *blocks,
                                size_t cap_blocks, uint32_t *nblocks, uint32_t *splits,
                                size_t cap_splits, uint32_t *nsplits, uint32_t *nlong_wg);
/* number of work blocks / split rows of a plan (for reporting) */
int intspmv_plan_info(const intspmv_plan *plan, uint32_t *nblocks, uint32_t *nsplit_rows,
                      uint32_t *nlong_rows);

#ifdef __cplusplus
}
#endif
#endif /* INTSPMV_H */
