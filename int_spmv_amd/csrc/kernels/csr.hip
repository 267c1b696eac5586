// csr.hip -- CSR kernels (replacing cf_CSR, intSpmv_main.c:382-396).
//
// spmv_csr_ordered: one lane per row, fma in CSR order from +0.0 (the
//   reference's sequential loop); bit-identical to it, not bandwidth-optimal
//   for rows of very different lengths.
// spmv_csr: plan blocks of <= 512 entries.  Phase 1 streams the block's
//   values and columns with coalesced loads, gathers x[col] and stages both in
//   LDS; phase 2 gives each row to one lane, which runs the same fma chain
//   from LDS -- so every row of a normal block has exactly the ordered bits.
//   Rows longer than 256 entries are streamed by one wave each (split over
//   several waves beyond 8192) and reduced by a wave tree, within the FP64
//   tolerance.  With an exact plan (intspmv_csr_plan_exact) long rows are not
//   split: the wave stages 512 entries at a time in LDS and one lane runs the
//   row's fma chain over them in CSR order, so every row has cf_CSR's bits.
#include "common.hpp"
#include "reduce.hpp"

namespace ispv {

#ifndef ISPV_CSR_UMAX
#define ISPV_CSR_UMAX 512
#endif
#ifndef ISPV_CSR_LONG
#define ISPV_CSR_LONG 256
#endif
constexpr uint32_t CSR_UMAX = ISPV_CSR_UMAX;
constexpr uint32_t CSR_RMAX = 256;
constexpr uint32_t CSR_LONG = ISPV_CSR_LONG;
constexpr uint32_t CSR_LONG_MAX = 8192; // longest row one wave streams alone

__global__ __launch_bounds__(BLOCK_THREADS) void csr_ordered_kernel(
    uint32_t nrows, const double *__restrict__ val, const uint32_t *__restrict__ col,
    const uint32_t *__restrict__ row, const double *__restrict__ x, double *__restrict__ y)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double acc = 0.0;
    for (uint32_t k = row[r], e = row[r + 1]; k < e; k++) acc = fma(val[k], x[col[k]], acc);
    y[r] = acc;
}

// PAIRS (nnz >= 2): phase 1 reads aligned pairs of entries (16 bytes of values
// and 8 of columns per lane, half the load instructions of one entry per lane).
// A pair never reaches past the arrays: pair starts are clamped to the last
// whole pair, and with odd nnz the last entry is fetched on its own (one
// uniform load) and substituted.
template <bool PAIRS, bool EXACT>
__global__ __launch_bounds__(BLOCK_THREADS) void csr_block_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const double *__restrict__ val, const uint32_t *__restrict__ col,
    const uint32_t *__restrict__ row, const double *__restrict__ x, double *__restrict__ y,
    uint32_t nrows, double *partial, uint32_t nnz)
{
    __shared__ double s_v[WAVES][CSR_UMAX];
    __shared__ double s_x[WAVES][CSR_UMAX];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    const uint32_t nu = d.uend - d.ubeg;
    constexpr int U = 8;
    if (EXACT && d.k != 0) {
        // long row, never split: chunks of CSR_UMAX entries staged in LDS
        // (value, gathered x), then lane 0 continues the row's chain over them
        double acc = 0.0;
        for (uint32_t c0 = 0; c0 < nu; c0 += CSR_UMAX) {
            const uint32_t m = nu - c0 < CSR_UMAX ? nu - c0 : CSR_UMAX;
            // the chunk's loads, then its gathers, then its stores, each all
            // issued together (branch-free: lanes past the chunk re-read its
            // last entry and store nothing)
            constexpr int NJ = CSR_UMAX / WAVE;
            double v[NJ], xv[NJ];
            uint32_t c[NJ];
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const uint32_t i = j * WAVE + lane < m ? j * WAVE + lane : m - 1;
                v[j] = ntload(val + d.ubeg + c0 + i);
                c[j] = ntload(col + d.ubeg + c0 + i);
            }
#pragma unroll
            for (int j = 0; j < NJ; j++) xv[j] = x[c[j]];
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                const uint32_t i = j * WAVE + lane;
                if (i < m) {
                    s_v[w][i] = v[j];
                    s_x[w][i] = xv[j];
                }
            }
            wave_lds_sync();
            if (lane == 0)
                in_order_chunks<16, d2v>(0, m, [&](uint32_t i) { return d2v{s_v[w][i], s_x[w][i]}; },
                                         [&](d2v e) { acc = fma(e[0], e[1], acc); });
            wave_lds_sync(); // the next chunk overwrites the staging
        }
        if (lane == 0 && d.rbeg < nrows) y[d.rbeg] = acc;
        return;
    }
    if (d.k != 0) {
        // long row: lane-strided fma chains in registers, one wave tree
        double acc = 0.0;
        for (uint32_t base = 0; base < nu; base += WAVE * U) {
            double v[U];
            uint32_t c[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t i = base + j * WAVE + lane;
                i = i < nu ? i : nu - 1;
                v[j] = ntload(val + d.ubeg + i);
                c[j] = ntload(col + d.ubeg + i);
            }
            // all gathers issued before the chain (clamped lanes re-read a
            // valid column), so they are in flight together
            double xv[U];
#pragma unroll
            for (int j = 0; j < U; j++) xv[j] = x[c[j]];
#pragma unroll
            for (int j = 0; j < U; j++)
                if (base + j * WAVE + lane < nu) acc = fma(v[j], xv[j], acc);
        }
        finish_long<double>(wave_sum_f64(acc), d, bid, y, nrows, partial);
        return;
    }
    // this lane's first row and its bounds, fetched before the stream
    const uint32_t r0 = d.rbeg + lane;
    uint32_t ra = 0, rb = 0;
    if (r0 < d.rend) {
        ra = row[r0];
        rb = row[r0 + 1];
    }
    if constexpr (PAIRS) {
        constexpr int U2 = U / 2;
        const uint32_t pb = d.ubeg & ~1u;             // first pair start (even)
        const uint32_t np = (d.uend - pb + 1) / 2;    // pairs touching the block
        const uint32_t gmax = (nnz - 2) & ~1u;        // last pair inside the arrays
        double tv = 0.0;
        uint32_t tc = 0;
        if ((nnz & 1) && d.uend == nnz) { // uniform: this block holds the lone last entry
            tv = val[nnz - 1];
            tc = col[nnz - 1];
        }
        for (uint32_t base = 0; base < np; base += WAVE * U2) {
            d2v v[U2];
            u2v c[U2];
#pragma unroll
            for (int j = 0; j < U2; j++) {
                uint32_t q = base + j * WAVE + lane;
                q = q < np ? q : np - 1; // branch-free: re-read the last pair
                const uint32_t g = pb + 2 * q, gl = g < gmax ? g : gmax;
                v[j] = ntload(reinterpret_cast<const d2v *>(val + gl));
                c[j] = ntload(reinterpret_cast<const u2v *>(col + gl));
                if (g != gl) { // only the odd last entry: its pair is not read
                    v[j][0] = tv;
                    c[j][0] = tc;
                }
            }
            double x0[U2], x1[U2];
#pragma unroll
            for (int j = 0; j < U2; j++) {
                x0[j] = x[c[j][0]];
                x1[j] = x[c[j][1]];
            }
#pragma unroll
            for (int j = 0; j < U2; j++) {
                const uint32_t q = base + j * WAVE + lane, g = pb + 2 * q;
                if (q < np && g >= d.ubeg) {
                    s_v[w][g - d.ubeg] = v[j][0];
                    s_x[w][g - d.ubeg] = x0[j];
                }
                if (q < np && g + 1 < d.uend) {
                    s_v[w][g + 1 - d.ubeg] = v[j][1];
                    s_x[w][g + 1 - d.ubeg] = x1[j];
                }
            }
        }
    } else {
        for (uint32_t base = 0; base < nu; base += WAVE * U) {
            double v[U];
            uint32_t c[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t i = base + j * WAVE + lane;
                i = i < nu ? i : nu - 1; // branch-free: re-read the last entry
                v[j] = ntload(val + d.ubeg + i);
                c[j] = ntload(col + d.ubeg + i);
            }
            double xv[U];
#pragma unroll
            for (int j = 0; j < U; j++) xv[j] = x[c[j]];
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t i = base + j * WAVE + lane;
                if (i < nu) {
                    s_v[w][i] = v[j];
                    s_x[w][i] = xv[j];
                }
            }
        }
    }
    wave_lds_sync();
    for (uint32_t r = r0; r < d.rend; r += WAVE) {
        uint32_t a = (r == r0 ? ra : row[r]) - d.ubeg, b = (r == r0 ? rb : row[r + 1]) - d.ubeg;
        double acc = 0.0;
        in_order_chunks<8, d2v>(a, b, [&](uint32_t i) { return d2v{s_v[w][i], s_x[w][i]}; },
                                [&](d2v e) { acc = fma(e[0], e[1], acc); });
        y[r] = acc;
    }
}

} // namespace ispv

using namespace ispv;

extern "C" int intspmv_csr_plan(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *row_host,
                                intspmv_plan **plan)
{
    if (!ctx || !plan || !row_host || row_host[0] != 0) return INTSPMV_ERR_ARG;
    return ispv_plan_build(ctx, PLAN_CSR, nrows, row_host, CSR_UMAX, CSR_RMAX, CSR_LONG,
                           CSR_LONG_MAX, plan);
}

extern "C" int intspmv_csr_plan_exact(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *row_host,
                                      intspmv_plan **plan)
{
    if (!ctx || !plan || !row_host || row_host[0] != 0) return INTSPMV_ERR_ARG;
    // long rows in one piece each: rows of up to 2^31 entries
    for (uint32_t r = 0; r < nrows; r++)
        if (row_host[r + 1] < row_host[r] || row_host[r + 1] - row_host[r] >= 0x80000000u)
            return INTSPMV_ERR_ARG;
    int st = ispv_plan_build(ctx, PLAN_CSR, nrows, row_host, CSR_UMAX, CSR_RMAX, CSR_LONG,
                             0x80000000u, plan);
    if (st) return st;
    (*plan)->exact = 1;
    return INTSPMV_OK;
}

extern "C" int intspmv_csr_plan_grouped(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *row_host,
                                        uint32_t ngroups, const uint32_t *group_rows,
                                        intspmv_plan **plan)
{
    if (!ctx || !plan || !row_host || row_host[0] != 0 || ngroups == 0) return INTSPMV_ERR_ARG;
    return ispv_plan_build(ctx, PLAN_CSR, nrows, row_host, CSR_UMAX, CSR_RMAX, CSR_LONG,
                           CSR_LONG_MAX, plan, ngroups, group_rows);
}

extern "C" int spmv_csr_ordered(intspmv_ctx *ctx, uint32_t nrows, const double *val,
                                const uint32_t *col, const uint32_t *row, const double *x,
                                double *y)
{
    int st = ispv_enter(ctx);
    if (st) return st;
    if (nrows == 0) return INTSPMV_OK;
    if (!row || !x || !y) return INTSPMV_ERR_ARG;
    dim3 grid((nrows + BLOCK_THREADS - 1) / BLOCK_THREADS);
    csr_ordered_kernel<<<grid, BLOCK_THREADS, 0, ctx->stream>>>(nrows, val, col, row, x, y);
    return ispv_launch_check(ctx);
}

extern "C" int spmv_csr(intspmv_ctx *ctx, const intspmv_plan *plan, const double *val,
                        const uint32_t *col, const uint32_t *row, const double *x, double *y)
{
    if (!plan || plan->kind != PLAN_CSR || plan->ctx != ctx) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    if (plan->nblocks == 0) return INTSPMV_OK;
    if (!row || !x || !y || (plan->nunits && (!val || !col))) return INTSPMV_ERR_ARG;
    const bool pairs = plan->nunits >= 2 && (uintptr_t)val % 16 == 0 && (uintptr_t)col % 8 == 0;
#define ISPV_CSR(P, E)                                                                          \
    csr_block_kernel<P, E><<<ispv_grid(plan), BLOCK_THREADS, 0, ctx->stream>>>(                  \
        plan->d_blocks, plan->nlong_wg, plan->nblocks, val, col, row, x, y, plan->nrows,        \
        plan->d_partial, plan->nunits)
    if (plan->exact) {
        if (pairs) ISPV_CSR(true, true);
        else ISPV_CSR(false, true);
    } else {
        if (pairs) ISPV_CSR(true, false);
        else ISPV_CSR(false, false);
    }
#undef ISPV_CSR
    st = ispv_launch_check(ctx);
    if (st) return st;
    return ispv_split_fixup(plan, y, 0);
}
