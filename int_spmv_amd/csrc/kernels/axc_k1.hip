// axc_k1.hip -- AXC and K1 kernels.
//
// AXC (int_axc, intSpmv_main.c:711-732): rows of 16-element bricks
//   [8 values | 8 embedded x] starting at brp[r].  Phase 1 forms each brick's
//   reduce tree with four lanes per brick (as the AXT th==1 kernel); phase 2
//   gives each row to one lane, which adds its brick sums in order from +0.0
//   -- the reference's `sum = sum + red`.
// K1 (int_k1 / int_k1p, intSpmv_main.c:943-1066): chunk c holds 8 rows of the
//   length-sorted order, element k of lane l at chp[c] + 8k + l.  Eight lanes
//   per chunk, one fma chain per lane in k order (the reference's lane-wise
//   vtSum).  Chunks with more than K1_WIDE steps get a whole wave each (and
//   are cut into K1_PIECE-step pieces that meet in the plan scratch); the
//   others are packed into blocks of consecutive chunks (at most K1_BLK
//   slots, contiguous in val/col): a wave streams a block with 16-byte pair
//   loads, gathers x and stages both in LDS, then each lane runs one row's
//   chain from LDS.  The result goes to res[8c + l] (int_k1) or is
//   scattered to y[permi[8c + l]] for the valid lanes (int_k1p).
#include <stdlib.h>

#include <vector>

#include "common.hpp"
#include "reduce.hpp"

namespace ispv {

// Blocks of 64 bricks (8 KB, all in flight at once), rows above 48 bricks on
// waves of their own; matrices with many empty rows take 128 / 64.
constexpr uint32_t AXC_UMAX = 128; // LDS capacity (bricks per block)
constexpr uint32_t AXC_UMAX_DENSE = 64, AXC_LONG_DENSE = 48;
constexpr uint32_t AXC_RMAX = 256;
constexpr uint32_t AXC_LONG = 64;
#ifndef ISPV_K1_WIDE
#define ISPV_K1_WIDE 64
#endif
constexpr uint32_t K1_WIDE = ISPV_K1_WIDE; // chunks with more steps get a whole wave
constexpr uint32_t K1_PIECE = 512; // steps per wide item (longer chunks split)
constexpr uint32_t K1_BLK = 512;   // slots per staged block (a multiple of 128)
constexpr uint32_t K1_BLK_CHUNKS = 64; // chunks per staged block
static_assert(K1_BLK % 128 == 0 && 8 * K1_WIDE <= K1_BLK, "a narrow chunk fits a staged block");

__global__ __launch_bounds__(BLOCK_THREADS) void axc_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const double *__restrict__ ax, const uint32_t *__restrict__ brp, double *__restrict__ y,
    uint32_t nrows, double *partial)
{
    __shared__ double s_val[WAVES][AXC_UMAX];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    // full-line mapping (as the AXT th==1 kernel): each load instruction
    // covers 8 whole 128-byte bricks; part p = l%8 holds values 2p,2p+1 (p<4)
    // or x (p>=4), and a DPP row shift pairs them
    const uint32_t nu = d.uend - d.ubeg, p = lane & 7;
    constexpr int U = 8;
    if (d.k != 0) {
        double acc = 0.0;
        for (uint32_t base = 0; base < nu; base += 8 * 2 * U) {
            d2v v[2 * U];
#pragma unroll
            for (int j = 0; j < 2 * U; j++) {
                uint32_t t = base + j * 8 + lane / 8;
                size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
                v[j] = ntload(reinterpret_cast<const d2v *>(ax + T * 16 + 2 * p));
            }
            loads_issued();
#pragma unroll
            for (int j = 0; j < 2 * U; j++) {
                double x0 = dpp_f64<DPP_ROW_SHL4>(v[j][0]), x1 = dpp_f64<DPP_ROW_SHL4>(v[j][1]);
                if (base + j * 8 + lane / 8 < nu && p < 4) acc = acc + v[j][0] * x0 + v[j][1] * x1;
            }
        }
        finish_long<double>(wave_sum_f64(acc), d, bid, y, nrows, partial);
        return;
    }
    // this lane's first row bounds, fetched before the stream
    const uint32_t r0 = d.rbeg + lane;
    uint32_t ra = 0, rb = 0;
    if (r0 < d.rend) {
        ra = brp[r0];
        rb = brp[r0 + 1];
    }
    // rounds of UN loads (a dense-matrix block); the rounds of a larger block
    // (matrices with many empty rows) are all in flight together, and a round
    // past the block's bricks (every round of a block of empty rows, whose
    // clamp brick nu - 1 does not exist) loads nothing
    constexpr int UN = AXC_UMAX_DENSE / 8, NR = AXC_UMAX / AXC_UMAX_DENSE;
    d2v vr[NR][UN];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (nu <= 8 * UN * r) break; // uniform
#pragma unroll
        for (int j = 0; j < UN; j++) {
            uint32_t t = 8 * UN * r + j * 8 + lane / 8;
            size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
            vr[r][j] = ntload(reinterpret_cast<const d2v *>(ax + T * 16 + 2 * p));
        }
    }
    loads_issued();
#pragma unroll
    for (int r = 0; r < NR; r++) {
        if (nu <= 8 * UN * r) break; // uniform
        const uint32_t base = 8 * UN * r;
        d2v (&v)[UN] = vr[r];
#pragma unroll
        for (int j = 0; j < UN; j++) {
            const bool ok = base + j * 8 + lane / 8 < nu && p < 4;
            double x0 = dpp_f64<DPP_ROW_SHL4>(v[j][0]), x1 = dpp_f64<DPP_ROW_SHL4>(v[j][1]);
            double p0 = ok ? v[j][0] * x0 : 0.0, p1 = ok ? v[j][1] * x1 : 0.0;
            p0 = p0 + dpp_f64<DPP_QUAD_XOR2>(p0);
            p1 = p1 + dpp_f64<DPP_QUAD_XOR2>(p1);
            p0 = p0 + dpp_f64<DPP_QUAD_XOR1>(p0);
            p1 = p1 + dpp_f64<DPP_QUAD_XOR1>(p1);
            uint32_t t = base + j * 8 + lane / 8;
            if (t < nu && p == 0) s_val[w][t] = p0 + p1;
        }
    }
    wave_lds_sync();
    for (uint32_t r = r0; r < d.rend; r += WAVE) {
        uint32_t a = (r == r0 ? ra : brp[r]) / 16 - d.ubeg, b = (r == r0 ? rb : brp[r + 1]) / 16 - d.ubeg;
        double sum = 0.0;
        in_order_chunks<8, double>(a, b, [&](uint32_t i) { return s_val[w][i]; },
                                   [&](double v) { sum = sum + v; });
        y[r] = sum;
    }
}

// K1 work item, wide: chunk c, steps [kbeg, kend); k > 1 marks pieces
// b0..b0+k-1.  Staged block (after the wide items): chunks [c, kbeg), slots
// [kend, b0) of val/col.
struct K1Item {
    uint32_t c, kbeg, kend, b0, k, pad0, pad1, pad2;
};

// One wave per item.  Wide items: all 64 lanes on one chunk, lane (l%8 = row
// of the chunk, l/8 = step slice) runs the steps k = kbeg + l/8 (mod 8), so
// every load instruction reads 64 consecutive entries, and the eight slices
// are added by a tree at the end (long chunks, where eight lanes alone would
// serialise thousands of dependent gathers).  Staged blocks: see the file
// comment; each row's fma chain runs over its steps in order from +0.0 -- the
// reference's lane-wise vtSum, padding terms (0 * x[0]) included.
__global__ __launch_bounds__(BLOCK_THREADS) void k1_kernel(
    const K1Item *__restrict__ items, uint32_t nitems, uint32_t nwide,
    const double *__restrict__ val, const uint32_t *__restrict__ col,
    const uint32_t *__restrict__ chp, const uint32_t *__restrict__ permi,
    const double *__restrict__ x, double *__restrict__ y, uint32_t nrows, double *partial)
{
    __shared__ double s_v[WAVES][K1_BLK];
    __shared__ double s_x[WAVES][K1_BLK];
    const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
    const uint32_t wid = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + w);
    if (wid >= nitems) return;
    const uint32_t l = lane % 8;
    if (wid < nwide) {
        constexpr int U = 8;
        const K1Item d = items[wid];
        double acc = 0.0;
        const uint32_t base = chp[d.c] + l, s0 = lane / 8;
        for (uint32_t k = d.kbeg + s0; k < d.kend; k += 8 * U) {
            double v[U];
            uint32_t cc[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t kk = k + 8 * j < d.kend ? k + 8 * j : d.kend - 1;
                v[j] = ntload(val + base + 8 * kk);
                cc[j] = ntload(col + base + 8 * kk);
            }
            double xv[U];
#pragma unroll
            for (int j = 0; j < U; j++) xv[j] = x[cc[j]];
#pragma unroll
            for (int j = 0; j < U; j++)
                if (k + 8 * j < d.kend) acc = fma(v[j], xv[j], acc);
        }
        acc += shfl_xor_f64(acc, 8);
        acc += shfl_xor_f64(acc, 16);
        acc += shfl_xor_f64(acc, 32);
        if (lane >= 8) return;
        const uint32_t s = d.c * 8 + l;
        if (d.k <= 1) {
            if (!permi) y[s] = acc;
            else if (s < nrows) y[permi[s]] = acc;
        } else {
            partial[(size_t)wid * 8 + l] = acc; // combined by k1_fixup_kernel
        }
        return;
    }
    const K1Item b = items[wid];
    const uint32_t c0 = b.c, c1 = b.kbeg, s0 = b.kend, np = (b.b0 - s0) / 2;
    // phase 1: the block's slots as aligned pairs (s0 and the slot count are
    // multiples of 8), x gathered, both staged in LDS
    constexpr int U2 = K1_BLK / (2 * WAVE);
    if (np != 0) { // uniform
        d2v v[U2];
        u2v cc[U2];
#pragma unroll
        for (int j = 0; j < U2; j++) {
            uint32_t q = j * WAVE + lane;
            q = q < np ? q : np - 1; // branch-free: re-read the last pair
            v[j] = ntload(reinterpret_cast<const d2v *>(val + s0 + 2 * q));
            cc[j] = ntload(reinterpret_cast<const u2v *>(col + s0 + 2 * q));
        }
        double x0[U2], x1[U2];
#pragma unroll
        for (int j = 0; j < U2; j++) {
            x0[j] = x[cc[j][0]];
            x1[j] = x[cc[j][1]];
        }
#pragma unroll
        for (int j = 0; j < U2; j++) {
            const uint32_t q = j * WAVE + lane;
            if (q < np) {
                s_v[w][2 * q] = v[j][0];
                s_v[w][2 * q + 1] = v[j][1];
                s_x[w][2 * q] = x0[j];
                s_x[w][2 * q + 1] = x1[j];
            }
        }
    }
    wave_lds_sync();
    // phase 2: row l of chunk c is slots chp[c] + 8k + l, k < nmc[c]
    for (uint32_t r = lane; r < 8 * (c1 - c0); r += WAVE) {
        const uint32_t c = c0 + r / 8, lr = r % 8;
        const uint32_t a = chp[c], e = c + 1 < c1 ? chp[c + 1] : b.b0;
        const uint32_t o = a - s0 + lr, n = (e - a) / 8;
        double acc = 0.0;
        in_order_chunks<8, d2v>(0, n, [&](uint32_t k) { return d2v{s_v[w][o + 8 * k], s_x[w][o + 8 * k]}; },
                                [&](d2v t) { acc = fma(t[0], t[1], acc); });
        const uint32_t s = c * 8 + lr;
        if (!permi) y[s] = acc;
        else if (s < nrows) y[permi[s]] = acc;
    }
}

// Split K1 chunks: 8 threads per chunk add the pieces' lane partials in
// piece order and store like the main kernel.
__global__ void k1_fixup_kernel(const SplitDesc *__restrict__ splits, uint32_t nsplit,
                                const double *__restrict__ partial,
                                const uint32_t *__restrict__ permi, double *__restrict__ y,
                                uint32_t nrows)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, i = t / 8, l = t % 8;
    if (i >= nsplit) return;
    const SplitDesc sp = splits[i];
    double tot = 0.0;
    for (uint32_t j = 0; j < sp.k; j++) tot = tot + partial[(size_t)(sp.b0 + j) * 8 + l];
    uint32_t s = sp.row * 8 + l;
    if (!permi) y[s] = tot;
    else if (s < nrows) y[permi[s]] = tot;
}

__global__ void k1_order_kernel(uint32_t nrows, const uint32_t *__restrict__ permi,
                                const double *__restrict__ res, double *__restrict__ y)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrows) y[permi[i]] = res[i];
}

} // namespace ispv

using namespace ispv;

extern "C" int intspmv_axc_plan(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *brp_host,
                                intspmv_plan **plan)
{
    if (!ctx || !plan || !brp_host || brp_host[0] != 0) return INTSPMV_ERR_ARG;
    uint32_t *uptr = (uint32_t *)malloc(((size_t)nrows + 1) * sizeof(uint32_t));
    if (!uptr) return INTSPMV_ERR_NOMEM;
    for (uint32_t r = 0; r <= nrows; r++) {
        if (brp_host[r] % 16) {
            free(uptr);
            return INTSPMV_ERR_ARG;
        }
        uptr[r] = brp_host[r] / 16;
    }
    const bool sparse_rows = ispv_many_empty_rows(uptr, nrows);
    int st = ispv_plan_build(ctx, PLAN_AXC, nrows, uptr, sparse_rows ? AXC_UMAX : AXC_UMAX_DENSE,
                             AXC_RMAX, sparse_rows ? AXC_LONG : AXC_LONG_DENSE, 1024, plan);
    free(uptr);
    return st;
}

extern "C" int spmv_axc(intspmv_ctx *ctx, const intspmv_plan *plan, const double *ax,
                        const uint32_t *brp, double *y)
{
    if (!plan || plan->kind != PLAN_AXC || plan->ctx != ctx) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    if (plan->nblocks == 0) return INTSPMV_OK;
    if (!brp || !y || (plan->nunits && !ax)) return INTSPMV_ERR_ARG;
    axc_kernel<<<ispv_grid(plan), BLOCK_THREADS, 0, ctx->stream>>>(
        plan->d_blocks, plan->nlong_wg, plan->nblocks, ax, brp, y, plan->nrows, plan->d_partial);
    st = ispv_launch_check(ctx);
    if (st) return st;
    return ispv_split_fixup(plan, y, 0);
}

extern "C" int intspmv_k1_plan(intspmv_ctx *ctx, uint32_t nrows, uint32_t chunkNum,
                               const uint32_t *nmc_host, intspmv_plan **plan)
{
    if (!ctx || !plan || (chunkNum && !nmc_host)) return INTSPMV_ERR_ARG;
    if (chunkNum != (nrows + 7) / 8) return INTSPMV_ERR_ARG;
    *plan = nullptr;
    // chunks are in length order (longest first): long chunks become wide
    // items of at most K1_PIECE steps, the rest narrow items
    std::vector<K1Item> items;
    std::vector<SplitDesc> splits;
    uint32_t nwide = 0;
    for (uint32_t c = 0; c < chunkNum; c++) {
        uint32_t n = nmc_host[c];
        if (n <= K1_WIDE) continue;
        uint32_t np = (n + K1_PIECE - 1) / K1_PIECE, per = (n + np - 1) / np;
        uint32_t b0 = (uint32_t)items.size();
        for (uint32_t q = 0; q < np; q++) {
            uint32_t a = q * per, b = a + per > n ? n : a + per;
            items.push_back({c, a, b, b0, np, 0, 0, 0});
        }
        if (np > 1) splits.push_back({c, b0, np, 0});
    }
    nwide = (uint32_t)items.size();
    // runs of the other chunks in staged blocks of at most K1_BLK slots and
    // K1_BLK_CHUNKS chunks (chunk c starts at slot chp[c] = 8 * the steps of
    // the chunks before it; a narrow chunk, <= 8 * K1_WIDE slots, always fits)
    {
        uint32_t c = 0, slot = 0;
        while (c < chunkNum) {
            if (nmc_host[c] > K1_WIDE) {
                slot += 8 * nmc_host[c++];
                continue;
            }
            const uint32_t c0 = c, s0 = slot;
            while (c < chunkNum && c - c0 < K1_BLK_CHUNKS && nmc_host[c] <= K1_WIDE &&
                   slot + 8 * nmc_host[c] - s0 <= K1_BLK)
                slot += 8 * nmc_host[c++];
            items.push_back({c0, c, s0, slot, 0, 0, 0, 0});
        }
    }
    intspmv_plan *p = (intspmv_plan *)calloc(1, sizeof(intspmv_plan));
    if (!p) return INTSPMV_ERR_NOMEM;
    p->ctx = ctx;
    p->kind = PLAN_K1;
    p->nrows = nrows;
    p->chunkNum = chunkNum;
    p->nblocks = (uint32_t)items.size();
    p->nunits = 0; // the longest chunk's steps: 0 means val/col are never read
    for (uint32_t c = 0; c < chunkNum; c++) p->nunits = nmc_host[c] > p->nunits ? nmc_host[c] : p->nunits;
    p->nlong_wg = nwide; // number of wide (one-item) waves
    p->nsplit_rows = (uint32_t)splits.size();
    int st = ispv_plan_upload(ctx, p, items.data(), sizeof(K1Item), items.size(), 8,
                              splits.data(), splits.size());
    if (st) {
        intspmv_plan_destroy(p);
        return st;
    }
    *plan = p;
    return INTSPMV_OK;
}

extern "C" int spmv_k1(intspmv_ctx *ctx, const intspmv_plan *plan, const double *val,
                       const uint32_t *col, const uint32_t *chp, const uint32_t *nmc,
                       const uint32_t *permi, const double *x, double *y)
{
    (void)nmc; // lengths are carried by the plan's work items
    if (!plan || plan->kind != PLAN_K1 || plan->ctx != ctx) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    if (plan->nblocks == 0) return INTSPMV_OK;
    // a matrix without entries (every chunk 0 steps) reads no val/col; the
    // staged blocks read val and col in 16- and 8-byte pairs
    if (!chp || !x || !y || (plan->nunits && (!val || !col))) return INTSPMV_ERR_ARG;
    if ((uintptr_t)val % 16 != 0 || (uintptr_t)col % 8 != 0) return INTSPMV_ERR_ARG;
    const uint32_t nwide = plan->nlong_wg;
    dim3 grid((plan->nblocks + WAVES - 1) / WAVES); // one wave per item
    k1_kernel<<<grid, BLOCK_THREADS, 0, ctx->stream>>>(
        (const K1Item *)plan->d_blocks, plan->nblocks, nwide, val, col, chp, permi, x, y,
        plan->nrows, plan->d_partial);
    st = ispv_launch_check(ctx);
    if (st || plan->nsplit_rows == 0) return st;
    k1_fixup_kernel<<<(plan->nsplit_rows * 8 + 255) / 256, 256, 0, ctx->stream>>>(
        plan->d_splits, plan->nsplit_rows, plan->d_partial, permi, y, plan->nrows);
    return ispv_launch_check(ctx);
}

extern "C" int spmv_k1_order(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *permi,
                             const double *res, double *y)
{
    int st = ispv_enter(ctx);
    if (st) return st;
    if (nrows == 0) return INTSPMV_OK;
    if (!permi || !res || !y) return INTSPMV_ERR_ARG;
    k1_order_kernel<<<(nrows + 255) / 256, 256, 0, ctx->stream>>>(nrows, permi, res, y);
    return ispv_launch_check(ctx);
}
