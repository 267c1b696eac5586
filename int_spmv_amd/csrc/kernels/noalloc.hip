// noalloc.hip -- spmv_axt_h1_noalloc: the reference-layout product for tile
// height 1 (int_axt_noc_h1, intSpmv_main.c:1385-1404) with no plan and no
// allocation, for callers that cannot allocate device memory.
//
// Work is cut without looking at the matrix: wave c takes tiles
// [64c, 64c + 64) (8 KB of FP64 tiles, the size of the plan kernel's blocks),
// forms each tile's reduce tree as axt_h1_kernel does, and adds the tile sums
// of each row (a run of equal sec words) in tile order from +0.0.  A row that
// lies inside its chunk is written directly.  A row that crosses chunk
// boundaries leaves one partial per chunk in the scratch array the caller's
// context points at (abi.hip): part[2c] for the chunk's first row when it
// continues from chunk c-1, part[2c+1] for its last row when it starts in
// chunk c and continues into c+1.  A second kernel walks each such row from
// its first chunk and adds the partials in chunk order.  Rows without tiles
// are zeroed by the chunk that sees the gap before the next row (or after the
// last row, in the last chunk).
#include "common.hpp"
#include "reduce.hpp"

namespace ispv {

constexpr uint32_t NA_CHUNK = 64; // tiles per wave
constexpr uint32_t NA_NONE = 0xFFFFFFFFu;

__global__ __launch_bounds__(BLOCK_THREADS) void axt_h1_noalloc_kernel(
    const double *__restrict__ ax, const uint32_t *__restrict__ sec, double *__restrict__ y,
    uint32_t nrows, uint32_t tileN, uint32_t nchunks, double *__restrict__ part)
{
    __shared__ double s_val[WAVES][NA_CHUNK];
    __shared__ uint32_t s_key[WAVES][NA_CHUNK];
    __shared__ uint16_t s_seg[WAVES][NA_CHUNK + 1];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t c = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * WAVES + w);
    if (c >= nchunks) return;
    const uint32_t T0 = c * NA_CHUNK, left = tileN - T0;
    const uint32_t nu = left < NA_CHUNK ? left : NA_CHUNK;
    if ((uint32_t)lane < nu) s_key[w][lane] = ntload(sec + T0 + lane);
    // 8 loads of 8 whole 128-byte tiles; lane part p = l%8: values 2p, 2p+1
    // (p < 4) or the x pair 2(p-4), 2(p-4)+1 (p >= 4), brought to the value
    // lanes by a DPP row shift; the quads then form the tree
    // ((p0+p4)+(p2+p6)) + ((p1+p5)+(p3+p7)).
    const uint32_t p = lane & 7;
    d2v v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint32_t t = j * 8 + lane / 8;
        const size_t T = (size_t)T0 + (t < nu ? t : nu - 1);
        v[j] = ntload(reinterpret_cast<const d2v *>(ax + T * 16 + 2 * p));
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint32_t t = j * 8 + lane / 8;
        const bool ok = t < nu && p < 4;
        double x0 = dpp_f64<DPP_ROW_SHL4>(v[j][0]), x1 = dpp_f64<DPP_ROW_SHL4>(v[j][1]);
        double p0 = ok ? v[j][0] * x0 : 0.0, p1 = ok ? v[j][1] * x1 : 0.0;
        p0 = p0 + dpp_f64<DPP_QUAD_XOR2>(p0);
        p1 = p1 + dpp_f64<DPP_QUAD_XOR2>(p1);
        p0 = p0 + dpp_f64<DPP_QUAD_XOR1>(p0);
        p1 = p1 + dpp_f64<DPP_QUAD_XOR1>(p1);
        if (t < nu && p == 0) s_val[w][t] = p0 + p1;
    }
    wave_lds_sync();
    const uint32_t nseg = find_segments(s_key[w], s_seg[w], nu); // >= 1: nu >= 1
    const uint32_t prev = T0 ? sec[T0 - 1] : NA_NONE;
    const uint32_t next = T0 + nu < tileN ? sec[T0 + nu] : NA_NONE;
    for (uint32_t s = lane; s < nseg; s += WAVE) {
        const uint32_t a = s_seg[w][s], b = s_seg[w][s + 1], r = s_key[w][a];
        double acc = 0.0;
        in_order_chunks<8, double>(a, b, [&](uint32_t i) { return s_val[w][i]; },
                                   [&](double t) { acc = acc + t; });
        if (s == 0 && prev == r) part[2 * (size_t)c] = acc;
        else if (s == nseg - 1 && next == r) part[2 * (size_t)c + 1] = acc;
        else if (r < nrows) y[r] = acc;
    }
    // rows without tiles: the gap before each run's row, and in the last chunk
    // the rows after the last run
    for (uint32_t s = 0; s <= nseg; s++) {
        uint32_t lo, hi;
        if (s < nseg) {
            const uint32_t r = s_key[w][s_seg[w][s]];
            lo = s ? s_key[w][s_seg[w][s - 1]] + 1 : (T0 ? prev + 1 : 0);
            hi = r < nrows ? r : nrows;
        } else {
            if (c + 1 != nchunks) break;
            lo = s_key[w][s_seg[w][nseg - 1]] + 1;
            hi = nrows;
        }
        for (uint32_t z = lo + lane; z < hi; z += WAVE) y[z] = 0.0;
    }
}

// One thread per chunk s whose last row starts in it and continues into
// chunk s+1: adds that row's partials in chunk order and writes the row.
__global__ __launch_bounds__(256) void axt_h1_noalloc_fixup(const uint32_t *__restrict__ sec,
                                                            double *__restrict__ y, uint32_t nrows,
                                                            uint32_t tileN, uint32_t nchunks,
                                                            const double *__restrict__ part)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s + 1 >= nchunks) return; // the last chunk continues nowhere
    const uint32_t first = s * NA_CHUNK, last = first + NA_CHUNK - 1; // chunk s is full
    const uint32_t r = sec[last];
    if (sec[last + 1] != r) return;                     // its last row ends in it
    if (s > 0 && sec[first] == r && sec[first - 1] == r) return; // began in an earlier chunk
    double tot = part[2 * (size_t)s + 1];
    for (uint32_t c = s + 1; c < nchunks; c++) {
        tot = tot + part[2 * (size_t)c];
        const uint32_t lc = c * NA_CHUNK + NA_CHUNK - 1 < tileN ? c * NA_CHUNK + NA_CHUNK - 1 : tileN - 1;
        if (sec[lc] != r || c + 1 >= nchunks || sec[lc + 1] != r) break;
    }
    if (r < nrows) y[r] = tot;
}

__global__ __launch_bounds__(256) void fill_zero_kernel(double *__restrict__ y, uint32_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = 0.0;
}

} // namespace ispv

using namespace ispv;

extern "C" int spmv_axt_h1_noalloc(intspmv_ctx *ctx, uint32_t nrows, uint32_t tileN,
                                   const double *ax, const uint32_t *sec, double *y)
{
    if (!ctx) return INTSPMV_ERR_ARG;
    if ((nrows && !y) || (tileN && (!ax || !sec)) || (tileN && !nrows)) return INTSPMV_ERR_ARG;
    if (tileN > INTSPMV_NOALLOC_MAX_TILES) return INTSPMV_ERR_TOO_LARGE;
    if (nrows == 0) return INTSPMV_OK;
    int st = ispv_enter(ctx);
    if (st) return st;
    if (!ctx->noalloc_scratch) return INTSPMV_ERR_HIP;
    hipStream_t s = ctx->stream;
    if (tileN == 0) {
        fill_zero_kernel<<<(nrows + 255) / 256, 256, 0, s>>>(y, nrows);
        return ispv_launch_check(ctx);
    }
    const uint32_t nchunks = (tileN + NA_CHUNK - 1) / NA_CHUNK; // <= scratch / 2 doubles
    axt_h1_noalloc_kernel<<<(nchunks + WAVES - 1) / WAVES, BLOCK_THREADS, 0, s>>>(
        ax, sec, y, nrows, tileN, nchunks, ctx->noalloc_scratch);
    st = ispv_launch_check(ctx);
    if (st || nchunks < 2) return st;
    axt_h1_noalloc_fixup<<<(nchunks + 255) / 256, 256, 0, s>>>(sec, y, nrows, tileN, nchunks,
                                                             ctx->noalloc_scratch);
    return ispv_launch_check(ctx);
}
