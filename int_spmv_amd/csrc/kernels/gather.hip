// gather.hip -- AXT with a column map: values and columns in slot order, x as
// a separate (n x k, row-major) block of right-hand sides.
//
// The reference embeds x in the tiles (intSpmv_main.c:1171, 1207), so a new x
// needs a rebuild and several right-hand sides or A^T products cannot reuse
// the arrays.  The column-map form keeps the same tiles, slot order and sec
// words (getArraysAxSecAXT_NOC*, intSpmv_main.c:1152-1221) with
//   aval[s] = value of slot s, acol[s] = its column (INTSPMV_ACOL_PAD if the
//   slot is padding), slot s of tile T, brick-row R, lane C = T*th*8 + R*8 + C,
// and gathers x[acol[s]*ldx + j].  It serves
//   spmv_axt_gather      Y = A X for 1..32 right-hand sides,
//   the adjoint          X_bar = A^T Y_bar through the same kernel on the map
//                        of A^T (built once, intspmv_axt_transpose_map),
//   spmv_axt_grad_vals   A_bar[s] = y_bar[row(s)] * x[acol[s]] on the pattern,
//   spmv_axt_refresh_x   rewriting the embedded x of the reference layout on
//                        the device instead of rebuilding on the host.
#include "common.hpp"
#include "gather_block.hpp"
#include "reduce.hpp"

namespace ispv {

#ifndef ISPV_G_UG1
#define ISPV_G_UG1 8 // one right-hand side: 8 groups of 16 tiles, 16 gathers per lane in flight
#endif

template <typename VT, typename XT, typename AT, int KC, bool H1, bool XB, bool PV,
          bool GR = false>
__global__ __launch_bounds__(BLOCK_THREADS) void axt_gather_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const VT *__restrict__ aval, const uint32_t *__restrict__ acol,
    const uint32_t *__restrict__ sec, const XT *__restrict__ x, uint32_t ldx,
    AT *__restrict__ y, uint32_t ldy, uint32_t j0, uint32_t kv, uint32_t th, uint32_t nrows,
    double *partial, const uint32_t *__restrict__ vperm, const XT *__restrict__ xg = nullptr,
    VT *__restrict__ gvt = nullptr)
{
    __shared__ GatherLds<AT, KC> lds[WAVES];
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    GatherLds<AT, KC> &L = lds[threadIdx.x / WAVE];
    gather_block<VT, XT, AT, KC, H1, XB, ISPV_G_UG1, PV, 2, GR>(
        d, aval, acol, sec, x, ldx, y, ldy, j0, kv, th, nrows, partial, L.val, L.key, L.seg,
        vperm, xg, gvt);
}

// Wide passes (tile height 1) over the whole plan: the long-row blocks
// (workgroups [0, nlong_wg), one wave each, see gather_long_wide, UL tiles
// per group in flight) and the normal blocks (see gather_block_wide), in one
// launch as the narrow passes.
template <typename VT, typename XT, typename AT, int KW, bool XV, int UL>
__global__ __launch_bounds__(BLOCK_THREADS) void axt_gather_wide_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const VT *__restrict__ aval, const uint32_t *__restrict__ acol,
    const uint32_t *__restrict__ sec, const XT *__restrict__ x, uint32_t ldx,
    AT *__restrict__ y, uint32_t ldy, uint32_t j0, uint32_t kv, uint32_t nrows,
    double *partial)
{
    __shared__ uint32_t s_key[WAVES][G_UMAX];
    __shared__ uint16_t s_seg[WAVES][G_UMAX + 1];
    __shared__ WideLds<AT, KW, XT> s_wide[WAVES];
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    if (blockIdx.x < nlong_wg) {
        gather_long_wide<VT, XT, AT, KW, XV, UL>(d, aval, acol, x, ldx, y, ldy, j0, kv, nrows,
                                                 partial);
        return;
    }
    const uint32_t w = threadIdx.x / WAVE;
    gather_block_wide<VT, XT, AT, KW, XV>(d, aval, acol, sec, x, ldx, y, ldy, j0, kv, nrows,
                                          s_key[w], s_seg[w], s_wide[w]);
}

// A_bar[s] = y_bar[row(s)] * x[acol[s]] (0 on padding).  row(s) is the sec
// word of the slot's tile (th == 1) or column (th > 1).  Two slots per thread.
template <typename VT, typename XT, typename AT>
__global__ __launch_bounds__(256) void axt_grad_vals_kernel(
    uint32_t nslots, uint32_t th, const uint32_t *__restrict__ acol,
    const uint32_t *__restrict__ sec, const XT *__restrict__ x, const AT *__restrict__ ybar,
    VT *__restrict__ vbar)
{
    const size_t s = 2 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= nslots) return;
    const u2v cc = ntload(reinterpret_cast<const u2v *>(acol + s));
    const size_t T = s / (8 * th), C = s % 8;
    uint32_t r0, r1;
    if (th == 1) {
        r0 = r1 = sec[T];
    } else {
        const u2v kk = *reinterpret_cast<const u2v *>(sec + 8 * T + C);
        r0 = kk[0];
        r1 = kk[1];
    }
    AT g0 = cc[0] != PAD ? ybar[r0] * (AT)x[cc[0]] : (AT)0;
    AT g1 = cc[1] != PAD ? ybar[r1] * (AT)x[cc[1]] : (AT)0;
    if constexpr (sizeof(VT) == 8) {
        d2v o;
        o[0] = g0;
        o[1] = g1;
        __builtin_nontemporal_store(o, reinterpret_cast<d2v *>(vbar + s));
    } else {
        f2v o;
        o[0] = (float)g0;
        o[1] = (float)g1;
        __builtin_nontemporal_store(o, reinterpret_cast<f2v *>(vbar + s));
    }
}

// Embedded (reference) layout: A_bar[s] = y_bar[row(s)] * (x half of slot s).
__global__ __launch_bounds__(256) void axt_grad_vals_embedded_kernel(
    uint32_t nslots, uint32_t th, const double *__restrict__ ax, const uint32_t *__restrict__ sec,
    const double *__restrict__ ybar, double *__restrict__ vbar)
{
    const size_t s = 2 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= nslots) return;
    const size_t T = s / (8 * th), w = s % (8 * th), R = w / 8, C = w % 8;
    const d2v xx = ntload(reinterpret_cast<const d2v *>(ax + (T * th + R) * 16 + 8 + C));
    uint32_t r0, r1;
    if (th == 1) {
        r0 = r1 = sec[T];
    } else {
        const u2v kk = *reinterpret_cast<const u2v *>(sec + 8 * T + C);
        r0 = kk[0];
        r1 = kk[1];
    }
    d2v o;
    o[0] = ybar[r0] * xx[0]; // padding slots hold x = 0
    o[1] = ybar[r1] * xx[1];
    __builtin_nontemporal_store(o, reinterpret_cast<d2v *>(vbar + s));
}

// dst[s] = src[perm[s]] (0 where perm is PAD): moves values between the slot
// orders of A and A^T.
template <typename VT>
__global__ __launch_bounds__(256) void permute_vals_kernel(uint32_t n, const uint32_t *__restrict__ perm,
                                                           const VT *__restrict__ src,
                                                           VT *__restrict__ dst)
{
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t p = ntload(perm + s);
    dst[s] = p != PAD ? src[p] : (VT)0;
}

// Embedded x refresh: the x half of every slot becomes x[acol[s]], 0 on
// padding (as the builder leaves it).  Two slots per thread: one 8-byte
// column load and one 16-byte store (slot counts are multiples of 8).
__global__ __launch_bounds__(256) void refresh_x_kernel(uint32_t nslots, uint32_t th,
                                                        const uint32_t *__restrict__ acol,
                                                        const double *__restrict__ x,
                                                        double *__restrict__ ax)
{
    const size_t s = 2 * ((size_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= nslots) return;
    const u2v c = ntload(reinterpret_cast<const u2v *>(acol + s));
    const bool ok0 = c[0] != PAD, ok1 = c[1] != PAD;
    const double x0 = x[ok0 ? c[0] : 0], x1 = x[ok1 ? c[1] : 0]; // row 0 read for padding
    d2v o;
    o[0] = ok0 ? x0 : 0.0;
    o[1] = ok1 ? x1 : 0.0;
    const size_t T = s / (8 * th), w = s % (8 * th);
    *reinterpret_cast<d2v *>(ax + (T * th + w / 8) * 16 + 8 + w % 8) = o;
}

// Right-hand sides per wide pass for the rest of k (0: take narrow passes).
template <typename XT>
__host__ __device__ constexpr uint32_t wide_width(uint32_t rem)
{
    if (rem < wide_min_kc<XT>() || wide_min_kc<XT>() > 32) return 0;
    uint32_t kw = wide_min_kc<XT>();
    while (kw < rem && kw < 32) kw *= 2;
    return kw;
}

// Narrow passes of kc right-hand sides from j0 over blocks [b0, b0 + nb) of
// the plan's list (the long section is the first nlong_wg * WAVES blocks).
template <typename VT, typename XT, typename AT, bool PV>
static void launch_narrow(const intspmv_plan *p, const BlockDesc *blocks, uint32_t nlong_wg,
                          uint32_t nb, dim3 grid, const void *aval, const uint32_t *acol,
                          const uint32_t *sec, const void *x, uint32_t ldx, void *y,
                          uint32_t ldy, uint32_t j0, uint32_t kc, uint32_t kv,
                          const uint32_t *vperm)
{
    const uint32_t th = p->axt.tileH;
    hipStream_t s = p->ctx->stream;
    dim3 blk(BLOCK_THREADS);
    // buffer-resource gathers for passes of fewer right-hand sides than one
    // 16-byte load holds (with more, the plain-load kernel measured faster)
    const bool xb = xbuf_fits(p->nrows, ldx, sizeof(XT)) && kc < 16 / sizeof(XT);
#define ISPV_GX(KC, H1, XB)                                                                     \
    axt_gather_kernel<VT, XT, AT, KC, H1, XB, PV><<<grid, blk, 0, s>>>(                         \
        blocks, nlong_wg, nb, (const VT *)aval, acol, sec, (const XT *)x, ldx, (AT *)y, ldy, j0, \
        kv, th, p->nrows, p->d_partial, vperm)
#define ISPV_G(KC)                                                                              \
    if (kc == KC) {                                                                             \
        if constexpr (KC < 16 / sizeof(XT)) {                                                   \
            if (th == 1) {                                                                      \
                if (xb) ISPV_GX(KC, true, true);                                                \
                else ISPV_GX(KC, true, false);                                                  \
            } else {                                                                            \
                if (xb) ISPV_GX(KC, false, true);                                               \
                else ISPV_GX(KC, false, false);                                                 \
            }                                                                                   \
        } else {                                                                                \
            if (th == 1) ISPV_GX(KC, true, false);                                              \
            else ISPV_GX(KC, false, false);                                                     \
        }                                                                                       \
    }
    ISPV_G(1)
    ISPV_G(2)
    ISPV_G(4)
    ISPV_G(8)
    if constexpr (sizeof(AT) == 4) {
        ISPV_G(16)
    }
#undef ISPV_G
#undef ISPV_GX
}

template <typename VT, typename XT, typename AT, bool PV = false>
static int launch_gather(const intspmv_plan *p, const void *aval, const uint32_t *acol,
                         const uint32_t *sec, const void *x, uint32_t ldx, void *y, uint32_t ldy,
                         uint32_t k, uint32_t kcf, const uint32_t *vperm = nullptr)
{
    const uint32_t nb = p->nblocks, th = p->axt.tileH;
    if (nb == 0) return INTSPMV_OK;
    hipStream_t s = p->ctx->stream;
    const uint32_t nl = p->nlong_wg; // long section: blocks [0, nl * WAVES)
    constexpr uint32_t V = 16 / sizeof(XT);
    const bool wide_ok = th == 1 && !PV;
    // 16-byte loads of x rows: rows and pass starts aligned to V values
    const bool xv = ldx % V == 0 && (uintptr_t)x % 16 == 0;
    for (uint32_t j0 = 0; j0 < k;) {
        const uint32_t rem = k - j0;
        uint32_t kc = kcf;
        if (!kc) {
            kc = wide_ok ? wide_width<XT>(rem) : 0;
            if (!kc)
                kc = (rem >= 16 && sizeof(AT) == 4) ? 16
                   : (rem >= 8 ? 8 : (rem >= 4 ? 4 : (rem >= 2 ? 2 : 1)));
        }
        const uint32_t kv = rem < kc ? rem : kc;
        if (wide_ok && kc >= wide_min_kc<XT>()) {
            // one wide pass over the long rows and the normal blocks.  FP32 /
            // bf16: a plan without long rows takes the kernel with one long-row
            // tile per group in flight, whose registers leave more waves per
            // SIMD (the long-row path sets them: FP32 stencil 128^3 k = 32
            // 617 us with one against 647 us with two); FP64 needs as many
            // registers for the normal blocks, so it always takes two.
            dim3 grid(ispv_grid(p)), blk(BLOCK_THREADS);
#define ISPV_GWU(KW, XVB, ULV)                                                                  \
    axt_gather_wide_kernel<VT, XT, AT, KW, XVB, ULV><<<grid, blk, 0, s>>>(                      \
        p->d_blocks, nl, nb, (const VT *)aval, acol, sec, (const XT *)x, ldx, (AT *)y, ldy, j0, \
        kv, p->nrows, p->d_partial)
#define ISPV_GW(KW)                                                                             \
    if (kc == KW) {                                                                             \
        if constexpr (sizeof(XT) == 8) {                                                        \
            if (xv) ISPV_GWU(KW, true, wide_long_u<XT>());                                      \
            else ISPV_GWU(KW, false, wide_long_u<XT>());                                        \
        } else if (nl) {                                                                        \
            if (xv) ISPV_GWU(KW, true, wide_long_u<XT>());                                      \
            else ISPV_GWU(KW, false, wide_long_u<XT>());                                        \
        } else {                                                                                \
            if (xv) ISPV_GWU(KW, true, 1);                                                      \
            else ISPV_GWU(KW, false, 1);                                                        \
        }                                                                                       \
    }
            if constexpr (V == 2) {
                ISPV_GW(4)
            }
            ISPV_GW(8)
            ISPV_GW(16)
            ISPV_GW(32)
#undef ISPV_GW
#undef ISPV_GWU
            int st = ispv_launch_check(p->ctx);
            if (st) return st;
        } else {
            launch_narrow<VT, XT, AT, PV>(p, p->d_blocks, nl, nb, dim3(ispv_grid(p)), aval, acol,
                                          sec, x, ldx, y, ldy, j0, kc, kv, vperm);
            int st = ispv_launch_check(p->ctx);
            if (st) return st;
        }
        j0 += kc;
    }
    return ispv_split_fixup(p, y, sizeof(AT) == 4, k, ldy, 0);
}

// Slots of the tile columns no block covers -- the unoccupied columns
// [colN, tileN * 8) at the end of the last tile (th > 1) -- are padding: 0.
template <typename VT>
__global__ __launch_bounds__(64) void tail_columns_zero_kernel(VT *__restrict__ g, uint32_t tileN,
                                                               uint32_t th, uint32_t colN)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x, c = colN + i / th, R = i % th;
    if (c >= tileN * 8) return;
    g[((size_t)(c / 8) * th + R) * 8 + c % 8] = (VT)0;
}

// X_bar = A^T Y_bar (one right-hand side) with A's value gradient in this
// map's slot order written in the same pass (GR).
template <typename VT, typename XT, typename AT>
static int launch_gather_grad(const intspmv_plan *p, const void *aval, const uint32_t *acol,
                              const uint32_t *sec, const void *ybar, void *xbar, const void *x,
                              void *gvt)
{
    const uint32_t nb = p->nblocks, th = p->axt.tileH;
    if (nb == 0) return INTSPMV_OK;
    dim3 grid(ispv_grid(p)), blk(BLOCK_THREADS);
    hipStream_t s = p->ctx->stream;
    const bool xb = xbuf_fits(p->nrows, 1, sizeof(XT));
#define ISPV_GG(H1, XB)                                                                         \
    axt_gather_kernel<VT, XT, AT, 1, H1, XB, false, true><<<grid, blk, 0, s>>>(                 \
        p->d_blocks, p->nlong_wg, nb, (const VT *)aval, acol, sec, (const XT *)ybar, 1,          \
        (AT *)xbar, 1, 0, 1, th, p->nrows, p->d_partial, nullptr, (const XT *)x, (VT *)gvt)
    if (th == 1) {
        if (xb) ISPV_GG(true, true);
        else ISPV_GG(true, false);
    } else {
        if (xb) ISPV_GG(false, true);
        else ISPV_GG(false, false);
    }
#undef ISPV_GG
    int st = ispv_launch_check(p->ctx);
    if (st) return st;
    const uint32_t colN = p->axt.colN, cols = p->axt.tileN * 8;
    if (th > 1 && colN < cols) {
        tail_columns_zero_kernel<VT><<<((cols - colN) * th + 63) / 64, 64, 0, s>>>(
            (VT *)gvt, p->axt.tileN, th, colN);
        st = ispv_launch_check(p->ctx);
        if (st) return st;
    }
    return ispv_split_fixup(p, xbar, sizeof(AT) == 4, 1, 1, 0);
}

} // namespace ispv

using namespace ispv;

static bool gather_plan_ok(const intspmv_ctx *ctx, const intspmv_plan *plan)
{
    return ctx && plan && plan->kind == PLAN_AXT && plan->ctx == ctx &&
           plan->axt.mode == INTSPMV_NOC;
}

// kc: 0 (automatic), 1, 2, 4, 8 (narrow passes), 16 (narrow for FP32 / BF16
// values, or wide), 32 (wide); wide passes need tile height 1 and values read
// directly (not through a permutation), and at least two lanes per x row.
static bool gather_kc_ok(const intspmv_plan *plan, uint32_t kc, bool permuted)
{
    if (kc == 0 || kc == 1 || kc == 2 || kc == 4 || kc == 8) return true;
    const bool f64x = plan->axt.x_type == INTSPMV_F64;
    const uint32_t wmin = f64x ? wide_min_kc<double>() : wide_min_kc<float>();
    const bool wide = plan->axt.tileH == 1 && !permuted && kc >= wmin;
    if (kc == 16) return wide || !f64x;
    return kc == 32 && wide;
}

extern "C" int spmv_axt_gather(intspmv_ctx *ctx, const intspmv_plan *plan, const void *aval,
                               const uint32_t *acol, const uint32_t *sec, const void *x,
                               uint32_t ldx, void *y, uint32_t ldy, uint32_t k, uint32_t kc)
{
    // the column-map blocks stage at most G_UMAX units: a plan built for the
    // reference-layout kernels with larger blocks is refused
    if (!gather_plan_ok(ctx, plan) || plan->umax > G_UMAX || k == 0 || k > PARTIAL_STRIDE ||
        ldx < k || ldy < k || !gather_kc_ok(plan, kc, false))
        return INTSPMV_ERR_ARG;
    if (plan->nrows && (!x || !y || (plan->nunits && (!aval || !acol || !sec))))
        return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    const intspmv_axt_desc &D = plan->axt;
    if (D.val_type == INTSPMV_F64 && D.x_type == INTSPMV_F64)
        return launch_gather<double, double, double>(plan, aval, acol, sec, x, ldx, y, ldy, k, kc);
    if (D.val_type == INTSPMV_F32 && D.x_type == INTSPMV_F32)
        return launch_gather<float, float, float>(plan, aval, acol, sec, x, ldx, y, ldy, k, kc);
    if (D.val_type == INTSPMV_BF16 && D.x_type == INTSPMV_F32)
        return launch_gather<uint16_t, float, float>(plan, aval, acol, sec, x, ldx, y, ldy, k, kc);
    return INTSPMV_ERR_UNSUPPORTED;
}

// Times `reps` products of k right-hand sides under a plan (after one
// untimed), on the context's stream; microseconds per product in *us.
static int time_gather(intspmv_ctx *ctx, const intspmv_plan *plan, const void *aval,
                       const uint32_t *acol, const uint32_t *sec, const void *x, void *y,
                       uint32_t k, uint32_t reps, float *us)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    int st = e == hipSuccess ? spmv_axt_gather(ctx, plan, aval, acol, sec, x, k, y, k, k, 0)
                             : INTSPMV_ERR_HIP;
    if (!st && hipEventRecord(e0, ctx->stream) != hipSuccess) st = INTSPMV_ERR_HIP;
    for (uint32_t r = 0; r < reps && !st; r++)
        st = spmv_axt_gather(ctx, plan, aval, acol, sec, x, k, y, k, k, 0);
    float ms = 0.0f;
    if (!st && (hipEventRecord(e1, ctx->stream) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = INTSPMV_ERR_HIP;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!st) *us = ms * 1e3f / (float)reps;
    return st;
}

extern "C" int intspmv_axt_plan_tuned(intspmv_ctx *ctx, const intspmv_axt_desc *desc,
                                      const uint32_t *con_host, const void *aval,
                                      const uint32_t *acol, const uint32_t *sec, uint32_t k,
                                      uint32_t reps, intspmv_plan **plan, int *grouped,
                                      float *us_plain, float *us_grouped)
{
    if (!plan || !desc || !ctx || k == 0 || k > PARTIAL_STRIDE || reps == 0 ||
        desc->layout != INTSPMV_AXT_COLMAP || desc->mode != INTSPMV_NOC)
        return INTSPMV_ERR_ARG;
    *plan = nullptr;
    intspmv_plan *pp = nullptr, *pg = nullptr;
    int st = intspmv_axt_plan(ctx, desc, con_host, &pp);
    if (!st) st = intspmv_axt_plan_grouped(ctx, desc, con_host, NUM_XCD, nullptr, &pg);
    if (st) {
        intspmv_plan_destroy(pp);
        return st;
    }
    // scratch X (zeros) and Y, n x k of the value type's accumulation type
    const size_t esz = desc->x_type == INTSPMV_F64 ? 8 : 4;
    const size_t bytes = ((size_t)desc->nrows * k + 1) * esz;
    void *x = nullptr, *y = nullptr;
    hipError_t e = hipMalloc(&x, bytes);
    if (e == hipSuccess) e = hipMalloc(&y, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(x, 0, bytes, ctx->stream);
    float tp = 0.0f, tg = 0.0f;
    st = e == hipSuccess ? INTSPMV_OK : INTSPMV_ERR_NOMEM;
    if (!st) st = time_gather(ctx, pp, aval, acol, sec, x, y, k, reps, &tp);
    if (!st) st = time_gather(ctx, pg, aval, acol, sec, x, y, k, reps, &tg);
    (void)hipStreamSynchronize(ctx->stream);
    if (x) (void)hipFree(x);
    if (y) (void)hipFree(y);
    if (st) {
        intspmv_plan_destroy(pp);
        intspmv_plan_destroy(pg);
        return st;
    }
    const bool g = tg < tp;
    intspmv_plan_destroy(g ? pp : pg);
    *plan = g ? pg : pp;
    if (grouped) *grouped = g;
    if (us_plain) *us_plain = tp;
    if (us_grouped) *us_grouped = tg;
    return INTSPMV_OK;
}

extern "C" int spmv_axt_gather_permuted(intspmv_ctx *ctx, const intspmv_plan *plan,
                                        const void *vals, const uint32_t *vperm,
                                        const uint32_t *acol, const uint32_t *sec, const void *x,
                                        uint32_t ldx, void *y, uint32_t ldy, uint32_t k,
                                        uint32_t kc)
{
    if (!gather_plan_ok(ctx, plan) || plan->umax > G_UMAX || k == 0 || k > PARTIAL_STRIDE ||
        ldx < k || ldy < k || !gather_kc_ok(plan, kc, true))
        return INTSPMV_ERR_ARG;
    if (plan->nrows && (!x || !y || (plan->nunits && (!vals || !vperm || !acol || !sec))))
        return INTSPMV_ERR_ARG;
    if (plan->nunits && (uintptr_t)vperm % 8 != 0) return INTSPMV_ERR_ARG; // pairs of indexes
    int st = ispv_enter(ctx);
    if (st) return st;
    const intspmv_axt_desc &D = plan->axt;
    if (D.val_type == INTSPMV_F64 && D.x_type == INTSPMV_F64)
        return launch_gather<double, double, double, true>(plan, vals, acol, sec, x, ldx, y, ldy,
                                                           k, kc, vperm);
    if (D.val_type == INTSPMV_F32 && D.x_type == INTSPMV_F32)
        return launch_gather<float, float, float, true>(plan, vals, acol, sec, x, ldx, y, ldy, k,
                                                        kc, vperm);
    return INTSPMV_ERR_UNSUPPORTED;
}

extern "C" int spmv_axt_gather_grad(intspmv_ctx *ctx, const intspmv_plan *plan,
                                    const void *aval, const uint32_t *acol, const uint32_t *sec,
                                    const void *ybar, void *xbar, const void *x, void *vbar)
{
    if (!gather_plan_ok(ctx, plan) || plan->umax > G_UMAX) return INTSPMV_ERR_ARG;
    if (plan->nrows && (!ybar || !xbar || !x || (plan->nunits && (!aval || !acol || !sec || !vbar))))
        return INTSPMV_ERR_ARG;
    if (plan->nunits && ((uintptr_t)vbar % (plan->axt.val_type == INTSPMV_F64 ? 16 : 8)))
        return INTSPMV_ERR_ARG; // pairs of slots are stored together
    int st = ispv_enter(ctx);
    if (st) return st;
    const intspmv_axt_desc &D = plan->axt;
    if (D.val_type == INTSPMV_F64 && D.x_type == INTSPMV_F64)
        return launch_gather_grad<double, double, double>(plan, aval, acol, sec, ybar, xbar, x, vbar);
    if (D.val_type == INTSPMV_F32 && D.x_type == INTSPMV_F32)
        return launch_gather_grad<float, float, float>(plan, aval, acol, sec, ybar, xbar, x, vbar);
    return INTSPMV_ERR_UNSUPPORTED;
}

extern "C" int spmv_axt_grad_vals(intspmv_ctx *ctx, const intspmv_plan *plan,
                                  const uint32_t *acol, const uint32_t *sec, const void *x,
                                  const void *ybar, void *vbar)
{
    if (!gather_plan_ok(ctx, plan)) return INTSPMV_ERR_ARG;
    const intspmv_axt_desc &D = plan->axt;
    const uint32_t nslots = D.lenAX / 2;
    if (nslots == 0) return INTSPMV_OK;
    if (!acol || !sec || !x || !ybar || !vbar) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    dim3 grid((nslots / 2 + 255) / 256);
    if (D.val_type == INTSPMV_F64 && D.x_type == INTSPMV_F64)
        axt_grad_vals_kernel<double, double, double><<<grid, 256, 0, ctx->stream>>>(
            nslots, D.tileH, acol, sec, (const double *)x, (const double *)ybar, (double *)vbar);
    else if (D.val_type == INTSPMV_F32 && D.x_type == INTSPMV_F32)
        axt_grad_vals_kernel<float, float, float><<<grid, 256, 0, ctx->stream>>>(
            nslots, D.tileH, acol, sec, (const float *)x, (const float *)ybar, (float *)vbar);
    else
        return INTSPMV_ERR_UNSUPPORTED;
    return ispv_launch_check(ctx);
}

extern "C" int spmv_axt_grad_vals_embedded(intspmv_ctx *ctx, const intspmv_plan *plan,
                                           const double *ax, const uint32_t *sec,
                                           const double *ybar, double *vbar)
{
    if (!gather_plan_ok(ctx, plan)) return INTSPMV_ERR_ARG;
    const intspmv_axt_desc &D = plan->axt;
    if (D.val_type != INTSPMV_F64 || D.layout != INTSPMV_AXT_INTERLEAVED)
        return INTSPMV_ERR_UNSUPPORTED;
    const uint32_t nslots = D.lenAX / 2;
    if (nslots == 0) return INTSPMV_OK;
    if (!ax || !sec || !ybar || !vbar) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    axt_grad_vals_embedded_kernel<<<(nslots / 2 + 255) / 256, 256, 0, ctx->stream>>>(
        nslots, D.tileH, ax, sec, ybar, vbar);
    return ispv_launch_check(ctx);
}

extern "C" int spmv_permute_vals(intspmv_ctx *ctx, uint32_t n, int val_type,
                                 const uint32_t *perm, const void *src, void *dst)
{
    int st = ispv_enter(ctx);
    if (st) return st;
    if (n == 0) return INTSPMV_OK;
    if (!perm || !src || !dst) return INTSPMV_ERR_ARG;
    dim3 grid((n + 255) / 256);
    if (val_type == INTSPMV_F64)
        permute_vals_kernel<double><<<grid, 256, 0, ctx->stream>>>(n, perm, (const double *)src,
                                                                   (double *)dst);
    else if (val_type == INTSPMV_F32)
        permute_vals_kernel<float><<<grid, 256, 0, ctx->stream>>>(n, perm, (const float *)src,
                                                                  (float *)dst);
    else if (val_type == INTSPMV_BF16)
        permute_vals_kernel<uint16_t><<<grid, 256, 0, ctx->stream>>>(
            n, perm, (const uint16_t *)src, (uint16_t *)dst);
    else
        return INTSPMV_ERR_ARG;
    return ispv_launch_check(ctx);
}

extern "C" int spmv_axt_refresh_x(intspmv_ctx *ctx, const intspmv_plan *plan,
                                  const uint32_t *acol, const double *x, double *ax)
{
    if (!ctx || !plan || plan->kind != PLAN_AXT || plan->ctx != ctx) return INTSPMV_ERR_ARG;
    const intspmv_axt_desc &D = plan->axt;
    if (D.val_type != INTSPMV_F64 || D.x_type != INTSPMV_F64 ||
        D.layout != INTSPMV_AXT_INTERLEAVED)
        return INTSPMV_ERR_UNSUPPORTED;
    const uint32_t nslots = D.lenAX / 2;
    if (nslots == 0) return INTSPMV_OK;
    if (!acol || !x || !ax) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    if (nslots % 8 || (uintptr_t)acol % 8 || (uintptr_t)ax % 16) return INTSPMV_ERR_ARG;
    refresh_x_kernel<<<(nslots / 2 + 255) / 256, 256, 0, ctx->stream>>>(nslots, D.tileH, acol, x, ax);
    return ispv_launch_check(ctx);
}
