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
#include "reduce.hpp"

namespace ispv {

constexpr uint32_t G_UMAX = ISPV_AXT_UMAX; // units per normal block (shared AXT plan)
constexpr uint32_t PAD = 0xFFFFFFFFu;

template <typename T>
struct GPair;
template <>
struct GPair<double> {
    __device__ __forceinline__ static void ld(const double *p, double &a, double &b)
    {
        d2v v = ntload(reinterpret_cast<const d2v *>(p));
        a = v[0];
        b = v[1];
    }
};
template <>
struct GPair<float> {
    __device__ __forceinline__ static void ld(const float *p, float &a, float &b)
    {
        f2v v = ntload(reinterpret_cast<const f2v *>(p));
        a = v[0];
        b = v[1];
    }
};
template <>
struct GPair<uint16_t> {
    __device__ __forceinline__ static void ld(const uint16_t *p, float &a, float &b)
    {
        uint32_t v = ntload(reinterpret_cast<const uint32_t *>(p));
        a = __uint_as_float(v << 16);
        b = __uint_as_float(v & 0xFFFF0000u);
    }
};
template <typename VT>
using GLoad = typename std::conditional<std::is_same<VT, uint16_t>::value, float, VT>::type;

template <typename AT>
__device__ __forceinline__ AT gfma(AT a, AT b, AT c);
template <>
__device__ __forceinline__ double gfma<double>(double a, double b, double c) { return fma(a, b, c); }
template <>
__device__ __forceinline__ float gfma<float>(float a, float b, float c) { return fmaf(a, b, c); }

// x values of one slot for KC right-hand sides starting at j0 (0 for padding)
template <typename XT, typename AT, int KC>
__device__ __forceinline__ void gather_x(const XT *x, uint32_t c, uint32_t ldx, uint32_t j0,
                                         uint32_t kv, AT (&xv)[KC])
{
#pragma unroll
    for (int j = 0; j < KC; j++)
        xv[j] = (c != PAD && (uint32_t)j < kv) ? (AT)x[(size_t)c * ldx + j0 + j] : (AT)0;
}

// One pass over KC right-hand sides.  th == 1: four lanes per tile, the
// reference's reduce tree per right-hand side; th > 1: two columns per lane,
// fma chains over the brick-rows.  Normal blocks stage unit sums in LDS and
// give each row to one lane; long rows keep per-lane sums.
template <typename VT, typename XT, typename AT, int KC, bool H1>
__global__ __launch_bounds__(BLOCK_THREADS) void axt_gather_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const VT *__restrict__ aval, const uint32_t *__restrict__ acol,
    const uint32_t *__restrict__ sec, const XT *__restrict__ x, uint32_t ldx,
    AT *__restrict__ y, uint32_t ldy, uint32_t j0, uint32_t kv, uint32_t th, uint32_t nrows,
    double *partial)
{
    __shared__ AT s_val[WAVES][G_UMAX * KC];
    __shared__ uint32_t s_key[WAVES][G_UMAX];
    __shared__ uint16_t s_seg[WAVES][G_UMAX + 1];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    const uint32_t nu = d.uend - d.ubeg, q = lane & 3, C = 2 * q;
    AT lacc[KC];
#pragma unroll
    for (int j = 0; j < KC; j++) lacc[j] = 0;
    // tiles of this block; UG steps of 16 tiles are loaded before their
    // gathers are issued, so a lane keeps 2*UG gathers in flight
    constexpr int UG = KC >= 8 ? 1 : (KC >= 4 ? 2 : 4);
    const uint32_t T0 = H1 ? d.ubeg : d.ubeg / 8, T1 = H1 ? d.uend : (d.uend + 7) / 8;
    for (uint32_t Tbb = T0; Tbb < T1; Tbb += 16 * UG) {
        AT a0[UG][KC], a1[UG][KC];
#pragma unroll
        for (int u = 0; u < UG; u++)
#pragma unroll
            for (int j = 0; j < KC; j++) a0[u][j] = a1[u][j] = 0;
        for (uint32_t R = 0; R < th; R++) {
            GLoad<VT> v0[UG], v1[UG];
            u2v cc[UG];
#pragma unroll
            for (int u = 0; u < UG; u++) {
                const uint32_t Tr = Tbb + 16 * u + lane / 4;
                const size_t T = Tr < T1 ? Tr : T1 - 1;
                const size_t sidx = (T * th + R) * 8 + C;
                GPair<VT>::ld(aval + sidx, v0[u], v1[u]);
                cc[u] = ntload(reinterpret_cast<const u2v *>(acol + sidx));
            }
#pragma unroll
            for (int u = 0; u < UG; u++) {
                AT x0[KC], x1[KC];
                gather_x<XT, AT, KC>(x, cc[u][0], ldx, j0, kv, x0);
                gather_x<XT, AT, KC>(x, cc[u][1], ldx, j0, kv, x1);
#pragma unroll
                for (int j = 0; j < KC; j++) {
                    if (H1) {
                        a0[u][j] = (AT)v0[u] * x0[j];
                        a1[u][j] = (AT)v1[u] * x1[j];
                    } else {
                        a0[u][j] = gfma<AT>((AT)v0[u], x0[j], a0[u][j]);
                        a1[u][j] = gfma<AT>((AT)v1[u], x1[j], a1[u][j]);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UG; u++) {
            const uint32_t Tr = Tbb + 16 * u + lane / 4;
            const size_t T = Tr < T1 ? Tr : T1 - 1;
            if (H1) {
                const bool ok = Tr < T1;
                if (d.k != 0) {
                    if (ok)
#pragma unroll
                        for (int j = 0; j < KC; j++) lacc[j] = lacc[j] + (a0[u][j] + a1[u][j]);
                    continue;
                }
                const uint32_t key = ntload(sec + T);
#pragma unroll
                for (int j = 0; j < KC; j++) {
                    AT p0 = a0[u][j], p1 = a1[u][j];
                    p0 = p0 + dpp_any<AT, DPP_QUAD_XOR2>(p0);
                    p1 = p1 + dpp_any<AT, DPP_QUAD_XOR2>(p1);
                    p0 = p0 + dpp_any<AT, DPP_QUAD_XOR1>(p0);
                    p1 = p1 + dpp_any<AT, DPP_QUAD_XOR1>(p1);
                    if (ok && q == 0) s_val[w][(Tr - d.ubeg) * KC + j] = p0 + p1;
                }
                if (ok && q == 0) s_key[w][Tr - d.ubeg] = key;
            } else {
                const uint32_t u0 = 8 * Tr + C;
                const bool ok0 = Tr < T1 && u0 >= d.ubeg && u0 < d.uend;
                const bool ok1 = Tr < T1 && u0 + 1 >= d.ubeg && u0 + 1 < d.uend;
                if (d.k != 0) {
#pragma unroll
                    for (int j = 0; j < KC; j++) {
                        if (ok0) lacc[j] = lacc[j] + a0[u][j];
                        if (ok1) lacc[j] = lacc[j] + a1[u][j];
                    }
                    continue;
                }
                const u2v kk = ntload(reinterpret_cast<const u2v *>(sec + 8 * T + C));
#pragma unroll
                for (int j = 0; j < KC; j++) {
                    if (ok0) s_val[w][(u0 - d.ubeg) * KC + j] = a0[u][j];
                    if (ok1) s_val[w][(u0 + 1 - d.ubeg) * KC + j] = a1[u][j];
                }
                if (ok0) s_key[w][u0 - d.ubeg] = kk[0];
                if (ok1) s_key[w][u0 + 1 - d.ubeg] = kk[1];
            }
        }
    }
    if (d.k != 0) {
#pragma unroll
        for (int j = 0; j < KC; j++) {
            AT t = wave_sum<AT>(lacc[j]);
            if (lane == 0 && (uint32_t)j < kv) {
                if (d.k == 1) {
                    if (d.rbeg < nrows) y[(size_t)d.rbeg * ldy + j0 + j] = t;
                } else {
                    partial[(size_t)d.piece * PARTIAL_STRIDE + j0 + j] = (double)t;
                }
            }
        }
        return;
    }
    wave_lds_sync();
    const uint32_t nseg = find_segments(s_key[w], s_seg[w], nu);
    for (uint32_t sg = lane; sg < nseg; sg += WAVE) {
        const uint32_t a = s_seg[w][sg], b = s_seg[w][sg + 1], r = s_key[w][a];
        for (int j = 0; j < KC; j++) {
            if ((uint32_t)j >= kv) break;
            AT acc = 0;
            for (uint32_t i = a; i < b; i++) acc = acc + s_val[w][i * KC + j];
            if (r < nrows) y[(size_t)r * ldy + j0 + j] = acc;
        }
        // empty rows before this run (and after the last run) of the block
        const uint32_t lo = (sg == 0) ? d.rbeg : s_key[w][s_seg[w][sg - 1]] + 1;
        for (uint32_t z = lo; z < r && z < nrows; z++)
            for (uint32_t j = 0; j < kv; j++) y[(size_t)z * ldy + j0 + j] = 0;
        if (sg == nseg - 1)
            for (uint32_t z = r + 1; z < d.rend && z < nrows; z++)
                for (uint32_t j = 0; j < kv; j++) y[(size_t)z * ldy + j0 + j] = 0;
    }
    if (nseg == 0)
        for (uint32_t z = d.rbeg + lane; z < d.rend && z < nrows; z += WAVE)
            for (uint32_t j = 0; j < kv; j++) y[(size_t)z * ldy + j0 + j] = 0;
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

// Embedded x refresh: the x half of every occupied slot becomes x[acol[s]].
__global__ __launch_bounds__(256) void refresh_x_kernel(uint32_t nslots, uint32_t th,
                                                        const uint32_t *__restrict__ acol,
                                                        const double *__restrict__ x,
                                                        double *__restrict__ ax)
{
    const size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const uint32_t c = ntload(acol + s);
    if (c == PAD) return;
    const size_t T = s / (8 * th), w = s % (8 * th);
    ax[(T * th + w / 8) * 16 + 8 + w % 8] = x[c];
}

template <typename VT, typename XT, typename AT>
static int launch_gather(const intspmv_plan *p, const void *aval, const uint32_t *acol,
                         const uint32_t *sec, const void *x, uint32_t ldx, void *y, uint32_t ldy,
                         uint32_t k)
{
    const uint32_t nb = p->nblocks, th = p->axt.tileH;
    if (nb == 0) return INTSPMV_OK;
    dim3 grid(ispv_grid(p)), blk(BLOCK_THREADS);
    hipStream_t s = p->ctx->stream;
    const uint32_t nl = p->nlong_wg;
    for (uint32_t j0 = 0; j0 < k;) {
        const uint32_t rem = k - j0;
        const uint32_t kc = rem >= 8 ? 8 : (rem >= 4 ? 4 : (rem >= 2 ? 2 : 1));
#define ISPV_G(KC)                                                                              \
    if (kc == KC) {                                                                             \
        if (th == 1)                                                                            \
            axt_gather_kernel<VT, XT, AT, KC, true><<<grid, blk, 0, s>>>(                       \
                p->d_blocks, nl, nb, (const VT *)aval, acol, sec, (const XT *)x, ldx, (AT *)y,  \
                ldy, j0, KC, th, p->nrows, p->d_partial);                                       \
        else                                                                                    \
            axt_gather_kernel<VT, XT, AT, KC, false><<<grid, blk, 0, s>>>(                      \
                p->d_blocks, nl, nb, (const VT *)aval, acol, sec, (const XT *)x, ldx, (AT *)y,  \
                ldy, j0, KC, th, p->nrows, p->d_partial);                                       \
    }
        ISPV_G(1)
        ISPV_G(2)
        ISPV_G(4)
        ISPV_G(8)
#undef ISPV_G
        int st = ispv_launch_check(p->ctx);
        if (st) return st;
        j0 += kc;
    }
    return ispv_split_fixup(p, y, sizeof(AT) == 4, k, ldy, 0);
}

} // namespace ispv

using namespace ispv;

static bool gather_plan_ok(const intspmv_ctx *ctx, const intspmv_plan *plan)
{
    return ctx && plan && plan->kind == PLAN_AXT && plan->ctx == ctx &&
           plan->axt.mode == INTSPMV_NOC;
}

extern "C" int spmv_axt_gather(intspmv_ctx *ctx, const intspmv_plan *plan, const void *aval,
                               const uint32_t *acol, const uint32_t *sec, const void *x,
                               uint32_t ldx, void *y, uint32_t ldy, uint32_t k)
{
    if (!gather_plan_ok(ctx, plan) || k == 0 || k > PARTIAL_STRIDE || ldx < k || ldy < k)
        return INTSPMV_ERR_ARG;
    if (plan->nrows && (!x || !y || (plan->nunits && (!aval || !acol || !sec))))
        return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    const intspmv_axt_desc &D = plan->axt;
    if (D.val_type == INTSPMV_F64 && D.x_type == INTSPMV_F64)
        return launch_gather<double, double, double>(plan, aval, acol, sec, x, ldx, y, ldy, k);
    if (D.val_type == INTSPMV_F32 && D.x_type == INTSPMV_F32)
        return launch_gather<float, float, float>(plan, aval, acol, sec, x, ldx, y, ldy, k);
    if (D.val_type == INTSPMV_BF16 && D.x_type == INTSPMV_F32)
        return launch_gather<uint16_t, float, float>(plan, aval, acol, sec, x, ldx, y, ldy, k);
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
    refresh_x_kernel<<<(nslots + 255) / 256, 256, 0, ctx->stream>>>(nslots, D.tileH, acol, x, ax);
    return ispv_launch_check(ctx);
}
