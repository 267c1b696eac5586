// batch.hip -- many AXT matrices, 1..32 right-hand sides each, in one
// persistent launch that can be captured once in a hipGraph and replayed.
//
// Each matrix is in column-map form (gather.hip) with its own tile height,
// its own plan blocks and its own right-hand-side blocking kc (chosen per
// matrix by the autotuner).  The host flattens everything into one list of
// work items (matrix, block, first right-hand side), longest first; a grid
// sized to the resident capacity of the device walks the list with a static
// stride (wave w takes items w, w + W, ...), so the schedule is fixed and the
// results are deterministic.  Rows split across items meet in a batch-wide
// partial-sum scratch that one fix-up kernel reduces afterwards.  The reference
// has no counterpart (it times one matrix, one x).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "gather_block.hpp"
#include "reduce.hpp"

namespace ispv {

struct BatchMat {
    const void *aval;
    const uint32_t *acol, *sec;
    const void *x;
    void *y;
    uint32_t nrows, th, k, ldx, ldy, kc, piece_off;
    uint32_t xv; // x rows 16-byte aligned (wide passes load them whole)
};
struct BatchItem {
    uint32_t mat, block, j0, pad;
};
struct BatchSplit {
    uint32_t mat, row, piece, k;
};

} // namespace ispv

struct intspmv_batch {
    intspmv_ctx *ctx;
    int device; // the context's device, so destroying needs nothing of the context
    int val_type, x_type;
    uint32_t nmat, nitems, nblocks, nsplit, grid;
    uint32_t nnarrow, grid_w; // items [0, nnarrow) narrow passes, the rest wide; wide grid
    bool xb; // every x spans less than 2 GiB: buffer-resource gathers
    ispv::BatchMat *d_mats;
    ispv::BatchItem *d_items;
    ispv::BlockDesc *d_blocks;
    ispv::BatchSplit *d_splits;
    double *d_partial;
    uint32_t *d_xcd; // item ranges of the XCDs: narrow [xcd[i], xcd[i+1]), wide from xcd + 9
    hipGraph_t graph;
    hipGraphExec_t exec;
    hipStream_t cap_stream;
    hipStream_t aux;             // the wide kernel's stream (default priority)
    hipEvent_t fork, join;       // orders aux after / before the caller's stream
};


namespace ispv {

// Groups of 16 tiles in flight per step in narrow passes of one and of four
// right-hand sides.  The narrow and the wide passes run as two kernels (PATHS
// 1 and 2) so that each keeps its own registers: in one kernel more groups in
// flight sped the k = 1 and k = 4 matrices of the 256-matrix batch
// (396 -> 316 us, 587 -> 487 us) but slowed the k = 32 ones (852 -> 990 us).
#ifndef ISPV_B_UG1
#define ISPV_B_UG1 8
#endif
#ifndef ISPV_B_UG4
#define ISPV_B_UG4 4
#endif
// FP64 sums keep 4 and 2 (twice the registers per value: 8 and 4 would take
// the narrow kernel to 220 VGPRs, two waves per SIMD)
template <typename AT>
constexpr int batch_ug1() { return sizeof(AT) == 8 ? 4 : ISPV_B_UG1; }
template <typename AT>
constexpr int batch_ug4() { return sizeof(AT) == 8 ? 2 : ISPV_B_UG4; }
enum { BATCH_NARROW = 1, BATCH_WIDE = 2 };
template <typename VT, typename XT, typename AT, bool XB, int PATHS>
__global__ __launch_bounds__(BLOCK_THREADS) void batch_kernel(
    const BatchItem *__restrict__ items, const uint32_t *__restrict__ xcd_range,
    const BlockDesc *__restrict__ blocks, const BatchMat *__restrict__ mats, double *partial)
{
    __shared__ GatherLds<AT, 8> lds[WAVES];
    const uint32_t w = threadIdx.x / WAVE;
    // Workgroups are dealt round-robin to the XCDs (grid a multiple of
    // NUM_XCD): the waves of XCD x sweep its own contiguous item range
    // together, so neighbouring blocks of a matrix -- which gather the same
    // rows of x -- run at the same time on the L2 they share.
    const uint32_t xcd = blockIdx.x % NUM_XCD;
    const uint32_t total = gridDim.x / NUM_XCD * WAVES, end = xcd_range[xcd + 1];
    for (uint32_t it = __builtin_amdgcn_readfirstlane(xcd_range[xcd] +
                                                      (blockIdx.x / NUM_XCD) * WAVES + w);
         it < end; it += total) {
        const BatchItem item = items[it];
        const BatchMat md = mats[item.mat];
        const BlockDesc d = blocks[item.block];
        double *part = partial + (size_t)md.piece_off * PARTIAL_STRIDE;
        const VT *av = (const VT *)md.aval;
        const XT *xx = (const XT *)md.x;
        AT *yy = (AT *)md.y;
        GatherLds<AT, 8> &L = lds[w];
        // PATHS = BATCH_WIDE: wide passes (tile height 1: normal blocks as
        // gather_block_wide, keeping a round in the wave's narrow-pass value
        // area; long rows as gather_long_wide, one tile per group in flight so
        // that they add no registers -- the same bits as more); BATCH_NARROW:
        // narrow passes of at most 8 right-hand sides
        static_assert(sizeof(WideLds<AT, 32, XT>) <= sizeof(L.val) &&
                          sizeof(WideLds<AT, 8, XT>) <= sizeof(L.val),
                      "a wide round fits the wave's LDS");
        if constexpr (PATHS & BATCH_WIDE) {
            const uint32_t kv = md.k - item.j0 < md.kc ? md.k - item.j0 : md.kc;
#define ISPV_BW(KW)                                                                             \
    if (md.kc == KW) {                                                                          \
        if (d.k != 0) {                                                                         \
            if (md.xv)                                                                          \
                gather_long_wide<VT, XT, AT, KW, true, 1>(d, av, md.acol, xx, md.ldx, yy, md.ldy,  \
                                                       item.j0, kv, md.nrows, part);            \
            else                                                                                \
                gather_long_wide<VT, XT, AT, KW, false, 1>(d, av, md.acol, xx, md.ldx, yy, md.ldy, \
                                                        item.j0, kv, md.nrows, part);           \
        } else if (md.xv)                                                                       \
            gather_block_wide<VT, XT, AT, KW, true>(d, av, md.acol, md.sec, xx, md.ldx, yy,     \
                                                    md.ldy, item.j0, kv, md.nrows, L.key,       \
                                                    L.seg,                                      \
                                                    *reinterpret_cast<WideLds<AT, KW, XT> *>(L.val)); \
        else                                                                                    \
            gather_block_wide<VT, XT, AT, KW, false>(d, av, md.acol, md.sec, xx, md.ldx, yy,    \
                                                     md.ldy, item.j0, kv, md.nrows, L.key,      \
                                                     L.seg,                                     \
                                                     *reinterpret_cast<WideLds<AT, KW, XT> *>(L.val)); \
    }
            if constexpr (sizeof(XT) == 8) {
                ISPV_BW(4)
            }
            ISPV_BW(8)
            ISPV_BW(16)
            ISPV_BW(32)
#undef ISPV_BW
        }
        if constexpr (PATHS & BATCH_NARROW) {
        const uint32_t kcn = md.kc > 8 ? 8 : md.kc;
        const uint32_t kv = md.k - item.j0 < kcn ? md.k - item.j0 : kcn;
#define ISPV_B(KC)                                                                              \
    if (kcn == KC) {                                                                            \
        if (md.th == 1)                                                                         \
            gather_block<VT, XT, AT, KC, true, XB, batch_ug1<AT>(), false, batch_ug4<AT>()>(d, av, md.acol, md.sec, xx, md.ldx, yy, md.ldy,  \
                                               item.j0, kv, 1, md.nrows, part, L.val, L.key, L.seg); \
        else                                                                                    \
            gather_block<VT, XT, AT, KC, false, XB, batch_ug1<AT>(), false, batch_ug4<AT>()>(d, av, md.acol, md.sec, xx, md.ldx, yy, md.ldy, \
                                                item.j0, kv, md.th, md.nrows, part, L.val, L.key,  \
                                                L.seg);                                         \
    }
        ISPV_B(1)
        ISPV_B(2)
        ISPV_B(4)
        ISPV_B(8)
#undef ISPV_B
        }
        wave_lds_sync(); // the next item reuses this wave's LDS
    }
}

template <typename YT>
__global__ void batch_fixup_kernel(const BatchSplit *__restrict__ splits, uint32_t nsplit,
                                   const BatchMat *__restrict__ mats,
                                   const double *__restrict__ partial)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, i = t / PARTIAL_STRIDE,
                   j = t % PARTIAL_STRIDE;
    if (i >= nsplit) return;
    const BatchSplit sp = splits[i];
    const BatchMat md = mats[sp.mat];
    if (j >= md.k) return;
    double acc = 0.0;
    for (uint32_t q = 0; q < sp.k; q++)
        acc = acc + partial[(size_t)(md.piece_off + sp.piece + q) * PARTIAL_STRIDE + j];
    if (sp.row < md.nrows) ((YT *)md.y)[(size_t)sp.row * md.ldy + j] = (YT)acc;
}

} // namespace ispv

using namespace ispv;


extern "C" int intspmv_batch_create(intspmv_ctx *ctx, int val_type, int x_type, uint32_t nmat,
                                    const intspmv_batch_matrix *mats, intspmv_batch **out)
{
    if (!ctx || !out || (nmat && !mats)) return INTSPMV_ERR_ARG;
    *out = nullptr;
    const bool f64 = val_type == INTSPMV_F64 && x_type == INTSPMV_F64;
    const bool f32 = val_type == INTSPMV_F32 && x_type == INTSPMV_F32;
    const bool b16 = val_type == INTSPMV_BF16 && x_type == INTSPMV_F32;
    if (!f64 && !f32 && !b16) return INTSPMV_ERR_UNSUPPORTED;
    std::vector<BatchMat> hm(nmat);
    std::vector<BlockDesc> blocks;
    std::vector<BatchSplit> splits;
    struct Cand {
        BatchItem item;
        uint64_t cost;
    };
    std::vector<Cand> cand;
    uint32_t piece_base = 0;
    bool xb = true;
    for (uint32_t i = 0; i < nmat; i++) {
        const intspmv_batch_matrix &M = mats[i];
        const uint32_t th = M.tileH;
        if (th == 0 || M.k == 0 || M.k > PARTIAL_STRIDE || M.ldx < M.k || M.ldy < M.k ||
            (M.nrows && (!M.con_host || !M.x || !M.y)))
            return INTSPMV_ERR_ARG;
        // kc 1, 2, 4, 8: narrow passes; at tile height 1, kc >= 2 * (16 / sizeof
        // x) up to 32: wide passes; 0: as spmv_axt_gather chooses
        const uint32_t vx = x_type == INTSPMV_F64 ? 2 : 4;
        const uint32_t wmin = x_type == INTSPMV_F64 ? wide_min_kc<double>() : wide_min_kc<float>();
        uint32_t kc = M.kc;
        if (!kc && th == 1 && M.k >= wmin && wmin <= 32) {
            kc = wmin;
            while (kc < M.k && kc < 32) kc *= 2;
        }
        if (!kc) kc = M.k >= 8 ? 8 : (M.k >= 4 ? 4 : (M.k >= 2 ? 2 : 1));
        const bool wide = th == 1 && kc >= wmin;
        if (kc != 1 && kc != 2 && kc != 4 && kc != 8 && !(wide && (kc == 16 || kc == 32)))
            return INTSPMV_ERR_ARG;
        const uint32_t umax = M.umax ? M.umax : ISPV_AXT_UMAX;
        if (umax > ISPV_AXT_UMAX || umax < 8) return INTSPMV_ERR_ARG;
        const uint32_t ths = 8 * th;
        if ((uint64_t)M.tileN * 2 * ths != M.lenAX) return INTSPMV_ERR_ARG;
        const uint32_t nunits = th == 1 ? M.tileN : M.colN;
        std::vector<uint32_t> uptr((size_t)M.nrows + 1);
        for (uint32_t r = 0; r < M.nrows; r++) {
            uint32_t c = M.con_host[r], T = c / (2 * ths), wpos = c % (2 * ths);
            uptr[r] = th == 1 ? T : T * 8 + wpos % 16;
            if (wpos % 16 >= 8 || uptr[r] > nunits || (r && uptr[r] < uptr[r - 1]))
                return INTSPMV_ERR_ARG;
        }
        uptr[M.nrows] = nunits;
        uint32_t long_units = th == 1 ? umax / 2 : (512 / th < 4 ? 4 : (512 / th > umax / 2 ? umax / 2 : 512 / th));
        std::vector<BlockDesc> mb;
        std::vector<SplitDesc> ms;
        uint32_t nlw = 0, nlong = 0, npieces = 0;
        int st = ispv_build_blocks(M.nrows, uptr.data(), umax, BLOCK_ROWS_MAX, long_units,
                                   th == 1 ? GATHER_H1_LONG_MAX : 8192 / th, mb, ms, nlw, nlong,
                                   npieces);
        if (st) return st;
        const uint32_t xv = M.ldx % vx == 0 && (uintptr_t)M.x % 16 == 0;
        hm[i] = {M.aval, M.acol, M.sec, M.x, M.y, M.nrows, th, M.k, M.ldx, M.ldy, kc, piece_base, xv};
        xb = xb && xbuf_fits(M.nrows, M.ldx, x_type == INTSPMV_F64 ? 8 : 4);
        for (const BlockDesc &d : mb) {
            const uint32_t bi = (uint32_t)blocks.size();
            blocks.push_back(d);
            if (d.rbeg == 0xFFFFFFFFu) continue; // XCD padding of the per-matrix plan
            const uint64_t slots = (uint64_t)(d.uend - d.ubeg) * (th == 1 ? 8 : th);
            for (uint32_t j0 = 0; j0 < M.k; j0 += kc) {
                // modelled bytes of the item (as the autotuner's model): the
                // stream plus the cache lines of its x gathers
                const uint32_t kv = kc < M.k - j0 ? kc : M.k - j0;
                const uint64_t xs = x_type == INTSPMV_F64 ? 8 : 4;
                const uint64_t xline = wide ? (kv * xs + 63) / 64 * 64 : (kv * xs + 15) / 16 * 64;
                cand.push_back({{i, bi, j0, 0}, slots * (8 + xline) + 256});
            }
        }
        for (const SplitDesc &sp : ms) splits.push_back({i, sp.row, sp.b0, sp.k});
        piece_base += npieces;
    }
    // Items stay in (matrix, block, first RHS) order: the narrow passes first,
    // then the wide ones (tile height 1 with a wide kc), each
    // list cut into contiguous per-XCD ranges of equal cost.
    std::vector<BatchItem> items;
    std::vector<uint32_t> xcd_range(2 * (NUM_XCD + 1));
    uint32_t nnarrow = 0;
    uint64_t cost_n = 0, cost_w = 0;
    for (int pass = 0; pass < 2; pass++) {
        std::vector<const Cand *> sel;
        uint64_t cost_all = 0;
        for (const Cand &c : cand) {
            const BatchMat &md = hm[c.item.mat];
            const uint32_t wmin =
                x_type == INTSPMV_F64 ? wide_min_kc<double>() : wide_min_kc<float>();
            const bool w = md.th == 1 && md.kc >= wmin;
            if (w == (pass == 1)) {
                sel.push_back(&c);
                cost_all += c.cost;
            }
        }
        const uint32_t base = (uint32_t)items.size();
        uint32_t *xr = &xcd_range[pass * (NUM_XCD + 1)];
        for (int x = 0; x <= NUM_XCD; x++) xr[x] = base + (uint32_t)sel.size();
        xr[0] = base;
        uint64_t acc = 0;
        int x = 1;
        for (size_t i = 0; i < sel.size(); i++) {
            items.push_back(sel[i]->item);
            acc += sel[i]->cost;
            while (x < NUM_XCD && acc * NUM_XCD >= cost_all * (uint64_t)x)
                xr[x++] = base + (uint32_t)(i + 1);
        }
        if (pass == 0) nnarrow = (uint32_t)items.size();
        (pass == 0 ? cost_n : cost_w) = cost_all;
    }

    intspmv_batch *b = (intspmv_batch *)calloc(1, sizeof(intspmv_batch));
    if (!b) return INTSPMV_ERR_NOMEM;
    b->ctx = ctx;
    b->device = ctx->device;
    b->val_type = val_type;
    b->x_type = x_type;
    b->nmat = nmat;
    b->nitems = (uint32_t)items.size();
    b->nnarrow = nnarrow;
    b->nblocks = (uint32_t)blocks.size();
    b->nsplit = (uint32_t)splits.size();
    int st = ispv_enter(ctx);
    if (st) {
        free(b);
        return st;
    }
    int dev = ctx->device, cus = 0, per_n = 0, per_w = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    b->xb = xb;
#define ISPV_OCC(VT, XT, AT)                                                                    \
    do {                                                                                        \
        if (xb)                                                                                 \
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                                   \
                &per_n, batch_kernel<VT, XT, AT, true, BATCH_NARROW>, BLOCK_THREADS, 0);        \
        else                                                                                    \
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                                   \
                &per_n, batch_kernel<VT, XT, AT, false, BATCH_NARROW>, BLOCK_THREADS, 0);       \
        if (e == hipSuccess)                                                                    \
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(                                   \
                &per_w, batch_kernel<VT, XT, AT, false, BATCH_WIDE>, BLOCK_THREADS, 0);         \
    } while (0)
    if (e == hipSuccess) {
        if (f64) ISPV_OCC(double, double, double);
        else if (f32) ISPV_OCC(float, float, float);
        else ISPV_OCC(uint16_t, float, float);
    }
#undef ISPV_OCC
    if (e != hipSuccess || cus <= 0) {
        ctx->last_err = e;
        free(b);
        return INTSPMV_ERR_HIP;
    }
    // grids: the resident capacity (whole rounds over the XCDs), no more
    // waves than items
    auto grid_for = [&](int per_cu, uint32_t n) {
        per_cu = per_cu < 1 ? 1 : per_cu;
        uint32_t want = (n + WAVES - 1) / WAVES;
        uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>((uint32_t)(cus * per_cu), want));
        return (g + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    };
    // The two kernels run at the same time (fork / join on a second stream),
    // each on a share of the device in proportion to its modelled bytes, so
    // both finish together instead of one after the other.
    const double fn = (cost_n + cost_w) ? (double)cost_n / (double)(cost_n + cost_w) : 0.0;
    b->grid = grid_for((int)std::max(1.0, per_n * fn + 0.5), nnarrow);
    b->grid_w = grid_for((int)std::max(1.0, per_w * (1.0 - fn) + 0.5), b->nitems - nnarrow);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join, hipEventDisableTiming);
    auto up = [&](void **dst, const void *src, size_t bytes) {
        if (e != hipSuccess) return;
        e = hipMalloc(dst, bytes ? bytes : 16);
        if (e == hipSuccess && bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    };
    up((void **)&b->d_mats, hm.data(), hm.size() * sizeof(BatchMat));
    up((void **)&b->d_items, items.data(), items.size() * sizeof(BatchItem));
    up((void **)&b->d_xcd, xcd_range.data(), xcd_range.size() * sizeof(uint32_t));
    up((void **)&b->d_blocks, blocks.data(), blocks.size() * sizeof(BlockDesc));
    up((void **)&b->d_splits, splits.data(), splits.size() * sizeof(BatchSplit));
    if (e == hipSuccess)
        e = hipMalloc(&b->d_partial, (size_t)(piece_base ? piece_base : 1) * PARTIAL_STRIDE * sizeof(double));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        ctx->last_err = e;
        intspmv_batch_destroy(b);
        return INTSPMV_ERR_NOMEM;
    }
    *out = b;
    return INTSPMV_OK;
}

static int batch_enqueue(const intspmv_batch *b, hipStream_t s)
{
    const bool both = b->nnarrow && b->nitems > b->nnarrow;
    hipStream_t sw = both ? b->aux : s; // the wide kernel's stream
    if (both && (hipEventRecord(b->fork, s) != hipSuccess ||
                 hipStreamWaitEvent(b->aux, b->fork, 0) != hipSuccess))
        return INTSPMV_ERR_HIP;
    if (b->nitems) {
        const uint32_t *xw = b->d_xcd + NUM_XCD + 1;
#define ISPV_BL(VT, XT, AT)                                                                     \
    if (b->nnarrow) {                                                                           \
        if (b->xb)                                                                              \
            batch_kernel<VT, XT, AT, true, BATCH_NARROW><<<b->grid, BLOCK_THREADS, 0, s>>>(      \
                b->d_items, b->d_xcd, b->d_blocks, b->d_mats, b->d_partial);                     \
        else                                                                                    \
            batch_kernel<VT, XT, AT, false, BATCH_NARROW><<<b->grid, BLOCK_THREADS, 0, s>>>(     \
                b->d_items, b->d_xcd, b->d_blocks, b->d_mats, b->d_partial);                     \
    }                                                                                           \
    if (b->nitems > b->nnarrow)                                                                 \
        batch_kernel<VT, XT, AT, false, BATCH_WIDE><<<b->grid_w, BLOCK_THREADS, 0, sw>>>(        \
            b->d_items, xw, b->d_blocks, b->d_mats, b->d_partial)
        if (b->val_type == INTSPMV_F64) {
            ISPV_BL(double, double, double);
        } else if (b->val_type == INTSPMV_F32) {
            ISPV_BL(float, float, float);
        } else {
            ISPV_BL(uint16_t, float, float);
        }
#undef ISPV_BL
    }
    if (both && (hipEventRecord(b->join, b->aux) != hipSuccess ||
                 hipStreamWaitEvent(s, b->join, 0) != hipSuccess))
        return INTSPMV_ERR_HIP;
    if (b->nsplit) {
        dim3 g((b->nsplit * PARTIAL_STRIDE + 255) / 256);
        if (b->val_type == INTSPMV_F64)
            batch_fixup_kernel<double><<<g, 256, 0, s>>>(b->d_splits, b->nsplit, b->d_mats, b->d_partial);
        else
            batch_fixup_kernel<float><<<g, 256, 0, s>>>(b->d_splits, b->nsplit, b->d_mats, b->d_partial);
    }
    return hipGetLastError() == hipSuccess ? INTSPMV_OK : INTSPMV_ERR_HIP;
}

extern "C" int intspmv_batch_launch(intspmv_ctx *ctx, const intspmv_batch *b)
{
    if (!ctx || !b || b->ctx != ctx) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    st = batch_enqueue(b, ctx->stream);
    if (st) ctx->last_err = hipGetLastError();
    return st;
}

extern "C" int intspmv_batch_graph_capture(intspmv_ctx *ctx, intspmv_batch *b)
{
    if (!ctx || !b || b->ctx != ctx) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    if (b->exec) return INTSPMV_OK;
    hipError_t e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        int est = batch_enqueue(b, b->cap_stream);
        hipError_t e2 = hipStreamEndCapture(b->cap_stream, &b->graph);
        e = est ? hipErrorUnknown : e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        ctx->last_err = e;
        return INTSPMV_ERR_HIP;
    }
    return INTSPMV_OK;
}

extern "C" int intspmv_batch_graph_launch(intspmv_ctx *ctx, const intspmv_batch *b)
{
    if (!ctx || !b || b->ctx != ctx || !b->exec) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    hipError_t e = hipGraphLaunch(b->exec, ctx->stream);
    if (e != hipSuccess) {
        ctx->last_err = e;
        return INTSPMV_ERR_HIP;
    }
    return INTSPMV_OK;
}

extern "C" int intspmv_batch_info(const intspmv_batch *b, uint32_t *nitems, uint32_t *grid,
                                  uint32_t *nsplit)
{
    if (!b) return INTSPMV_ERR_ARG;
    if (nitems) *nitems = b->nitems;
    if (grid) *grid = b->grid;
    if (nsplit) *nsplit = b->nsplit;
    return INTSPMV_OK;
}

extern "C" int intspmv_batch_destroy(intspmv_batch *b)
{
    if (!b) return INTSPMV_OK;
    (void)hipSetDevice(b->device); // not through the context: it may be gone already
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->aux) (void)hipStreamDestroy(b->aux);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    (void)hipFree(b->d_mats);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_blocks);
    (void)hipFree(b->d_splits);
    (void)hipFree(b->d_partial);
    (void)hipFree(b->d_xcd);
    free(b);
    return INTSPMV_OK;
}
