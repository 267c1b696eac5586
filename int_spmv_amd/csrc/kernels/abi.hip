// abi.hip -- context, memory, error and plan entry points of libintspmv.
//
// The plan is the host-side analysis that replaces the reference's per-call
// OpenMP dynamic scheduling (schedule(dynamic) over rows / tiles / chunks,
// intSpmv_main.c:717, 948, 1391, 1451): rows are packed once into work blocks
// of bounded size, one per wave, so every launch is a static, balanced sweep.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "common.hpp"

// ---- scratch of the allocation-free entry point ---------------------------
// spmv_axt_h1_noalloc keeps the partial sums of rows that cross its fixed tile
// chunks here (two doubles per chunk of 64 tiles), so that it never allocates.
// A file-scope device array: every device the library runs on holds its own
// copy, at its own address, until the process ends.  In host code the name
// `scratch` is not a device address: the address on each device comes from
// hipGetSymbolAddress, taken once per device (with that device current) when
// the first context on it is created, and only that pointer is used.
static __device__ __align__(256) double scratch[INTSPMV_NOALLOC_SCRATCH_BYTES / sizeof(double)];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MAX_DEVICES];

// The calling thread's current device must be `device`.
static int scratch_address(int device, double **out, hipError_t *err)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= MAX_DEVICES) return INTSPMV_ERR_UNSUPPORTED;
    if (!g_scratch_ptr[device]) {
        double *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            *err = e;
            return INTSPMV_ERR_HIP;
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return INTSPMV_OK;
}

extern "C" const char *intspmv_strerror(int st)
{
    switch (st) {
    case INTSPMV_OK: return "ok";
    case INTSPMV_ERR_ARG: return "invalid argument";
    case INTSPMV_ERR_HIP: return "HIP runtime error";
    case INTSPMV_ERR_NOMEM: return "out of memory";
    case INTSPMV_ERR_UNSUPPORTED: return "unsupported combination";
    case INTSPMV_ERR_OVERFLOW: return "size exceeds the 32-bit index range";
    case INTSPMV_ERR_NODEVICE: return "no HIP device";
    case INTSPMV_ERR_TOO_LARGE:
        return "matrix exceeds the allocation-free entry point's scratch "
               "(INTSPMV_NOALLOC_MAX_TILES tiles)";
    default: return "unknown status";
    }
}

int ispv_enter(intspmv_ctx *ctx)
{
    if (!ctx) return INTSPMV_ERR_ARG;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != ctx->device) e = hipSetDevice(ctx->device);
    if (e != hipSuccess) {
        ctx->last_err = e;
        return INTSPMV_ERR_HIP;
    }
    return INTSPMV_OK;
}

int ispv_launch_check(intspmv_ctx *ctx)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->last_err = e;
        return INTSPMV_ERR_HIP;
    }
    return INTSPMV_OK;
}

extern "C" int intspmv_device_count(int *count)
{
    if (!count) return INTSPMV_ERR_ARG;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        return INTSPMV_ERR_NODEVICE;
    }
    return INTSPMV_OK;
}

extern "C" int intspmv_create(intspmv_ctx **ctx, int device)
{
    if (!ctx) return INTSPMV_ERR_ARG;
    *ctx = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return INTSPMV_ERR_NODEVICE;
    if (device < 0 || device >= n) return INTSPMV_ERR_ARG;
    intspmv_ctx *c = (intspmv_ctx *)calloc(1, sizeof(intspmv_ctx));
    if (!c) return INTSPMV_ERR_NOMEM;
    c->device = device;
    c->stream = nullptr;
    if (hipSetDevice(device) != hipSuccess) {
        free(c);
        return INTSPMV_ERR_HIP;
    }
    hipError_t e = hipSuccess;
    int st = scratch_address(device, &c->noalloc_scratch, &e);
    if (st) {
        free(c);
        return st;
    }
    *ctx = c;
    return INTSPMV_OK;
}

extern "C" int intspmv_destroy(intspmv_ctx *ctx)
{
    free(ctx);
    return INTSPMV_OK;
}

extern "C" int intspmv_set_stream(intspmv_ctx *ctx, void *stream)
{
    if (!ctx) return INTSPMV_ERR_ARG;
    ctx->stream = (hipStream_t)stream;
    return INTSPMV_OK;
}

extern "C" int intspmv_sync(intspmv_ctx *ctx)
{
    int st = ispv_enter(ctx);
    if (st) return st;
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx->last_err = e;
        return INTSPMV_ERR_HIP;
    }
    return INTSPMV_OK;
}

extern "C" const char *intspmv_last_hip_error(intspmv_ctx *ctx)
{
    return ctx ? hipGetErrorString(ctx->last_err) : "no context";
}

extern "C" int intspmv_malloc(intspmv_ctx *ctx, void **dptr, size_t bytes)
{
    int st = ispv_enter(ctx);
    if (st) return st;
    if (!dptr) return INTSPMV_ERR_ARG;
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        ctx->last_err = e;
        *dptr = nullptr;
        return INTSPMV_ERR_NOMEM;
    }
    return INTSPMV_OK;
}

extern "C" int intspmv_free(intspmv_ctx *ctx, void *dptr)
{
    int st = ispv_enter(ctx);
    if (st) return st;
    if (dptr) {
        hipError_t e = hipFree(dptr);
        if (e != hipSuccess) {
            ctx->last_err = e;
            return INTSPMV_ERR_HIP;
        }
    }
    return INTSPMV_OK;
}

static int copy(intspmv_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind k)
{
    int st = ispv_enter(ctx);
    if (st) return st;
    if (bytes == 0) return INTSPMV_OK;
    if (!dst || !src) return INTSPMV_ERR_ARG;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, k, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx->last_err = e;
        return INTSPMV_ERR_HIP;
    }
    return INTSPMV_OK;
}

extern "C" int intspmv_memcpy_htod(intspmv_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    return copy(ctx, dst, src, bytes, hipMemcpyHostToDevice);
}

extern "C" int intspmv_memcpy_dtoh(intspmv_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    return copy(ctx, dst, src, bytes, hipMemcpyDeviceToHost);
}

// Split-row fix-up: one thread per (split row, right-hand side) adds the
// row's piece partials in piece order (deterministic) and stores the result
// at y[row * ldy + j0 + j].
template <typename YT>
__global__ void split_fixup_kernel(const ispv::SplitDesc *__restrict__ splits, uint32_t nsplit,
                                   const double *__restrict__ partial, YT *__restrict__ y,
                                   uint32_t nrows, uint32_t k, uint32_t ldy, uint32_t j0)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, i = t / k, j = t % k;
    if (i >= nsplit) return;
    const ispv::SplitDesc s = splits[i];
    double acc = 0.0;
    for (uint32_t q = 0; q < s.k; q++) acc = acc + partial[(size_t)(s.b0 + q) * ispv::PARTIAL_STRIDE + j];
    if (s.row < nrows) y[(size_t)s.row * ldy + j0 + j] = (YT)acc;
}

int ispv_split_fixup(const intspmv_plan *p, void *y, int y_is_f32, uint32_t k, uint32_t ldy,
                     uint32_t j0)
{
    if (p->nsplit_rows == 0) return INTSPMV_OK;
    dim3 grid((p->nsplit_rows * k + 255) / 256);
    if (y_is_f32)
        split_fixup_kernel<float><<<grid, 256, 0, p->ctx->stream>>>(
            p->d_splits, p->nsplit_rows, p->d_partial, (float *)y, p->nrows, k, ldy, j0);
    else
        split_fixup_kernel<double><<<grid, 256, 0, p->ctx->stream>>>(
            p->d_splits, p->nsplit_rows, p->d_partial, (double *)y, p->nrows, k, ldy, j0);
    return ispv_launch_check(p->ctx);
}

// Allocates and fills the device side of a plan whose host fields are set.
int ispv_plan_upload(intspmv_ctx *ctx, intspmv_plan *p, const void *blocks, size_t block_bytes,
                     size_t nblocks, size_t partial_per_block, const ispv::SplitDesc *splits,
                     size_t nsplit)
{
    int st = ispv_enter(ctx);
    if (st) return st;
    size_t nb = nblocks ? nblocks : 1, ns = nsplit ? nsplit : 1;
    hipError_t e = hipMalloc(&p->d_blocks, nb * block_bytes);
    size_t npart = partial_per_block ? nb * partial_per_block
                                     : (size_t)(p->npieces ? p->npieces : 1) * ispv::PARTIAL_STRIDE;
    if (e == hipSuccess) e = hipMalloc(&p->d_partial, npart * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&p->d_splits, ns * sizeof(ispv::SplitDesc));
    if (e == hipSuccess && nblocks)
        e = hipMemcpy(p->d_blocks, blocks, nblocks * block_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && nsplit)
        e = hipMemcpy(p->d_splits, splits, nsplit * sizeof(ispv::SplitDesc),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        ctx->last_err = e;
        return INTSPMV_ERR_NOMEM;
    }
    return INTSPMV_OK;
}

// Packs rows [r0, r1) into blocks.  uptr[r]..uptr[r+1] are row r's units
// (monotone).  Rows with at most long_units units are packed greedily into
// normal blocks of at most umax units and rmax rows.  Longer rows become long
// blocks of one wave each; rows longer than long_max are split into
// ceil(len / long_max) near-equal pieces (k = pieces, pad1 = index of the
// piece within its row until the pieces are numbered).
static void collect_blocks(uint32_t r0, uint32_t r1, const uint32_t *uptr, uint32_t umax,
                           uint32_t rmax, uint32_t long_units, uint32_t long_max,
                           std::vector<ispv::BlockDesc> &normal, std::vector<ispv::BlockDesc> &lng,
                           uint32_t &nlong)
{
    uint32_t r = r0;
    while (r < r1) {
        uint32_t len = uptr[r + 1] - uptr[r];
        if (len > long_units) {
            uint32_t np = (uint32_t)(((uint64_t)len + long_max - 1) / long_max);
            if (np == 1) {
                lng.push_back({uptr[r], uptr[r + 1], r, r + 1, 0, 1, 0, 0});
                nlong++;
            } else {
                uint64_t per = (len + (uint64_t)np - 1) / np;
                for (uint32_t q = 0; q < np; q++) {
                    uint64_t a = uptr[r] + q * per, b = a + per;
                    if (b > uptr[r + 1]) b = uptr[r + 1];
                    lng.push_back({(uint32_t)a, (uint32_t)b, r, r + 1, 0, np, 0, q});
                }
            }
            r++;
            continue;
        }
        uint32_t rs = r, units = 0;
        while (r < r1 && r - rs < rmax) {
            uint32_t l = uptr[r + 1] - uptr[r];
            if (l > long_units || units + l > umax) break;
            units += l;
            r++;
        }
        normal.push_back({uptr[rs], uptr[r], rs, r, 0, 0, 0, 0});
    }
}

// Longest first; the pieces of one split row stay together and in order.
static void sort_long(std::vector<ispv::BlockDesc> &lng)
{
    std::stable_sort(lng.begin(), lng.end(), [](const ispv::BlockDesc &a, const ispv::BlockDesc &b) {
        uint32_t la = a.k > 1 ? 0xFFFFFFFFu - a.rbeg : a.uend - a.ubeg;
        uint32_t lb = b.k > 1 ? 0xFFFFFFFFu - b.rbeg : b.uend - b.ubeg;
        return la > lb;
    });
}

// Numbers the split pieces of a sorted long-block list in order (each
// piece's slot in the partial-sum scratch, continuing from npieces) and
// appends the split rows.
static int number_pieces(std::vector<ispv::BlockDesc> &blocks, std::vector<ispv::SplitDesc> &splits,
                         uint32_t &npieces)
{
    for (size_t i = 0; i < blocks.size(); i++)
        if (blocks[i].k > 1 && blocks[i].pad1 == 0 && blocks[i].rbeg != 0xFFFFFFFFu) {
            // the k pieces of this row follow in order (stable sort, equal keys)
            const uint32_t k = blocks[i].k;
            if (i + k > blocks.size()) return INTSPMV_ERR_ARG;
            splits.push_back({blocks[i].rbeg, npieces, k, 0});
            for (uint32_t q = 0; q < k; q++) {
                if (blocks[i + q].rbeg != blocks[i].rbeg || blocks[i + q].pad1 != q)
                    return INTSPMV_ERR_ARG;
                blocks[i + q].b0 = (uint32_t)i;
                blocks[i + q].piece = npieces++;
                blocks[i + q].pad1 = 0;
            }
            i += k - 1;
        }
    return INTSPMV_OK;
}

static const ispv::BlockDesc LONG_FILL = {0, 0, 0xFFFFFFFFu, 0xFFFFFFFFu, 0, 1, 0, 0};

static bool blocks_args_ok(uint32_t nrows, const uint32_t *uptr, uint32_t umax, uint32_t rmax,
                           uint32_t long_units, uint32_t long_max)
{
    if (!uptr || umax == 0 || rmax == 0 || long_units >= umax || long_max <= long_units)
        return false;
    for (uint32_t r = 0; r < nrows; r++)
        if (uptr[r + 1] < uptr[r]) return false;
    return true;
}

// Packs all rows into blocks (collect_blocks).  The long blocks come first,
// longest first, padded to whole XCD rounds of workgroups; the normal blocks
// follow in row order.  Pure host code.
int ispv_build_blocks(uint32_t nrows, const uint32_t *uptr, uint32_t umax, uint32_t rmax,
                      uint32_t long_units, uint32_t long_max,
                      std::vector<ispv::BlockDesc> &blocks,
                      std::vector<ispv::SplitDesc> &splits, uint32_t &nlong_wg, uint32_t &nlong,
                      uint32_t &npieces)
{
    if (!blocks_args_ok(nrows, uptr, umax, rmax, long_units, long_max)) return INTSPMV_ERR_ARG;
    std::vector<ispv::BlockDesc> normal, lng;
    nlong = 0;
    collect_blocks(0, nrows, uptr, umax, rmax, long_units, long_max, normal, lng, nlong);
    sort_long(lng);
    const uint32_t per_round = ispv::WAVES * ispv::NUM_XCD;
    size_t nl = lng.empty() ? 0 : (lng.size() + per_round - 1) / per_round * per_round;
    while (lng.size() < nl) lng.push_back(LONG_FILL);
    splits.clear();
    npieces = 0;
    int st = number_pieces(lng, splits, npieces);
    if (st) return st;
    blocks = lng;
    blocks.insert(blocks.end(), normal.begin(), normal.end());
    nlong_wg = (uint32_t)(nl / ispv::WAVES);
    return INTSPMV_OK;
}

// Blocks of rows in groups pinned to XCDs.  Group g is rows
// [gbeg[g], gbeg[g+1]) and runs on XCD g % 8: blocks never straddle two
// groups, and the block list is laid out for block_index() so that
// workgroups on XCD x (hardware ids x, x+8, ...) take only blocks of its
// groups.  Long section: round r's workgroup 8r + x holds the next WAVES
// long blocks of XCD x (longest first over its groups); normal section: XCD
// x's normal blocks (its groups in order) padded with empty blocks to the
// longest XCD list, so the contiguous range block_index() hands XCD x is
// exactly its own list.  gbeg == NULL: ngroups groups of about equal units.
// Used where the x values a group of rows gathers should stay in one XCD's
// L2 (banded matrices whose long rows would otherwise land on any XCD).
int ispv_build_blocks_grouped(uint32_t nrows, const uint32_t *uptr, uint32_t umax, uint32_t rmax,
                              uint32_t long_units, uint32_t long_max, uint32_t ngroups,
                              const uint32_t *gbeg, std::vector<ispv::BlockDesc> &blocks,
                              std::vector<ispv::SplitDesc> &splits, uint32_t &nlong_wg,
                              uint32_t &nlong, uint32_t &npieces)
{
    using ispv::NUM_XCD;
    using ispv::WAVES;
    if (!blocks_args_ok(nrows, uptr, umax, rmax, long_units, long_max) || ngroups == 0)
        return INTSPMV_ERR_ARG;
    std::vector<uint32_t> even;
    if (!gbeg) {
        // ngroups groups of about equal units, cut at row starts
        even.resize(ngroups + 1);
        const uint64_t tot = uptr[nrows];
        uint32_t r = 0;
        even[0] = 0;
        for (uint32_t g = 1; g < ngroups; g++) {
            const uint64_t want = tot * g / ngroups;
            while (r < nrows && uptr[r] < want) r++;
            even[g] = r;
        }
        even[ngroups] = nrows;
        gbeg = even.data();
    }
    if (gbeg[0] != 0 || gbeg[ngroups] != nrows) return INTSPMV_ERR_ARG;
    for (uint32_t g = 0; g < ngroups; g++)
        if (gbeg[g + 1] < gbeg[g]) return INTSPMV_ERR_ARG;
    std::vector<ispv::BlockDesc> normal[NUM_XCD], lng[NUM_XCD];
    nlong = 0;
    for (uint32_t g = 0; g < ngroups; g++)
        collect_blocks(gbeg[g], gbeg[g + 1], uptr, umax, rmax, long_units, long_max,
                       normal[g % NUM_XCD], lng[g % NUM_XCD], nlong);
    size_t rounds = 0, nmax = 0;
    splits.clear();
    npieces = 0;
    for (int x = 0; x < NUM_XCD; x++) {
        sort_long(lng[x]);
        int st = number_pieces(lng[x], splits, npieces);
        if (st) return st;
        rounds = std::max(rounds, (lng[x].size() + WAVES - 1) / WAVES);
        nmax = std::max(nmax, normal[x].size());
    }
    nmax = (nmax + WAVES - 1) / WAVES * WAVES;
    blocks.clear();
    for (size_t r = 0; r < rounds; r++)
        for (int x = 0; x < NUM_XCD; x++)
            for (int w = 0; w < WAVES; w++) {
                const size_t i = r * WAVES + w;
                blocks.push_back(i < lng[x].size() ? lng[x][i] : LONG_FILL);
            }
    for (int x = 0; x < NUM_XCD; x++) {
        const uint32_t u = normal[x].empty() ? 0 : normal[x].back().uend;
        const uint32_t rr = normal[x].empty() ? 0 : normal[x].back().rend;
        blocks.insert(blocks.end(), normal[x].begin(), normal[x].end());
        for (size_t i = normal[x].size(); i < nmax; i++) blocks.push_back({u, u, rr, rr, 0, 0, 0, 0});
    }
    nlong_wg = (uint32_t)(rounds * NUM_XCD);
    return INTSPMV_OK;
}

extern "C" int intspmv_plan_layout(uint32_t nrows, const uint32_t *uptr, uint32_t umax,
                                   uint32_t rmax, uint32_t long_units, uint32_t long_max,
                                   uint32_t *blocks, size_t cap_blocks, uint32_t *nblocks,
                                   uint32_t *splits, size_t cap_splits, uint32_t *nsplits,
                                   uint32_t *nlong_wg)
{
    std::vector<ispv::BlockDesc> b;
    std::vector<ispv::SplitDesc> sp;
    uint32_t nlw = 0, nlong = 0, npieces = 0;
    int st = ispv_build_blocks(nrows, uptr, umax, rmax, long_units, long_max, b, sp, nlw, nlong,
                          npieces);
    if (st) return st;
    if (nblocks) *nblocks = (uint32_t)b.size();
    if (nsplits) *nsplits = (uint32_t)sp.size();
    if (nlong_wg) *nlong_wg = nlw;
    if (blocks && cap_blocks >= b.size())
        memcpy(blocks, b.data(), b.size() * sizeof(ispv::BlockDesc));
    if (splits && cap_splits >= sp.size())
        memcpy(splits, sp.data(), sp.size() * sizeof(ispv::SplitDesc));
    return INTSPMV_OK;
}

extern "C" int intspmv_plan_layout_grouped(uint32_t nrows, const uint32_t *uptr, uint32_t umax,
                                           uint32_t rmax, uint32_t long_units, uint32_t long_max,
                                           uint32_t ngroups, const uint32_t *group_rows,
                                           uint32_t *blocks, size_t cap_blocks, uint32_t *nblocks,
                                           uint32_t *splits, size_t cap_splits, uint32_t *nsplits,
                                           uint32_t *nlong_wg)
{
    std::vector<ispv::BlockDesc> b;
    std::vector<ispv::SplitDesc> sp;
    uint32_t nlw = 0, nlong = 0, npieces = 0;
    int st = ispv_build_blocks_grouped(nrows, uptr, umax, rmax, long_units, long_max, ngroups,
                                       group_rows, b, sp, nlw, nlong, npieces);
    if (st) return st;
    if (nblocks) *nblocks = (uint32_t)b.size();
    if (nsplits) *nsplits = (uint32_t)sp.size();
    if (nlong_wg) *nlong_wg = nlw;
    if (blocks && cap_blocks >= b.size())
        memcpy(blocks, b.data(), b.size() * sizeof(ispv::BlockDesc));
    if (splits && cap_splits >= sp.size())
        memcpy(splits, sp.data(), sp.size() * sizeof(ispv::SplitDesc));
    return INTSPMV_OK;
}

int ispv_plan_build(intspmv_ctx *ctx, int kind, uint32_t nrows, const uint32_t *uptr,
                    uint32_t umax, uint32_t rmax, uint32_t long_units, uint32_t long_max,
                    intspmv_plan **out, uint32_t ngroups, const uint32_t *gbeg)
{
    if (!out) return INTSPMV_ERR_ARG;
    *out = nullptr;
    std::vector<ispv::BlockDesc> blocks;
    std::vector<ispv::SplitDesc> splits;
    uint32_t nlw = 0, nlong = 0, npieces = 0;
    int st = ngroups ? ispv_build_blocks_grouped(nrows, uptr, umax, rmax, long_units, long_max,
                                                 ngroups, gbeg, blocks, splits, nlw, nlong,
                                                 npieces)
                     : ispv_build_blocks(nrows, uptr, umax, rmax, long_units, long_max, blocks,
                                         splits, nlw, nlong, npieces);
    if (st) return st;
    intspmv_plan *p = (intspmv_plan *)calloc(1, sizeof(intspmv_plan));
    if (!p) return INTSPMV_ERR_NOMEM;
    p->ctx = ctx;
    p->device = ctx->device;
    p->kind = kind;
    p->nrows = nrows;
    p->nunits = nrows ? uptr[nrows] : 0;
    p->nblocks = (uint32_t)blocks.size();
    p->nlong_wg = nlw;
    p->nsplit_rows = (uint32_t)splits.size();
    p->nlong_rows = nlong;
    p->umax = umax;
    p->rmax = rmax;
    p->long_units = long_units;
    p->npieces = npieces;
    st = ispv_plan_upload(ctx, p, blocks.data(), sizeof(ispv::BlockDesc), blocks.size(), 0,
                          splits.data(), splits.size());
    if (st) {
        intspmv_plan_destroy(p);
        return st;
    }
    *out = p;
    return INTSPMV_OK;
}

extern "C" int intspmv_plan_destroy(intspmv_plan *p)
{
    if (!p) return INTSPMV_OK;
    (void)hipSetDevice(p->device); // not through the context: it may be gone already
    if (p->d_blocks) (void)hipFree(p->d_blocks);
    if (p->d_partial) (void)hipFree(p->d_partial);
    if (p->d_splits) (void)hipFree(p->d_splits);
    free(p);
    return INTSPMV_OK;
}

extern "C" int intspmv_plan_info(const intspmv_plan *p, uint32_t *nblocks, uint32_t *nsplit,
                                 uint32_t *nlong)
{
    if (!p) return INTSPMV_ERR_ARG;
    if (nblocks) *nblocks = p->nblocks;
    if (nsplit) *nsplit = p->nsplit_rows;
    if (nlong) *nlong = p->nlong_rows;
    return INTSPMV_OK;
}
