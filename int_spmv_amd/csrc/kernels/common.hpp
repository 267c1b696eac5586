// common.hpp -- shared device helpers and host-side structures of libintspmv.
//
// Execution model (all kernels): a workgroup is WAVES wave64 wavefronts; each
// wave owns one work block of the plan and never synchronises with the other
// waves of its workgroup, so per-wave LDS regions are private and only
// wave-level ordering fences are used.  Work blocks are handed out so that
// the eight XCDs of the MI355X each stream a contiguous range of blocks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "intspmv.h"

namespace ispv {

constexpr int WAVE = 64;
constexpr int WAVES = 4;            // waves per workgroup
constexpr int BLOCK_THREADS = WAVE * WAVES;
constexpr int NUM_XCD = 8;          // MI355X: 8 XCDs, workgroups dispatched round-robin

// AXT units (tiles for th == 1, tile columns for th > 1) per normal block;
// the embedded and column-map kernels share one plan, so one constant.
#ifndef ISPV_AXT_UMAX
#define ISPV_AXT_UMAX 128
#endif

// One unit of work for one wave.  Units are the format's reduction atoms
// (CSR entries, AXC bricks, AXT tiles for th==1, AXT tile columns for th>1,
// AXT COM slots).  Rows [rbeg, rend) own units [ubeg, uend).
//   k == 0 : normal block, whole rows, each summed by one lane in unit order;
//   k == 1 : one long row, streamed by one wave into per-lane registers and
//            reduced by a wave tree, written directly;
//   k >= 2 : piece of a long row split over k consecutive blocks b0..b0+k-1;
//            each piece leaves its partial sum in the plan's scratch and the
//            split fix-up kernel adds them in piece order.
//   piece   (k >= 2): index of this piece among all split pieces of the plan
//            (its slot in the partial-sum scratch).
struct BlockDesc {
    uint32_t ubeg, uend, rbeg, rend;
    uint32_t b0, k, piece, pad1;
};
// Partial sums of split pieces: PARTIAL_STRIDE slots per piece, one per
// right-hand side (multi-RHS products use up to 32).
constexpr uint32_t PARTIAL_STRIDE = 32;
// Column-map plans at tile height 1 (single matrix and batch alike, so both
// give the same bits): rows longer than this many tiles are split into pieces
// of one wave each.  Measured on MI355X (FP64 k = 1 / k = 8, FP32 k = 32, us):
// power-law 2^20 rows 221 / 449 / 781 at 1024, 221 / 441 / 747 at 256,
// 205 / 429 / 731 at 128; R-MAT 2^22 rows 490 / 1066 / 2121 at 1024,
// 515 / 1127 / 2242 at 256, 529 / 1165 / 2291 at 128.
#ifndef ISPV_GATHER_H1_LONG_MAX
#define ISPV_GATHER_H1_LONG_MAX 1024
#endif
constexpr uint32_t GATHER_H1_LONG_MAX = ISPV_GATHER_H1_LONG_MAX;

// Maps the flat workgroup id so that XCD x (which receives ids x, x+8, ...)
// gets a contiguous range [x*q + min(x, r), ...) of logical workgroups.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t g, uint32_t n)
{
    uint32_t q = n / NUM_XCD, r = n % NUM_XCD, x = g % NUM_XCD, l = g / NUM_XCD;
    return x * q + (x < r ? x : r) + l;
}

// Block index of this wave.  The first nlong_wg workgroups hold the long-row
// blocks (longest first, dispatched round-robin over the XCDs so they start
// early everywhere); the rest are the normal blocks, XCD-remapped.  nlong_wg
// is a multiple of NUM_XCD so both sections keep hardware id % 8 == XCD.
__device__ __forceinline__ uint32_t block_index(uint32_t nlong_wg)
{
    uint32_t g = blockIdx.x;
    uint32_t lg = g < nlong_wg ? g : nlong_wg + xcd_remap(g - nlong_wg, gridDim.x - nlong_wg);
    // wave-uniform by construction; readfirstlane lets the block descriptor
    // come in through scalar loads
    return __builtin_amdgcn_readfirstlane(lg * WAVES + threadIdx.x / WAVE);
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// ---- streaming loads -------------------------------------------------------
// Matrix arrays are read exactly once per product: nontemporal loads keep
// them from displacing the gathered / reused data in L2 and the MALL.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef uint32_t u2v __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ T ntload(const T *p) { return __builtin_nontemporal_load(p); }

// Keeps the loads issued above the call ahead of every instruction below it.
// The scheduler otherwise hoists an early consumer between them, and that
// consumer's wait for the first load holds back the loads after it.
__device__ __forceinline__ void loads_issued() { __builtin_amdgcn_sched_barrier(0); }

// Orders this wave's LDS accesses before/after the call (no cross-wave sync).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// ---- cross-lane moves (DPP within a quad, ds_bpermute otherwise) ----------
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <typename T, int CTRL>
__device__ __forceinline__ T dpp_any(T v)
{
    if constexpr (sizeof(T) == 8) return dpp_f64<CTRL>(v);
    else return dpp_f32<CTRL>(v);
}
constexpr int DPP_QUAD_XOR1 = 0xB1; // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E; // quad_perm [2,3,0,1]
constexpr int DPP_ROW_SHL4 = 0x104; // lane i reads lane i+4 of its row of 16
constexpr int DPP_ROW_SHL2 = 0x102; // lane i reads lane i+2 of its row of 16
constexpr int DPP_ROW_SHL8 = 0x108; // lane i reads lane i+8 of its row of 16
constexpr int DPP_ROW_SHR4 = 0x114; // lane i reads lane i-4 of its row of 16

__device__ __forceinline__ double shfl_xor_f64(double v, int m) { return __shfl_xor(v, m, WAVE); }
__device__ __forceinline__ float shfl_xor_f32(float v, int m) { return __shfl_xor(v, m, WAVE); }

__device__ __forceinline__ double wave_sum_f64(double v)
{
    v += dpp_f64<DPP_QUAD_XOR1>(v);
    v += dpp_f64<DPP_QUAD_XOR2>(v);
#pragma unroll
    for (int m = 4; m < WAVE; m <<= 1) v += shfl_xor_f64(v, m);
    return v;
}
__device__ __forceinline__ float wave_sum_f32(float v)
{
    v += dpp_f32<DPP_QUAD_XOR1>(v);
    v += dpp_f32<DPP_QUAD_XOR2>(v);
#pragma unroll
    for (int m = 4; m < WAVE; m <<= 1) v += shfl_xor_f32(v, m);
    return v;
}

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// ---- bf16 ----------------------------------------------------------------
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---- split rows ----------------------------------------------------------
// A row longer than one block is cut into k pieces b0..b0+k-1; each piece
// writes its partial sum to partial[block] and a second, small kernel adds
// the k partials in piece order once the main kernel has finished (the kernel
// boundary orders the writes; no device-wide fences in the streaming kernel).
// b0: piece index of the row's first piece (pieces are consecutive).
struct SplitDesc {
    uint32_t row, b0, k, pad;
};

} // namespace ispv

// ---- host-side objects -----------------------------------------------------
struct intspmv_ctx {
    int device;
    hipStream_t stream;
    hipError_t last_err;
    double *noalloc_scratch; // this device's scratch array (spmv_axt_h1_noalloc)
};

enum PlanKind { PLAN_CSR = 1, PLAN_AXC = 2, PLAN_AXT = 3, PLAN_K1 = 4 };

struct intspmv_plan {
    intspmv_ctx *ctx;
    int device; // the context's device, so destroying needs nothing of the context
    int kind;
    uint32_t nrows, nunits, nblocks, nsplit_rows, nlong_rows;
    uint32_t nlong_wg;  // workgroups of the long-row section (multiple of 8)
    uint32_t umax, rmax, long_units;
    ispv::BlockDesc *d_blocks;
    double *d_partial;          // [npieces * PARTIAL_STRIDE] (K1: [nitems * 8])
    uint32_t npieces;
    ispv::SplitDesc *d_splits;  // [nsplit_rows]
    uint32_t chunkNum;  // K1
    uint32_t exact;     // CSR: long rows unsplit, each one fma chain in CSR order
    intspmv_axt_desc axt;
};

// internal helpers implemented in abi.hip
int ispv_enter(intspmv_ctx *ctx);
int ispv_launch_check(intspmv_ctx *ctx);
int ispv_plan_upload(intspmv_ctx *ctx, intspmv_plan *p, const void *blocks, size_t block_bytes,
                     size_t nblocks, size_t partial_per_block, const ispv::SplitDesc *splits,
                     size_t nsplit);
int ispv_split_fixup(const intspmv_plan *p, void *y, int y_is_f32, uint32_t k = 1,
                     uint32_t ldy = 1, uint32_t j0 = 0);
// ngroups > 0: rows in groups [gbeg[g], gbeg[g+1]) pinned to XCD g % 8
// (ispv_build_blocks_grouped).
int ispv_plan_build(intspmv_ctx *ctx, int kind, uint32_t nrows, const uint32_t *uptr,
                    uint32_t umax, uint32_t rmax, uint32_t long_units, uint32_t long_max,
                    intspmv_plan **out, uint32_t ngroups = 0, const uint32_t *gbeg = nullptr);
#include <vector>
// host-only block builder (abi.hip), shared by the single-matrix and batch plans
// True when more than a quarter of the rows own no unit (uptr[r] == uptr[r+1]).
// Such matrices amortise the per-block row bookkeeping (row ranges, zero
// fill) better with larger blocks.
inline bool ispv_many_empty_rows(const uint32_t *uptr, uint32_t nrows)
{
    uint64_t empty = 0;
    for (uint32_t r = 0; r < nrows; r++) empty += uptr[r + 1] == uptr[r];
    return empty * 4 > nrows;
}
int ispv_build_blocks(uint32_t nrows, const uint32_t *uptr, uint32_t umax, uint32_t rmax,
                      uint32_t long_units, uint32_t long_max, std::vector<ispv::BlockDesc> &blocks,
                      std::vector<ispv::SplitDesc> &splits, uint32_t &nlong_wg, uint32_t &nlong,
                      uint32_t &npieces);
int ispv_build_blocks_grouped(uint32_t nrows, const uint32_t *uptr, uint32_t umax, uint32_t rmax,
                              uint32_t long_units, uint32_t long_max, uint32_t ngroups,
                              const uint32_t *gbeg, std::vector<ispv::BlockDesc> &blocks,
                              std::vector<ispv::SplitDesc> &splits, uint32_t &nlong_wg,
                              uint32_t &nlong, uint32_t &npieces);
// grid size of a block plan
static inline unsigned ispv_grid(const intspmv_plan *p)
{
    return p->nlong_wg + (p->nblocks - p->nlong_wg * ispv::WAVES + ispv::WAVES - 1) / ispv::WAVES;
}
