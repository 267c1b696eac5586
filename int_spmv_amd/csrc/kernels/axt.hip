// axt.hip -- AXT NOC kernels (embedded x) and the spmv_axt entry point.
//
// Reference kernels replaced: int_axt_noc_h1 (intSpmv_main.c:1385-1404) and
// int_axt_noc (intSpmv_main.c:1443-1474).  Layout (getArraysAxSecAXT_NOC*,
// intSpmv_main.c:1152-1221): tile T of height th holds th brick-rows of
// 8 values followed by the 8 embedded x values; lane C of the tile is a
// column; for th == 1 sec[T] is the row of the whole tile, for th > 1
// sec[8T+C] is the row of column C.
//
// Wave mapping on gfx950: four lanes per tile, each lane streaming two
// adjacent lanes of the tile (16-byte loads of values and of x for FP64), so a
// wave covers 16 tiles per step with fully used 64-byte runs.
//   th == 1: the tile sum is the reference's reduce tree
//            ((p0+p4)+(p2+p6)) + ((p1+p5)+(p3+p7)), formed with two DPP
//            quad exchanges (xor 2, xor 1).
//   th  > 1: each lane keeps two column sums, one fused multiply-add per
//            brick-row, in brick-row order (the reference's lane-wise
//            vertical sum).
// Unit sums go to LDS and phase 2 (reduce.hpp) forms the rows.
#include "common.hpp"
#include "reduce.hpp"

namespace ispv {

#ifndef ISPV_H1_STEPS
#define ISPV_H1_STEPS 4
#endif
// FP64 reference layout, th = 1: blocks of 64 tiles (8 KB), all in flight at
// once, and rows above 48 tiles on waves of their own (measured against 128 /
// 64: 85.3 -> 83.8 us on the power-law matrix, 173.7 -> 168.0 us on the
// stencil); matrices with many empty rows keep 128 / 64.
#ifndef ISPV_H1_F64_UMAX
#define ISPV_H1_F64_UMAX 64
#endif
#ifndef ISPV_H1_F64_LONG
#define ISPV_H1_F64_LONG 48
#endif
#ifndef ISPV_TH2_UMAX
#define ISPV_TH2_UMAX 256 // th = 2, FP64: 256 columns = 8 KB per block
#endif
constexpr uint32_t TH2_UMAX = ISPV_TH2_UMAX;
// FP32 and bf16 embedded-x storage at th = 2, 4: blocks of RED_TH_UMAX columns
// (256: 4-8 KB per block rather than 2-4 KB); the kernels size their LDS for it.
#ifndef ISPV_RED_TH_UMAX
#define ISPV_RED_TH_UMAX 256
#endif
__host__ __device__ constexpr uint32_t red_umax(int th)
{
    return (th == 2 || th == 4) ? ISPV_RED_TH_UMAX : ISPV_AXT_UMAX;
}
#ifndef ISPV_TALL_UMAX
#define ISPV_TALL_UMAX 32 // th >= 16, FP64 reference layout, dense rows (32 columns of 2+ KB)
#endif
#ifndef ISPV_TH8_UMAX
#define ISPV_TH8_UMAX 64 // th = 8, FP64 and FP32 reference layouts
#endif
#ifndef ISPV_H1_RED_UMAX
#define ISPV_H1_RED_UMAX 128 // th = 1, FP32 / bf16 storage and the column-map form
#endif
constexpr uint32_t AXT_UMAX = ISPV_AXT_UMAX; // units (tiles or columns) per block
constexpr uint32_t AXT_RMAX = 256;  // rows per block
static_assert(AXT_RMAX <= BLOCK_ROWS_MAX, "keyed blocks map their rows in a 256-bit LDS map");
// A long row is streamed by one wave up to this many slots (128 KB of FP64
// interleaved tiles); longer rows are split across waves.
#ifndef ISPV_LONG_SLOTS
#define ISPV_LONG_SLOTS 8192
#endif
#ifndef ISPV_H1_LONG
#define ISPV_H1_LONG 64
#endif
constexpr uint32_t LONG_SLOTS = ISPV_LONG_SLOTS;

template <typename T>
struct Pair;
template <>
struct Pair<double> {
    __device__ __forceinline__ static void ld(const double *p, double &a, double &b)
    {
        d2v v = ntload(reinterpret_cast<const d2v *>(p));
        a = v[0];
        b = v[1];
    }
};
template <>
struct Pair<float> {
    __device__ __forceinline__ static void ld(const float *p, float &a, float &b)
    {
        f2v v = ntload(reinterpret_cast<const f2v *>(p));
        a = v[0];
        b = v[1];
    }
};
template <>
struct Pair<uint16_t> { // bf16 storage
    __device__ __forceinline__ static void ld(const uint16_t *p, float &a, float &b)
    {
        uint32_t v = ntload(reinterpret_cast<const uint32_t *>(p));
        a = __uint_as_float(v << 16);
        b = __uint_as_float(v & 0xFFFF0000u);
    }
};

template <typename VT>
using LoadT = typename std::conditional<std::is_same<VT, uint16_t>::value, float, VT>::type;

template <typename AT>
__device__ __forceinline__ AT fma_acc(AT a, AT b, AT c);
template <>
__device__ __forceinline__ double fma_acc<double>(double a, double b, double c) { return fma(a, b, c); }
template <>
__device__ __forceinline__ float fma_acc<float>(float a, float b, float c) { return fmaf(a, b, c); }

template <typename AT>
__device__ __forceinline__ AT quad_xor2(AT v);
template <>
__device__ __forceinline__ double quad_xor2<double>(double v) { return dpp_f64<DPP_QUAD_XOR2>(v); }
template <>
__device__ __forceinline__ float quad_xor2<float>(float v) { return dpp_f32<DPP_QUAD_XOR2>(v); }
template <typename AT>
__device__ __forceinline__ AT quad_xor1(AT v);
template <>
__device__ __forceinline__ double quad_xor1<double>(double v) { return dpp_f64<DPP_QUAD_XOR1>(v); }
template <>
__device__ __forceinline__ float quad_xor1<float>(float v) { return dpp_f32<DPP_QUAD_XOR1>(v); }

// Element index of value slot (T, R, C) and of its x; interleaved: one array
// with 16 elements per brick-row, planar: two arrays with 8 per brick-row.
template <bool PLANAR>
__device__ __forceinline__ size_t vidx(size_t T, uint32_t th, uint32_t R, uint32_t C)
{
    return PLANAR ? (T * th + R) * 8 + C : (T * th + R) * 16 + C;
}
template <bool PLANAR>
__device__ __forceinline__ size_t xidx(size_t T, uint32_t th, uint32_t R, uint32_t C)
{
    return PLANAR ? (T * th + R) * 8 + C : (T * th + R) * 16 + 8 + C;
}

// Long row of th==1 tiles [ubeg, uend): every lane keeps its own products'
// running sum; one wave tree at the end.  UL steps of 16 tiles in flight.
__device__ __forceinline__ float4 f4load(const float *p)
{
    typedef float f4v __attribute__((ext_vector_type(4)));
    f4v v = ntload(reinterpret_cast<const f4v *>(p));
    return make_float4(v[0], v[1], v[2], v[3]);
}
template <int CTRL>
__device__ __forceinline__ float4 f4dpp(float4 v)
{
    return make_float4(dpp_f32<CTRL>(v.x), dpp_f32<CTRL>(v.y), dpp_f32<CTRL>(v.z), dpp_f32<CTRL>(v.w));
}
// Sum of one planar tile (8 values, 8 x) read whole by one lane.
template <typename VT, typename XT, typename AT>
__device__ __forceinline__ AT planar_tile(const VT *av, const XT *xp, size_t T)
{
    AT v[8], x[8];
    if constexpr (sizeof(VT) == 2) {
        typedef uint32_t u4v __attribute__((ext_vector_type(4)));
        u4v b = ntload(reinterpret_cast<const u4v *>(av + T * 8));
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v[2 * i] = __uint_as_float(b[i] << 16);
            v[2 * i + 1] = __uint_as_float(b[i] & 0xFFFF0000u);
        }
    } else if constexpr (sizeof(VT) == 4) {
        float4 a = f4load(reinterpret_cast<const float *>(av) + T * 8);
        float4 b = f4load(reinterpret_cast<const float *>(av) + T * 8 + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
            LoadT<VT> a, b;
            Pair<VT>::ld(av + T * 8 + i, a, b);
            v[i] = (AT)a;
            v[i + 1] = (AT)b;
        }
    }
    if constexpr (sizeof(XT) == 4) {
        float4 a = f4load(reinterpret_cast<const float *>(xp) + T * 8);
        float4 b = f4load(reinterpret_cast<const float *>(xp) + T * 8 + 4);
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
        x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
    } else {
#pragma unroll
        for (int i = 0; i < 8; i += 2) {
            LoadT<XT> a, b;
            Pair<XT>::ld(xp + T * 8 + i, a, b);
            x[i] = (AT)a;
            x[i + 1] = (AT)b;
        }
    }
    // same halving tree as the reference's reduce (any order is allowed for
    // the reduced-precision storage types)
    AT p[8];
#pragma unroll
    for (int i = 0; i < 8; i++) p[i] = v[i] * x[i];
    return ((p[0] + p[4]) + (p[2] + p[6])) + ((p[1] + p[5]) + (p[3] + p[7]));
}

#ifndef ISPV_H1_LONG_STEPS
#define ISPV_H1_LONG_STEPS 8
#endif
template <typename VT, typename XT, typename AT, bool PLANAR>
__device__ __forceinline__ AT h1_long(const VT *av, const XT *xb, const BlockDesc &d)
{
    constexpr int UL = ISPV_H1_LONG_STEPS;
    const int lane = lane_id();
    if constexpr (!PLANAR && sizeof(VT) == 8) {
        // full-line mapping as in the normal path: 8 tiles per load
        const uint32_t nu = d.uend - d.ubeg, p = lane & 7;
        double acc = 0.0;
        for (uint32_t base = 0; base < nu; base += 8 * 2 * UL) {
            d2v v[2 * UL];
#pragma unroll
            for (int j = 0; j < 2 * UL; j++) {
                uint32_t t = base + j * 8 + lane / 8;
                size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
                v[j] = ntload(reinterpret_cast<const d2v *>(av + T * 16 + 2 * p));
            }
            loads_issued();
#pragma unroll
            for (int j = 0; j < 2 * UL; j++) {
                double x0 = dpp_f64<DPP_ROW_SHL4>(v[j][0]), x1 = dpp_f64<DPP_ROW_SHL4>(v[j][1]);
                if (base + j * 8 + lane / 8 < nu && p < 4) acc = acc + v[j][0] * x0 + v[j][1] * x1;
            }
        }
        return (AT)wave_sum_f64(acc);
    }
    if constexpr (!PLANAR && sizeof(VT) == 4) {
        // FP32 interleaved: a 128-byte line holds two 64-byte tiles; part
        // p = l%8 of line l/8: values (p%4 < 2) or x (p%4 >= 2), 4 each
        const uint32_t nu = d.uend - d.ubeg, p = lane & 7;
        LaneSum<float> acc;
        for (uint32_t base = 0; base < nu; base += 16 * UL) {
            float4 v[UL];
#pragma unroll
            for (int j = 0; j < UL; j++) {
                uint32_t t = base + j * 16 + 2 * (lane / 8) + p / 4;
                size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
                v[j] = f4load(reinterpret_cast<const float *>(av) + T * 16 + 4 * (p % 4));
            }
            loads_issued();
#pragma unroll
            for (int j = 0; j < UL; j++) {
                float4 xx = f4dpp<DPP_ROW_SHL2>(v[j]);
                if (base + j * 16 + 2 * (lane / 8) + p / 4 < nu && p % 4 < 2)
                    acc.add(v[j].x * xx.x + v[j].y * xx.y + (v[j].z * xx.z + v[j].w * xx.w));
            }
        }
        return (AT)wave_sum_f32(acc.value());
    }
    if constexpr (PLANAR) {
        // planar: one lane per tile, the tile's 8 values and 8 x whole
        const uint32_t nu = d.uend - d.ubeg;
        LaneSum<AT> acc;
        for (uint32_t base = 0; base < nu; base += 64 * UL / 2) {
            AT ts[UL / 2];
#pragma unroll
            for (int j = 0; j < UL / 2; j++) {
                uint32_t t = base + j * 64 + lane;
                ts[j] = planar_tile<VT, XT, AT>(av, xb, (size_t)d.ubeg + (t < nu ? t : nu - 1));
            }
#pragma unroll
            for (int j = 0; j < UL / 2; j++)
                if (base + j * 64 + lane < nu) acc.add(ts[j]);
        }
        return wave_sum<AT>(acc.value());
    }
    const uint32_t nu = d.uend - d.ubeg, C = 2 * (lane & 3);
    LaneSum<AT> acc;
    for (uint32_t base = 0; base < nu; base += 16 * UL) {
        LoadT<VT> v0[UL], v1[UL];
        LoadT<XT> x0[UL], x1[UL];
#pragma unroll
        for (int j = 0; j < UL; j++) {
            uint32_t t = base + j * 16 + lane / 4;
            size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
            Pair<VT>::ld(av + vidx<PLANAR>(T, 1, 0, C), v0[j], v1[j]);
            Pair<XT>::ld(xb + xidx<PLANAR>(T, 1, 0, C), x0[j], x1[j]);
        }
#pragma unroll
        for (int j = 0; j < UL; j++)
            if (base + j * 16 + lane / 4 < nu) {
                acc.add((AT)v0[j] * (AT)x0[j]);
                acc.add((AT)v1[j] * (AT)x1[j]);
            }
    }
    return wave_sum<AT>(acc.value());
}

// A normal block's sec words (one per unit, nu <= UMAX), one lane per unit:
// load() issues one or two coalesced loads, store() puts them in LDS.  The
// kernels load the keys before their value loads and store them after the
// values are consumed, so that no wait for the keys holds the value loads
// back (the keys arrive first: loads return in order).
template <uint32_t UMAX = AXT_UMAX>
struct BlockKeys {
    static_assert(UMAX % WAVE == 0, "key staging");
    uint32_t kk[UMAX / WAVE];
    __device__ __forceinline__ void load(const uint32_t *sec, const BlockDesc &d, uint32_t nu)
    {
        if (nu == 0) return;
        const int lane = lane_id();
        kk[0] = ntload(sec + d.ubeg + (lane < (int)nu ? lane : nu - 1));
#pragma unroll
        for (int j = 1; j < (int)(UMAX / WAVE); j++) {
            const uint32_t t = j * WAVE + lane;
            if (j * WAVE < (int)nu) kk[j] = ntload(sec + d.ubeg + (t < nu ? t : nu - 1)); // uniform
        }
    }
    __device__ __forceinline__ void store(uint32_t nu, uint32_t *sk) const
    {
        const int lane = lane_id();
#pragma unroll
        for (int j = 0; j < (int)(UMAX / WAVE); j++)
            if (j * WAVE + lane < nu) sk[j * WAVE + lane] = kk[j];
    }
};

#ifdef ISPV_WAVES_PER_EU
#define ISPV_OCC __attribute__((amdgpu_waves_per_eu(ISPV_WAVES_PER_EU)))
#else
#define ISPV_OCC
#endif
template <typename VT, typename XT, typename AT, bool PLANAR>
__global__ __launch_bounds__(BLOCK_THREADS) ISPV_OCC void axt_h1_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const VT *__restrict__ av, const XT *__restrict__ axv, const uint32_t *__restrict__ sec,
    AT *__restrict__ y, uint32_t nrows, double *partial)
{
    __shared__ AT s_val[WAVES][AXT_UMAX];
    __shared__ uint32_t s_key[WAVES][AXT_UMAX];
    __shared__ uint16_t s_seg[WAVES][AXT_UMAX + 1];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    const XT *xb = PLANAR ? axv : (const XT *)av;
    if (d.k != 0) {
        finish_long<AT>(h1_long<VT, XT, AT, PLANAR>(av, xb, d), d, bid, y, nrows, partial);
        return;
    }
    const uint32_t nu = d.uend - d.ubeg;
    if constexpr (!PLANAR && sizeof(VT) == 8) {
        // FP64 reference layout: each load instruction covers 8 whole 128-byte
        // tiles (1 KB, full cache lines).  Lane part p = l%8 holds values
        // 2p,2p+1 (p<4) or x 2(p-4),2(p-4)+1 (p>=4); one DPP row shift brings
        // the x pair to the value lanes, whose quads then form the tree.  A
        // 64-tile block is in flight at once (8 loads).
        // Blocks of up to AXT_UMAX tiles (plans of matrices with many empty
        // rows) have all their rounds of U loads in flight together; a round
        // past the block's tiles (every round of a block of empty rows, whose
        // clamp tile nu - 1 does not exist) loads nothing.
        const uint32_t p = lane & 7;
        constexpr int U = ISPV_H1_F64_UMAX < 64 ? ISPV_H1_F64_UMAX / 8 : 8;
        constexpr int NR = AXT_UMAX / (8 * U);
        BlockKeys<AXT_UMAX> keys;
        keys.load(sec, d, nu);
        d2v v[NR][U];
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (nu <= 8 * U * r) break; // uniform
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t t = 8 * U * r + j * 8 + lane / 8;
                size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
                v[r][j] = ntload(reinterpret_cast<const d2v *>(av + T * 16 + 2 * p));
            }
        }
        loads_issued();
#pragma unroll
        for (int r = 0; r < NR; r++) {
            if (nu <= 8 * U * r) break; // uniform
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t t = 8 * U * r + j * 8 + lane / 8;
                const bool ok = t < nu && p < 4;
                double x0 = dpp_f64<DPP_ROW_SHL4>(v[r][j][0]), x1 = dpp_f64<DPP_ROW_SHL4>(v[r][j][1]);
                double p0 = ok ? v[r][j][0] * x0 : 0.0, p1 = ok ? v[r][j][1] * x1 : 0.0;
                p0 = p0 + dpp_f64<DPP_QUAD_XOR2>(p0);
                p1 = p1 + dpp_f64<DPP_QUAD_XOR2>(p1);
                p0 = p0 + dpp_f64<DPP_QUAD_XOR1>(p0);
                p1 = p1 + dpp_f64<DPP_QUAD_XOR1>(p1);
                if (t < nu && p == 0) s_val[w][t] = (AT)(p0 + p1);
            }
        }
        keys.store(nu, s_key[w]);
        wave_lds_sync();
        reduce_keyed<AT>(s_val[w], s_key[w], s_seg[w], nu, d, bid, y, nrows, partial);
        return;
    }
    if constexpr (!PLANAR && sizeof(VT) == 4) {
        // FP32 interleaved, two 64-byte tiles per 128-byte line (see h1_long)
        const uint32_t p = lane & 7;
        constexpr int U = 2 * ISPV_H1_STEPS;
        BlockKeys<AXT_UMAX> keys;
        keys.load(sec, d, nu);
        for (uint32_t base = 0; base < nu; base += 16 * U) {
            float4 v[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t t = base + j * 16 + 2 * (lane / 8) + p / 4;
                size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
                v[j] = f4load(reinterpret_cast<const float *>(av) + T * 16 + 4 * (p % 4));
            }
            loads_issued();
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t t = base + j * 16 + 2 * (lane / 8) + p / 4;
                float4 xx = f4dpp<DPP_ROW_SHL2>(v[j]);
                float ps = (v[j].x * xx.x + v[j].y * xx.y) + (v[j].z * xx.z + v[j].w * xx.w);
                ps = ps + dpp_f32<DPP_QUAD_XOR1>(ps); // values 0-3 + values 4-7
                if (t < nu && p % 4 == 0) s_val[w][t] = (AT)ps;
            }
        }
        keys.store(nu, s_key[w]);
        wave_lds_sync();
        reduce_keyed<AT>(s_val[w], s_key[w], s_seg[w], nu, d, bid, y, nrows, partial);
        return;
    }
    if constexpr (PLANAR) {
        // planar storage: one lane per tile
        constexpr int U = ISPV_H1_STEPS;
        BlockKeys<AXT_UMAX> keys;
        keys.load(sec, d, nu);
        for (uint32_t base = 0; base < nu; base += 64 * U) {
            AT ts[U];
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t t = base + j * 64 + lane;
                size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
                ts[j] = planar_tile<VT, XT, AT>(av, xb, T);
            }
#pragma unroll
            for (int j = 0; j < U; j++) {
                uint32_t t = base + j * 64 + lane;
                if (t < nu) s_val[w][t] = ts[j];
            }
        }
        keys.store(nu, s_key[w]);
        wave_lds_sync();
        reduce_keyed<AT>(s_val[w], s_key[w], s_seg[w], nu, d, bid, y, nrows, partial);
        return;
    }
    const uint32_t q = lane & 3, C = 2 * q;
    constexpr int U = ISPV_H1_STEPS; // steps of 16 tiles in flight per lane
    for (uint32_t base = 0; base < nu; base += 16 * U) {
        LoadT<VT> v0[U], v1[U];
        LoadT<XT> x0[U], x1[U];
        uint32_t key[U];
        // branch-free loads: lanes past the block re-read its last tile and
        // are masked afterwards, so all loads of the batch issue back to back
#pragma unroll
        for (int j = 0; j < U; j++) {
            uint32_t t = base + j * 16 + lane / 4;
            size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
            Pair<VT>::ld(av + vidx<PLANAR>(T, 1, 0, C), v0[j], v1[j]);
            Pair<XT>::ld(xb + xidx<PLANAR>(T, 1, 0, C), x0[j], x1[j]);
            key[j] = ntload(sec + T);
        }
        loads_issued();
#pragma unroll
        for (int j = 0; j < U; j++) {
            const bool ok = base + j * 16 + lane / 4 < nu;
            AT p0 = ok ? (AT)v0[j] * (AT)x0[j] : (AT)0, p1 = ok ? (AT)v1[j] * (AT)x1[j] : (AT)0;
            p0 = p0 + quad_xor2<AT>(p0);
            p1 = p1 + quad_xor2<AT>(p1);
            p0 = p0 + quad_xor1<AT>(p0);
            p1 = p1 + quad_xor1<AT>(p1);
            AT s = p0 + p1;
            uint32_t t = base + j * 16 + lane / 4;
            if (t < nu && q == 0) {
                s_val[w][t] = s;
                s_key[w][t] = key[j];
            }
        }
    }
    wave_lds_sync();
    reduce_keyed<AT>(s_val[w], s_key[w], s_seg[w], nu, d, bid, y, nrows, partial);
}

// Column sums of the tile group starting at tile Tb for this lane's two
// columns (fma over the th brick-rows in order).
template <typename VT, typename XT, typename AT, bool PLANAR, int TH>
__device__ __forceinline__ void noc_columns(const VT *av, const XT *xb, uint32_t T, uint32_t th,
                                            uint32_t C, AT &a0, AT &a1)
{
    a0 = 0;
    a1 = 0;
    if constexpr (TH > 0) {
        LoadT<VT> v0[TH], v1[TH];
        LoadT<XT> x0[TH], x1[TH];
#pragma unroll
        for (int R = 0; R < TH; R++) {
            Pair<VT>::ld(av + vidx<PLANAR>(T, TH, R, C), v0[R], v1[R]);
            Pair<XT>::ld(xb + xidx<PLANAR>(T, TH, R, C), x0[R], x1[R]);
        }
#pragma unroll
        for (int R = 0; R < TH; R++) {
            a0 = fma_acc<AT>((AT)v0[R], (AT)x0[R], a0);
            a1 = fma_acc<AT>((AT)v1[R], (AT)x1[R], a1);
        }
    } else {
        for (uint32_t R = 0; R < th; R++) {
            LoadT<VT> v0, v1;
            LoadT<XT> x0, x1;
            Pair<VT>::ld(av + vidx<PLANAR>(T, th, R, C), v0, v1);
            Pair<XT>::ld(xb + xidx<PLANAR>(T, th, R, C), x0, x1);
            a0 = fma_acc<AT>((AT)v0, (AT)x0, a0);
            a1 = fma_acc<AT>((AT)v1, (AT)x1, a1);
        }
    }
}

template <typename VT, typename XT, typename AT, bool PLANAR, int TH>
__global__ __launch_bounds__(BLOCK_THREADS) void axt_noc_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const VT *__restrict__ av, const XT *__restrict__ axv, const uint32_t *__restrict__ sec,
    AT *__restrict__ y, uint32_t nrows, double *partial, uint32_t th_rt)
{
    __shared__ AT s_val[WAVES][AXT_UMAX];
    __shared__ uint32_t s_key[WAVES][AXT_UMAX];
    __shared__ uint16_t s_seg[WAVES][AXT_UMAX + 1];
    const uint32_t th = TH ? TH : th_rt;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    const uint32_t nu = d.uend - d.ubeg;
    const uint32_t C = 2 * (lane & 3);
    const uint32_t T0 = d.ubeg / 8, T1 = (d.uend + 7) / 8;
    const XT *xb = PLANAR ? axv : (const XT *)av;
    LaneSum<AT> lacc; // long rows
    BlockKeys<AXT_UMAX> keys;
    if (d.k == 0) keys.load(sec, d, nu);
    for (uint32_t Tb = T0; Tb < T1; Tb += 16) {
        const uint32_t Tr = Tb + lane / 4, u0 = 8 * Tr + C;
        const bool ok0 = u0 >= d.ubeg && u0 < d.uend;
        const bool ok1 = u0 + 1 >= d.ubeg && u0 + 1 < d.uend;
        // lanes past the block re-read its last tile (results discarded) so
        // that the loads stay branch-free
        const uint32_t T = Tr < T1 ? Tr : T1 - 1;
        AT a0, a1;
        noc_columns<VT, XT, AT, PLANAR, TH>(av, xb, T, th, C, a0, a1);
        if (d.k != 0) {
            if (ok0) lacc.add(a0);
            if (ok1) lacc.add(a1);
            continue;
        }
        if (ok0) s_val[w][u0 - d.ubeg] = a0;
        if (ok1) s_val[w][u0 + 1 - d.ubeg] = a1;
    }
    if (d.k != 0) {
        finish_long<AT>(wave_sum<AT>(lacc.value()), d, bid, y, nrows, partial);
        return;
    }
    keys.store(nu, s_key[w]);
    wave_lds_sync();
    reduce_keyed<AT>(s_val[w], s_key[w], s_seg[w], nu, d, bid, y, nrows, partial);
}

// AXT COM (tile height 1, exact CSR order, int_axt_com_h1,
// intSpmv_main.c:1517-1566).  Slot s of the matrix is CSR entry s; sec[s] is a
// head word (row << log | o) at the first slot of a row and at every slot
// with s % bs == 0, and the segment it heads runs to the next head.  A head is
// recognised by sec[s] != 0 or s % bs == 0: the second test also catches
// row 0 with o == 0, whose word is 0 and which the reference kernel skips.
// Blocks start at row starts, so a block's rows are the runs between heads
// whose row differs from the previous head's, found for all slots at once by
// ballots of the head flags.  Each row is then one lane's fused multiply-add
// chain over its slots in CSR order from +0.0 -- the same bits as cf_CSR.
#ifndef ISPV_COM_UMAX
#define ISPV_COM_UMAX 256
#endif
constexpr uint32_t COM_UMAX = ISPV_COM_UMAX;
// tiles per load step: 5 x 8 = 40 covers a block of COM_UMAX slots at any
// alignment (at most 33 tiles), so a normal block is one step in flight
constexpr int COM_U = (COM_UMAX + 7) / 64 + 1;
constexpr uint32_t COM_NOHEAD = 0xFFFFFFFFu;
#ifndef ISPV_COM_CHAIN
#define ISPV_COM_CHAIN 8
#endif
constexpr int COM_CHAIN = ISPV_COM_CHAIN;

template <typename AT>
struct alignas(2 * sizeof(AT)) ComSlot {
    AT v, x;
};

template <typename VT, typename XT, typename AT>
__global__ __launch_bounds__(BLOCK_THREADS) void axt_com_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks, const VT *__restrict__ ax,
    const uint32_t *__restrict__ sec, AT *__restrict__ y, uint32_t nrows, double *partial,
    uint32_t bs, uint32_t lg)
{
    __shared__ ComSlot<AT> s_vx[WAVES][COM_UMAX];
    __shared__ uint32_t s_key[WAVES][COM_UMAX];
    __shared__ uint16_t s_seg[WAVES][COM_UMAX + 1];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    const uint32_t nu = d.uend - d.ubeg;
    const uint32_t T0 = d.ubeg / 8, T1 = (d.uend + 7) / 8;
    // full-line mapping: 8 whole tiles per load, lane part p = l%8 reads
    // doubles 2p, 2p+1 of its tile (values of slots 2p, 2p+1 for p < 4, the
    // x of slots 2p-8, 2p-7 for p >= 4)
    const uint32_t p = lane & 7;
    if (d.k != 0) {
        // long rows: lane-strided partials of slots 2p, 2p+1 (p < 4), whose x
        // a DPP row shift brings; FP64 an fma chain, FP32 a compensated
        // (Neumaier) sum of the products
        AT acc = 0;
        LaneSum<AT> cacc;
        const uint32_t Cq = 2 * (p & 3);
        for (uint32_t Tb = T0; Tb < T1; Tb += 8 * COM_U) {
            LoadT<VT> v0[COM_U], v1[COM_U];
#pragma unroll
            for (int u = 0; u < COM_U; u++) {
                const uint32_t Tr = Tb + 8 * u + lane / 8;
                const uint32_t T = Tr < T1 ? Tr : T1 - 1;
                Pair<VT>::ld(ax + (size_t)T * 16 + 2 * p, v0[u], v1[u]);
            }
#pragma unroll
            for (int u = 0; u < COM_U; u++) {
                const uint32_t Tr = Tb + 8 * u + lane / 8, u0 = 8 * Tr + Cq;
                const AT x0 = (AT)dpp_any<LoadT<VT>, DPP_ROW_SHL4>(v0[u]);
                const AT x1 = (AT)dpp_any<LoadT<VT>, DPP_ROW_SHL4>(v1[u]);
                const bool ok0 = p < 4 && Tr < T1 && u0 >= d.ubeg && u0 < d.uend;
                const bool ok1 = p < 4 && Tr < T1 && u0 + 1 >= d.ubeg && u0 + 1 < d.uend;
                if constexpr (sizeof(AT) == 4) {
                    if (ok0) cacc.add((AT)v0[u] * x0);
                    if (ok1) cacc.add((AT)v1[u] * x1);
                } else {
                    if (ok0) acc = fma_acc<AT>((AT)v0[u], x0, acc);
                    if (ok1) acc = fma_acc<AT>((AT)v1[u], x1, acc);
                }
            }
        }
        if constexpr (sizeof(AT) == 4) acc = cacc.value();
        finish_long<AT>(wave_sum<AT>(acc), d, bid, y, nrows, partial);
        return;
    }
    // Normal block: at most COM_UMAX slots, all loaded in one step.  Each
    // lane stages one slot: part p < 4 slot 2p (its own value, the x of part
    // p + 4 by a left row shift), part p >= 4 slot 2p - 7 (its own x, the
    // value of part p - 4 by a right row shift).  bs is a power of two.  A
    // block without slots (T0 == T1, e.g. empty rows at the start) loads
    // nothing: its clamp tile T1 - 1 may not exist.
    if (T0 < T1) {
        const uint32_t sl = p < 4 ? 2 * p : 2 * p - 7, bm = bs - 1;
        LoadT<VT> v0[COM_U], v1[COM_U];
        uint32_t hw[COM_U];
#pragma unroll
        for (int u = 0; u < COM_U; u++) {
            const uint32_t Tr = T0 + 8 * u + lane / 8;
            const uint32_t T = Tr < T1 ? Tr : T1 - 1;
            Pair<VT>::ld(ax + (size_t)T * 16 + 2 * p, v0[u], v1[u]);
            hw[u] = ntload(sec + 8 * (size_t)T + sl);
        }
        loads_issued();
#pragma unroll
        for (int u = 0; u < COM_U; u++) {
            // slots past the block (clamped tiles included) give i >= nu
            const uint32_t s = 8 * (T0 + 8 * u + lane / 8) + sl, i = s - d.ubeg;
            // keeps the head-word load with the others (the compiler would
            // otherwise sink it into the store below, one more round trip)
            asm volatile("" : "+v"(hw[u]));
            const LoadT<VT> xl = dpp_any<LoadT<VT>, DPP_ROW_SHL4>(v0[u]);
            const LoadT<VT> vr = dpp_any<LoadT<VT>, DPP_ROW_SHR4>(v1[u]);
            if (i < nu) {
                s_vx[w][i] = {(AT)(p < 4 ? v0[u] : vr), (AT)(p < 4 ? xl : v1[u])};
                s_key[w][i] = (hw[u] != 0 || (s & bm) == 0) ? hw[u] >> lg : COM_NOHEAD;
            }
        }
    }
    wave_lds_sync();
    // Rows: a head whose row differs from the previous head's starts a row
    // (heads at multiples of bs inside a row repeat its row).  Slot i = 64c +
    // lane; all reads first, then the segment starts by ballot.
    uint32_t nseg = 0;
    {
        constexpr int NC = COM_UMAX / WAVE;
        uint32_t hk[NC], prow[NC];
        uint64_t hm[NC];
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const uint32_t i = 64 * c + lane;
            hk[c] = i < nu ? s_key[w][i] : COM_NOHEAD;
            hm[c] = __ballot(hk[c] != COM_NOHEAD);
        }
        int32_t last = -1; // last head in the chunks before c (wave-uniform)
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const uint64_t m = hm[c] & lanemask_lt();
            const int32_t pos = m ? 64 * c + 63 - __builtin_clzll(m) : last;
            prow[c] = pos >= 0 ? s_key[w][pos] : COM_NOHEAD;
            if (hm[c]) last = 64 * c + 63 - __builtin_clzll(hm[c]);
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const bool start = hk[c] != COM_NOHEAD && hk[c] != prow[c];
            const uint64_t sm = __ballot(start);
            if (start) s_seg[w][nseg + __popcll(sm & lanemask_lt())] = (uint16_t)(64 * c + lane);
            nseg += __popcll(sm);
        }
        if (lane == 0) s_seg[w][nseg] = (uint16_t)nu;
    }
    wave_lds_sync();
    for (uint32_t s = lane; s < nseg; s += WAVE) {
        AT a = 0;
        const uint32_t b = s_seg[w][s + 1];
        uint32_t i = s_seg[w][s];
        // COM_CHAIN slots read before their fmas, so LDS latency is paid once
        // per COM_CHAIN (a masked last chunk measured slower here)
        for (; i + COM_CHAIN <= b; i += COM_CHAIN) {
            ComSlot<AT> e[COM_CHAIN];
#pragma unroll
            for (int q = 0; q < COM_CHAIN; q++) e[q] = s_vx[w][i + q];
#pragma unroll
            for (int q = 0; q < COM_CHAIN; q++) a = fma_acc<AT>(e[q].v, e[q].x, a);
        }
        for (; i < b; i++) {
            const ComSlot<AT> e = s_vx[w][i];
            a = fma_acc<AT>(e.v, e.x, a);
        }
        const uint32_t r = s_key[w][s_seg[w][s]];
        if (r < nrows) y[r] = a;
    }
    zero_empty_rows(s_key[w], s_seg[w], nseg, d, nrows, [&](uint32_t z) { y[z] = (AT)0; });
}

// th > 1, FP64 reference layout, full-line mapping: lane (tile Tb + l/8,
// part p = l%8) reads 16 bytes of each brick-row of its tile, so every load
// instruction covers 8 whole 128-byte brick-rows; a DPP row shift brings the
// x pair to the value lanes (p < 4), which run the fma chains of columns
// 2p, 2p+1 over the brick-rows in order.
template <int TH>
__global__ __launch_bounds__(BLOCK_THREADS) void axt_noc_fl_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const double *__restrict__ ax, const uint32_t *__restrict__ sec, double *__restrict__ y,
    uint32_t nrows, double *partial, uint32_t th_rt)
{
    static_assert(TH <= 16 || TH % 8 == 0, "tall tiles load brick-rows in chunks of eight");
    // th = 2 blocks hold 256 columns (8 KB, like th = 4's 128): capacity per height
    constexpr uint32_t UM = TH == 2 ? TH2_UMAX : AXT_UMAX;
    __shared__ double s_val[WAVES][UM];
    __shared__ uint32_t s_key[WAVES][UM];
    __shared__ uint16_t s_seg[WAVES][UM + 1];
    const uint32_t th = TH ? TH : th_rt;
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    const uint32_t nu = d.uend - d.ubeg, p = lane & 7, C = 2 * (p & 3);
    const uint32_t T0 = d.ubeg / 8, T1 = (d.uend + 7) / 8;
    double lacc = 0.0; // long rows
    BlockKeys<UM> keys; // unit u = column 8T+C has key sec[u]
    if (d.k == 0) keys.load(sec, d, nu);
    // G groups of 8 tiles per step, so that a normal block is in flight at
    // once: th = 2: 32 tiles; th = 4: 16 tiles (th = 8, 16: 8 tiles per step)
    constexpr int G = TH == 2 ? (int)(TH2_UMAX / 64) : (TH == 0 || TH >= 8) ? 1 : 2;
    for (uint32_t Tb = T0; Tb < T1; Tb += 8 * G) {
        double a0[G], a1[G];
        if constexpr (TH > 0 && TH <= 16) {
            d2v v[G][TH];
#pragma unroll
            for (int g = 0; g < G; g++) {
                const uint32_t Tr = Tb + 8 * g + lane / 8;
                const size_t T = Tr < T1 ? Tr : T1 - 1;
#pragma unroll
                for (int R = 0; R < TH; R++)
                    v[g][R] = ntload(reinterpret_cast<const d2v *>(ax + (T * TH + R) * 16 + 2 * p));
            }
            loads_issued();
#pragma unroll
            for (int g = 0; g < G; g++) {
                a0[g] = 0.0;
                a1[g] = 0.0;
#pragma unroll
                for (int R = 0; R < TH; R++) {
                    double x0 = dpp_f64<DPP_ROW_SHL4>(v[g][R][0]), x1 = dpp_f64<DPP_ROW_SHL4>(v[g][R][1]);
                    a0[g] = fma(v[g][R][0], x0, a0[g]);
                    a1[g] = fma(v[g][R][1], x1, a1[g]);
                }
            }
        } else {
            // tall tiles: brick-rows in chunks of eight loads issued together
            // (the fma chain still runs over the brick-rows in order)
            const uint32_t Tr = Tb + lane / 8;
            const size_t T = Tr < T1 ? Tr : T1 - 1;
            a0[0] = 0.0;
            a1[0] = 0.0;
            for (uint32_t R0 = 0; R0 < th; R0 += 8) {
                d2v v[8];
#pragma unroll
                for (int r = 0; r < 8; r++)
                    if (TH > 0 || R0 + r < th) // whole chunks when TH (a multiple of 8) is known
                        v[r] = ntload(reinterpret_cast<const d2v *>(ax + (T * th + R0 + r) * 16 + 2 * p));
                loads_issued();
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    if (TH == 0 && R0 + r >= th) break;
                    const double x0 = dpp_f64<DPP_ROW_SHL4>(v[r][0]), x1 = dpp_f64<DPP_ROW_SHL4>(v[r][1]);
                    a0[0] = fma(v[r][0], x0, a0[0]);
                    a1[0] = fma(v[r][1], x1, a1[0]);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const uint32_t Tr = Tb + 8 * g + lane / 8, u0 = 8 * Tr + C;
            const bool ok0 = p < 4 && Tr < T1 && u0 >= d.ubeg && u0 < d.uend;
            const bool ok1 = p < 4 && Tr < T1 && u0 + 1 >= d.ubeg && u0 + 1 < d.uend;
            if (d.k != 0) {
                if (ok0) lacc = lacc + a0[g];
                if (ok1) lacc = lacc + a1[g];
            } else {
                if (ok0) s_val[w][u0 - d.ubeg] = a0[g];
                if (ok1) s_val[w][u0 + 1 - d.ubeg] = a1[g];
            }
        }
    }
    if (d.k != 0) {
        finish_long<double>(wave_sum_f64(lacc), d, bid, y, nrows, partial);
        return;
    }
    keys.store(nu, s_key[w]);
    wave_lds_sync();
    reduce_keyed<double>(s_val[w], s_key[w], s_seg[w], nu, d, bid, y, nrows, partial);
}

// th > 1 (even), FP32 reference layout, full-line mapping: a 128-byte line
// holds brick-rows R, R+1 of one tile (64 bytes each).  Lane (tile Tb + l/8,
// part p = l%8) reads 16 bytes: values (p%4 < 2) or x (p%4 >= 2) of columns
// 4(p%4 % 2).. of brick-row R + p/4.  Lanes p = 0, 1 run the fma chains of
// their four columns over the brick-rows in order, taking the x of row R and
// the values and x of row R+1 from the other lanes of the line by DPP.
template <int TH>
__global__ __launch_bounds__(BLOCK_THREADS) void axt_noc_fl32_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const float *__restrict__ ax, const uint32_t *__restrict__ sec, float *__restrict__ y,
    uint32_t nrows, double *partial)
{
    static_assert(TH >= 2 && TH % 2 == 0, "brick-row pairs");
    constexpr uint32_t UM = red_umax(TH); // th 2, 4: 256-column blocks
    __shared__ float s_val[WAVES][UM];
    __shared__ uint32_t s_key[WAVES][UM];
    __shared__ uint16_t s_seg[WAVES][UM + 1];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    const uint32_t nu = d.uend - d.ubeg, p = lane & 7;
    const uint32_t T0 = d.ubeg / 8, T1 = (d.uend + 7) / 8;
    LaneSum<float> lacc; // long rows
    BlockKeys<UM> keys;
    if (d.k == 0) keys.load(sec, d, nu);
    constexpr int L = TH / 2; // lines per tile
    // tile groups of 8 tiles per step: the whole block in flight (th 8: 8-tile blocks)
    constexpr int G = L >= 4 ? 1 : (int)(UM / 64);
    for (uint32_t Tb = T0; Tb < T1; Tb += 8 * G) {
        float4 v[G][L];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const uint32_t Tr = Tb + 8 * g + lane / 8;
            const size_t T = Tr < T1 ? Tr : T1 - 1;
#pragma unroll
            for (int i = 0; i < L; i++)
                v[g][i] = f4load(ax + (T * TH + 2 * i + p / 4) * 16 + 4 * (p % 4));
        }
        loads_issued();
#pragma unroll
        for (int g = 0; g < G; g++) {
            float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int i = 0; i < L; i++) {
                const float4 x0 = f4dpp<DPP_ROW_SHL2>(v[g][i]); // x of row 2i
                const float4 v1 = f4dpp<DPP_ROW_SHL4>(v[g][i]); // values of row 2i+1
                const float4 x1 = f4dpp<DPP_ROW_SHL4>(x0);      // x of row 2i+1
                a.x = fmaf(v[g][i].x, x0.x, a.x);
                a.y = fmaf(v[g][i].y, x0.y, a.y);
                a.z = fmaf(v[g][i].z, x0.z, a.z);
                a.w = fmaf(v[g][i].w, x0.w, a.w);
                a.x = fmaf(v1.x, x1.x, a.x);
                a.y = fmaf(v1.y, x1.y, a.y);
                a.z = fmaf(v1.z, x1.z, a.z);
                a.w = fmaf(v1.w, x1.w, a.w);
            }
            const uint32_t Tr = Tb + 8 * g + lane / 8, u0 = 8 * Tr + 4 * p;
            if (p < 2 && Tr < T1) {
                const float c[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const uint32_t u = u0 + q;
                    if (u < d.ubeg || u >= d.uend) continue;
                    if (d.k != 0) lacc.add(c[q]);
                    else s_val[w][u - d.ubeg] = c[q];
                }
            }
        }
    }
    if (d.k != 0) {
        finish_long<float>(wave_sum_f32(lacc.value()), d, bid, y, nrows, partial);
        return;
    }
    keys.store(nu, s_key[w]);
    wave_lds_sync();
    reduce_keyed<float>(s_val[w], s_key[w], s_seg[w], nu, d, bid, y, nrows, partial);
}

#ifndef ISPV_PLANAR_STEPS
#define ISPV_PLANAR_STEPS 4 // tile height 1
#endif
#ifndef ISPV_PLANAR_STEPS_TH
#define ISPV_PLANAR_STEPS_TH 2 // tile height > 1
#endif
// Four stored values (bf16 or FP32) of the value plane: load4_raw reads
// them as stored, as_floats widens them.  Apart, so that the widening of a
// load's bf16 bits comes after all of a step's loads are issued.
template <typename VT>
using Raw4 = typename std::conditional<sizeof(VT) == 2, u2v, float4>::type;
template <typename VT>
__device__ __forceinline__ Raw4<VT> load4_raw(const VT *p)
{
    if constexpr (sizeof(VT) == 2) return ntload(reinterpret_cast<const u2v *>(p));
    else return f4load(reinterpret_cast<const float *>(p));
}
template <typename VT>
__device__ __forceinline__ float4 as_floats(const Raw4<VT> &b)
{
    if constexpr (sizeof(VT) == 2)
        return make_float4(__uint_as_float(b[0] << 16), __uint_as_float(b[0] & 0xFFFF0000u),
                           __uint_as_float(b[1] << 16), __uint_as_float(b[1] & 0xFFFF0000u));
    else return b;
}

// Planar storage (value plane of bf16 or FP32, x plane of FP32), tile height
// TH in {1, 2, 4, 8, 16}.  Lane l of load step j takes half-brick-row
// h = H0 + 64 j + l, i.e. the four slots 4h..4h+3 (columns 4(h%2).. of
// brick-row h/2), so every load instruction reads contiguous memory: 64 x 4
// values and 64 x 16 bytes of x.  TH = 1: the two halves of a tile meet by a
// quad exchange.  TH > 1: the 2 TH lanes of a tile are aligned in a row of
// 16 (H0 is a multiple of 2 TH) and row shifts sum each column over the
// brick-rows into the tile's first two lanes (TH = 16 adds the second row
// with one shuffle).  FP32 accumulation; long rows Kahan-Babuska-Neumaier.
template <typename VT, int TH>
__global__ __launch_bounds__(BLOCK_THREADS) void axt_planar_kernel(
    const BlockDesc *__restrict__ blocks, uint32_t nlong_wg, uint32_t nblocks,
    const VT *__restrict__ av, const float *__restrict__ xp, const uint32_t *__restrict__ sec,
    float *__restrict__ y, uint32_t nrows, double *partial)
{
    static_assert(TH == 1 || TH == 2 || TH == 4 || TH == 8 || TH == 16, "tile height");
    constexpr uint32_t UM = red_umax(TH);
    __shared__ float s_val[WAVES][UM];
    __shared__ uint32_t s_key[WAVES][UM];
    __shared__ uint16_t s_seg[WAVES][UM + 1];
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const uint32_t bid = block_index(nlong_wg);
    if (bid >= nblocks) return;
    const BlockDesc d = blocks[bid];
    const uint32_t nu = d.uend - d.ubeg;
    const size_t T0 = TH == 1 ? d.ubeg : d.ubeg / 8, T1 = TH == 1 ? d.uend : (d.uend + 7) / 8;
    const size_t H0 = T0 * 2 * TH, H1 = T1 * 2 * TH;
    LaneSum<float> lacc; // long rows
    BlockKeys<UM> keys;
    if (d.k == 0) keys.load(sec, d, nu);
    // load steps of 64 half-brick-rows: th 2, 4 take their whole block at once
    constexpr int U = TH == 1 ? ISPV_PLANAR_STEPS
                    : (TH == 2 || TH == 4) ? (int)(UM * TH / 256) : ISPV_PLANAR_STEPS_TH;
    for (size_t hb = H0; hb < H1; hb += 64 * U) {
        Raw4<VT> vb[U];
        float4 x[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
            const size_t h = hb + 64 * j + lane, hc = h < H1 ? h : H1 - 1;
            vb[j] = load4_raw<VT>(av + hc * 4);
            x[j] = f4load(xp + hc * 4);
        }
        loads_issued();
#pragma unroll
        for (int j = 0; j < U; j++) {
            const size_t h = hb + 64 * j + lane;
            const float4 v = as_floats<VT>(vb[j]);
            float c[4] = {v.x * x[j].x, v.y * x[j].y, v.z * x[j].z, v.w * x[j].w};
            if constexpr (TH == 1) {
                float t = (c[0] + c[1]) + (c[2] + c[3]);
                t = t + dpp_f32<DPP_QUAD_XOR1>(t);
                if (h < H1 && (lane & 1) == 0) {
                    const uint32_t u = (uint32_t)(h / 2);
                    if (d.k != 0) lacc.add(t);
                    else s_val[w][u - d.ubeg] = t;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    c[q] = c[q] + dpp_f32<DPP_ROW_SHL2>(c[q]);
                    if constexpr (TH >= 4) c[q] = c[q] + dpp_f32<DPP_ROW_SHL4>(c[q]);
                    if constexpr (TH >= 8) c[q] = c[q] + dpp_f32<DPP_ROW_SHL8>(c[q]);
                    if constexpr (TH >= 16) c[q] = c[q] + __shfl_down(c[q], 16, WAVE);
                }
                if (h < H1 && (lane & (2 * TH - 2)) == 0) { // brick-row 0 of the tile
                    const uint32_t u0 = (uint32_t)(8 * (h / (2 * TH)) + 4 * (h & 1));
#pragma unroll
                    for (int q = 0; q < 4; q++) {
                        const uint32_t u = u0 + q;
                        if (u < d.ubeg || u >= d.uend) continue;
                        if (d.k != 0) lacc.add(c[q]);
                        else s_val[w][u - d.ubeg] = c[q];
                    }
                }
            }
        }
    }
    if (d.k != 0) {
        finish_long<float>(wave_sum_f32(lacc.value()), d, bid, y, nrows, partial);
        return;
    }
    keys.store(nu, s_key[w]);
    wave_lds_sync();
    reduce_keyed<float>(s_val[w], s_key[w], s_seg[w], nu, d, bid, y, nrows, partial);
}

template <typename VT, typename XT, typename AT, bool PLANAR>
static int launch_noc(const intspmv_plan *p, const void *ax, const void *xp, const uint32_t *sec,
                      void *y)
{
    const uint32_t nb = p->nblocks, th = p->axt.tileH;
    if (nb == 0) return INTSPMV_OK;
    dim3 grid(ispv_grid(p)), blk(BLOCK_THREADS);
    const uint32_t nl = p->nlong_wg;
    hipStream_t s = p->ctx->stream;
    auto a = (const VT *)ax;
    auto x = (const XT *)xp;
    auto yy = (AT *)y;
    if (p->axt.mode == INTSPMV_COM) {
        if constexpr (!PLANAR) {
            axt_com_kernel<VT, XT, AT><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows,
                                                           p->d_partial, p->axt.bs, p->axt.log);
        } else {
            return INTSPMV_ERR_UNSUPPORTED;
        }
    } else if (PLANAR && sizeof(VT) <= 4 && sizeof(XT) == 4 && sizeof(AT) == 4 &&
               (th == 1 || th == 2 || th == 4 || th == 8 || th == 16)) {
        if constexpr (PLANAR && sizeof(VT) <= 4 && sizeof(XT) == 4 && sizeof(AT) == 4) {
            auto xf = (const float *)xp;
            auto yf = (float *)y;
            switch (th) {
            case 1: axt_planar_kernel<VT, 1><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, xf, sec, yf, p->nrows, p->d_partial); break;
            case 2: axt_planar_kernel<VT, 2><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, xf, sec, yf, p->nrows, p->d_partial); break;
            case 4: axt_planar_kernel<VT, 4><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, xf, sec, yf, p->nrows, p->d_partial); break;
            case 8: axt_planar_kernel<VT, 8><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, xf, sec, yf, p->nrows, p->d_partial); break;
            default: axt_planar_kernel<VT, 16><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, xf, sec, yf, p->nrows, p->d_partial); break;
            }
        }
    } else if (th == 1) {
        axt_h1_kernel<VT, XT, AT, PLANAR><<<grid, blk, 0, s>>>(
            p->d_blocks, nl, nb, a, x, sec, yy, p->nrows, p->d_partial);
    } else if constexpr (!PLANAR && sizeof(VT) == 4 && sizeof(XT) == 4) {
        switch (th) {
        case 2: axt_noc_fl32_kernel<2><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial); break;
        case 4: axt_noc_fl32_kernel<4><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial); break;
        case 8: axt_noc_fl32_kernel<8><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial); break;
        case 16: axt_noc_fl32_kernel<16><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial); break;
        default:
            axt_noc_kernel<VT, XT, AT, PLANAR, 0><<<grid, blk, 0, s>>>(
                p->d_blocks, nl, nb, a, x, sec, yy, p->nrows, p->d_partial, th);
        }
    } else if constexpr (!PLANAR && sizeof(VT) == 8) {
        switch (th) {
        case 2: axt_noc_fl_kernel<2><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial, th); break;
        case 4: axt_noc_fl_kernel<4><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial, th); break;
        case 8: axt_noc_fl_kernel<8><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial, th); break;
        case 16: axt_noc_fl_kernel<16><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial, th); break;
        case 32: axt_noc_fl_kernel<32><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial, th); break;
        case 48: axt_noc_fl_kernel<48><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial, th); break;
        default: axt_noc_fl_kernel<0><<<grid, blk, 0, s>>>(p->d_blocks, nl, nb, a, sec, yy, p->nrows, p->d_partial, th);
        }
    } else {
#define ISPV_TH(N)                                                                            \
    case N:                                                                                   \
        axt_noc_kernel<VT, XT, AT, PLANAR, N><<<grid, blk, 0, s>>>(                           \
            p->d_blocks, nl, nb, a, x, sec, yy, p->nrows, p->d_partial, th);        \
        break;
        switch (th) {
            ISPV_TH(2)
            ISPV_TH(4)
            ISPV_TH(8)
            ISPV_TH(16)
        default:
            axt_noc_kernel<VT, XT, AT, PLANAR, 0><<<grid, blk, 0, s>>>(
                p->d_blocks, nl, nb, a, x, sec, yy, p->nrows, p->d_partial, th);
        }
#undef ISPV_TH
    }
    int st = ispv_launch_check(p->ctx);
    if (st) return st;
    return ispv_split_fixup(p, y, sizeof(AT) == 4);
}

} // namespace ispv

using namespace ispv;

static int axt_plan_impl(intspmv_ctx *ctx, const intspmv_axt_desc *desc, const uint32_t *con_host,
                         uint32_t ngroups, const uint32_t *gbeg, intspmv_plan **plan)
{
    if (!ctx || !desc || !plan || (desc->nrows && !con_host)) return INTSPMV_ERR_ARG;
    const intspmv_axt_desc &D = *desc;
    if (D.tileHW != 8 || D.tileH == 0) return INTSPMV_ERR_ARG;
    if (D.mode != INTSPMV_NOC && D.mode != INTSPMV_COM) return INTSPMV_ERR_ARG;
    if (D.layout != INTSPMV_AXT_INTERLEAVED && D.layout != INTSPMV_AXT_PLANAR &&
        D.layout != INTSPMV_AXT_COLMAP)
        return INTSPMV_ERR_ARG;
    if (D.mode == INTSPMV_COM) {
        // units are slots = CSR positions; con[r] = (p/8)*16 + p%8 with p the
        // first CSR position of row r (getArraysLenAXT with positions = rl)
        if (D.tileH != 1 || D.bs < 2 || D.bs > 512 || (D.bs & (D.bs - 1)) || (1u << D.log) != D.bs)
            return INTSPMV_ERR_ARG;
        if (D.colN != D.nnz || D.lenSEC != D.tileN * 8 || D.lenAX != D.tileN * 16 ||
            D.nnz > D.lenSEC)
            return INTSPMV_ERR_ARG;
        uint32_t *uptr = (uint32_t *)malloc(((size_t)D.nrows + 1) * sizeof(uint32_t));
        if (!uptr) return INTSPMV_ERR_NOMEM;
        for (uint32_t r = 0; r < D.nrows; r++) {
            uint32_t c = con_host[r];
            uptr[r] = (c / 16) * 8 + c % 16;
            if (c % 16 >= 8 || uptr[r] > D.nnz || (r && uptr[r] < uptr[r - 1])) {
                free(uptr);
                return INTSPMV_ERR_ARG;
            }
        }
        uptr[D.nrows] = D.nnz;
        int st = ispv_plan_build(ctx, PLAN_AXT, D.nrows, uptr, COM_UMAX, AXT_RMAX, COM_UMAX / 2,
                                 LONG_SLOTS, plan, ngroups, gbeg);
        free(uptr);
        if (st) return st;
        (*plan)->axt = D;
        return INTSPMV_OK;
    }
    const uint32_t th = D.tileH, ths = 8 * th;
    // unit start of each row from con (getArraysLenAXT, intSpmv_main.c:1130-1134):
    // con = T*2ths + R*16 + C; NOC rows start at brick-row 0 of a column.
    const uint32_t nunits = (th == 1) ? D.colN / 8 : D.colN;
    if (th == 1 && (D.colN % 8 != 0 || nunits != D.tileN)) return INTSPMV_ERR_ARG;
    if (th > 1 && (D.colN > D.tileN * 8 || D.lenSEC != D.tileN * 8)) return INTSPMV_ERR_ARG;
    if ((uint64_t)D.tileN * 2 * ths != D.lenAX) return INTSPMV_ERR_ARG;
    uint32_t *uptr = (uint32_t *)malloc(((size_t)D.nrows + 1) * sizeof(uint32_t));
    if (!uptr) return INTSPMV_ERR_NOMEM;
    for (uint32_t r = 0; r < D.nrows; r++) {
        uint32_t c = con_host[r], T = c / (2 * ths), w = c % (2 * ths);
        uint32_t col = w % 16;
        uptr[r] = (th == 1) ? T : T * 8 + col;
        if (col >= 8 || w / 16 != 0 || uptr[r] > nunits || (r && uptr[r] < uptr[r - 1])) {
            free(uptr);
            return INTSPMV_ERR_ARG;
        }
    }
    uptr[D.nrows] = nunits;
    if (D.nrows && uptr[D.nrows - 1] > nunits) {
        free(uptr);
        return INTSPMV_ERR_ARG;
    }
    // block size and long-row threshold per kernel (see the kernels' notes)
    const bool h1_f64 = th == 1 && D.layout == INTSPMV_AXT_INTERLEAVED &&
                        D.val_type == INTSPMV_F64 && D.x_type == INTSPMV_F64;
    // Smaller blocks keep a whole block in one load step where that measured
    // faster (FP64 th = 1: 64 tiles; th = 8: 64 columns = 8 KB); matrices with
    // many empty rows keep 128-unit blocks.
    const bool dense_rows = !ispv_many_empty_rows(uptr, D.nrows);
    const bool small = h1_f64 && dense_rows;
    // the embedded-x kernels only (column-map plans keep blocks of at most
    // AXT_UMAX units, what the column-map kernels stage)
    const bool fl64 = D.layout == INTSPMV_AXT_INTERLEAVED && D.val_type == INTSPMV_F64 &&
                      D.x_type == INTSPMV_F64 && D.mode == INTSPMV_NOC;
    const bool fl32 = D.layout == INTSPMV_AXT_INTERLEAVED && D.val_type == INTSPMV_F32 &&
                      D.x_type == INTSPMV_F32 && D.mode == INTSPMV_NOC;
    const bool planar32 = D.layout == INTSPMV_AXT_PLANAR && D.val_type != INTSPMV_F64 &&
                          D.x_type == INTSPMV_F32 && D.mode == INTSPMV_NOC;
    const uint32_t umax = small ? ISPV_H1_F64_UMAX : th == 1 ? ISPV_H1_RED_UMAX
                        : (th == 2 && fl64) ? TH2_UMAX
                        : (th == 8 && dense_rows) ? ISPV_TH8_UMAX
                        : (th >= 16 && fl64 && dense_rows) ? ISPV_TALL_UMAX
                        : (fl32 || planar32) ? red_umax(th) : AXT_UMAX;
    uint32_t long_units = th == 1 ? (small ? ISPV_H1_F64_LONG : ISPV_H1_LONG)
                                  : (512 / th < 4 ? 4 : (512 / th > umax / 2 ? umax / 2 : 512 / th));
    const uint32_t long_max = th > 1 ? LONG_SLOTS / th
                            : D.layout == INTSPMV_AXT_COLMAP ? GATHER_H1_LONG_MAX
                                                             : LONG_SLOTS / 8;
    int st = ispv_plan_build(ctx, PLAN_AXT, D.nrows, uptr, umax, AXT_RMAX, long_units, long_max,
                             plan, ngroups, gbeg);
    free(uptr);
    if (st) return st;
    (*plan)->axt = D;
    return INTSPMV_OK;
}

extern "C" int intspmv_axt_plan(intspmv_ctx *ctx, const intspmv_axt_desc *desc,
                                const uint32_t *con_host, intspmv_plan **plan)
{
    return axt_plan_impl(ctx, desc, con_host, 0, nullptr, plan);
}

extern "C" int intspmv_axt_plan_grouped(intspmv_ctx *ctx, const intspmv_axt_desc *desc,
                                        const uint32_t *con_host, uint32_t ngroups,
                                        const uint32_t *group_rows, intspmv_plan **plan)
{
    if (ngroups == 0 || !desc || desc->layout != INTSPMV_AXT_COLMAP || desc->mode != INTSPMV_NOC)
        return INTSPMV_ERR_ARG;
    return axt_plan_impl(ctx, desc, con_host, ngroups, group_rows, plan);
}

extern "C" int spmv_axt(intspmv_ctx *ctx, const intspmv_plan *plan, const void *ax,
                        const void *xplane, const uint32_t *sec, void *y)
{
    if (!ctx || !plan || plan->kind != PLAN_AXT || plan->ctx != ctx) return INTSPMV_ERR_ARG;
    const intspmv_axt_desc &D = plan->axt;
    if (plan->nrows && (!y || (plan->nunits && (!ax || !sec)))) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    if (D.layout == INTSPMV_AXT_COLMAP) return INTSPMV_ERR_ARG; // a column-map plan
    const bool planar = D.layout == INTSPMV_AXT_PLANAR;
    if (planar && plan->nunits && !xplane) return INTSPMV_ERR_ARG;
    if (!planar) {
        if (D.val_type == INTSPMV_F64 && D.x_type == INTSPMV_F64)
            return launch_noc<double, double, double, false>(plan, ax, nullptr, sec, y);
        if (D.val_type == INTSPMV_F32 && D.x_type == INTSPMV_F32)
            return launch_noc<float, float, float, false>(plan, ax, nullptr, sec, y);
        return INTSPMV_ERR_UNSUPPORTED;
    }
    if (D.val_type == INTSPMV_F64 && D.x_type == INTSPMV_F64)
        return launch_noc<double, double, double, true>(plan, ax, xplane, sec, y);
    if (D.val_type == INTSPMV_F32 && D.x_type == INTSPMV_F32)
        return launch_noc<float, float, float, true>(plan, ax, xplane, sec, y);
    if (D.val_type == INTSPMV_BF16 && D.x_type == INTSPMV_F32)
        return launch_noc<uint16_t, float, float, true>(plan, ax, xplane, sec, y);
    return INTSPMV_ERR_UNSUPPORTED;
}
