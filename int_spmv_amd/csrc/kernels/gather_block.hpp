// gather_block.hpp -- the per-block body of the column-map AXT product,
// shared by the single-matrix kernel (gather.hip) and the batched persistent
// kernel (batch.hip).
#pragma once
#include <type_traits>
#include <utility>

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

// Single x values may be read through a buffer resource when x spans less
// than 2 GiB (XB, chosen on the host: xbuf_fits): a padding slot then loads
// from an offset past the resource's range, which the hardware answers with
// zeros and no memory request.  Otherwise, and for the 16-byte loads of
// several right-hand sides (measured faster as plain loads), padding reads
// row 0 and is masked.
constexpr uint32_t XB_OOB = 0x80000000u; // past every range xbuf_fits accepts
constexpr int XB_DWORD3 = 0x00020000;    // buffer resource word 3 for gfx9-family raw loads
__host__ __device__ inline bool xbuf_fits(uint32_t nrows, uint32_t ldx, uint32_t xsize)
{
    return (uint64_t)nrows * ldx * xsize < XB_OOB;
}
template <typename XT>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_xbuf(const XT *x, uint32_t nrows,
                                                            uint32_t ldx)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<XT *>(x), 0,
                                             (int)((uint64_t)nrows * ldx * sizeof(XT)), XB_DWORD3);
}

// x values of one slot for KC right-hand sides starting at j0 (0 for padding),
// without per-lane branches.  When every value is wanted (kv == KC) and the row
// offset c*ldx + j0 is a multiple of the vector width (uniform), the KC values
// come in 16-byte loads; otherwise one load per value.
template <typename XT, typename AT, int KC, bool XB>
__device__ __forceinline__ void gather_x(const XT *x, __amdgpu_buffer_rsrc_t xr, uint32_t c,
                                         uint32_t ldx, uint32_t j0, uint32_t kv, AT (&xv)[KC])
{
    constexpr int VW = 16 / sizeof(XT); // elements per 16-byte load
    constexpr uint32_t SZ = sizeof(XT);
    const bool ok = c != PAD;
    const XT *p = x + (size_t)(ok ? c : 0) * ldx + j0;
    if constexpr (KC >= VW && KC % VW == 0) {
        if (kv == (uint32_t)KC && ((ldx | j0) % VW) == 0) {
#pragma unroll
            for (int j = 0; j < KC; j += VW) {
                if constexpr (SZ == 8) {
                    d2v v = *reinterpret_cast<const d2v *>(p + j);
                    xv[j] = ok ? (AT)v[0] : (AT)0;
                    xv[j + 1] = ok ? (AT)v[1] : (AT)0;
                } else {
                    typedef float f4v __attribute__((ext_vector_type(4)));
                    f4v v = *reinterpret_cast<const f4v *>(p + j);
#pragma unroll
                    for (int e = 0; e < 4; e++) xv[j + e] = ok ? (AT)v[e] : (AT)0;
                }
            }
            return;
        }
    }
    if constexpr (XB) {
        // every real offset is below the range (< 2 GiB), padding's is past it
        const uint32_t off = (c * ldx + j0) * SZ;
#pragma unroll
        for (int j = 0; j < KC; j++) {
            const uint32_t o = (ok && (uint32_t)j < kv) ? off + j * SZ : XB_OOB;
            if constexpr (SZ == 8)
                xv[j] = (AT)__builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(xr, o, 0, 0));
            else
                xv[j] = (AT)__int_as_float(__builtin_amdgcn_raw_buffer_load_b32(xr, o, 0, 0));
        }
    } else {
        // indexes clamped to the row's first kv values
#pragma unroll
        for (int j = 0; j < KC; j++) {
            const XT v = p[(uint32_t)j < kv ? j : 0];
            xv[j] = (ok && (uint32_t)j < kv) ? (AT)v : (AT)0;
        }
    }
}

// One pass over KC right-hand sides.  th == 1: four lanes per tile, the
// reference's reduce tree per right-hand side; th > 1: two columns per lane,
// fma chains over the brick-rows.  Normal blocks stage unit sums in LDS and
// give each row to one lane; long rows keep per-lane sums (FP32: with a
// Neumaier compensation term, as the embedded-x kernels' long rows).
// LDS of one wave for the column-map block code, sized for KC right-hand sides.
template <typename AT, int KC>
struct GatherLds {
    AT val[G_UMAX * KC];
    uint32_t key[G_UMAX];
    uint16_t seg[G_UMAX + 1];
};

// UG1 / UG4: groups of 16 tiles in flight per step for passes of one / four
// right-hand sides (the kernels choose them by measurement).
// PV: slot values read through a permutation, value of slot s =
// aval[vperm[s]] (0 where vperm[s] is PAD) -- Aᵀ's map reading A's values.
// GR (one right-hand side, the transposed map of a matrix A): the same pass
// also writes the value gradient of A in this map's slot order,
// gvt[s] = x[acol[s]] * xg[row(s)] (0 on padding) -- with x = y_bar and
// xg = the forward's x, A_bar of the entry slot s holds, the product of
// spmv_axt_grad_vals in the same operand order, so the same bits.
template <typename VT, typename XT, typename AT, int KC, bool H1, bool XB, int UG1 = 4,
          bool PV = false, int UG4 = 2, bool GR = false>
__device__ __forceinline__ void gather_block(const BlockDesc &d, const VT *__restrict__ aval,
                                             const uint32_t *__restrict__ acol,
                                             const uint32_t *__restrict__ sec,
                                             const XT *__restrict__ x, uint32_t ldx,
                                             AT *__restrict__ y, uint32_t ldy, uint32_t j0,
                                             uint32_t kv, uint32_t th, uint32_t nrows,
                                             double *partial, AT *sv, uint32_t *sk,
                                             uint16_t *sg, const uint32_t *__restrict__ vperm = nullptr,
                                             const XT *__restrict__ xg = nullptr,
                                             VT *__restrict__ gvt = nullptr)
{
    static_assert(!GR || (KC == 1 && !std::is_same<VT, uint16_t>::value),
                  "the value gradient: one right-hand side, FP64 / FP32 values");
    const int lane = lane_id();
    const uint32_t nu = d.uend - d.ubeg, q = lane & 3, C = 2 * q;
    const __amdgpu_buffer_rsrc_t xr = make_xbuf(x, nrows, ldx); // unused unless XB
    LaneSum<AT> lacc[KC]; // long rows; FP32 sums compensated (Neumaier)
    // tiles of this block; UG steps of 16 tiles are loaded before their
    // gathers are issued, so a lane keeps 2*UG gathers in flight
    // (th > 1: a block holds at most G_UMAX = 128 tile columns, i.e. 16
    // tiles, 17 when it starts inside a tile, so two groups cover it; more
    // groups would only reload the block's last tile, clamped)
    constexpr int UGH = KC >= 8 ? 1 : (KC >= 4 ? UG4 : (KC == 1 ? UG1 : 4));
    constexpr int UG = H1 ? UGH : (UGH < 2 ? UGH : 2);
    static_assert(H1 || G_UMAX <= 16 * 8, "th > 1: a block spans at most 17 tiles");
    const uint32_t T0 = H1 ? d.ubeg : d.ubeg / 8, T1 = H1 ? d.uend : (d.uend + 7) / 8;
    for (uint32_t Tbb = T0; Tbb < T1; Tbb += 16 * UG) {
        AT a0[UG][KC], a1[UG][KC];
        u2v kk[UG]; // row keys: the tile's (th == 1) or the lane's two columns'
#pragma unroll
        for (int u = 0; u < UG; u++) {
#pragma unroll
            for (int j = 0; j < KC; j++) a0[u][j] = a1[u][j] = 0;
            const uint32_t Tr = Tbb + 16 * u + lane / 4;
            const size_t T = Tr < T1 ? Tr : T1 - 1;
            if (H1) kk[u][0] = ntload(sec + T);
            else kk[u] = ntload(reinterpret_cast<const u2v *>(sec + 8 * T + C));
        }
        // th == 1 (H1): one brick-row, known here so that the value loads
        // issue with the column loads instead of after the gathers
        for (uint32_t R = 0; R < (H1 ? 1u : th); R++) {
            GLoad<VT> v0[UG], v1[UG];
            u2v cc[UG];
#pragma unroll
            for (int u = 0; u < UG; u++) {
                const uint32_t Tr = Tbb + 16 * u + lane / 4;
                const size_t T = Tr < T1 ? Tr : T1 - 1;
                const size_t sidx = (T * th + R) * 8 + C;
                if constexpr (PV) {
                    static_assert(!std::is_same<VT, uint16_t>::value, "FP64 / FP32 values");
                    const u2v pp = ntload(reinterpret_cast<const u2v *>(vperm + sidx));
                    v0[u] = pp[0] != PAD ? aval[pp[0]] : (VT)0;
                    v1[u] = pp[1] != PAD ? aval[pp[1]] : (VT)0;
                } else {
                    GPair<VT>::ld(aval + sidx, v0[u], v1[u]);
                }
                cc[u] = ntload(reinterpret_cast<const u2v *>(acol + sidx));
            }
#pragma unroll
            for (int u = 0; u < UG; u++) {
                AT x0[KC], x1[KC];
                gather_x<XT, AT, KC, XB>(x, xr, cc[u][0], ldx, j0, kv, x0);
                gather_x<XT, AT, KC, XB>(x, xr, cc[u][1], ldx, j0, kv, x1);
                if constexpr (GR) {
                    const uint32_t Tr = Tbb + 16 * u + lane / 4;
                    const size_t T = Tr < T1 ? Tr : T1 - 1;
                    const size_t sidx = (T * th + R) * 8 + C;
                    // slots of this block only (th > 1 blocks may share a tile)
                    const uint32_t u0 = 8 * Tr + C;
                    const bool ok0 = Tr < T1 && (H1 || (u0 >= d.ubeg && u0 < d.uend));
                    const bool ok1 = Tr < T1 && (H1 || (u0 + 1 >= d.ubeg && u0 + 1 < d.uend));
                    const uint32_t r0 = kk[u][0], r1 = H1 ? kk[u][0] : kk[u][1];
                    const AT xa = r0 < nrows ? (AT)xg[r0] : (AT)0;
                    const AT xb = r1 < nrows ? (AT)xg[r1] : (AT)0;
                    const AT g0 = cc[u][0] != PAD ? x0[0] * xa : (AT)0;
                    const AT g1 = cc[u][1] != PAD ? x1[0] * xb : (AT)0;
                    if (ok0 && ok1) {
                        if constexpr (sizeof(VT) == 8) {
                            d2v o;
                            o[0] = g0;
                            o[1] = g1;
                            __builtin_nontemporal_store(o, reinterpret_cast<d2v *>(gvt + sidx));
                        } else {
                            f2v o;
                            o[0] = g0;
                            o[1] = g1;
                            __builtin_nontemporal_store(o, reinterpret_cast<f2v *>(gvt + sidx));
                        }
                    } else {
                        if (ok0) gvt[sidx] = (VT)g0;
                        if (ok1) gvt[sidx + 1] = (VT)g1;
                    }
                }
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
            if (H1) {
                const bool ok = Tr < T1;
                if (d.k != 0) {
                    if (ok)
#pragma unroll
                        for (int j = 0; j < KC; j++) lacc[j].add(a0[u][j] + a1[u][j]);
                    continue;
                }
                const uint32_t key = kk[u][0];
                AT ts[KC];
#pragma unroll
                for (int j = 0; j < KC; j++) {
                    AT p0 = a0[u][j], p1 = a1[u][j];
                    p0 = p0 + dpp_any<AT, DPP_QUAD_XOR2>(p0);
                    p1 = p1 + dpp_any<AT, DPP_QUAD_XOR2>(p1);
                    p0 = p0 + dpp_any<AT, DPP_QUAD_XOR1>(p0);
                    p1 = p1 + dpp_any<AT, DPP_QUAD_XOR1>(p1);
                    ts[j] = p0 + p1;
                }
                if (ok && q == 0) { // one branch for the tile's KC sums and its key
#pragma unroll
                    for (int j = 0; j < KC; j++) sv[(Tr - d.ubeg) * KC + j] = ts[j];
                    sk[Tr - d.ubeg] = key;
                }
            } else {
                const uint32_t u0 = 8 * Tr + C;
                const bool ok0 = Tr < T1 && u0 >= d.ubeg && u0 < d.uend;
                const bool ok1 = Tr < T1 && u0 + 1 >= d.ubeg && u0 + 1 < d.uend;
                if (d.k != 0) {
#pragma unroll
                    for (int j = 0; j < KC; j++) {
                        if (ok0) lacc[j].add(a0[u][j]);
                        if (ok1) lacc[j].add(a1[u][j]);
                    }
                    continue;
                }
                if (ok0) { // one branch per column for its KC sums and key
#pragma unroll
                    for (int j = 0; j < KC; j++) sv[(u0 - d.ubeg) * KC + j] = a0[u][j];
                    sk[u0 - d.ubeg] = kk[u][0];
                }
                if (ok1) {
#pragma unroll
                    for (int j = 0; j < KC; j++) sv[(u0 + 1 - d.ubeg) * KC + j] = a1[u][j];
                    sk[u0 + 1 - d.ubeg] = kk[u][1];
                }
            }
        }
    }
    if (d.k != 0) {
#pragma unroll
        for (int j = 0; j < KC; j++) {
            AT t = wave_sum<AT>(lacc[j].value());
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
    const uint32_t nseg = find_segments(sk, sg, nu);
    for (uint32_t si = lane; si < nseg; si += WAVE) {
        const uint32_t a = sg[si], b = sg[si + 1], r = sk[a];
        // all KC sums of the row at once (units in order for each of them)
        AT acc[KC];
#pragma unroll
        for (int j = 0; j < KC; j++) acc[j] = 0;
        if constexpr (KC <= 4) {
            struct U {
                AT v[KC];
            };
            in_order_chunks<8 / KC, U>(
                a, b,
                [&](uint32_t i) {
                    U u;
#pragma unroll
                    for (int j = 0; j < KC; j++) u.v[j] = sv[i * KC + j];
                    return u;
                },
                [&](const U &u) {
#pragma unroll
                    for (int j = 0; j < KC; j++) acc[j] = acc[j] + u.v[j];
                });
        } else {
            for (uint32_t i = a; i < b; i++)
#pragma unroll
                for (int j = 0; j < KC; j++) acc[j] = acc[j] + sv[i * KC + j];
        }
        if (r < nrows)
#pragma unroll
            for (int j = 0; j < KC; j++)
                if ((uint32_t)j < kv) y[(size_t)r * ldy + j0 + j] = acc[j];
    }
    zero_empty_rows(sk, sg, nseg, d, nrows, [&](uint32_t z) {
        for (uint32_t j = 0; j < kv; j++) y[(size_t)z * ldy + j0 + j] = (AT)0;
    });
}

// ---- wide passes: lanes along the right-hand sides ------------------------
// For a pass of KW right-hand sides over a normal block of tile height 1.  A
// group of G = KW / V lanes (V = values per 16-byte load of x: 4 FP32, 2 FP64)
// takes one tile at a time; lane l of a group owns right-hand sides
// j0 + V l .. j0 + V l + V - 1.  For each slot the group's lanes read the
// slot's x row of KW values together (one cache line for 32 FP32 or 16 FP64
// values) instead of every lane reading its own slot's row, and the block's
// values and columns are read once for all KW right-hand sides.  Tiles are
// dealt round-robin to the 64 / G groups, wide_u per group per round; a
// round's tile sums (the reference's tree of the eight products, per
// right-hand side) go to LDS, and one lane per right-hand side then adds them
// to its running row sum in tile order, carrying the row across rounds --
// tiles of a row added in order from +0.0, the bits of the narrow passes.
// Tiles per group per round: two for FP64 x (V = 2), one for FP32 x, whose
// four values per slot and lane fill the registers of a second tile's
// gathers (measured: FP64 stencil k = 8 420 against 491 us with one; FP32
// stencil k = 32 607 against 765 us with two).
#ifdef ISPV_WIDE_U
template <typename XT>
constexpr int wide_u() { return ISPV_WIDE_U; }
#else
template <typename XT>
constexpr int wide_u() { return sizeof(XT) == 8 ? 2 : 1; }
#endif
// Smallest group (lanes per x row) that takes wide passes: a pass of kc
// right-hand sides at tile height 1 is wide when kc >= ISPV_WIDE_MIN_G * V.
#ifndef ISPV_WIDE_MIN_G
#define ISPV_WIDE_MIN_G 2
#endif
template <typename XT>
__host__ __device__ constexpr uint32_t wide_min_kc()
{
    return ISPV_WIDE_MIN_G * (16 / sizeof(XT));
}
template <int KW, typename XT>
struct WideShape {
    static constexpr int V = 16 / (int)sizeof(XT);
    static constexpr int U = wide_u<XT>();
    static constexpr int G = KW / V, NG = WAVE / G, R = U * NG; // tiles per round
    static_assert(G >= 2 && WAVE % G == 0 && R <= WAVE, "wide passes: 2+ lanes per row");
};
// LDS of one wave for a wide pass: a round's tile sums sv[R][KW], the carried
// row sums carry[KW] and the round's run starts rs[R + 1].
template <typename AT, int KW, typename XT>
struct WideLds {
    AT sv[WideShape<KW, XT>::R * KW];
    AT carry[KW];
    uint16_t rs[WideShape<KW, XT>::R + 1];
};

// The eight slot values of tile T as AT.
template <typename VT, typename AT>
__device__ __forceinline__ void tile_values(const VT *aval, size_t T, AT (&v)[8])
{
    if constexpr (sizeof(VT) == 8) {
#pragma unroll
        for (int q = 0; q < 8; q += 2) {
            const d2v t = ntload(reinterpret_cast<const d2v *>(aval + T * 8 + q));
            v[q] = (AT)t[0];
            v[q + 1] = (AT)t[1];
        }
    } else if constexpr (sizeof(VT) == 4) {
        typedef float f4v __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int q = 0; q < 8; q += 4) {
            const f4v t = ntload(reinterpret_cast<const f4v *>(aval + T * 8 + q));
#pragma unroll
            for (int e = 0; e < 4; e++) v[q + e] = (AT)t[e];
        }
    } else {
        typedef uint32_t u4v __attribute__((ext_vector_type(4)));
        const u4v t = ntload(reinterpret_cast<const u4v *>(aval + T * 8));
#pragma unroll
        for (int e = 0; e < 4; e++) {
            v[2 * e] = (AT)__uint_as_float(t[e] << 16);
            v[2 * e + 1] = (AT)__uint_as_float(t[e] & 0xFFFF0000u);
        }
    }
}

typedef uint32_t u4v __attribute__((ext_vector_type(4)));

// ds_swizzle in bit mode over each 32-lane half: every lane reads lane
// (lane & AND) | OR.  Groups of G lanes (G a power of two <= 32): AND keeps
// the group's bits, OR names the lane within the group.
template <int G, int SRC>
__device__ __forceinline__ uint32_t group_bcast(uint32_t v)
{
    static_assert(G >= 2 && G <= 32 && (G & (G - 1)) == 0 && SRC < G, "lane within the group");
    constexpr int AND = 0x1F & ~(G - 1);
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, AND | (SRC << 5));
}
template <int G, int SRC, typename AT>
__device__ __forceinline__ AT group_bcast_val(AT v)
{
    if constexpr (sizeof(AT) == 8) {
        const uint32_t lo = group_bcast<G, SRC>((uint32_t)__double2loint(v));
        const uint32_t hi = group_bcast<G, SRC>((uint32_t)__double2hiint(v));
        return __hiloint2double((int)hi, (int)lo);
    } else {
        return __uint_as_float(group_bcast<G, SRC>(__float_as_uint(v)));
    }
}

template <int G, int P, typename AT, int... Q>
__device__ __forceinline__ void bcast_slots(std::integer_sequence<int, Q...>, const AT (&vo)[P],
                                            const uint32_t (&co)[P], AT (&v)[8],
                                            uint32_t (&c)[8])
{
    ((c[Q] = group_bcast<G, Q / P>(co[Q % P]), v[Q] = group_bcast_val<G, Q / P, AT>(vo[Q % P])),
     ...);
}

// The eight values and columns of tile T in every lane of a group of G lanes,
// each slot read from memory by one lane only: lane i of the group loads
// slots P i .. P i + P - 1 (P = 8 / G; for G > 8, lanes i and i + 8 the same
// slot) and the group exchanges them by lane swizzles.  The same operands as
// tile_values and a load of all eight columns per lane, with an eighth of
// the loads' lane traffic.
template <typename VT, typename AT, int G>
__device__ __forceinline__ void tile_slots_group(const VT *__restrict__ aval,
                                                 const uint32_t *__restrict__ acol, size_t T,
                                                 uint32_t li, AT (&v)[8], uint32_t (&c)[8])
{
    constexpr int P = G >= 8 ? 1 : 8 / G;
    const size_t s0 = T * 8 + (G >= 8 ? (li & 7u) : li * P);
    AT vo[P];
    uint32_t co[P];
    if constexpr (P == 1) {
        co[0] = ntload(acol + s0);
        if constexpr (std::is_same<VT, uint16_t>::value) vo[0] = bf16_to_f32(ntload(aval + s0));
        else vo[0] = (AT)ntload(aval + s0);
    } else {
        typedef uint32_t uPv __attribute__((ext_vector_type(P)));
        const uPv cc = ntload(reinterpret_cast<const uPv *>(acol + s0));
#pragma unroll
        for (int p = 0; p < P; p++) co[p] = cc[p];
        if constexpr (std::is_same<VT, uint16_t>::value) {
#pragma unroll
            for (int p = 0; p < P; p++) vo[p] = bf16_to_f32(ntload(aval + s0 + p));
        } else {
            typedef VT vPv __attribute__((ext_vector_type(P)));
            const vPv vv = ntload(reinterpret_cast<const vPv *>(aval + s0));
#pragma unroll
            for (int p = 0; p < P; p++) vo[p] = (AT)vv[p];
        }
    }
    bcast_slots<(G > 8 ? 8 : G), P>(std::make_integer_sequence<int, 8>{}, vo, co, v, c);
}

// V values of x row c from right-hand side j (0 for padding and for the
// values at or past cnt = right-hand sides left for this lane), without
// branches, so that the eight gathers of a tile stay in flight together.
// XV (rows and j aligned to V, ldx a multiple of V): one 16-byte load, in
// bounds since j < k <= ldx makes j + V <= ldx; a lane with nothing left
// (cnt <= 0) reads the row's first V values instead, and padding row 0.
// Otherwise one load per value, indexes clamped the same way.
template <typename XT, typename AT, int V, bool XV>
__device__ __forceinline__ void wide_x(const XT *x, uint32_t c, uint32_t ldx, uint32_t j, int cnt,
                                       AT (&o)[V])
{
    const bool ok = c != PAD;
    const XT *p = x + (size_t)(ok ? c : 0) * ldx + (cnt > 0 ? j : 0);
    if constexpr (XV) {
        if constexpr (sizeof(XT) == 8) {
            const d2v t = *reinterpret_cast<const d2v *>(p);
#pragma unroll
            for (int e = 0; e < V; e++) o[e] = (ok && e < cnt) ? (AT)t[e] : (AT)0;
        } else {
            typedef float f4v __attribute__((ext_vector_type(4)));
            const f4v t = *reinterpret_cast<const f4v *>(p);
#pragma unroll
            for (int e = 0; e < V; e++) o[e] = (ok && e < cnt) ? (AT)t[e] : (AT)0;
        }
    } else {
#pragma unroll
        for (int e = 0; e < V; e++) {
            const XT t = p[e < cnt ? e : 0];
            o[e] = (ok && e < cnt) ? (AT)t : (AT)0;
        }
    }
}

template <typename VT, typename XT, typename AT, int KW, bool XV>
__device__ __forceinline__ void gather_block_wide(const BlockDesc &d, const VT *__restrict__ aval,
                                                  const uint32_t *__restrict__ acol,
                                                  const uint32_t *__restrict__ sec,
                                                  const XT *__restrict__ x, uint32_t ldx,
                                                  AT *__restrict__ y, uint32_t ldy, uint32_t j0,
                                                  uint32_t kv, uint32_t nrows, uint32_t *sk,
                                                  uint16_t *sg, WideLds<AT, KW, XT> &W)
{
    using S = WideShape<KW, XT>;
    constexpr int V = S::V, G = S::G, NG = S::NG, R = S::R, U = S::U;
    const int lane = lane_id();
    const uint32_t grp = lane / G, jl = (lane % G) * V;
    const int cnt = (int)kv - (int)jl; // right-hand sides left for this lane (<= 0: none)
    const uint32_t nu = d.uend - d.ubeg;
    for (uint32_t i = lane; i < nu; i += WAVE) sk[i] = ntload(sec + d.ubeg + i);
    wave_lds_sync();
    const uint32_t nseg = find_segments(sk, sg, nu);
    for (uint32_t b0 = 0; b0 < nu; b0 += R) {
        // tile sums of this round: group grp takes tiles b0 + grp + NG u
        AT ts[U][V];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t t = b0 + grp + NG * u;
            const size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
            AT v[8];
            uint32_t c[8];
            tile_slots_group<VT, AT, G>(aval, acol, T, lane % G, v, c);
            AT xv[8][V];
#pragma unroll
            for (int q = 0; q < 8; q++) wide_x<XT, AT, V, XV>(x, c[q], ldx, j0 + jl, cnt, xv[q]);
#pragma unroll
            for (int e = 0; e < V; e++) {
                AT pr[8];
#pragma unroll
                for (int q = 0; q < 8; q++) pr[q] = v[q] * xv[q][e];
                ts[u][e] = ((pr[0] + pr[4]) + (pr[2] + pr[6])) + ((pr[1] + pr[5]) + (pr[3] + pr[7]));
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (b0 + grp + NG * u < nu)
#pragma unroll
                for (int e = 0; e < V; e++) W.sv[(grp + NG * u) * KW + jl + e] = ts[u][e];
        // runs of equal rows among the round's tiles (a run at the round's
        // start may continue the previous round's row, the last may go on)
        const uint32_t here = nu - b0 < (uint32_t)R ? nu - b0 : (uint32_t)R;
        const uint32_t t = b0 + lane;
        const bool start = (uint32_t)lane < here && (lane == 0 || sk[t] != sk[t - 1]);
        const uint64_t m = __ballot(start);
        if (start) W.rs[__popcll(m & lanemask_lt())] = (uint16_t)lane;
        const uint32_t nr = __popcll(m);
        if (lane == 0) W.rs[nr] = (uint16_t)here;
        const bool first_cont = b0 > 0 && sk[b0] == sk[b0 - 1];
        const bool last_cont = b0 + here < nu && sk[b0 + here] == sk[b0 + here - 1];
        wave_lds_sync();
        // one lane per (run, right-hand side): the run's tile sums in order,
        // from the carried sum when the run continues a row, from +0.0 else
        for (uint32_t p = lane; p < nr * KW; p += WAVE) {
            const uint32_t ri = p / KW, j = p % KW, a = W.rs[ri], b = W.rs[ri + 1];
            AT acc = (ri == 0 && first_cont) ? W.carry[j] : (AT)0;
            for (uint32_t i = a; i < b; i++) acc = acc + W.sv[i * KW + j];
            if (ri + 1 == nr && last_cont) {
                W.carry[j] = acc;
            } else {
                const uint32_t r = sk[b0 + a];
                if (r < nrows && j < kv) y[(size_t)r * ldy + j0 + j] = acc;
            }
        }
        wave_lds_sync(); // the next round overwrites sv and rs
    }
    zero_empty_rows(sk, sg, nseg, d, nrows, [&](uint32_t z) {
        for (uint32_t j = 0; j < kv; j++) y[(size_t)z * ldy + j0 + j] = (AT)0;
    });
}

// Tiles per group in flight in a wide pass over a long row (the tile order of
// each lane's sums does not depend on it).  Measured with the normal blocks in
// the same kernel (us, power-law 2^20 rows FP32 k = 32 / R-MAT 2^22 rows FP32
// k = 8): 518.5 / 799.0 with one, 500.0 / 784.5 with two; FP64 unchanged.
#ifndef ISPV_WIDE_LONG_U
#define ISPV_WIDE_LONG_U 2
#endif
template <typename XT>
constexpr int wide_long_u() { return ISPV_WIDE_LONG_U; }

// A wide pass over a long-row block (d.k >= 1, tile height 1: every tile of
// the block is the row's, so no keys are read).  Group grp takes tiles
// grp, grp + NG, grp + 2 NG, ... in that order and each lane keeps a running
// sum per right-hand side of the tiles' tree sums (FP32: compensated); the NG
// groups' sums then meet in a butterfly over the lane bits above the group,
// and group 0 writes y (d.k == 1) or the piece's partial sums (d.k >= 2).
template <typename VT, typename XT, typename AT, int KW, bool XV, int UL = wide_long_u<XT>()>
__device__ __forceinline__ void gather_long_wide(const BlockDesc &d, const VT *__restrict__ aval,
                                                 const uint32_t *__restrict__ acol,
                                                 const XT *__restrict__ x, uint32_t ldx,
                                                 AT *__restrict__ y, uint32_t ldy, uint32_t j0,
                                                 uint32_t kv, uint32_t nrows, double *partial)
{
    using S = WideShape<KW, XT>;
    constexpr int V = S::V, G = S::G, NG = S::NG;
    const int lane = lane_id();
    const uint32_t grp = lane / G, jl = (lane % G) * V;
    const int cnt = (int)kv - (int)jl; // right-hand sides left for this lane (<= 0: none)
    const uint32_t nu = d.uend - d.ubeg;
    if (nu == 0) return; // XCD padding of the plan's long section
    LaneSum<AT> acc[V];
    for (uint32_t b0 = 0; b0 < nu; b0 += UL * NG) {
        AT ts[UL][V];
#pragma unroll
        for (int u = 0; u < UL; u++) {
            const uint32_t t = b0 + grp + NG * u;
            const size_t T = (size_t)d.ubeg + (t < nu ? t : nu - 1);
            AT v[8];
            tile_values<VT, AT>(aval, T, v);
            const u4v c0 = ntload(reinterpret_cast<const u4v *>(acol + T * 8));
            const u4v c1 = ntload(reinterpret_cast<const u4v *>(acol + T * 8 + 4));
            AT xv[8][V];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                wide_x<XT, AT, V, XV>(x, c0[q], ldx, j0 + jl, cnt, xv[q]);
                wide_x<XT, AT, V, XV>(x, c1[q], ldx, j0 + jl, cnt, xv[q + 4]);
            }
#pragma unroll
            for (int e = 0; e < V; e++) {
                AT pr[8];
#pragma unroll
                for (int q = 0; q < 8; q++) pr[q] = v[q] * xv[q][e];
                ts[u][e] = ((pr[0] + pr[4]) + (pr[2] + pr[6])) + ((pr[1] + pr[5]) + (pr[3] + pr[7]));
            }
        }
#pragma unroll
        for (int u = 0; u < UL; u++)
            if (b0 + grp + NG * u < nu)
#pragma unroll
                for (int e = 0; e < V; e++) acc[e].add(ts[u][e]);
    }
#pragma unroll
    for (int e = 0; e < V; e++) {
        AT s = acc[e].value();
#pragma unroll
        for (int m = G; m < WAVE; m <<= 1) s += __shfl_xor(s, m, WAVE);
        if (grp == 0 && e < cnt) {
            if (d.k == 1) {
                if (d.rbeg < nrows) y[(size_t)d.rbeg * ldy + j0 + jl + e] = s;
            } else {
                partial[(size_t)d.piece * PARTIAL_STRIDE + j0 + jl + e] = (double)s;
            }
        }
    }
}

} // namespace ispv
