// reduce.hpp -- phase 2 of the block kernels: turning the per-unit sums a
// wave staged in LDS into row results.
//
//   keyed:  each unit carries its row (AXT sec word); runs of equal keys are
//           rows.  One lane per row adds the row's unit sums in unit order
//           from +0.0, which is the order the reference's atomic
//           accumulation takes when run serially (intSpmv_main.c:1400-1401,
//           1469-1470).  Rows of the block with no unit are written as 0.
//   rowptr: the format's row pointer gives each row's units (CSR row[],
//           AXC brp[]); one lane per row, same left-to-right order.
//   long:   the block is one row: per-lane register sums and a wave tree; a
//           row split over several blocks meets in the plan's scratch.
#pragma once
#include "common.hpp"

namespace ispv {

template <typename AT>
__device__ __forceinline__ AT wave_sum(AT v);
template <>
__device__ __forceinline__ double wave_sum<double>(double v) { return wave_sum_f64(v); }
template <>
__device__ __forceinline__ float wave_sum<float>(float v) { return wave_sum_f32(v); }

// Compensated (Kahan-Babuska-Neumaier) running sum, used for FP32 rows.
struct Neumaier {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float v)
    {
        float t = s + v;
        if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

// Writes the long-row result (k == 1) or this piece's partial (k >= 2).
template <typename AT>
__device__ __forceinline__ void finish_long(AT total, const BlockDesc &d, uint32_t bid, AT *y,
                                            uint32_t nrows, double *partial)
{
    (void)bid;
    if (lane_id() != 0) return;
    if (d.k == 1) {
        if (d.rbeg < nrows) y[d.rbeg] = total;
    } else {
        partial[(size_t)d.piece * PARTIAL_STRIDE] = (double)total;
    }
}

// Per-lane running sum of a long row.  FP64 adds plainly; FP32 keeps a
// Kahan-Babuska-Neumaier compensation term, so rows of thousands of entries
// keep FP32 accumulation accurate.
template <typename AT>
struct LaneSum {
    AT s = 0;
    __device__ __forceinline__ void add(AT v) { s = s + v; }
    __device__ __forceinline__ AT value() const { return s; }
};
template <>
struct LaneSum<float> {
    Neumaier k;
    __device__ __forceinline__ void add(float v) { k.add(v); }
    __device__ __forceinline__ float value() const { return k.value(); }
};

// Calls use(get(i)) for i = a .. b-1 in order, with the reads of CH indexes
// issued together and the last b-a mod CH one by one, so a row's chain
// waits for LDS about once per CH entries instead of once per entry.  The
// order of use() is unchanged.  (Reading a whole last chunk, masked, measured
// slower.)
template <int CH, typename T, typename Get, typename Use>
__device__ __forceinline__ void in_order_chunks(uint32_t a, uint32_t b, Get get, Use use)
{
    uint32_t i = a;
    for (; i + CH <= b; i += CH) {
        T t[CH];
#pragma unroll
        for (int q = 0; q < CH; q++) t[q] = get(i + q);
#pragma unroll
        for (int q = 0; q < CH; q++) use(t[q]);
    }
    for (; i < b; i++) use(get(i));
}

// Finds the runs of equal keys (rows) of a normal block: sseg[0..nseg) are
// the first units of the runs, sseg[nseg] = nu.  Returns nseg.
__device__ __forceinline__ uint32_t find_segments(const uint32_t *sk, uint16_t *sseg, uint32_t nu)
{
    const int lane = lane_id();
    uint32_t nseg = 0;
    for (uint32_t b = 0; b < nu; b += WAVE) {
        uint32_t p = b + lane;
        bool head = p < nu && (p == 0 || sk[p] != sk[p - 1]);
        uint64_t m = __ballot(head);
        if (head) sseg[nseg + __popcll(m & lanemask_lt())] = (uint16_t)p;
        nseg += __popcll(m);
    }
    if (lane == 0) sseg[nseg] = (uint16_t)nu;
    wave_lds_sync();
    return nseg;
}

// Rows per normal block are capped at this (plan rmax) for the keyed kernels.
constexpr uint32_t BLOCK_ROWS_MAX = 256;

// Calls zero_row(r) for every row r of the block [rbeg, rend) that heads no
// run (a row without units).  A 256-bit map in LDS marks the rows the runs
// write; consecutive lanes then take consecutive rows, so the zeros are
// written coalesced however long a gap of empty rows is.  Nothing happens
// when every row of the block has a run (the common case).
template <typename F>
__device__ __forceinline__ void zero_empty_rows(const uint32_t *sk, const uint16_t *sseg,
                                                uint32_t nseg, const BlockDesc &d,
                                                uint32_t nrows, F zero_row)
{
    __shared__ uint32_t rowmap[WAVES][BLOCK_ROWS_MAX / 32];
    const uint32_t nr = d.rend - d.rbeg;
    if (nseg >= nr) return; // wave-uniform
    const int lane = lane_id();
    uint32_t *map = rowmap[threadIdx.x / WAVE];
    if (lane < (int)(BLOCK_ROWS_MAX / 32)) map[lane] = 0;
    wave_lds_sync();
    for (uint32_t s = lane; s < nseg; s += WAVE) {
        const uint32_t o = sk[sseg[s]] - d.rbeg;
        if (o < nr && o < BLOCK_ROWS_MAX) atomicOr(&map[o >> 5], 1u << (o & 31));
    }
    wave_lds_sync();
    for (uint32_t z = lane; z < nr && z < BLOCK_ROWS_MAX; z += WAVE)
        if (!((map[z >> 5] >> (z & 31)) & 1u) && d.rbeg + z < nrows) zero_row(d.rbeg + z);
}

// Keyed segmentation.  sseg needs room for nu + 1 entries.
template <typename AT>
__device__ __forceinline__ void reduce_keyed(const AT *sv, const uint32_t *sk, uint16_t *sseg,
                                             uint32_t nu, const BlockDesc &d, uint32_t bid, AT *y,
                                             uint32_t nrows, double *partial)
{
    const int lane = lane_id();
    uint32_t nseg = 0;
    for (uint32_t b = 0; b < nu; b += WAVE) {
        uint32_t p = b + lane;
        bool head = p < nu && (p == 0 || sk[p] != sk[p - 1]);
        uint64_t m = __ballot(head);
        if (head) sseg[nseg + __popcll(m & lanemask_lt())] = (uint16_t)p;
        nseg += __popcll(m);
    }
    if (lane == 0) sseg[nseg] = (uint16_t)nu;
    wave_lds_sync();
    for (uint32_t s = lane; s < nseg; s += WAVE) {
        uint32_t a = sseg[s], b = sseg[s + 1];
        AT acc = 0;
        in_order_chunks<8, AT>(a, b, [&](uint32_t i) { return sv[i]; },
                               [&](AT v) { acc = acc + v; });
        uint32_t r = sk[a];
        if (r < nrows) y[r] = acc;
    }
    zero_empty_rows(sk, sseg, nseg, d, nrows, [&](uint32_t z) { y[z] = (AT)0; });
}

} // namespace ispv
