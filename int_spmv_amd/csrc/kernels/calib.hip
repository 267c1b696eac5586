// calib.hip -- ceilings of the memory system that bound the gather-bound
// products, measured on the device the library runs on.
//
// The column-map, CSR and K1 products read x at each entry's column: one
// request per entry that no other lane of the instruction shares, unless the
// columns are close.  Their rate is then set by how many such requests the
// L1 / L2 / fabric path serves per second, not by HBM bandwidth.
// intspmv_gather_rate times a kernel that does nothing else: each thread
// streams a 4-byte index (as acol is streamed) and gathers one element at it,
// eight gathers in flight per thread, from a buffer of a given span.  A span
// inside one XCD's L2 (4 MB) gives the L2 request ceiling, the span of a
// product's x the ceiling for uniformly placed columns.
// intspmv_stream_rate times the other ceiling, that of the streaming products:
// a plain read of a buffer of a given size with 16-byte nontemporal loads,
// every load instruction covering whole lines (the best of 1 to 8 KB per wave).
// No reference counterpart: measurement utilities for bench.py.
#include <algorithm>

#include "common.hpp"

namespace ispv {

__device__ __forceinline__ uint32_t mix32(uint32_t v)
{
    // a 32-bit finaliser (Murmur3 fmix32): well-spread indexes from a counter
    v ^= v >> 16;
    v *= 0x85EBCA6Bu;
    v ^= v >> 13;
    v *= 0xC2B2AE35u;
    v ^= v >> 16;
    return v;
}

__global__ __launch_bounds__(256) void calib_fill_kernel(uint32_t *__restrict__ idx, size_t count,
                                                         uint32_t nelem, uint32_t seed)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) idx[i] = mix32((uint32_t)i * 2654435761u ^ seed) % nelem;
}

template <typename T>
__device__ __forceinline__ double calib_sum(const T &v) { return (double)v; }
template <>
__device__ __forceinline__ double calib_sum<d2v>(const d2v &v) { return v[0] + v[1]; }

// Eight gathers in flight per thread; one output per thread keeps them live.
template <typename T>
__global__ __launch_bounds__(256) void calib_gather_kernel(const uint32_t *__restrict__ idx,
                                                           const T *__restrict__ buf, size_t count,
                                                           double *__restrict__ out)
{
    constexpr int U = 8;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (size_t i = t; i < count; i += stride * U) {
        uint32_t c[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const size_t j = i + u * stride;
            c[u] = j < count ? ntload(idx + j) : 0u;
        }
        T v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = buf[c[u]];
#pragma unroll
        for (int u = 0; u < U; u++) acc += calib_sum(v[u]);
    }
    out[t] = acc;
}

// Reads n16 16-byte words the way the streaming products read their blocks:
// each wave one contiguous chunk of U 1-KB loads (64 lanes x 16 B), all in
// flight, workgroups XCD-remapped so each XCD reads a contiguous range.  The
// sum is stored only if it is NaN (never, for the zero-filled buffer), which
// keeps the loads live without write traffic.
template <int U>
__global__ __launch_bounds__(256) void calib_stream_kernel(const d2v *__restrict__ buf, size_t n16,
                                                           double *__restrict__ out)
{
    const uint32_t g = xcd_remap(blockIdx.x, gridDim.x);
    const size_t chunk = (size_t)g * WAVES + threadIdx.x / WAVE;
    const size_t base = chunk * (U * WAVE) + threadIdx.x % WAVE;
    d2v v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const size_t j = base + (size_t)u * WAVE;
        v[u] = j < n16 ? ntload(buf + j) : d2v{0.0, 0.0};
    }
    loads_issued();
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < U; u++) acc += v[u][0] + v[u][1];
    if (acc != acc) out[chunk] = acc;
}

} // namespace ispv

extern "C" int intspmv_stream_rate(intspmv_ctx *ctx, size_t bytes, uint32_t reps,
                                   double *bytes_per_s)
{
    using namespace ispv;
    if (!ctx || !bytes_per_s || reps == 0 || bytes < 16) return INTSPMV_ERR_ARG;
    const size_t n16 = bytes / 16;
    // one workgroup per WAVES chunks of the smallest chunk size (1 KB)
    if ((n16 + WAVES * WAVE - 1) / (WAVES * WAVE) > 0x7FFFFFFFu) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    void *buf = nullptr;
    double *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s = ctx->stream;
    const size_t max_chunks = (n16 + WAVE - 1) / WAVE;
    hipError_t e = hipMalloc(&buf, n16 * 16);
    if (e == hipSuccess) e = hipMalloc(&out, (max_chunks + WAVES) * sizeof(double));
    if (e == hipSuccess) e = hipMemsetAsync(buf, 0, n16 * 16, s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // the best of four chunk sizes per wave (1, 2, 4, 8 KB): which one the
    // memory system serves fastest varies with the size read
    double best = 0.0;
    for (int U = 1; U <= 8 && e == hipSuccess; U *= 2) {
        const size_t per_wg = (size_t)U * WAVE * WAVES;
        const unsigned grid = (unsigned)((n16 + per_wg - 1) / per_wg);
        auto launch = [&]() {
            if (U == 1) calib_stream_kernel<1><<<grid, BLOCK_THREADS, 0, s>>>((const d2v *)buf, n16, out);
            else if (U == 2) calib_stream_kernel<2><<<grid, BLOCK_THREADS, 0, s>>>((const d2v *)buf, n16, out);
            else if (U == 4) calib_stream_kernel<4><<<grid, BLOCK_THREADS, 0, s>>>((const d2v *)buf, n16, out);
            else calib_stream_kernel<8><<<grid, BLOCK_THREADS, 0, s>>>((const d2v *)buf, n16, out);
            return hipGetLastError();
        };
        for (int w = 0; w < 2 && e == hipSuccess; w++) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (uint32_t r = 0; r < reps && e == hipSuccess; r++) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0.0f) best = std::max(best, (double)(n16 * 16) * reps / (ms * 1e-3));
    }
    if (e == hipSuccess) *bytes_per_s = best;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    if (buf) (void)hipFree(buf);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) {
        ctx->last_err = e;
        return e == hipErrorOutOfMemory ? INTSPMV_ERR_NOMEM : INTSPMV_ERR_HIP;
    }
    return INTSPMV_OK;
}

extern "C" int intspmv_gather_rate(intspmv_ctx *ctx, size_t span_bytes, uint32_t elem_bytes,
                                   size_t count, uint32_t reps, double *gathers_per_s)
{
    using namespace ispv;
    if (!ctx || !gathers_per_s || reps == 0 || count == 0 || count > 0xFFFFFFFFu ||
        (elem_bytes != 4 && elem_bytes != 8 && elem_bytes != 16) || span_bytes < elem_bytes ||
        span_bytes / elem_bytes > 0xFFFFFFFFu)
        return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    const uint32_t nelem = (uint32_t)(span_bytes / elem_bytes);
    const unsigned grid = 256 * 8, block = 256;
    uint32_t *idx = nullptr;
    void *buf = nullptr;
    double *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s = ctx->stream;
    hipError_t e = hipMalloc(&idx, count * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&buf, (size_t)nelem * elem_bytes);
    if (e == hipSuccess) e = hipMalloc(&out, (size_t)grid * block * sizeof(double));
    if (e == hipSuccess) e = hipMemsetAsync(buf, 0, (size_t)nelem * elem_bytes, s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        calib_fill_kernel<<<(unsigned)((count + 255) / 256), 256, 0, s>>>(idx, count, nelem, 12345u);
        e = hipGetLastError();
    }
    auto launch = [&]() {
        if (elem_bytes == 4)
            calib_gather_kernel<float><<<grid, block, 0, s>>>(idx, (const float *)buf, count, out);
        else if (elem_bytes == 8)
            calib_gather_kernel<double><<<grid, block, 0, s>>>(idx, (const double *)buf, count, out);
        else
            calib_gather_kernel<d2v><<<grid, block, 0, s>>>(idx, (const d2v *)buf, count, out);
        return hipGetLastError();
    };
    for (int w = 0; w < 2 && e == hipSuccess; w++) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (uint32_t r = 0; r < reps && e == hipSuccess; r++) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) *gathers_per_s = ms > 0.0f ? (double)count * reps / (ms * 1e-3) : 0.0;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    if (idx) (void)hipFree(idx);
    if (buf) (void)hipFree(buf);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) {
        ctx->last_err = e;
        return e == hipErrorOutOfMemory ? INTSPMV_ERR_NOMEM : INTSPMV_ERR_HIP;
    }
    return INTSPMV_OK;
}
