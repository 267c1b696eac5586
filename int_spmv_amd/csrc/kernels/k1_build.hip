// k1_build.hip -- the K1 conversion (getFormatDataK1, intSpmv_main.c:841-905)
// on the device, from a CSR already in device memory, so that a caller whose
// matrix lives on the GPU never round-trips it through the host.
//
//   permi   rows by length, longest first, ties in row order: a stable radix
//           sort (rocPRIM) of the keys 0xFFFFFFFF - rl (orderFunction /
//           getArrayPermiK1, 775-796; the host builder's stable counting
//           sort gives the same order); entries past nrows are 0
//   nmc     length of each chunk's first (longest) row; chp the exclusive
//           prefix of 8 * nmc (getArraysNmcChpK1, 800-813)
//   val/col element k of the chunk's row l at chp[c] + 8k + l, padding 0
//           (getArraysValColK1, 817-837): one thread per output slot finds
//           its chunk by binary search over chp
//
// The arrays equal the host builder's bit for bit.
#include <string.h> // rocPRIM's headers use memset without including it

#include <algorithm>

#include <rocprim/rocprim.hpp>

#include "common.hpp"

namespace ispv {

__global__ void k1_keys_kernel(uint32_t n, const uint32_t *__restrict__ row,
                               uint32_t *__restrict__ key, uint32_t *__restrict__ idx)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    key[r] = 0xFFFFFFFFu - (row[r + 1] - row[r]);
    idx[r] = r;
}

__global__ void k1_nmc_kernel(uint32_t chunkNum, const uint32_t *__restrict__ key_sorted,
                              uint32_t *__restrict__ nmc, uint64_t *__restrict__ w8)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= chunkNum) return;
    const uint32_t l = 0xFFFFFFFFu - key_sorted[(size_t)8 * c];
    nmc[c] = l;
    w8[c] = 8ull * l;
}

__global__ void k1_chp_kernel(uint32_t chunkNum, const uint64_t *__restrict__ pref,
                              uint32_t *__restrict__ chp)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < chunkNum) chp[c] = (uint32_t)pref[c];
}

__global__ void k1_fill_kernel(uint32_t *__restrict__ p, uint32_t n, uint32_t v)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k1_scatter_kernel(uint32_t chunkNum, uint32_t nrows, uint32_t lenVC,
                                  const uint32_t *__restrict__ chp,
                                  const uint32_t *__restrict__ permi,
                                  const uint32_t *__restrict__ row,
                                  const uint32_t *__restrict__ col,
                                  const double *__restrict__ val, double *__restrict__ val_out,
                                  uint32_t *__restrict__ col_out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= lenVC) return;
    // last chunk with chp[c] <= p (chunks of length 0 sit at the end, at lenVC)
    uint32_t lo = 0, hi = chunkNum - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (chp[mid] <= p) lo = mid;
        else hi = mid - 1;
    }
    const uint32_t off = p - chp[lo], k = off / 8, s = 8 * lo + off % 8;
    double v = 0.0;
    uint32_t cc = 0;
    if (s < nrows) {
        const uint32_t r = permi[s], a = row[r];
        if (k < row[r + 1] - a) {
            v = val[a + k];
            cc = col[a + k];
        }
    }
    val_out[p] = v;
    col_out[p] = cc;
}

} // namespace ispv

using namespace ispv;

extern "C" int intspmv_k1_build_device(intspmv_ctx *ctx, uint32_t nrows, const uint32_t *row,
                                       const uint32_t *col, const double *val, uint32_t *permi,
                                       uint32_t *nmc, uint32_t *chp, double *val_out,
                                       uint32_t *col_out, uint32_t cap_vc, uint32_t *lenVC)
{
    if (!ctx || !lenVC || !permi || (nrows && (!row || !nmc || !chp))) return INTSPMV_ERR_ARG;
    if ((val_out == nullptr) != (col_out == nullptr)) return INTSPMV_ERR_ARG;
    int st = ispv_enter(ctx);
    if (st) return st;
    hipStream_t s = ctx->stream;
    const uint32_t chunkNum = (nrows + 7) / 8, plen = (chunkNum + 1) * 8;
    *lenVC = 0;
    uint32_t *key = nullptr, *key_sorted = nullptr, *idx = nullptr;
    uint64_t *w8 = nullptr, *pref = nullptr;
    void *tmp = nullptr;
    size_t tmp_sort = 0, tmp_scan = 0;
    hipError_t e = hipSuccess;
    auto grid = [](uint32_t n) { return dim3((n + 255) / 256 ? (n + 255) / 256 : 1); };
    // permi: zero, then the sorted rows in front
    k1_fill_kernel<<<grid(plen), 256, 0, s>>>(permi, plen, 0u);
    if (nrows) {
        e = hipMalloc(&key, (size_t)nrows * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMalloc(&key_sorted, (size_t)nrows * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMalloc(&idx, (size_t)nrows * sizeof(uint32_t));
        if (e == hipSuccess) e = hipMalloc(&w8, (size_t)chunkNum * sizeof(uint64_t));
        if (e == hipSuccess) e = hipMalloc(&pref, (size_t)chunkNum * sizeof(uint64_t));
        if (e == hipSuccess)
            e = rocprim::radix_sort_pairs(nullptr, tmp_sort, key, key_sorted, idx, permi, nrows, 0,
                                          32, s);
        if (e == hipSuccess)
            e = rocprim::exclusive_scan(nullptr, tmp_scan, w8, pref, (uint64_t)0, chunkNum,
                                        rocprim::plus<uint64_t>(), s);
        if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(std::max(tmp_sort, tmp_scan), 16));
        if (e == hipSuccess) {
            k1_keys_kernel<<<grid(nrows), 256, 0, s>>>(nrows, row, key, idx);
            e = rocprim::radix_sort_pairs(tmp, tmp_sort, key, key_sorted, idx, permi, nrows, 0, 32,
                                          s);
        }
        if (e == hipSuccess) {
            k1_nmc_kernel<<<grid(chunkNum), 256, 0, s>>>(chunkNum, key_sorted, nmc, w8);
            e = rocprim::exclusive_scan(tmp, tmp_scan, w8, pref, (uint64_t)0, chunkNum,
                                        rocprim::plus<uint64_t>(), s);
        }
        uint64_t last[2] = {0, 0};
        if (e == hipSuccess) {
            k1_chp_kernel<<<grid(chunkNum), 256, 0, s>>>(chunkNum, pref, chp);
            e = hipMemcpyAsync(&last[0], pref + chunkNum - 1, sizeof(uint64_t),
                               hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipMemcpyAsync(&last[1], w8 + chunkNum - 1, sizeof(uint64_t),
                                   hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
        }
        const uint64_t total = last[0] + last[1];
        if (e == hipSuccess && total > 0xFFFFFFFFull) st = INTSPMV_ERR_OVERFLOW;
        if (e == hipSuccess && !st) {
            *lenVC = (uint32_t)total;
            if (val_out) {
                if (cap_vc < *lenVC) st = INTSPMV_ERR_OVERFLOW;
                else if (*lenVC)
                    k1_scatter_kernel<<<grid(*lenVC), 256, 0, s>>>(chunkNum, nrows, *lenVC, chp,
                                                                   permi, row, col, val, val_out,
                                                                   col_out);
            }
        }
        if (e == hipSuccess) e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s); // before the temporaries go
    }
    (void)hipFree(key);
    (void)hipFree(key_sorted);
    (void)hipFree(idx);
    (void)hipFree(w8);
    (void)hipFree(pref);
    (void)hipFree(tmp);
    if (e != hipSuccess) {
        ctx->last_err = e;
        return INTSPMV_ERR_HIP;
    }
    return st;
}
