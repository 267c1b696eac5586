/*
 * tune.c -- the model half of the AXT autotuner.
 *
 * Candidate layouts are ranked by the bytes a column-map product moves,
 * computed exactly from the padded lengths (getArraysLenAXT,
 * intSpmv_main.c:1118-1148) without building any array, with x priced by
 * the cache lines its gathers touch (x_cost).  The device half
 * (int_spmv_amd/tune.py) times the best few candidates and keeps the fastest.
 */
#include "intspmv_host.h"

#include <stdlib.h>
#include <string.h>

/* x bytes one stored entry costs per pass.  Gathered rows are scattered, so
 * each request is priced as the 64-byte line it touches: a narrow pass loads
 * kc * sx bytes of the entry's x row per lane in 16-byte requests, one line
 * each; a wide pass (tile height 1, kc >= 2 * 16 / sx: the lanes of a group
 * share the row) reads the kc * sx bytes once, in whole lines.  Padding slots
 * gather nothing. */
static uint64_t x_cost(uint32_t th, uint32_t kc, uint32_t sx)
{
    const uint32_t v = 16 / sx;
    if (th == 1 && kc >= 2 * v) return ((uint64_t)kc * sx + 63) / 64 * 64;
    return ((uint64_t)kc * sx + 15) / 16 * 64;
}

int intspmv_tune_axt(const intspmv_csr *m, uint32_t k, uint32_t sv, uint32_t sx,
                     intspmv_tune_choice *out, uint32_t cap, uint32_t *n)
{
    if (!m || !m->rl || k == 0 || k > 32 || !n || (sx != 4 && sx != 8)) return INTSPMV_HOST_EARG;
    static const uint32_t ths[] = {1, 2, 4, 8, 16};
    static const uint32_t kcs[] = {1, 2, 4, 8, 16, 32};
    intspmv_tune_choice all[30];
    uint32_t cnt = 0;
    uint32_t *con = (uint32_t *)malloc((m->nrows ? m->nrows : 1) * sizeof(uint32_t));
    if (!con) return INTSPMV_HOST_ENOMEM;
    for (int a = 0; a < 5; a++) {
        uint32_t tileN, lenAX, lenSEC, colN;
        if (intspmv_axt_lengths(m, 8, ths[a], "NOC", con, &tileN, &lenAX, &lenSEC, &colN)) continue;
        for (int b = 0; b < 6; b++) {
            const uint32_t kc = kcs[b], v = 16 / sx;
            const int wide = ths[a] == 1 && kc >= 2 * v;
            if (kc > 1 && kc / 2 >= k) continue;  /* a narrower width covers k */
            if (kc > 8 && !wide) continue;        /* narrow passes take at most 8 */
            uint64_t passes = (k + kc - 1) / kc, slots = lenAX / 2;
            uint64_t bytes = passes * (slots * (sv + 4) + (uint64_t)m->nnz * x_cost(ths[a], kc, sx) +
                                       4ull * lenSEC) +
                             (uint64_t)k * m->nrows * sx;
            all[cnt++] = (intspmv_tune_choice){ths[a], kc, slots, bytes};
        }
    }
    free(con);
    /* insertion sort by bytes, ties by smaller th then larger kc */
    for (uint32_t i = 1; i < cnt; i++) {
        intspmv_tune_choice t = all[i];
        uint32_t j = i;
        while (j > 0 && (all[j - 1].bytes > t.bytes ||
                         (all[j - 1].bytes == t.bytes && all[j - 1].tileH > t.tileH))) {
            all[j] = all[j - 1];
            j--;
        }
        all[j] = t;
    }
    *n = cnt;
    if (out) memcpy(out, all, (cap < cnt ? cap : cnt) * sizeof(intspmv_tune_choice));
    return INTSPMV_HOST_OK;
}
