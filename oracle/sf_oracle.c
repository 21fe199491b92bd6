/*
 * sf_oracle.c -- CPU oracle for the alignment hot path.  TEST INFRASTRUCTURE ONLY (see
 * sf_oracle.h): loaded by tests/, __graft_entry__.smoke() and bench.py's CPU baseline, never by
 * the product library.
 *
 * A scalar restatement of spearfish crates/map/src/align.rs.  Each function cites the lines it
 * follows.  Arithmetic is i32 as in the reference (built with -fwrapv so that C matches Rust's
 * release-mode wrapping should a value ever leave the range the reference assumes).
 */
#include "sf_oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define W 64
#define NEG SFO_NEG

static inline int32_t max32(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t min32(int32_t a, int32_t b) { return a < b ? a : b; }
static inline uint64_t sat_add_u(uint64_t a, uint64_t b) { return a > UINT64_MAX - b ? UINT64_MAX : a + b; }
static inline uint64_t sat_sub_u(uint64_t a, uint64_t b) { return a > b ? a - b : 0; }
static inline int32_t sat_add_i(int32_t a, int32_t b) {
    int64_t r = (int64_t)a + b;
    return r > INT32_MAX ? INT32_MAX : r < INT32_MIN ? INT32_MIN : (int32_t)r;
}
static inline int32_t sat_sub_i(int32_t a, int32_t b) {
    int64_t r = (int64_t)a - b;
    return r > INT32_MAX ? INT32_MAX : r < INT32_MIN ? INT32_MIN : (int32_t)r;
}

/* ------------------------------------------------------------------------------------------ */
/* Myers                                                                                       */
/* ------------------------------------------------------------------------------------------ */

/* align.rs:49-66 `block`. */
static inline int32_t block(uint64_t *pv, uint64_t *mv, uint64_t eq, int32_t hin, unsigned bit) {
    uint64_t neg = (uint64_t)(hin < 0), pos = (uint64_t)(hin > 0);
    uint64_t p = *pv, m = *mv;
    uint64_t xv = eq | m;
    eq |= neg;
    uint64_t xh = (((eq & p) + p) ^ p) | eq;
    uint64_t ph = m | ~(xh | p);
    uint64_t mh = p & xh;
    int32_t hout = (int32_t)((ph >> bit) & 1) - (int32_t)((mh >> bit) & 1);
    ph = (ph << 1) | pos;
    mh = (mh << 1) | neg;
    *pv = mh | ~(xv | ph);
    *mv = ph & xv;
    return hout;
}

/* align.rs:68-85 `Pattern::new`: peq[b * 4 + c]. */
static uint64_t *build_peq(const uint8_t *p, uint64_t m) {
    uint64_t blocks = (m + W - 1) / W;
    uint64_t *peq = (uint64_t *)calloc(blocks * 4 + 4, sizeof(uint64_t));
    for (uint64_t i = 0; i < m; i++) peq[(i / W) * 4 + (p[i] & 3)] |= 1ull << (i % W);
    return peq;
}

/* align.rs:92-168 `distance_with`. */
static uint32_t distance_with(const uint64_t *peq, uint64_t m, const uint8_t *t, uint64_t n, int mode,
                              uint32_t k_in) {
    if (m == 0) {
        uint32_t d = mode == SFO_GLOBAL ? (uint32_t)n : 0;
        return d <= k_in ? d : SFO_NONE;
    }
    uint32_t kc = k_in < (uint32_t)(m + n) ? k_in : (uint32_t)(m + n);
    int32_t k = (int32_t)kc;
    uint64_t blocks = (m + W - 1) / W;
    unsigned last_bit = (unsigned)((m - 1) % W);
#define ROWS(b) ((b) + 1 == blocks ? (int32_t)last_bit + 1 : W)
#define BIT(b) ((b) + 1 == blocks ? last_bit : (unsigned)(W - 1))
    int32_t top = mode == SFO_INFIX ? 0 : 1;
    uint64_t *pv = (uint64_t *)malloc(blocks * sizeof(uint64_t));
    uint64_t *mv = (uint64_t *)malloc(blocks * sizeof(uint64_t));
    int32_t *score = (int32_t *)malloc(blocks * sizeof(int32_t));
    for (uint64_t b = 0; b < blocks; b++) {
        pv[b] = UINT64_MAX;
        mv[b] = 0;
        score[b] = (int32_t)(b * W) + ROWS(b);
    }
    uint64_t kb = ((uint64_t)k + 1 + W - 1) / W;
    int64_t last = (int64_t)(blocks < kb ? blocks : kb) - 1;
    int32_t best = ((uint64_t)(last + 1) == blocks && score[blocks - 1] <= k) ? score[blocks - 1] : INT32_MAX;
    uint32_t result = SFO_NONE;
    int done = 0;
    for (uint64_t j = 0; j < n; j++) {
        unsigned c = t[j] & 3;
        int32_t hin = top;
        for (int64_t b = 0; b <= last; b++) {
            hin = block(&pv[b], &mv[b], peq[b * 4 + c], hin, BIT((uint64_t)b));
            score[b] += hin;
        }
        uint64_t lb = (uint64_t)last;
        if (lb + 1 < blocks && score[lb] - hin <= k && ((peq[(lb + 1) * 4 + c] & 1) != 0 || hin < 0)) {
            uint64_t nb = lb + 1;
            pv[nb] = UINT64_MAX;
            mv[nb] = 0;
            int32_t h = block(&pv[nb], &mv[nb], peq[nb * 4 + c], hin, BIT(nb));
            score[nb] = score[lb] - hin + ROWS(nb) + h;
            last += 1;
        } else {
            while (last >= 0 && score[last] >= k + W) last -= 1;
            if (last < 0) {
                if (mode == SFO_GLOBAL) result = SFO_NONE;
                else result = best <= k ? (uint32_t)best : SFO_NONE;
                done = 1;
                break;
            }
        }
        if ((uint64_t)(last + 1) == blocks) {
            int32_t s = score[blocks - 1];
            if (mode == SFO_GLOBAL) {
                if (j + 1 == n) best = s;
            } else {
                best = min32(best, s);
            }
        } else if (mode == SFO_GLOBAL && j + 1 == n) {
            best = INT32_MAX;
        }
    }
    if (!done) {
        if (mode == SFO_GLOBAL && n == 0) best = (int32_t)m;
        result = best <= k ? (uint32_t)best : SFO_NONE;
    }
    free(pv);
    free(mv);
    free(score);
#undef ROWS
#undef BIT
    return result;
}

uint32_t sfo_distance(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode,
                      uint32_t k) {
    uint64_t *peq = build_peq(pat, m);
    uint32_t d = distance_with(peq, m, txt, n, mode, k);
    free(peq);
    return d;
}

/* align.rs:552-653 `batch_lanes::<L>` with L = 8 (the AVX-512 width), texts in chunks of 8
 * (align.rs:205). */
#define LANES 8
static void batch_lanes8(const uint64_t *peq, uint64_t m, const uint8_t *const *txts, const uint64_t *lens_in,
                         uint64_t nt, int mode, uint32_t k_in, uint32_t *out) {
    uint64_t n_max = 0;
    for (uint64_t l = 0; l < nt; l++) n_max = lens_in[l] > n_max ? lens_in[l] : n_max;
    uint32_t kc = k_in < (uint32_t)(m + n_max) ? k_in : (uint32_t)(m + n_max);
    int32_t kk = (int32_t)kc;
    uint64_t blocks = (m + W - 1) / W;
    unsigned last_bit = (unsigned)((m - 1) % W);
#define ROWS(b) ((b) + 1 == blocks ? (int32_t)last_bit + 1 : W)
#define BIT(b) ((b) + 1 == blocks ? last_bit : (unsigned)(W - 1))
    uint64_t lens[LANES] = {0};
    for (uint64_t l = 0; l < nt; l++) lens[l] = lens_in[l];
    uint64_t *pv = (uint64_t *)malloc(blocks * LANES * sizeof(uint64_t));
    uint64_t *mv = (uint64_t *)malloc(blocks * LANES * sizeof(uint64_t));
    int32_t *score = (int32_t *)malloc(blocks * LANES * sizeof(int32_t));
    for (uint64_t b = 0; b < blocks; b++)
        for (int l = 0; l < LANES; l++) {
            pv[b * LANES + l] = UINT64_MAX;
            mv[b * LANES + l] = 0;
            score[b * LANES + l] = (int32_t)(b * W) + ROWS(b);
        }
    int32_t kw = kk + W;
    uint64_t kb = ((uint64_t)kk + 1 + W - 1) / W;
    int64_t last = (int64_t)(blocks < kb ? blocks : kb) - 1;
    int32_t best0 = ((uint64_t)(last + 1) == blocks && (int32_t)m <= kk) ? (int32_t)m : INT32_MAX;
    int32_t best[LANES];
    for (int l = 0; l < LANES; l++) {
        if (mode == SFO_GLOBAL) best[l] = lens[l] == 0 ? (int32_t)m : INT32_MAX;
        else best[l] = best0;
    }
    int32_t h0 = mode == SFO_INFIX ? 0 : 1;
    for (uint64_t j = 0; j < n_max; j++) {
        unsigned c[LANES];
        int live[LANES];
        for (int l = 0; l < LANES; l++) {
            live[l] = (uint64_t)l < nt && j < lens[l];
            c[l] = (live[l] ? txts[l][j] : 0) & 3; /* code 0 past a lane's text (align.rs:574-581) */
        }
        int32_t h[LANES];
        for (int l = 0; l < LANES; l++) h[l] = h0;
        for (int64_t b = 0; b <= last; b++)
            for (int l = 0; l < LANES; l++) {
                h[l] = block(&pv[b * LANES + l], &mv[b * LANES + l], peq[b * 4 + c[l]], h[l], BIT((uint64_t)b));
                score[b * LANES + l] += h[l];
            }
        uint64_t lb = (uint64_t)last;
        int added = 0;
        if (lb + 1 < blocks) {
            uint64_t nb = lb + 1;
            int32_t top[LANES];
            int want = 0;
            for (int l = 0; l < LANES; l++) {
                top[l] = score[lb * LANES + l] - h[l];
                if (live[l] && top[l] <= kk && ((peq[nb * 4 + c[l]] & 1) != 0 || h[l] < 0)) want = 1;
            }
            if (want) {
                for (int l = 0; l < LANES; l++) {
                    pv[nb * LANES + l] = UINT64_MAX;
                    mv[nb * LANES + l] = 0;
                    int32_t hn = block(&pv[nb * LANES + l], &mv[nb * LANES + l], peq[nb * 4 + c[l]], h[l], BIT(nb));
                    score[nb * LANES + l] = top[l] + ROWS(nb) + hn;
                }
                last += 1;
                added = 1;
            }
        }
        if (!added) {
            for (;;) {
                if (last < 0) break;
                int any = 0;
                for (int l = 0; l < LANES; l++)
                    if (live[l] && score[(uint64_t)last * LANES + l] < kw) any = 1;
                if (any) break;
                last -= 1;
            }
            if (last < 0) break;
        }
        if ((uint64_t)(last + 1) == blocks) {
            for (int l = 0; l < LANES; l++) {
                int32_t s = score[(blocks - 1) * LANES + l];
                if (mode == SFO_GLOBAL) {
                    if (j + 1 == lens[l]) best[l] = s;
                } else if (live[l]) {
                    best[l] = min32(best[l], s);
                }
            }
        }
    }
    for (uint64_t l = 0; l < nt; l++) {
        uint32_t klc = k_in < (uint32_t)(m + lens[l]) ? k_in : (uint32_t)(m + lens[l]);
        out[l] = best[l] <= (int32_t)klc ? (uint32_t)best[l] : SFO_NONE;
    }
    free(pv);
    free(mv);
    free(score);
#undef ROWS
#undef BIT
}

/* align.rs:170-200 `batch_distances` (the empty pattern goes through distance_with, line 198). */
void sfo_batch_distances(const uint8_t *pat, uint64_t m, const uint8_t *const *txts, const uint64_t *lens,
                         uint64_t nt, int mode, uint32_t k, uint32_t *out) {
    uint64_t *peq = build_peq(pat, m);
    if (m == 0) {
        for (uint64_t l = 0; l < nt; l++) out[l] = distance_with(peq, 0, txts[l], lens[l], mode, k);
    } else {
        for (uint64_t l0 = 0; l0 < nt; l0 += LANES) {
            uint64_t c = nt - l0 < LANES ? nt - l0 : LANES;
            batch_lanes8(peq, m, txts + l0, lens + l0, c, mode, k, out + l0);
        }
    }
    free(peq);
}

/* ---- the caller's re-capping (crates/map/src/lib.rs:851-911 `Mapper::banded`, and the test
 * definition `sequential`, lib.rs:3016-3034) ---- */

static int same_text(const uint8_t *a, uint64_t na, const uint8_t *b, uint64_t nb) {
    return na == nb && (na == 0 || memcmp(a, b, na) == 0);
}

static uint32_t sat_add_u32(uint32_t a, uint32_t b) { return a > UINT32_MAX - b ? UINT32_MAX : a + b; }

/* lib.rs:851-911: texts deduplicated in first-occurrence order (`distinct`: no repeats, no
 * hashing), distinct texts in batches of eight (batch_distances, lib.rs:891-897) at the threshold
 * k = best + cap in force at each batch's first text, each text's value re-capped in order at its
 * own threshold (lib.rs:898-904), a repeat taking its first occurrence's value re-capped
 * (lib.rs:884-889).  The repeat lookup is a linear scan here (the reference hashes; the same
 * first-occurrence map). */
void sfo_banded(const uint8_t *read, uint64_t m, const uint8_t *const *texts, const uint64_t *lens, uint64_t nt,
                int mode, uint32_t cap, int distinct, uint32_t *out) {
    uint64_t *id = (uint64_t *)malloc((nt ? nt : 1) * sizeof(uint64_t));
    uint64_t *at = (uint64_t *)malloc((nt ? nt : 1) * sizeof(uint64_t)); /* distinct u -> first text index */
    uint64_t nu = 0;
    for (uint64_t i = 0; i < nt; i++) {
        uint64_t u = nu;
        if (!distinct)
            for (uint64_t v = 0; v < nu; v++)
                if (same_text(texts[at[v]], lens[at[v]], texts[i], lens[i])) {
                    u = v;
                    break;
                }
        if (u == nu) at[nu++] = i;
        id[i] = u;
    }
    uint32_t *raw = (uint32_t *)malloc((nu ? nu : 1) * sizeof(uint32_t));
    uint32_t *val = (uint32_t *)calloc(nu ? nu : 1, sizeof(uint32_t));
    const uint8_t **bt = (const uint8_t **)malloc(8 * sizeof(uint8_t *));
    uint64_t bl[8];
    uint64_t nraw = 0;
    uint32_t best = UINT32_MAX;
    for (uint64_t i = 0; i < nt; i++) {
        const uint64_t u = id[i];
        const uint32_t k = sat_add_u32(best, cap);
        if (at[u] != i) {
            const uint32_t d = val[u];
            out[i] = d <= k ? d : sat_add_u32(k, 1);
            continue;
        }
        if (u == nraw) {
            const uint64_t c = nu - u < 8 ? nu - u : 8;
            for (uint64_t q = 0; q < c; q++) {
                bt[q] = texts[at[u + q]];
                bl[q] = lens[at[u + q]];
            }
            sfo_batch_distances(read, m, bt, bl, c, mode, k, raw + u);
            nraw += c;
        }
        const uint32_t d = (raw[u] != SFO_NONE && raw[u] <= k) ? raw[u] : sat_add_u32(k, 1);
        if (d < best) best = d;
        val[u] = d;
        out[i] = d;
    }
    free(id), free(at), free(raw), free(val), free(bt);
}

/* lib.rs:3016-3034 `sequential`: one distance_with per distinct text in order at k = best + cap
 * (None -> k + 1), a repeat its first occurrence's value re-capped at the current k. */
void sfo_sequential(const uint8_t *read, uint64_t m, const uint8_t *const *texts, const uint64_t *lens, uint64_t nt,
                    int mode, uint32_t cap, uint32_t *out) {
    uint32_t best = UINT32_MAX;
    for (uint64_t i = 0; i < nt; i++) {
        const uint32_t k = sat_add_u32(best, cap);
        uint64_t first = i;
        for (uint64_t j = 0; j < i; j++)
            if (same_text(texts[j], lens[j], texts[i], lens[i])) {
                first = j;
                break;
            }
        if (first != i) {
            const uint32_t d = out[first]; /* the first occurrence's value, as it was computed */
            /* (out[first] is the first occurrence's own result: lib.rs stores `seen[t] = d`) */
            out[i] = d <= k ? d : sat_add_u32(k, 1);
            continue;
        }
        const uint32_t r = sfo_distance(read, m, texts[i], lens[i], mode, k);
        const uint32_t d = r == SFO_NONE ? sat_add_u32(k, 1) : r;
        if (d < best) best = d;
        out[i] = d;
    }
}

/* D[m][j] for j = 0..n by a full bit-parallel pass (no cut-off): the last row of the
 * (mode-dependent) edit-distance matrix.  Own helper for sfo_locate. */
static void myers_lastrow(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, int mode, int32_t *row) {
    if (m == 0) {
        for (uint64_t j = 0; j <= n; j++) row[j] = mode == SFO_INFIX ? 0 : (int32_t)j;
        return;
    }
    uint64_t *peq = build_peq(p, m);
    uint64_t blocks = (m + W - 1) / W;
    unsigned last_bit = (unsigned)((m - 1) % W);
    uint64_t *pv = (uint64_t *)malloc(blocks * sizeof(uint64_t));
    uint64_t *mv = (uint64_t *)calloc(blocks, sizeof(uint64_t));
    for (uint64_t b = 0; b < blocks; b++) pv[b] = UINT64_MAX;
    int32_t s = (int32_t)m;
    row[0] = s;
    int32_t top = mode == SFO_INFIX ? 0 : 1;
    for (uint64_t j = 0; j < n; j++) {
        unsigned c = t[j] & 3;
        int32_t hin = top;
        for (uint64_t b = 0; b < blocks; b++)
            hin = block(&pv[b], &mv[b], peq[b * 4 + c], hin, b + 1 == blocks ? last_bit : W - 1);
        s += hin;
        row[j + 1] = s;
    }
    free(pv);
    free(mv);
    free(peq);
}

void sfo_locate(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode, uint32_t k,
                uint32_t *dist, uint32_t *end, uint32_t *start) {
    uint32_t d = sfo_distance(pat, m, txt, n, mode, k);
    *dist = d;
    if (d == SFO_NONE) {
        *end = *start = SFO_NONE;
        return;
    }
    if (mode == SFO_GLOBAL) {
        *end = (uint32_t)n;
        *start = 0;
        return;
    }
    int32_t *row = (int32_t *)malloc((n + 1) * sizeof(int32_t));
    myers_lastrow(pat, m, txt, n, mode, row);
    uint64_t e = 0;
    while (e <= n && row[e] != (int32_t)d) e++;
    *end = (uint32_t)e;
    if (mode == SFO_PREFIX) {
        *start = 0;
    } else {
        uint8_t *rp = (uint8_t *)malloc(m + 1), *rt = (uint8_t *)malloc(e + 1);
        for (uint64_t i = 0; i < m; i++) rp[i] = pat[m - 1 - i];
        for (uint64_t j = 0; j < e; j++) rt[j] = txt[e - 1 - j];
        myers_lastrow(rp, m, rt, e, SFO_PREFIX, row);
        uint64_t L = 0;
        while (L <= e && row[L] != (int32_t)d) L++;
        *start = (uint32_t)(e - L);
        free(rp);
        free(rt);
    }
    free(row);
}

/* ------------------------------------------------------------------------------------------ */
/* Affine                                                                                      */
/* ------------------------------------------------------------------------------------------ */

/* align.rs:984-995 `excl_prefix_max`. */
static inline void excl_prefix_max(int32_t *xs, uint64_t w, int32_t init) {
    int32_t g = init;
    for (uint64_t k = 0; k < w; k++) {
        int32_t next = max32(g, xs[k]);
        xs[k] = g;
        g = next;
    }
}

/* align.rs:1018-1072 `row_scalar`: hprev w+1 cells, fu w, tx w, nh/nf w+1 (the last gets the
 * NEG sentinel).  `eout` (own addition for the traceback) receives the horizontal-gap value of
 * each cell, gs[k] - ext*k after pass 2. */
static int32_t row_scalar(const int32_t *hprev, const int32_t *fu, const uint8_t *tx, int32_t *nh, int32_t *nf,
                          unsigned c, int32_t mat, int32_t mis, int32_t go, int32_t ext, int32_t e, int32_t floor,
                          uint64_t w, int32_t *hp, int32_t *gs, int32_t *eout) {
    nh[w] = NEG;
    nf[w] = NEG;
    for (uint64_t k = 0; k < w; k++) {
        int32_t fj = max32(max32(hprev[k + 1] - go, fu[k] - ext), NEG);
        int32_t s = c == (unsigned)(tx[k] & 3) ? mat : -mis;
        int32_t x = max32(max32(hprev[k] + s, fj), floor);
        hp[k] = x;
        gs[k] = x - go + ext * ((int32_t)k + 1);
        nf[k] = fj;
    }
    excl_prefix_max(gs, w, e);
    int32_t mx = NEG;
    for (uint64_t k = 0; k < w; k++) {
        int32_t v = max32(hp[k], gs[k] - ext * (int32_t)k);
        nh[k] = v;
        if (eout) eout[k] = gs[k] - ext * (int32_t)k;
        mx = max32(mx, v);
    }
    return mx;
}

/* One DP row: the signature of row_scalar, as the reference's `Rows::row` (align.rs:997-1016)
 * lets run_core take the scalar or the AVX-512 rows. */
typedef int32_t (*row_fn)(const int32_t *hprev, const int32_t *fu, const uint8_t *tx, int32_t *nh, int32_t *nf,
                          unsigned c, int32_t mat, int32_t mis, int32_t go, int32_t ext, int32_t e, int32_t floor,
                          uint64_t w, int32_t *hp, int32_t *gs, int32_t *eout);

#if defined(__x86_64__)
#include <immintrin.h>

/* The same row on 16 i32 lanes (the reference's AVX-512 rows, align.rs:1150-1267, restated; used
 * by the CPU baseline only).  Per 16 cells: the vertical and diagonal terms as in row_scalar, the
 * gap-open values gs = x - go + ext*(k + 1), their exclusive running max (an in-register prefix
 * max over shifts of 1, 2, 4 and 8 lanes, then one lane over with the carry from the cells
 * before), and H = max(x, that - ext*k).  Lanes past the row end are masked.  Same integer
 * operations, so the same results. */
__attribute__((target("avx512f,avx512bw,avx512vl"))) static int32_t
row_avx512(const int32_t *hprev, const int32_t *fu, const uint8_t *tx, int32_t *nh, int32_t *nf, unsigned c,
           int32_t mat, int32_t mis, int32_t go, int32_t ext, int32_t e, int32_t floor, uint64_t w, int32_t *hp,
           int32_t *gs, int32_t *eout) {
    (void)hp, (void)gs, (void)eout;
    nh[w] = NEG;
    nf[w] = NEG;
    const __m512i vneg = _mm512_set1_epi32(NEG), vfloor = _mm512_set1_epi32(floor);
    const __m512i vgo = _mm512_set1_epi32(go), vext = _mm512_set1_epi32(ext);
    const __m512i vmat = _mm512_set1_epi32(mat), vmis = _mm512_set1_epi32(-mis);
    const __m512i vc = _mm512_set1_epi32((int32_t)c), v3 = _mm512_set1_epi32(3);
    const __m512i lanes = _mm512_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
    __m512i vmx = vneg;
    int32_t carry = e; /* max of e and gs of every cell before this vector */
    for (uint64_t k = 0; k < w; k += 16) {
        const uint64_t rem = w - k;
        const __mmask16 mk = rem >= 16 ? (__mmask16)0xFFFF : (__mmask16)((1u << rem) - 1);
        const __m512i hd = _mm512_maskz_loadu_epi32(mk, hprev + k);
        const __m512i hu = _mm512_maskz_loadu_epi32(mk, hprev + k + 1);
        const __m512i fv = _mm512_maskz_loadu_epi32(mk, fu + k);
        const __m512i t = _mm512_and_si512(_mm512_cvtepu8_epi32(_mm_maskz_loadu_epi8(mk, tx + k)), v3);
        const __m512i fj = _mm512_max_epi32(_mm512_max_epi32(_mm512_sub_epi32(hu, vgo), _mm512_sub_epi32(fv, vext)),
                                            vneg);
        const __m512i sv = _mm512_mask_blend_epi32(_mm512_cmpeq_epi32_mask(t, vc), vmis, vmat);
        const __m512i x = _mm512_max_epi32(_mm512_max_epi32(_mm512_add_epi32(hd, sv), fj), vfloor);
        const __m512i ek = _mm512_mullo_epi32(_mm512_add_epi32(lanes, _mm512_set1_epi32((int32_t)k)), vext);
        const __m512i g = _mm512_mask_mov_epi32(vneg, mk, _mm512_add_epi32(_mm512_sub_epi32(x, vgo),
                                                                           _mm512_add_epi32(ek, vext)));
        __m512i p = _mm512_max_epi32(g, _mm512_alignr_epi32(g, vneg, 15));
        p = _mm512_max_epi32(p, _mm512_alignr_epi32(p, vneg, 14));
        p = _mm512_max_epi32(p, _mm512_alignr_epi32(p, vneg, 12));
        p = _mm512_max_epi32(p, _mm512_alignr_epi32(p, vneg, 8));
        const __m512i vcarry = _mm512_set1_epi32(carry);
        const __m512i excl = _mm512_max_epi32(_mm512_alignr_epi32(p, vcarry, 15), vcarry);
        const __m512i hv = _mm512_max_epi32(x, _mm512_sub_epi32(excl, ek));
        _mm512_mask_storeu_epi32(nh + k, mk, hv);
        _mm512_mask_storeu_epi32(nf + k, mk, fj);
        vmx = _mm512_mask_max_epi32(vmx, mk, vmx, hv);
        carry = max32(carry, _mm512_reduce_max_epi32(p));
    }
    return _mm512_reduce_max_epi32(vmx);
}

int sfo_avx512_available(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
           __builtin_cpu_supports("avx512vl");
}
#else
int sfo_avx512_available(void) { return 0; }
#endif

typedef struct {
    uint64_t below, above;
    int full;
    uint64_t n;
} band_t;

static band_t band_of(uint64_t m, uint64_t n, int mode, uint64_t band, uint64_t geom) {
    band_t b;
    b.full = mode == SFO_INFIX || band == SFO_UMAX;
    b.below = sat_add_u(band, sat_sub_u(m, geom));
    b.above = sat_add_u(band, sat_sub_u(geom, m));
    b.n = n;
    return b;
}

/* align.rs:2475-2481 `range`. */
static inline void range_of(const band_t *b, uint64_t i, uint64_t *lo, uint64_t *hi) {
    if (b->full) {
        *lo = 0;
        *hi = b->n;
    } else {
        *lo = sat_sub_u(i, b->below);
        uint64_t h = sat_add_u(i, b->above);
        *hi = h < b->n ? h : b->n;
    }
}

#if defined(__x86_64__)
/* The 16-bit form of the same row on 32 lanes, for passes whose scores provably fit (see narrow
 * below; the reference's run16_avx512, align.rs:1348-1877, restated as rows of i16 cells, not
 * its cache-line layout).  Saturating adds and subtractions keep every unreachable value at the
 * i16 floor, which orders below every real score as NEG does. */
#define NEG16 ((int16_t)-32768)
__attribute__((target("avx512f,avx512bw,avx512vl"))) static void
row16_avx512(const int16_t *hprev, const int16_t *fu, const uint8_t *tx, int16_t *nh, int16_t *nf, unsigned c,
             int16_t mat, int16_t mis, int16_t go, int16_t ext, int16_t e, uint64_t w) {
    nh[w] = NEG16;
    nf[w] = NEG16;
    const __m512i vneg = _mm512_set1_epi16(NEG16), vgo = _mm512_set1_epi16(go), vext = _mm512_set1_epi16(ext);
    const __m512i vmat = _mm512_set1_epi16(mat), vmis = _mm512_set1_epi16((int16_t)-mis);
    const __m512i vc = _mm512_set1_epi16((int16_t)c), v3 = _mm512_set1_epi16(3);
    int16_t lane_ix[32];
    for (int q = 0; q < 32; q++) lane_ix[q] = (int16_t)q;
    const __m512i lanes = _mm512_loadu_si512(lane_ix);
    /* Lane q reads lane q - d (d = 1, 2, 4, 8, 16) where q >= d. */
    __m512i back[5];
    __mmask32 from[5];
    for (int t = 0; t < 5; t++) {
        const int d = 1 << t;
        back[t] = _mm512_sub_epi16(lanes, _mm512_set1_epi16((int16_t)d));
        from[t] = (__mmask32)(0xFFFFFFFFu << d);
    }
    int16_t carry = e;
    for (uint64_t k = 0; k < w; k += 32) {
        const uint64_t rem = w - k;
        const __mmask32 mk = rem >= 32 ? (__mmask32)0xFFFFFFFFu : (__mmask32)((1u << rem) - 1);
        const __m512i hd = _mm512_maskz_loadu_epi16(mk, hprev + k);
        const __m512i hu = _mm512_maskz_loadu_epi16(mk, hprev + k + 1);
        const __m512i fv = _mm512_maskz_loadu_epi16(mk, fu + k);
        const __m512i t = _mm512_and_si512(_mm512_cvtepu8_epi16(_mm256_maskz_loadu_epi8(mk, tx + k)), v3);
        const __m512i fj = _mm512_max_epi16(_mm512_subs_epi16(hu, vgo), _mm512_subs_epi16(fv, vext));
        const __m512i sv = _mm512_mask_blend_epi16(_mm512_cmpeq_epi16_mask(t, vc), vmis, vmat);
        const __m512i x = _mm512_max_epi16(_mm512_adds_epi16(hd, sv), fj);
        const __m512i ek = _mm512_mullo_epi16(_mm512_add_epi16(lanes, _mm512_set1_epi16((int16_t)k)), vext);
        const __m512i g = _mm512_mask_mov_epi16(
            vneg, mk, _mm512_adds_epi16(_mm512_subs_epi16(x, vgo), _mm512_adds_epi16(ek, vext)));
        __m512i p = g;
        for (int q = 0; q < 5; q++)
            p = _mm512_max_epi16(p, _mm512_mask_permutexvar_epi16(vneg, from[q], back[q], p));
        const __m512i vcarry = _mm512_set1_epi16(carry);
        const __m512i excl = _mm512_max_epi16(_mm512_mask_permutexvar_epi16(vcarry, from[0], back[0], p), vcarry);
        const __m512i hv = _mm512_max_epi16(x, _mm512_subs_epi16(excl, ek));
        _mm512_mask_storeu_epi16(nh + k, mk, hv);
        _mm512_mask_storeu_epi16(nf + k, mk, fj);
        /* Lane 31 of the inclusive prefix max: the max over the vector (masked lanes are the floor). */
        const int16_t top = (int16_t)_mm_extract_epi16(_mm512_extracti32x4_epi32(p, 3), 7);
        carry = top > carry ? top : carry;
    }
}

/* align.rs:2381-2426 `narrow`: whether a banded Global/Prefix pass (no clip) fits 16-bit cells. */
static int narrow16(uint64_t m, int mode, const sfo_scoring *sc, uint64_t band, uint64_t geom) {
    if (mode == SFO_INFIX || band == SFO_UMAX) return 0;
    if (sc->mat < 0 || sc->mis < 0 || sc->open < 0 || sc->ext < 0) return 0;
    const uint64_t mat = (uint64_t)sc->mat, mis = (uint64_t)sc->mis, open = (uint64_t)sc->open,
                   ext = (uint64_t)sc->ext;
    const uint64_t below = sat_add_u(band, sat_sub_u(m, geom)), above = sat_add_u(band, sat_sub_u(geom, m));
    const uint64_t width = sat_add_u(sat_add_u(below, above), 66);
    const uint64_t gaps = sat_add_u(sat_add_u(below > above ? below : above, width), 1);
    const uint64_t low = sat_add_u(sat_add_u(sat_add_u(mis * m, open * 2), ext * gaps), mat);
    const uint64_t high = sat_add_u(mat * m, ext * width);
    return (low > high ? low : high) <= 32000 && width <= 30000;
}

/* run_core's score pass (Global or Prefix, no clip, cut or traceback) on 16-bit rows. */
__attribute__((target("avx512f,avx512bw,avx512vl"))) static void
run16_core(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode, const sfo_scoring *sc,
           uint64_t band, int32_t *score_out, uint64_t *col_out) {
    const int16_t mat = (int16_t)sc->mat, mis = (int16_t)sc->mis, ext = (int16_t)sc->ext;
    const int16_t go = (int16_t)(sc->open + sc->ext);
    band_t bd = band_of(m, n, mode, band, n);
    int16_t *h = (int16_t *)malloc((n + 2) * sizeof(int16_t));
    int16_t *f = (int16_t *)malloc((n + 2) * sizeof(int16_t));
    int16_t *nh = (int16_t *)malloc((n + 2) * sizeof(int16_t));
    int16_t *nf = (int16_t *)malloc((n + 2) * sizeof(int16_t));
    for (uint64_t j = 0; j < n + 2; j++) h[j] = f[j] = nh[j] = nf[j] = NEG16;
    uint64_t lo0, hi0;
    range_of(&bd, 0, &lo0, &hi0);
    for (uint64_t j = lo0; j <= hi0; j++) h[j] = j == 0 ? 0 : (int16_t)-(sc->open + sc->ext * (int32_t)j);
    uint64_t plo = lo0, phi = hi0;
    for (uint64_t i = 1; i <= m; i++) {
        uint64_t lo, hi;
        range_of(&bd, i, &lo, &hi);
        if (lo > hi) {
            *score_out = NEG;
            *col_out = 0;
            goto done;
        }
        uint64_t start = lo;
        int16_t e = NEG16;
        if (lo == 0) {
            const int16_t a = (int16_t)(h[0] - go < NEG16 ? NEG16 : h[0] - go);
            const int16_t b = (int16_t)(f[0] - ext < NEG16 ? NEG16 : f[0] - ext);
            const int16_t fj = a > b ? a : b;
            nh[0] = nf[0] = fj;
            e = (int16_t)(fj - go < NEG16 ? NEG16 : fj - go);
            start = 1;
        }
        if (start <= hi)
            row16_avx512(&h[start - 1], &f[start], &txt[start - 1], &nh[start], &nf[start], pat[i - 1] & 3u, mat, mis,
                         go, ext, e, hi + 1 - start);
        else
            nh[hi + 1] = nf[hi + 1] = NEG16;
        int16_t *t;
        t = h, h = nh, nh = t;
        t = f, f = nf, nf = t;
        plo = lo;
        phi = hi;
    }
    if (mode == SFO_GLOBAL) {
        *score_out = h[n];
        *col_out = n;
    } else {
        int32_t best = NEG;
        uint64_t col = 0;
        int first = 1;
        for (uint64_t j = plo; j <= phi; j++)
            if (first || h[j] > best) best = h[j], col = j, first = 0;
        *score_out = best;
        *col_out = col;
    }
done:
    free(h);
    free(f);
    free(nh);
    free(nf);
}
#endif

/* Optional per-row capture of the pass for the traceback. */
typedef struct {
    uint64_t *lo, *hi; /* per row 0..m */
    uint64_t *off;     /* start of row i's cells in h/f/e */
    int32_t *h, *f, *e;
} rows_t;

/* align.rs:2446-2666 `run_core::<ScalarRows>` (no resume, no snapshots). */
static void run_core(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode, const sfo_scoring *sc,
                     uint64_t band, uint64_t geom, int has_clip, int32_t xdrop, int clip_cell, uint64_t cut,
                     sfo_dp *out, rows_t *keep, row_fn row) {
    int32_t clip_best = 0;
    uint64_t ci = 0, cj = 0;
    int32_t cut_best = 0;
    int32_t mat = sc->mat, mis = sc->mis, ext = sc->ext;
    int32_t go = sc->open + sc->ext;
    band_t bd = band_of(m, n, mode, band, geom);
    int32_t *h = (int32_t *)malloc((n + 2) * sizeof(int32_t));
    int32_t *f = (int32_t *)malloc((n + 2) * sizeof(int32_t));
    int32_t *nh = (int32_t *)malloc((n + 2) * sizeof(int32_t));
    int32_t *nf = (int32_t *)malloc((n + 2) * sizeof(int32_t));
    int32_t *hp = (int32_t *)malloc((n + 2) * sizeof(int32_t));
    int32_t *gs = (int32_t *)malloc((n + 2) * sizeof(int32_t));
    for (uint64_t j = 0; j < n + 2; j++) h[j] = f[j] = nh[j] = nf[j] = NEG;
    uint64_t lo0, hi0;
    range_of(&bd, 0, &lo0, &hi0);
    for (uint64_t j = lo0; j <= hi0; j++)
        h[j] = (j == 0 || mode == SFO_INFIX) ? 0 : -(sc->open + sc->ext * (int32_t)j);
    for (uint64_t j = lo0; j <= hi0; j++) f[j] = NEG;
    if (lo0 > 0) h[lo0 - 1] = f[lo0 - 1] = NEG;
    h[hi0 + 1] = f[hi0 + 1] = NEG;
    if (keep) {
        keep->lo[0] = lo0;
        keep->hi[0] = hi0;
        keep->off[0] = 0;
        for (uint64_t j = lo0; j <= hi0; j++) {
            keep->h[j - lo0] = h[j];
            keep->f[j - lo0] = f[j];
            keep->e[j - lo0] = NEG;
        }
    }
    uint64_t plo = lo0, phi = hi0;
    for (uint64_t i = 1; i <= m; i++) {
        uint64_t lo, hi;
        range_of(&bd, i, &lo, &hi);
        if (lo > hi) {
            sfo_dp r = {NEG, 0, clip_best, ci, cj, n, i, cut_best, NEG};
            *out = r;
            if (keep) keep->lo[i] = 1, keep->hi[i] = 0;
            goto done;
        }
        unsigned c = pat[i - 1] & 3;
        uint64_t start = lo;
        int32_t e = NEG;
        int32_t row_max = NEG;
        int32_t *erow = NULL;
        if (keep) {
            keep->lo[i] = lo;
            keep->hi[i] = hi;
            keep->off[i] = keep->off[i - 1] + (keep->hi[i - 1] + 1 - keep->lo[i - 1]);
            erow = keep->e + keep->off[i];
            erow[0] = NEG;
        }
        if (lo == 0) {
            /* Column 0: only a vertical gap reaches it (align.rs:2564-2572). */
            int32_t fj = max32(h[0] - go, f[0] - ext);
            nh[0] = fj;
            nf[0] = fj;
            e = fj - go;
            start = 1;
            row_max = fj;
        }
        if (start <= hi) {
            int32_t mx = row(&h[start - 1], &f[start], &txt[start - 1], &nh[start], &nf[start], c, mat, mis, go, ext, e,
                             NEG, hi + 1 - start, hp, gs, erow ? erow + (start - lo) : NULL);
            row_max = max32(row_max, mx);
        } else {
            nh[hi + 1] = NEG;
            nf[hi + 1] = NEG;
        }
        if (cut != SFO_UMAX && lo <= cut) {
            int32_t row;
            if (hi <= cut) {
                row = row_max;
            } else {
                row = NEG;
                for (uint64_t j = lo; j <= cut; j++) row = max32(row, nh[j]);
            }
            cut_best = max32(cut_best, row);
        }
        if (has_clip) {
            int32_t row = row_max;
            if (row > clip_best) {
                clip_best = row;
                if (clip_cell) {
                    uint64_t j = lo;
                    while (nh[j] != row) j++;
                    ci = i;
                    cj = j;
                }
            }
            if (row < sat_sub_i(clip_best, xdrop)) {
                sfo_dp r = {NEG, 0, clip_best, ci, cj, hi, i, cut_best, NEG};
                *out = r;
                goto done;
            }
        }
        if (keep) {
            for (uint64_t j = lo; j <= hi; j++) {
                keep->h[keep->off[i] + j - lo] = nh[j];
                keep->f[keep->off[i] + j - lo] = nf[j];
            }
        }
        {
            int32_t *t;
            t = h, h = nh, nh = t;
            t = f, f = nf, nf = t;
        }
        plo = lo;
        phi = hi;
    }
    {
        int32_t score;
        uint64_t col;
        if (mode == SFO_GLOBAL) {
            score = h[n];
            col = n;
        } else {
            score = NEG;
            col = 0;
            int first = 1;
            for (uint64_t j = plo; j <= phi; j++)
                if (first || h[j] > score) score = h[j], col = j, first = 0;
        }
        int32_t cut_score = NEG;
        if (cut != SFO_UMAX && plo <= cut) {
            uint64_t top = phi < cut ? phi : cut;
            for (uint64_t j = plo; j <= top; j++) cut_score = max32(cut_score, h[j]);
        }
        sfo_dp r = {score, col, clip_best, ci, cj, phi, m, cut_best, cut_score};
        *out = r;
    }
done:
    free(h);
    free(f);
    free(nh);
    free(nf);
    free(hp);
    free(gs);
}

void sfo_affine_dp(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode, const sfo_scoring *sc,
                   uint64_t band, uint64_t geom, int has_clip, int32_t xdrop, int clip_cell, uint64_t cut,
                   sfo_dp *out) {
    run_core(pat, m, txt, n, mode, sc, band, geom, has_clip, xdrop, clip_cell, cut, out, NULL, row_scalar);
}

int32_t sfo_affine(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode, const sfo_scoring *sc,
                   uint64_t band) {
    sfo_dp d;
    run_core(pat, m, txt, n, mode, sc, band, n, 0, 0, 0, SFO_UMAX, &d, NULL, row_scalar);
    return d.score;
}

/* align.rs:710-729. */
int32_t sfo_affine_extend(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, const sfo_scoring *sc,
                          uint64_t band, int32_t end_bonus, int32_t xdrop) {
    sfo_dp d;
    run_core(pat, m, txt, n, SFO_PREFIX, sc, band, n, 1, xdrop, 0, SFO_UMAX, &d, NULL, row_scalar);
    return max32(d.clip_best, sat_add_i(d.score, end_bonus));
}

/* align.rs:957-982. */
void sfo_affine_extend_window(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, uint64_t window,
                              const sfo_scoring *sc, uint64_t band, int32_t end_bonus, int32_t xdrop, int64_t out4[4]) {
    sfo_dp d;
    run_core(pat, m, txt, n, SFO_PREFIX, sc, band, window > n ? window : n, 1, xdrop, 1, SFO_UMAX, &d, NULL,
             row_scalar);
    int32_t whole = sat_add_i(d.score, end_bonus);
    if (d.clip_best >= whole) {
        out4[0] = d.clip_best;
        out4[1] = (int64_t)d.clip_i;
        out4[2] = (int64_t)d.clip_j;
    } else {
        out4[0] = whole;
        out4[1] = (int64_t)m;
        out4[2] = (int64_t)d.score_col;
    }
    out4[3] = (int64_t)d.text_read;
}

/* align.rs:903-949, one text. */
static void extend_cut_bonus(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, const sfo_scoring *sc,
                             uint64_t band, int32_t end_bonus, int32_t xdrop, int diagonal, uint64_t window,
                             uint64_t cut, int track_cut, int bonus_within_cut, int64_t out3[3], row_fn row) {
    uint64_t geom = diagonal ? (m < n ? m : n) : (n > window ? n : window);
    sfo_dp d;
    run_core(pat, m, txt, n, SFO_PREFIX, sc, band, geom, 1, xdrop, track_cut && cut != SFO_UMAX, cut, &d, NULL, row);
    int32_t end;
    if (bonus_within_cut && cut != SFO_UMAX) end = max32(d.score, sat_add_i(d.cut_score, end_bonus));
    else end = sat_add_i(d.score, end_bonus);
    int32_t whole = max32(d.clip_best, end);
    uint64_t cols = d.clip_best >= end ? d.clip_j : d.score_col;
    int32_t cutv = cut == SFO_UMAX ? whole : max32(d.cut_best, sat_add_i(d.cut_score, end_bonus));
    out3[0] = whole;
    out3[1] = cutv;
    out3[2] = (int64_t)cols;
}

void sfo_affine_extend_cut_bonus(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, const sfo_scoring *sc,
                                 uint64_t band, int32_t end_bonus, int32_t xdrop, int diagonal, uint64_t window,
                                 uint64_t cut, int track_cut, int bonus_within_cut, int64_t out3[3]) {
    extend_cut_bonus(pat, m, txt, n, sc, band, end_bonus, xdrop, diagonal, window, cut, track_cut, bonus_within_cut,
                     out3, row_scalar);
}

/* align.rs:783-852 `local_band_core::<ScalarRows>`. */
int32_t sfo_local_band_ends(const uint8_t *read, uint64_t m, const uint8_t *txt, uint64_t n, const sfo_scoring *sc,
                            int64_t dlo, int64_t dhi, int32_t bonus0, int32_t bonus1) {
    if (m == 0 || dlo > dhi) return 0;
    int32_t go = sc->open + sc->ext, ext = sc->ext;
    uint64_t w = n + 2;
    int32_t *hp = (int32_t *)malloc(w * sizeof(int32_t)), *fp = (int32_t *)malloc(w * sizeof(int32_t));
    int32_t *h = (int32_t *)malloc(w * sizeof(int32_t)), *f = (int32_t *)malloc(w * sizeof(int32_t));
    int32_t *tmp1 = (int32_t *)malloc(w * sizeof(int32_t)), *tmp2 = (int32_t *)malloc(w * sizeof(int32_t));
    for (uint64_t j = 0; j < w; j++) hp[j] = fp[j] = h[j] = f[j] = NEG;
#define CLAMPI(x, a, b) ((x) < (a) ? (a) : (x) > (b) ? (b) : (x))
    int64_t np1 = (int64_t)n + 1;
    {
        int64_t lo = CLAMPI(dlo, 0, np1), hi = CLAMPI(dhi + 1, 0, np1);
        if (hi < lo) hi = lo;
        for (int64_t j = lo; j < hi; j++) hp[j] = bonus0;
    }
    int32_t best = 0;
    for (uint64_t i = 1; i <= m; i++) {
        int64_t lo = CLAMPI((int64_t)i + dlo, 0, np1), hi = CLAMPI((int64_t)i + dhi + 1, 0, np1);
        if (hi < lo) hi = lo;
        int32_t floor = i < m ? 0 : NEG;
        int64_t start = lo;
        int32_t e = NEG, rowmax = NEG;
        if (lo == 0 && hi > 0) {
            int32_t fj = max32(max32(hp[0] - go, fp[0] - ext), NEG);
            int32_t v = max32(fj, floor);
            h[0] = v;
            f[0] = fj;
            start = 1;
            e = v - go;
            rowmax = v;
        }
        if (start < hi) {
            int32_t r = row_scalar(&hp[start - 1], &fp[start], &txt[start - 1], &h[start], &f[start], read[i - 1],
                                   sc->mat, sc->mis, go, ext, e, floor, (uint64_t)(hi - start), tmp1, tmp2, NULL);
            rowmax = max32(rowmax, r);
        } else {
            h[hi] = NEG;
            f[hi] = NEG;
        }
        best = max32(best, rowmax);
        if (i == m) best = max32(best, sat_add_i(rowmax, bonus1));
        int32_t *t;
        t = hp, hp = h, h = t;
        t = fp, fp = f, f = t;
    }
#undef CLAMPI
    free(hp);
    free(fp);
    free(h);
    free(f);
    free(tmp1);
    free(tmp2);
    return best;
}

uint64_t sfo_affine_cells(uint64_t m, uint64_t n, int mode, uint64_t band) {
    band_t bd = band_of(m, n, mode, band, n);
    uint64_t cells = 0;
    for (uint64_t i = 1; i <= m; i++) {
        uint64_t lo, hi;
        range_of(&bd, i, &lo, &hi);
        if (lo > hi) break;
        cells += hi + 1 - lo;
    }
    return cells;
}

/* ------------------------------------------------------------------------------------------ */
/* Traceback (own definition; the reference has none)                                          */
/* ------------------------------------------------------------------------------------------ */
/*
 * The pass is run_core's (so H and F are the reference's values cell for cell); E is the
 * horizontal-gap value row_scalar's pass 2 gives each cell.  Cells outside a row's band are
 * NEG.  The walk starts at (m, score_col) in state H and emits ops last to first:
 *
 *   state H at (i, j):
 *     i == 0: Infix stops (text start j); otherwise j 'D's (row 0 is one gap from the origin).
 *     j == 0: i 'I's (column 0 is one vertical gap from the origin).
 *     H == H[i-1][j-1] + s(i, j)  -> '=' or 'X', to (i-1, j-1)       (diagonal first)
 *     else H == F[i][j]           -> state F                          (then vertical)
 *     else                        -> state E                          (H == E[i][j])
 *   state F at (i, j): 'I'; extend (stay F) iff F[i][j] > H[i-1][j] - go, else open (state H);
 *                      to (i-1, j).
 *   state E at (i, j): 'D'; extend (stay E) iff E[i][j] > H[i][j-1] - go, else open (state H);
 *                      to (i, j-1).
 *
 * Ties thus prefer the diagonal over a vertical gap over a horizontal one for H, and opening
 * over extending inside a gap.  Runs of one op are merged.
 */
typedef struct {
    uint32_t *buf;
    uint64_t n, cap;
    int overflow;
    unsigned op;
    uint32_t len;
} cig_t;

static void cig_push(cig_t *c, unsigned op, uint32_t len) {
    if (len == 0) return;
    if (c->len && c->op == op) {
        c->len += len;
        return;
    }
    if (c->len) {
        if (c->n < c->cap) c->buf[c->n] = (c->len << 4) | c->op;
        else c->overflow = 1;
        c->n++;
    }
    c->op = op;
    c->len = len;
}

static void cig_flush(cig_t *c) {
    if (c->len) {
        if (c->n < c->cap) c->buf[c->n] = (c->len << 4) | c->op;
        else c->overflow = 1;
        c->n++;
        c->len = 0;
    }
}

#define OP_I 1u
#define OP_D 2u
#define OP_EQ 7u
#define OP_X 8u

int64_t sfo_align(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode, const sfo_scoring *sc,
                  uint64_t band, uint32_t *cigar, uint64_t cap, int32_t *score, uint64_t *score_col,
                  uint64_t *tstart) {
    band_t bd = band_of(m, n, mode, band, n);
    uint64_t total = 0;
    for (uint64_t i = 0; i <= m; i++) {
        uint64_t lo, hi;
        range_of(&bd, i, &lo, &hi);
        if (lo <= hi) total += hi + 1 - lo;
    }
    rows_t keep;
    keep.lo = (uint64_t *)calloc(m + 1, sizeof(uint64_t));
    keep.hi = (uint64_t *)calloc(m + 1, sizeof(uint64_t));
    keep.off = (uint64_t *)calloc(m + 1, sizeof(uint64_t));
    keep.h = (int32_t *)malloc((total + 1) * sizeof(int32_t));
    keep.f = (int32_t *)malloc((total + 1) * sizeof(int32_t));
    keep.e = (int32_t *)malloc((total + 1) * sizeof(int32_t));
    sfo_dp d;
    run_core(pat, m, txt, n, mode, sc, band, n, 0, 0, 0, SFO_UMAX, &d, &keep, row_scalar);
    *score = d.score;
    *score_col = d.score_col;
    *tstart = 0;
    int64_t ret;
    if (d.last_row != m || d.score == NEG) {
        ret = -2;
        goto out;
    }
    {
        int32_t go = sc->open + sc->ext, ext = sc->ext;
#define INB(i, j) ((j) >= keep.lo[i] && (j) <= keep.hi[i])
#define HV(i, j) (INB(i, j) ? keep.h[keep.off[i] + (j) - keep.lo[i]] : NEG)
#define FV(i, j) (INB(i, j) ? keep.f[keep.off[i] + (j) - keep.lo[i]] : NEG)
#define EV(i, j) (INB(i, j) ? keep.e[keep.off[i] + (j) - keep.lo[i]] : NEG)
        /* Ops are collected last to first, then reversed. */
        uint32_t *rev = (uint32_t *)malloc((m + n + 2) * sizeof(uint32_t));
        uint64_t nr = 0;
        uint64_t i = m, j = d.score_col;
        int state = 0; /* 0 H, 1 F, 2 E */
        for (;;) {
            if (state == 0) {
                if (i == 0) {
                    if (mode == SFO_INFIX) {
                        *tstart = j;
                    } else {
                        for (uint64_t q = 0; q < j; q++) rev[nr++] = OP_D;
                    }
                    break;
                }
                if (j == 0) {
                    for (uint64_t q = 0; q < i; q++) rev[nr++] = OP_I;
                    break;
                }
                int32_t hv = HV(i, j);
                int same = (unsigned)(pat[i - 1] & 3) == (unsigned)(txt[j - 1] & 3);
                int32_t dg = HV(i - 1, j - 1) + (same ? sc->mat : -sc->mis);
                if (hv == dg) {
                    rev[nr++] = same ? OP_EQ : OP_X;
                    i--;
                    j--;
                } else if (hv == FV(i, j)) {
                    state = 1;
                } else {
                    state = 2;
                }
            } else if (state == 1) {
                rev[nr++] = OP_I;
                int ext_ = FV(i, j) > HV(i - 1, j) - go;
                i--;
                state = ext_ ? 1 : 0;
            } else {
                rev[nr++] = OP_D;
                int ext_ = EV(i, j) > HV(i, j - 1) - go;
                j--;
                state = ext_ ? 2 : 0;
            }
        }
        (void)ext;
        cig_t c = {cigar, 0, cap, 0, 0, 0};
        for (uint64_t q = nr; q-- > 0;) cig_push(&c, rev[q], 1);
        cig_flush(&c);
        free(rev);
        ret = c.overflow ? -1 : (int64_t)c.n;
#undef INB
#undef HV
#undef FV
#undef EV
    }
out:
    free(keep.lo);
    free(keep.hi);
    free(keep.off);
    free(keep.h);
    free(keep.f);
    free(keep.e);
    return ret;
}

/* ------------------------------------------------------------------------------------------ */
/* Naive definitions (the reference's test oracles, align.rs:2801-3014)                        */
/* ------------------------------------------------------------------------------------------ */

/* align.rs:2801-2827 `naive`. */
uint32_t sfo_naive_edit(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, int mode) {
    uint32_t *prev = (uint32_t *)malloc((m + 1) * sizeof(uint32_t)), *cur = (uint32_t *)malloc((m + 1) * sizeof(uint32_t));
    for (uint64_t i = 0; i <= m; i++) prev[i] = (uint32_t)i;
    uint32_t best = (mode == SFO_GLOBAL && n > 0) ? UINT32_MAX : prev[m];
    for (uint64_t j = 1; j <= n; j++) {
        cur[0] = mode == SFO_INFIX ? 0 : (uint32_t)j;
        for (uint64_t i = 1; i <= m; i++) {
            uint32_t sub = prev[i - 1] + (p[i - 1] != t[j - 1]);
            uint32_t v = sub < prev[i] + 1 ? sub : prev[i] + 1;
            cur[i] = v < cur[i - 1] + 1 ? v : cur[i - 1] + 1;
        }
        if (mode == SFO_GLOBAL) {
            if (j == n) best = cur[m];
        } else if (cur[m] < best) {
            best = cur[m];
        }
        uint32_t *x = prev;
        prev = cur;
        cur = x;
    }
    free(prev);
    free(cur);
    return best;
}

void sfo_naive_edit_lastrow(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, int mode, uint32_t *row) {
    uint32_t *prev = (uint32_t *)malloc((m + 1) * sizeof(uint32_t)), *cur = (uint32_t *)malloc((m + 1) * sizeof(uint32_t));
    for (uint64_t i = 0; i <= m; i++) prev[i] = (uint32_t)i;
    row[0] = prev[m];
    for (uint64_t j = 1; j <= n; j++) {
        cur[0] = mode == SFO_INFIX ? 0 : (uint32_t)j;
        for (uint64_t i = 1; i <= m; i++) {
            uint32_t sub = prev[i - 1] + (p[i - 1] != t[j - 1]);
            uint32_t v = sub < prev[i] + 1 ? sub : prev[i] + 1;
            cur[i] = v < cur[i - 1] + 1 ? v : cur[i - 1] + 1;
        }
        row[j] = cur[m];
        uint32_t *x = prev;
        prev = cur;
        cur = x;
    }
    free(prev);
    free(cur);
}

/* align.rs:2928-2957 `naive_affine`. */
int32_t sfo_naive_affine(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, int mode, const sfo_scoring *sc) {
    int32_t go = sc->open + sc->ext;
    uint64_t N = n + 1;
    int32_t *h = (int32_t *)malloc((m + 1) * N * sizeof(int32_t));
    int32_t *e = (int32_t *)malloc((m + 1) * N * sizeof(int32_t));
    int32_t *f = (int32_t *)malloc((m + 1) * N * sizeof(int32_t));
    for (uint64_t q = 0; q < (m + 1) * N; q++) h[q] = e[q] = f[q] = NEG;
    h[0] = 0;
    for (uint64_t j = 1; j <= n; j++) {
        h[j] = mode == SFO_INFIX ? 0 : -(sc->open + sc->ext * (int32_t)j);
        e[j] = h[j];
    }
    for (uint64_t i = 1; i <= m; i++) {
        h[i * N] = -(sc->open + sc->ext * (int32_t)i);
        f[i * N] = h[i * N];
        for (uint64_t j = 1; j <= n; j++) {
            e[i * N + j] = max32(h[i * N + j - 1] - go, e[i * N + j - 1] - sc->ext);
            f[i * N + j] = max32(h[(i - 1) * N + j] - go, f[(i - 1) * N + j] - sc->ext);
            int32_t s = p[i - 1] == t[j - 1] ? sc->mat : -sc->mis;
            h[i * N + j] = max32(max32(h[(i - 1) * N + j - 1] + s, e[i * N + j]), f[i * N + j]);
        }
    }
    int32_t r;
    if (mode == SFO_GLOBAL) {
        r = h[m * N + n];
    } else {
        r = h[m * N];
        for (uint64_t j = 1; j <= n; j++) r = max32(r, h[m * N + j]);
    }
    free(h);
    free(e);
    free(f);
    return r;
}

/* align.rs:2959-2965 `naive_extend`. */
int32_t sfo_naive_extend(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, const sfo_scoring *sc) {
    int32_t best = INT32_MIN;
    for (uint64_t i = 0; i <= m; i++)
        for (uint64_t j = 0; j <= n; j++) best = max32(best, sfo_naive_affine(p, i, t, j, SFO_GLOBAL, sc));
    return best;
}

/* align.rs:2969-2983 `naive_local`. */
int32_t sfo_naive_local(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, const sfo_scoring *sc, int32_t b0,
                        int32_t b1) {
    int32_t best = 0;
    for (uint64_t i0 = 0; i0 < m; i0++)
        for (uint64_t i1 = i0 + 1; i1 <= m; i1++) {
            int32_t ends = b0 * (i0 == 0) + b1 * (i1 == m);
            for (uint64_t j0 = 0; j0 <= n; j0++)
                for (uint64_t j1 = j0; j1 <= n; j1++) {
                    int32_t x = sfo_naive_affine(p + i0, i1 - i0, t + j0, j1 - j0, SFO_GLOBAL, sc);
                    best = max32(best, x + ends);
                }
        }
    return best;
}

/* align.rs:2987-3014 `naive_local_band`. */
int32_t sfo_naive_local_band(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, const sfo_scoring *sc,
                             int64_t dlo, int64_t dhi, int32_t b0, int32_t b1) {
    int32_t go = sc->open + sc->ext;
    uint64_t N = n + 1;
    int32_t *h = (int32_t *)malloc((m + 1) * N * sizeof(int32_t));
    int32_t *e = (int32_t *)malloc((m + 1) * N * sizeof(int32_t));
    int32_t *f = (int32_t *)malloc((m + 1) * N * sizeof(int32_t));
    for (uint64_t q = 0; q < (m + 1) * N; q++) h[q] = e[q] = f[q] = NEG;
#define INSIDE(i, j) ((int64_t)(j) - (int64_t)(i) >= dlo && (int64_t)(j) - (int64_t)(i) <= dhi)
    for (uint64_t j = 0; j <= n; j++)
        if (INSIDE(0, j)) h[j] = b0;
    int32_t best = 0;
    for (uint64_t i = 1; i <= m; i++)
        for (uint64_t j = 0; j <= n; j++) {
            if (!INSIDE(i, j)) continue;
            f[i * N + j] = max32(h[(i - 1) * N + j] - go, f[(i - 1) * N + j] - sc->ext);
            int32_t x = f[i * N + j];
            if (j > 0) {
                e[i * N + j] = max32(h[i * N + j - 1] - go, e[i * N + j - 1] - sc->ext);
                int32_t s = p[i - 1] == t[j - 1] ? sc->mat : -sc->mis;
                x = max32(max32(x, e[i * N + j]), h[(i - 1) * N + j - 1] + s);
            }
            h[i * N + j] = i < m ? max32(x, 0) : x;
            best = max32(best, h[i * N + j] + (i == m ? b1 : 0));
        }
#undef INSIDE
    free(h);
    free(e);
    free(f);
    return best;
}

/* ------------------------------------------------------------------------------------------ */
/* Threaded batch drivers (CPU baseline)                                                       */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    const uint8_t *pat, *txt;
    const uint64_t *poff, *toff;
    const uint32_t *plen, *tlen;
    uint64_t npairs;
    int mode;
    uint32_t k;
    const sfo_scoring *sc;
    const uint64_t *band;
    uint32_t *dist;
    int32_t *score;
    uint64_t *score_col;
    row_fn row;
    int narrow; /* take the 16-bit rows where narrow16 holds (AVX-512 baseline) */
    atomic_uint_fast64_t next;
} job_t;

#define CHUNK 16

static void *myers_worker(void *arg) {
    job_t *jb = (job_t *)arg;
    for (;;) {
        uint64_t i0 = atomic_fetch_add(&jb->next, CHUNK);
        if (i0 >= jb->npairs) break;
        uint64_t i1 = i0 + CHUNK < jb->npairs ? i0 + CHUNK : jb->npairs;
        for (uint64_t i = i0; i < i1; i++)
            jb->dist[i] = sfo_distance(jb->pat + jb->poff[i], jb->plen[i], jb->txt + jb->toff[i], jb->tlen[i], jb->mode,
                                       jb->k);
    }
    return NULL;
}

static void *affine_worker(void *arg) {
    job_t *jb = (job_t *)arg;
    for (;;) {
        uint64_t i0 = atomic_fetch_add(&jb->next, CHUNK);
        if (i0 >= jb->npairs) break;
        uint64_t i1 = i0 + CHUNK < jb->npairs ? i0 + CHUNK : jb->npairs;
        for (uint64_t i = i0; i < i1; i++) {
#if defined(__x86_64__)
            if (jb->narrow && narrow16(jb->plen[i], jb->mode, jb->sc, jb->band[i], jb->tlen[i])) {
                uint64_t col;
                run16_core(jb->pat + jb->poff[i], jb->plen[i], jb->txt + jb->toff[i], jb->tlen[i], jb->mode, jb->sc,
                           jb->band[i], &jb->score[i], &col);
                if (jb->score_col) jb->score_col[i] = col;
                continue;
            }
#endif
            sfo_dp d;
            run_core(jb->pat + jb->poff[i], jb->plen[i], jb->txt + jb->toff[i], jb->tlen[i], jb->mode, jb->sc,
                     jb->band[i], jb->tlen[i], 0, 0, 0, SFO_UMAX, &d, NULL, jb->row);
            jb->score[i] = d.score;
            if (jb->score_col) jb->score_col[i] = d.score_col;
        }
    }
    return NULL;
}

static void run_threads(job_t *jb, void *(*fn)(void *), int threads) {
    if (threads < 1) threads = 1;
    atomic_init(&jb->next, 0);
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int q = 0; q < threads; q++) pthread_create(&th[q], NULL, fn, jb);
    for (int q = 0; q < threads; q++) pthread_join(th[q], NULL);
    free(th);
}

void sfo_batch_myers(const uint8_t *pat, const uint64_t *poff, const uint32_t *plen, const uint8_t *txt,
                     const uint64_t *toff, const uint32_t *tlen, uint64_t npairs, int mode, uint32_t k, uint32_t *dist,
                     int threads) {
    job_t jb;
    memset(&jb, 0, sizeof jb);
    jb.pat = pat, jb.poff = poff, jb.plen = plen, jb.txt = txt, jb.toff = toff, jb.tlen = tlen;
    jb.npairs = npairs, jb.mode = mode, jb.k = k, jb.dist = dist;
    run_threads(&jb, myers_worker, threads);
}

void sfo_batch_affine(const uint8_t *pat, const uint64_t *poff, const uint32_t *plen, const uint8_t *txt,
                      const uint64_t *toff, const uint32_t *tlen, uint64_t npairs, int mode, const sfo_scoring *sc,
                      const uint64_t *band, int32_t *score, uint64_t *score_col, int threads) {
    job_t jb;
    memset(&jb, 0, sizeof jb);
    jb.pat = pat, jb.poff = poff, jb.plen = plen, jb.txt = txt, jb.toff = toff, jb.tlen = tlen;
    jb.npairs = npairs, jb.mode = mode, jb.sc = sc, jb.band = band, jb.score = score, jb.score_col = score_col;
    jb.row = row_scalar;
    run_threads(&jb, affine_worker, threads);
}

/* The same as the reference's x86 dispatch runs it (align.rs:2320-2355): 16-bit AVX-512 rows where
 * `narrow` holds, 32-bit AVX-512 rows (run_core::<Avx512Rows>) otherwise; -1 where the host has no
 * AVX-512. */
int sfo_batch_affine_avx512(const uint8_t *pat, const uint64_t *poff, const uint32_t *plen, const uint8_t *txt,
                            const uint64_t *toff, const uint32_t *tlen, uint64_t npairs, int mode,
                            const sfo_scoring *sc, const uint64_t *band, int32_t *score, uint64_t *score_col,
                            int threads) {
#if defined(__x86_64__)
    if (!sfo_avx512_available()) return -1;
    job_t jb;
    memset(&jb, 0, sizeof jb);
    jb.pat = pat, jb.poff = poff, jb.plen = plen, jb.txt = txt, jb.toff = toff, jb.tlen = tlen;
    jb.npairs = npairs, jb.mode = mode, jb.sc = sc, jb.band = band, jb.score = score, jb.score_col = score_col;
    jb.row = row_avx512;
    jb.narrow = 1;
    run_threads(&jb, affine_worker, threads);
    return 0;
#else
    return -1;
#endif
}

/* ---- the tie-break's segment calls (crates/map/src/lib.rs:1307-1492 `seg`), many at once ---- */

/* crates/index/src/lib.rs:92-100 `word_at` on u64 words (base i at bits 2 (i % 32) of word i / 32). */
static uint64_t word_at64(const uint64_t *w, uint64_t nwords, uint64_t i) {
    uint64_t q = i >> 5;
    unsigned s = 2 * (unsigned)(i & 31);
    uint64_t lo = q < nwords ? w[q] >> s : 0;
    if (s == 0) return lo;
    return lo | (q + 1 < nwords ? w[q + 1] << (64 - s) : 0);
}

/* crates/index/src/lib.rs:104-114 `packed_eq`. */
int sfo_packed_eq(const uint64_t *w, uint64_t nwords, uint64_t a, uint64_t b, uint32_t len) {
    uint32_t k = 0;
    for (; k + 32 <= len; k += 32)
        if (word_at64(w, nwords, a + k) != word_at64(w, nwords, b + k)) return 0;
    uint32_t r = len - k;
    return r == 0 || ((word_at64(w, nwords, a + k) ^ word_at64(w, nwords, b + k)) & ((1ull << (2 * r)) - 1)) == 0;
}

typedef struct {
    const uint64_t *words;
    uint64_t nwords;
    const uint8_t *reads;
    const sfo_tb_segment *segs;
    uint64_t nseg;
    const uint64_t *cand_off;
    const uint32_t *cand_len;
    const sfo_scoring *sc;
    uint64_t band;
    int32_t end_bonus, xdrop;
    int64_t *out;
    uint32_t *rep;
    row_fn row;
    int narrow;
    atomic_uint_fast64_t next;
} tb_job_t;

static void tb_segment(tb_job_t *jb, const sfo_tb_segment *g, uint8_t **buf, uint64_t *cap) {
    const int rev = (g->flags & SFO_SEG_REVERSE) != 0;
    /* lib.rs:1337-1345: each candidate's first equal candidate (packed comparison). */
    uint64_t longest = 0, need = g->read_len;
    for (uint32_t c = g->cand0; c < g->cand0 + g->ncand; c++) {
        uint32_t r = c;
        for (uint32_t d = g->cand0; d < c; d++)
            if (jb->rep[d] == d - g->cand0 && jb->cand_len[d] == jb->cand_len[c] &&
                sfo_packed_eq(jb->words, jb->nwords, jb->cand_off[d], jb->cand_off[c], jb->cand_len[c])) {
                r = d;
                break;
            }
        jb->rep[c] = r - g->cand0;
        if (jb->cand_len[c] > longest) longest = jb->cand_len[c];
        if (r == c && jb->cand_len[c] > need) need = jb->cand_len[c];
    }
    if (2 * need > *cap) {
        free(*buf);
        *cap = 2 * need + 64;
        *buf = (uint8_t *)malloc(*cap);
    }
    /* lib.rs:1347-1365: the read slice and each distinct text decoded, reversed for a left end. */
    uint8_t *rd = *buf, *tx = *buf + need;
    for (uint32_t k = 0; k < g->read_len; k++)
        rd[k] = jb->reads[g->read_off + (rev ? g->read_len - 1 - k : k)];
    const uint64_t window = g->window > longest ? g->window : longest; /* lib.rs:1446-1451 */
    for (uint32_t c = g->cand0; c < g->cand0 + g->ncand; c++) {
        int64_t *o = jb->out + 3 * (uint64_t)c;
        if (jb->rep[c] != c - g->cand0) {
            const int64_t *q = jb->out + 3 * (uint64_t)(g->cand0 + jb->rep[c]);
            o[0] = q[0], o[1] = q[1], o[2] = q[2];
            continue;
        }
        const uint64_t n = jb->cand_len[c], off = jb->cand_off[c];
        for (uint64_t k = 0; k < n; k++) {
            uint64_t b = off + (rev ? n - 1 - k : k);
            tx[k] = (uint8_t)((jb->words[b >> 5] >> (2 * (b & 31))) & 3);
        }
        const uint64_t m = g->read_len;
        if (g->flags & SFO_SEG_EXTEND) {
            /* align.rs:903-949 for one text. */
            extend_cut_bonus(rd, m, tx, n, jb->sc, jb->band, jb->end_bonus, jb->xdrop,
                             (g->flags & SFO_SEG_DIAGONAL) != 0, window, g->cut, (g->flags & SFO_SEG_TRACK_CUT) != 0,
                             (g->flags & SFO_SEG_BONUS_WITHIN_CUT) != 0, o, jb->row);
        } else {
            /* align.rs:854-861, mapped to (x, x, 0) (lib.rs:1464-1469). */
            int32_t x;
#if defined(__x86_64__)
            if (jb->narrow && narrow16(m, g->mode, jb->sc, jb->band, n)) {
                uint64_t col;
                run16_core(rd, m, tx, n, g->mode, jb->sc, jb->band, &x, &col);
            } else
#endif
            {
                sfo_dp d;
                run_core(rd, m, tx, n, g->mode, jb->sc, jb->band, n, 0, 0, 0, SFO_UMAX, &d, NULL, jb->row);
                x = d.score;
            }
            o[0] = o[1] = x;
            o[2] = 0;
        }
    }
}

static void *tb_worker(void *arg) {
    tb_job_t *jb = (tb_job_t *)arg;
    uint8_t *buf = NULL;
    uint64_t cap = 0;
    for (;;) {
        uint64_t s0 = atomic_fetch_add(&jb->next, 4);
        if (s0 >= jb->nseg) break;
        uint64_t s1 = s0 + 4 < jb->nseg ? s0 + 4 : jb->nseg;
        for (uint64_t s = s0; s < s1; s++) tb_segment(jb, &jb->segs[s], &buf, &cap);
    }
    free(buf);
    return NULL;
}

int sfo_tiebreak_segments(const uint64_t *words, uint64_t bases, const uint8_t *reads, const sfo_tb_segment *segs,
                          uint64_t nseg, const uint64_t *cand_off, const uint32_t *cand_len, const sfo_scoring *sc,
                          uint64_t band, int32_t end_bonus, int32_t xdrop, int64_t *out, uint32_t *rep, int threads,
                          int avx512) {
    tb_job_t jb;
    memset(&jb, 0, sizeof jb);
    jb.words = words, jb.nwords = (bases + 31) / 32, jb.reads = reads, jb.segs = segs, jb.nseg = nseg;
    jb.cand_off = cand_off, jb.cand_len = cand_len, jb.sc = sc, jb.band = band, jb.end_bonus = end_bonus;
    jb.xdrop = xdrop, jb.out = out, jb.rep = rep, jb.row = row_scalar;
#if defined(__x86_64__)
    if (avx512) {
        if (!sfo_avx512_available()) return -1;
        jb.row = row_avx512;
        jb.narrow = 1;
    }
#else
    if (avx512) return -1;
#endif
    atomic_init(&jb.next, 0);
    if (threads < 1) threads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
    for (int q = 0; q < threads; q++) pthread_create(&th[q], NULL, tb_worker, &jb);
    for (int q = 0; q < threads; q++) pthread_join(th[q], NULL);
    free(th);
    return 0;
}
