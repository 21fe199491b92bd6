// tiebreak.hip -- the tie-break's per-segment alignments in batches (sf_tiebreak_segments) over a
// device-resident transcript store in the index's own packing (sf_store_*).
//
// The reference's tie-break scores each read segment against its candidates' texts one call at a
// time (crates/map/src/lib.rs:1307-1492, the `seg` closure): candidates whose texts are equal
// (packed comparison, crates/index/src/lib.rs:135-137 `sequence_eq`) are aligned once; left read
// ends are aligned reversed; Prefix read ends with clipping take align.rs:903-949
// `affine_extend_sorted_cut_bonus`, other segments align.rs:854-861 `affine_sorted`.  Here many
// reads' segments go in one call, and no text leaves the device:
//   - the store keeps, in one device buffer of 2-bit words, the transcript bases, the same bases
//     reversed, and an area for the call's reads (forward and reversed), so every pattern and text
//     of the call -- reversed ones included -- is a range of that buffer;
//   - candidates are deduplicated per segment on the device (packed 64-bit equality);
//   - the distinct (segment, text) pairs run as batches over that buffer (sf_batch_create_resident):
//     the extensions in one pass with per-pair band layouts and cuts, the sorted scores one pass per
//     mode; every candidate takes its distinct text's result.
// The sorted calls' row sharing between texts with common prefixes (align.rs:2210-2315, `Snap`) is
// an optimization with results equal to separate calls (align.rs:2210-2213) and is not done: each
// distinct text is one pair.
#include <algorithm>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/sf_align.h"
#include "internal.hpp"

using namespace sf;

int sf_runtime_fail(int code, const char *msg); // runtime.hip
int sf_runtime_ensure();                        // runtime.hip

namespace {

// 32 bases of the packed store from base i (bases past the end read as 0): crates/index/src/lib.rs:92-100
// `word_at`, on u64 words with base i at bits 2 (i % 32) of word i / 32.
__device__ __forceinline__ uint64_t word_at(const uint64_t *w, uint64_t nwords, uint64_t i) {
    const uint64_t q = i >> 5;
    const uint32_t s = 2 * (uint32_t)(i & 31);
    const uint64_t lo = q < nwords ? w[q] >> s : 0;
    if (s == 0) return lo;
    return lo | (q + 1 < nwords ? w[q + 1] << (64 - s) : 0);
}

// crates/index/src/lib.rs:104-114 `packed_eq`: 32 bases at a time, the tail under a mask.
__device__ bool packed_eq(const uint64_t *w, uint64_t nwords, uint64_t a, uint64_t b, uint32_t len) {
    uint32_t k = 0;
    for (; k + 32 <= len; k += 32)
        if (word_at(w, nwords, a + k) != word_at(w, nwords, b + k)) return false;
    const uint32_t r = len - k;
    return r == 0 || ((word_at(w, nwords, a + k) ^ word_at(w, nwords, b + k)) & ((1ull << (2 * r)) - 1)) == 0;
}

__global__ __launch_bounds__(256) void codes_eq_kernel(const uint64_t *w, uint64_t nwords, const uint64_t *a,
                                                       const uint64_t *b, const uint32_t *len, uint64_t n, uint8_t *eq) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) eq[i] = packed_eq(w, nwords, a[i], b[i], len[i]) ? 1 : 0;
}

// Candidate c of a segment whose candidates start at first[c]: the first candidate of the segment
// with an equal text (itself if none), as the reference's `rep0` (lib.rs:1337-1345).
__global__ __launch_bounds__(256) void dedup_kernel(const uint64_t *w, uint64_t nwords, const uint64_t *off,
                                                    const uint32_t *len, const uint32_t *first, uint64_t n,
                                                    uint32_t *rep) {
    const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    uint32_t r = (uint32_t)c;
    for (uint32_t d = first[c]; d < c; ++d)
        if (len[d] == len[c] && packed_eq(w, nwords, off[d], off[c], len[c])) {
            r = d;
            break;
        }
    rep[c] = r;
}

// Code k of a sequence of 2-bit words.
__device__ __forceinline__ uint32_t code2(const uint32_t *w, uint64_t k) { return (w[k >> 4] >> (2 * (k & 15))) & 3u; }

// Word j of the reversed store: bases R[16j + k] = S[bases - 1 - 16j - k], zero past the end.
__global__ __launch_bounds__(256) void reverse_kernel(const uint32_t *fwd, uint64_t bases, uint32_t *rev,
                                                      uint64_t words) {
    const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= words) return;
    uint32_t x = 0;
    for (uint32_t k = 0; k < 16; ++k) {
        const uint64_t r = 16 * j + k;
        if (r < bases) x |= code2(fwd, bases - 1 - r) << (2 * k);
    }
    rev[j] = x;
}

// The call's reads (one code a byte, masked with 3) as 2-bit words, forward into fwd and reversed
// (base r of the reversed reads = base len - 1 - r) into rev; `words` words each.
__global__ __launch_bounds__(256) void pack_reads_kernel(const uint8_t *bytes, uint64_t len, uint32_t *fwd,
                                                         uint32_t *rev, uint64_t words) {
    const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= words) return;
    uint32_t f = 0, r = 0;
    for (uint32_t k = 0; k < 16; ++k) {
        const uint64_t i = 16 * j + k;
        if (i < len) {
            f |= (bytes[i] & 3u) << (2 * k);
            r |= (bytes[len - 1 - i] & 3u) << (2 * k);
        }
    }
    fwd[j] = f;
    rev[j] = r;
}

unsigned grid(uint64_t n) { return (unsigned)((n + 255) / 256); }

int dev(const char *what, hipError_t e) {
    char msg[200];
    std::snprintf(msg, sizeof msg, "%s: %s", what, hipGetErrorString(e));
    return sf_runtime_fail(SF_EDEVICE, msg);
}

int32_t sat_add(int32_t a, int32_t b) {
    const int64_t r = (int64_t)a + b;
    return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
}

// Grows a device array to at least n elements (contents not kept).
template <class T>
bool reserve(T *&p, uint64_t &cap, uint64_t n) {
    if (n <= cap && p) return true;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const uint64_t c = std::max<uint64_t>(n + n / 2, 256);
    if (hipMalloc((void **)&p, c * sizeof(T)) != hipSuccess) return false;
    cap = c;
    return true;
}

} // namespace

struct sf_store {
    uint64_t bases = 0;
    uint64_t nwords = 0; // u64 words of the store (2 nwords u32 words, the same bytes)
    // [store | reversed store | the call's reads, forward | reversed], u32 words; the reversed store
    // starts at base 32 nwords, the reads at base 64 nwords.
    uint32_t *d_seq = nullptr;
    uint64_t seq_words = 0; // allocated
    std::mutex mu;          // calls on one store run one at a time: they share its buffers
    hipStream_t stream = nullptr;
    uint8_t *d_bytes = nullptr; // the call's reads, one code a byte (or sf_store_codes_eq's results)
    uint64_t bytes_cap = 0;
    uint64_t *d_off = nullptr;
    uint32_t *d_len = nullptr, *d_first = nullptr, *d_rep = nullptr;
    uint64_t off_cap = 0, len_cap = 0, first_cap = 0, rep_cap = 0;

    uint64_t rev_base() const { return 32 * nwords; }
    uint64_t reads_base() const { return 64 * nwords; }
    const uint64_t *words64() const { return reinterpret_cast<const uint64_t *>(d_seq); }

    ~sf_store() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : {(void *)d_seq, (void *)d_bytes, (void *)d_off, (void *)d_len, (void *)d_first, (void *)d_rep})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Room for `read_words` u32 words of reads after the two store copies: a larger buffer takes
    // the store copies along (device to device).
    int reserve_reads(uint64_t read_words) {
        const uint64_t need = 4 * nwords + read_words;
        if (need <= seq_words) return SF_OK;
        const uint64_t cap = std::max<uint64_t>(need + read_words / 2, 4 * nwords + 1024);
        uint32_t *p = nullptr;
        if (hipMalloc((void **)&p, cap * 4) != hipSuccess) return sf_runtime_fail(SF_ENOMEM, "store reads area");
        hipError_t e = hipSuccess;
        if (d_seq && nwords) e = hipMemcpyAsync(p, d_seq, 4 * nwords * 4, hipMemcpyDeviceToDevice, stream);
        if (!e) e = hipStreamSynchronize(stream);
        if (e) {
            (void)hipFree(p);
            return dev("store reads area", e);
        }
        if (d_seq) (void)hipFree(d_seq);
        d_seq = p;
        seq_words = cap;
        return SF_OK;
    }
};

extern "C" {

int sf_store_create(const uint64_t *words, uint64_t bases, sf_store **out) {
    if (!out || (bases && !words)) return sf_runtime_fail(SF_EINVAL, "null argument");
    *out = nullptr;
    if (bases >= (1ull << 40)) return sf_runtime_fail(SF_ERANGE, "a store of 2^40 bases or more");
    if (const int st = sf_runtime_ensure()) return st;
    auto *s = new sf_store();
    s->bases = bases;
    s->nwords = (bases + 31) / 32;
    auto bail = [&](int st) {
        delete s;
        return st;
    };
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(sf_runtime_fail(SF_EDEVICE, "store stream"));
    if (const int st = s->reserve_reads(1024)) return bail(st);
    std::vector<uint64_t> w(words, words + s->nwords);
    if (s->nwords && bases % 32) w.back() &= (1ull << (2 * (bases % 32))) - 1; // bases past the end read as 0
    hipError_t e = hipSuccess;
    if (s->nwords) {
        e = hipMemcpyAsync(s->d_seq, w.data(), s->nwords * 8, hipMemcpyHostToDevice, s->stream);
        if (!e) {
            hipLaunchKernelGGL(reverse_kernel, dim3(grid(2 * s->nwords)), dim3(256), 0, s->stream, s->d_seq, bases,
                               s->d_seq + 2 * s->nwords, 2 * s->nwords);
            e = hipGetLastError();
        }
    }
    if (!e) e = hipStreamSynchronize(s->stream);
    if (e) return bail(dev("store upload", e));
    *out = s;
    return SF_OK;
}

void sf_store_destroy(sf_store *s) { delete s; }

int sf_store_codes_eq(sf_store *s, const uint64_t *a_off, const uint64_t *b_off, const uint32_t *len, uint64_t n,
                      uint8_t *eq) {
    if (!s || (n && (!a_off || !b_off || !len || !eq))) return sf_runtime_fail(SF_EINVAL, "null argument");
    for (uint64_t i = 0; i < n; ++i)
        if (a_off[i] > s->bases || len[i] > s->bases - a_off[i] || b_off[i] > s->bases || len[i] > s->bases - b_off[i])
            return sf_runtime_fail(SF_EINVAL, "a range lies outside the store");
    if (n == 0) return SF_OK;
    std::lock_guard<std::mutex> lk(s->mu);
    // Scratch: a and b offsets in d_off (2n), lengths in d_len, results in the bytes area.
    if (!reserve(s->d_off, s->off_cap, 2 * n) || !reserve(s->d_len, s->len_cap, n) ||
        !reserve(s->d_bytes, s->bytes_cap, n))
        return sf_runtime_fail(SF_ENOMEM, "sf_store_codes_eq");
    hipError_t e = hipMemcpyAsync(s->d_off, a_off, n * 8, hipMemcpyHostToDevice, s->stream);
    if (!e) e = hipMemcpyAsync(s->d_off + n, b_off, n * 8, hipMemcpyHostToDevice, s->stream);
    if (!e) e = hipMemcpyAsync(s->d_len, len, n * 4, hipMemcpyHostToDevice, s->stream);
    if (!e) {
        hipLaunchKernelGGL(codes_eq_kernel, dim3(grid(n)), dim3(256), 0, s->stream, s->words64(), s->nwords, s->d_off,
                           s->d_off + n, s->d_len, n, s->d_bytes);
        e = hipGetLastError();
    }
    if (!e) e = hipMemcpyAsync(eq, s->d_bytes, n, hipMemcpyDeviceToHost, s->stream);
    if (!e) e = hipStreamSynchronize(s->stream);
    return e ? dev("sf_store_codes_eq", e) : SF_OK;
}

} // extern "C"

// A tie-break call prepared on the device: the reads packed into a sequence buffer beside the store
// (and its reversed copy), each segment's distinct texts found, and the distinct (segment, text)
// pairs as resident batches over that buffer, one for the extensions and one per mode for the
// sorted scores.
struct TbGroup {
    std::vector<uint64_t> cand, pbase, tbase, geom, cut;
    std::vector<uint32_t> plen, tlen, seg;
    sf_batch *b = nullptr;
};

struct sf_tb_batch {
    sf_store *st = nullptr;
    uint32_t *own_seq = nullptr; // [store | reversed store | reads] of an sf_tb_batch (null: the store's)
    std::vector<sf_tb_segment> segs;
    std::vector<uint32_t> rep, first;
    uint64_t ncand = 0;
    TbGroup ext, srt[3];
    sf_tb_params p{};
    bool ran = false;
    // A batch that runs many times keeps its results' staging between calls: the extension
    // records in pinned host memory (a copy into pageable memory runs several times slower), the
    // per-candidate results' buffer allocated once.
    bool resident = false;
    sf_dp *pin_dp = nullptr;
    uint64_t pin_n = 0;
    std::vector<sf_extend_cut> res;
    ~sf_tb_batch() {
        for (TbGroup *g : {&ext, &srt[0], &srt[1], &srt[2]})
            if (g->b) sf_batch_destroy(g->b);
        if (own_seq) (void)hipFree(own_seq);
        if (pin_dp) (void)hipHostFree(pin_dp);
    }
};

namespace {

// Checks a call's arguments (every candidate in exactly one segment, every range inside its buffer)
// and gives each candidate's segment and its segment's first candidate, and each segment's longest
// text.
int tb_check(const sf_store *st, uint64_t reads_len, const sf_tb_segment *segs, uint64_t nseg,
             const uint64_t *cand_off, const uint32_t *cand_len, uint64_t ncand, std::vector<uint32_t> &seg_of,
             std::vector<uint32_t> &first, std::vector<uint64_t> &longest) {
    if (ncand >= UINT32_MAX) return sf_runtime_fail(SF_ERANGE, "2^32 - 1 candidates or more");
    if (reads_len >= (1ull << 40)) return sf_runtime_fail(SF_ERANGE, "2^40 read bases or more");
    seg_of.assign(ncand, UINT32_MAX);
    first.assign(ncand, 0);
    longest.assign(nseg, 0);
    for (uint64_t s = 0; s < nseg; ++s) {
        const sf_tb_segment &g = segs[s];
        if (g.read_off > reads_len || g.read_len > reads_len - g.read_off)
            return sf_runtime_fail(SF_EINVAL, "a segment's read slice lies outside the reads");
        if ((uint64_t)g.cand0 + g.ncand > ncand) return sf_runtime_fail(SF_EINVAL, "a segment's candidates lie outside");
        if (g.mode < SF_INFIX || g.mode > SF_GLOBAL || (g.flags & ~31u))
            return sf_runtime_fail(SF_EINVAL, "segment mode or flags");
        if ((g.flags & SF_SEG_EXTEND) && g.mode != SF_PREFIX)
            return sf_runtime_fail(SF_EINVAL, "an extension segment is Prefix (align.rs:903-949)");
        if (g.read_len >= (1u << 30)) return sf_runtime_fail(SF_ERANGE, "a read slice of 2^30 bases or more");
        for (uint32_t c = g.cand0; c < g.cand0 + g.ncand; ++c) {
            if (seg_of[c] != UINT32_MAX) return sf_runtime_fail(SF_EINVAL, "a candidate belongs to two segments");
            seg_of[c] = (uint32_t)s;
            first[c] = g.cand0;
        }
    }
    for (uint64_t c = 0; c < ncand; ++c) {
        if (seg_of[c] == UINT32_MAX) return sf_runtime_fail(SF_EINVAL, "a candidate belongs to no segment");
        if (cand_off[c] > st->bases || cand_len[c] > st->bases - cand_off[c])
            return sf_runtime_fail(SF_EINVAL, "a candidate lies outside the store");
        if (cand_len[c] >= (1u << 30)) return sf_runtime_fail(SF_ERANGE, "a text of 2^30 bases or more");
        longest[seg_of[c]] = std::max<uint64_t>(longest[seg_of[c]], cand_len[c]);
    }
    return SF_OK;
}

// Prepares w over `seq` (the store's buffer or a copy of its two store parts, with room for
// 2 ceil(reads_len / 16) words of reads after them), the store's mutex held: the reads packed
// there forward and reversed, the candidates deduplicated per segment (with the store's scratch),
// the distinct pairs grouped and their batches made.
int tb_prepare(sf_store *st, uint32_t *seq, const uint8_t *reads, uint64_t reads_len, const sf_tb_segment *segs,
               uint64_t nseg, const uint64_t *cand_off, const uint32_t *cand_len, uint64_t ncand, sf_tb_batch &w) {
    std::vector<uint32_t> seg_of;
    std::vector<uint64_t> longest;
    if (const int q = tb_check(st, reads_len, segs, nseg, cand_off, cand_len, ncand, seg_of, w.first, longest))
        return q;
    w.segs.assign(segs, segs + nseg);
    w.ncand = ncand;
    w.rep.assign(ncand, 0);
    if (ncand == 0) return SF_OK;
    const uint64_t rwords = (reads_len + 15) / 16;
    if (!reserve(st->d_bytes, st->bytes_cap, reads_len) || !reserve(st->d_off, st->off_cap, ncand) ||
        !reserve(st->d_len, st->len_cap, ncand) || !reserve(st->d_first, st->first_cap, ncand) ||
        !reserve(st->d_rep, st->rep_cap, ncand))
        return sf_runtime_fail(SF_ENOMEM, "tie-break scratch");
    uint32_t *d_rfwd = seq + 4 * st->nwords, *d_rrev = d_rfwd + rwords;
    const uint64_t rfwd_base = st->reads_base(), rrev_base = rfwd_base + 16 * rwords;
    {
        hipError_t e = reads_len ? hipMemcpyAsync(st->d_bytes, reads, reads_len, hipMemcpyHostToDevice, st->stream)
                                 : hipSuccess;
        if (!e && rwords) {
            hipLaunchKernelGGL(pack_reads_kernel, dim3(grid(rwords)), dim3(256), 0, st->stream, st->d_bytes, reads_len,
                               d_rfwd, d_rrev, rwords);
            e = hipGetLastError();
        }
        if (!e) e = hipMemcpyAsync(st->d_off, cand_off, ncand * 8, hipMemcpyHostToDevice, st->stream);
        if (!e) e = hipMemcpyAsync(st->d_len, cand_len, ncand * 4, hipMemcpyHostToDevice, st->stream);
        if (!e) e = hipMemcpyAsync(st->d_first, w.first.data(), ncand * 4, hipMemcpyHostToDevice, st->stream);
        if (!e) {
            hipLaunchKernelGGL(dedup_kernel, dim3(grid(ncand)), dim3(256), 0, st->stream, st->words64(), st->nwords,
                               st->d_off, st->d_len, st->d_first, ncand, st->d_rep);
            e = hipGetLastError();
        }
        if (!e) e = hipMemcpyAsync(w.rep.data(), st->d_rep, ncand * 4, hipMemcpyDeviceToHost, st->stream);
        if (!e) e = hipStreamSynchronize(st->stream);
        if (e) return dev("tie-break dedup", e);
    }
    // The distinct (segment, text) pairs as ranges of the buffer: the read slice and the text, both
    // reversed for a left read end (lib.rs:1347-1365).  An extension's band is laid out for
    // min(m, n) with `diagonal`, else for the larger of the segment's window and its longest text
    // (align.rs:916, lib.rs:1446-1451).
    for (uint64_t c = 0; c < ncand; ++c) {
        if (w.rep[c] != c) continue;
        const sf_tb_segment &g = segs[seg_of[c]];
        const bool rev = g.flags & SF_SEG_REVERSE;
        TbGroup &q = (g.flags & SF_SEG_EXTEND) ? w.ext : w.srt[g.mode];
        q.cand.push_back(c);
        q.seg.push_back(seg_of[c]);
        q.pbase.push_back(rev ? rrev_base + (reads_len - g.read_off - g.read_len) : rfwd_base + g.read_off);
        q.tbase.push_back(rev ? st->rev_base() + (st->bases - cand_off[c] - cand_len[c]) : cand_off[c]);
        q.plen.push_back(g.read_len);
        q.tlen.push_back(cand_len[c]);
        if (g.flags & SF_SEG_EXTEND) {
            q.geom.push_back((g.flags & SF_SEG_DIAGONAL) ? std::min<uint64_t>(g.read_len, cand_len[c])
                                                         : std::max<uint64_t>(longest[seg_of[c]], g.window));
            q.cut.push_back(g.cut);
        }
    }
    for (TbGroup *q : {&w.ext, &w.srt[0], &w.srt[1], &w.srt[2]})
        if (!q->cand.empty())
            if (const int r = sf_batch_create_resident(seq, q->cand.size(), q->plen.data(), q->tlen.data(),
                                                       q->pbase.data(), q->tbase.data(), &q->b))
                return r;
    return SF_OK;
}

// Queues the passes of every group (their batches' streams run side by side).
int tb_run(sf_tb_batch &w, const sf_tb_params *p) {
    w.p = *p;
    if (w.ext.b) {
        std::vector<uint64_t> band(w.ext.cand.size(), p->band);
        // align.rs:924: the best cell is tracked only with track_cut and a cut.  One pass serves
        // every pair: it tracks the cell if any pair needs it (tracking changes no score; the results
        // drop the cell where it is not kept).
        bool cell = false;
        for (size_t i = 0; i < w.ext.cand.size() && !cell; ++i)
            cell = (w.segs[w.ext.seg[i]].flags & SF_SEG_TRACK_CUT) && w.ext.cut[i] != UINT64_MAX;
        if (const int r = sf_batch_affine_dp(w.ext.b, SF_PREFIX, &p->sc, band.data(), w.ext.geom.data(), 1, p->xdrop,
                                             cell ? 1 : 0, w.ext.cut.data()))
            return r;
    }
    for (int mode = SF_INFIX; mode <= SF_GLOBAL; ++mode)
        if (w.srt[mode].b) {
            std::vector<uint64_t> band(w.srt[mode].cand.size(), p->band);
            if (const int r = sf_batch_affine(w.srt[mode].b, mode, &p->sc, band.data(), 0)) return r;
        }
    w.ran = true;
    return SF_OK;
}

// Waits for the passes and gives every candidate its distinct text's result: align.rs:919-948 per
// extension pair with its segment's flags, (x, x, 0) for a sorted score (lib.rs:1442-1445).
int tb_results(sf_tb_batch &w, sf_extend_cut *out, uint32_t *rep_out) {
    if (!w.ran) return sf_runtime_fail(SF_EINVAL, "no run on this tie-break batch");
    std::vector<sf_extend_cut> &res = w.res;
    res.resize(w.ncand);
    const sf_tb_params &p = w.p;
    if (w.ext.b) {
        const size_t n = w.ext.cand.size();
        std::vector<sf_dp> dv;
        sf_dp *d = nullptr;
        if (w.resident && w.pin_n < n) {
            if (w.pin_dp) (void)hipHostFree(w.pin_dp);
            w.pin_dp = nullptr;
            w.pin_n = 0;
            if (hipHostMalloc((void **)&w.pin_dp, n * sizeof(sf_dp), hipHostMallocDefault) == hipSuccess) w.pin_n = n;
            else w.pin_dp = nullptr; // (pageable staging below)
        }
        if (w.resident && w.pin_n >= n) {
            d = w.pin_dp;
        } else {
            dv.resize(n);
            d = dv.data();
        }
        if (const int r = sf_batch_dp_results(w.ext.b, d)) return r;
        for (size_t i = 0; i < n; ++i) {
            const sf_tb_segment &g = w.segs[w.ext.seg[i]];
            const bool has_cut = g.cut != UINT64_MAX, track = (g.flags & SF_SEG_TRACK_CUT) && has_cut;
            sf_dp x = d[i];
            if (!track) x.clip_i = x.clip_j = 0;
            const int32_t end = (g.flags & SF_SEG_BONUS_WITHIN_CUT) && has_cut
                                    ? std::max(x.score, sat_add(x.cut_score, p.end_bonus))
                                    : sat_add(x.score, p.end_bonus);
            const int32_t whole = std::max(x.clip_best, end);
            const uint32_t cols = x.clip_best >= end ? x.clip_j : x.score_col;
            const int32_t cutv = has_cut ? std::max(x.cut_best, sat_add(x.cut_score, p.end_bonus)) : whole;
            res[w.ext.cand[i]] = sf_extend_cut{whole, cutv, cols};
        }
    }
    for (int mode = SF_INFIX; mode <= SF_GLOBAL; ++mode) {
        TbGroup &q = w.srt[mode];
        if (!q.b) continue;
        std::vector<int32_t> sc(q.cand.size());
        if (const int r = sf_batch_affine_results(q.b, sc.data(), nullptr, nullptr, nullptr, 0, nullptr)) return r;
        for (size_t i = 0; i < q.cand.size(); ++i) res[q.cand[i]] = sf_extend_cut{sc[i], sc[i], 0u};
    }
    for (uint64_t c = 0; c < w.ncand; ++c) {
        if (out) out[c] = res[w.rep[c]];
        if (rep_out) rep_out[c] = w.rep[c] - w.first[c];
    }
    return SF_OK;
}

} // namespace

extern "C" {

int sf_tiebreak_segments(sf_store *st, const uint8_t *reads, uint64_t reads_len, const sf_tb_segment *segs,
                         uint64_t nseg, const uint64_t *cand_off, const uint32_t *cand_len, uint64_t ncand,
                         const sf_tb_params *p, sf_extend_cut *out, uint32_t *rep_out) {
    if (!st || !p || (nseg && !segs) || (ncand && (!cand_off || !cand_len || !out)) || (reads_len && !reads))
        return sf_runtime_fail(SF_EINVAL, "null argument");
    // On the store's own buffer: its reads area is the call's, so calls on one store take turns.
    std::lock_guard<std::mutex> lk(st->mu);
    if (const int q = st->reserve_reads(2 * ((reads_len + 15) / 16) + 1)) return q;
    sf_tb_batch w;
    w.st = st;
    int r = tb_prepare(st, st->d_seq, reads, reads_len, segs, nseg, cand_off, cand_len, ncand, w);
    if (!r && ncand) r = tb_run(w, p);
    if (!r && ncand) r = tb_results(w, out, rep_out);
    return r;
}

int sf_tb_batch_create(sf_store *st, const uint8_t *reads, uint64_t reads_len, const sf_tb_segment *segs,
                       uint64_t nseg, const uint64_t *cand_off, const uint32_t *cand_len, uint64_t ncand,
                       sf_tb_batch **out) {
    if (!out) return sf_runtime_fail(SF_EINVAL, "out is null");
    *out = nullptr;
    if (!st || (nseg && !segs) || (ncand && (!cand_off || !cand_len)) || (reads_len && !reads))
        return sf_runtime_fail(SF_EINVAL, "null argument");
    if (reads_len >= (1ull << 40)) return sf_runtime_fail(SF_ERANGE, "2^40 read bases or more");
    auto *w = new sf_tb_batch();
    w->resident = true;
    w->st = st;
    // Its own buffer: the store's two parts copied (device to device) and room for the reads.
    const uint64_t words = 4 * st->nwords + 2 * ((reads_len + 15) / 16) + 1;
    if (hipMalloc((void **)&w->own_seq, words * 4) != hipSuccess) {
        delete w;
        return sf_runtime_fail(SF_ENOMEM, "tie-break batch buffer");
    }
    std::lock_guard<std::mutex> lk(st->mu);
    hipError_t e = st->nwords ? hipMemcpyAsync(w->own_seq, st->d_seq, 4 * st->nwords * 4, hipMemcpyDeviceToDevice,
                                               st->stream)
                              : hipSuccess;
    if (e) {
        delete w;
        return dev("tie-break batch buffer", e);
    }
    if (const int r = tb_prepare(st, w->own_seq, reads, reads_len, segs, nseg, cand_off, cand_len, ncand, *w)) {
        (void)hipStreamSynchronize(st->stream);
        delete w;
        return r;
    }
    *out = w;
    return SF_OK;
}

int sf_tb_batch_run(sf_tb_batch *b, const sf_tb_params *p) {
    if (!b || !p) return sf_runtime_fail(SF_EINVAL, "null argument");
    return b->ncand ? tb_run(*b, p) : (b->ran = true, (int)SF_OK);
}

int sf_tb_batch_results(sf_tb_batch *b, sf_extend_cut *out, uint32_t *rep) {
    if (!b) return sf_runtime_fail(SF_EINVAL, "batch is null");
    if (b->ncand && !out) return sf_runtime_fail(SF_EINVAL, "out is null");
    return tb_results(*b, out, rep);
}

int sf_tb_batch_timing(sf_tb_batch *b, sf_tb_timing *t, int reset) {
    if (!b || !t) return sf_runtime_fail(SF_EINVAL, "null argument");
    *t = sf_tb_timing{};
    if (b->ext.b) {
        sf_timing x;
        if (const int r = sf_batch_timing(b->ext.b, &x)) return r;
        t->ext_ms = x.affine_ms, t->ext_cells = x.affine_cells;
        if (reset) sf_batch_reset_timing(b->ext.b);
    }
    for (TbGroup &q : b->srt)
        if (q.b) {
            sf_timing x;
            if (const int r = sf_batch_timing(q.b, &x)) return r;
            t->sorted_ms += x.affine_ms, t->sorted_cells += x.affine_cells;
            if (reset) sf_batch_reset_timing(q.b);
        }
    return SF_OK;
}

void sf_tb_batch_destroy(sf_tb_batch *b) { delete b; }

} // extern "C"
