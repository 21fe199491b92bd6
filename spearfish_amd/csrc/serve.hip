// serve.hip -- continuous-batching alignment server (sf_server_* in include/sf_align.h).
//
// Reads arrive one request at a time (a read plus its segments: read slices against windows of
// a transcript store that stays resident in HBM).  A scheduler thread keeps up to `slots`
// batches on the device at once, each on its own HIP stream.  Whenever a slot is free it admits
// the waiting requests, oldest first, while they fit the slot's share of the HBM budget (read
// bases and segment count), packs them into the slot's pinned staging memory, copies them over,
// launches the Myers and affine kernels for every segment and queues the copy of the results
// back; requests that arrive meanwhile wait only for the next free slot, so their work overlaps
// the batches already running.  A finished batch's results go straight into the caller's output
// arrays and the caller's wait returns.  Per request the latency from submit to completion is
// recorded; sf_server_get_stats reports p50/p99 next to the counts.
//
// Per segment: Myers Infix distance and end column (align.rs:92-168 semantics, k unbounded)
// and the banded affine Global score and end column (align.rs:699-701, the server's band and
// scoring) -- the kernels of sf_batch_*, each bit-exact against the same recurrences, so a
// segment's results do not depend on which other requests share its batch (the batch's size may
// pick another kernel for a Myers segment: the 16-lane pipeline for small batches).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include <sys/prctl.h>

#include "../../include/sf_align.h"
#include "internal.hpp"

using namespace sf;
using Clock = std::chrono::steady_clock;

int sf_runtime_fail(int code, const char *msg); // runtime.hip

hipError_t sf_load_code_objects() {
    for (hipError_t (*f)() : {sf_load_code_myers, sf_load_code_myers_mix, sf_load_code_affine, sf_load_code_affine_pk,
                              sf_load_code_affine_pk64, sf_load_code_affine_mix, sf_load_code_affine_mix1,
                              sf_load_code_serve_mix})
        if (const hipError_t e = f()) return e;
    return hipSuccess;
}
int sf_runtime_device();                        // runtime.hip
int sf_runtime_ensure();                        // runtime.hip

namespace {

// A request.  Requests are recycled through the server's pool (their vectors keep their
// capacity), so that a submission in steady state allocates nothing.
struct Req {
    uint64_t ticket = 0;
    std::vector<uint32_t> words; // the read, 2-bit packed
    std::vector<uint32_t> roff, rlen, tlen;
    std::vector<uint64_t> toff;
    uint32_t *dist = nullptr, *end = nullptr, *col = nullptr;
    int32_t *score = nullptr;
    Clock::time_point t_submit, t_done;
    int status = SF_OK;
    bool done = false;
    bool queued = false;    // waiting in `pending` (not yet taken into a batch)
    bool cancelled = false; // results are dropped (set under copy_mu, read by copy_out under it)
    uint32_t waiters = 0;   // threads inside sf_server_wait for this ticket

    void reset() {
        ticket = 0;
        words.clear(), roff.clear(), rlen.clear(), tlen.clear(), toff.clear();
        dist = end = col = nullptr;
        score = nullptr;
        status = SF_OK;
        done = queued = cancelled = false;
        waiters = 0;
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    uint64_t words_base = 0; // word offset of this slot's read area in the device sequence buffer
    // A batch's segment records, one block of 104 B a segment laid out by the batch's own segment
    // count (poff, toff, ba: u64; plen, tlen, Myers order, affine order: u32; then the pair
    // records in Myers and in affine launch order), the same layout in pinned host memory and on
    // the device.  A small batch's pair records are written here on the host, so one copy moves
    // all of it and no kernel builds them; a large one copies the first 40 B a segment and builds
    // them on the device.  Results: dist, end, score, col, each the batch's segment count long,
    // one copy back.
    static constexpr uint64_t kStageWords = 13;      // 104 B a segment
    static_assert(sizeof(PairDesc) == 32, "the stage block holds two 32 B pair records a segment");
    uint64_t *d_stage = nullptr, *h_stage = nullptr; // kStageWords x cap_seg words
    uint32_t *d_res = nullptr, *h_res = nullptr;     // 4 x cap_seg
    uint32_t *h_words = nullptr;                     // pinned: the batch's reads
    // views into the stage blocks for the current batch (set by views())
    uint64_t *d_poff = nullptr, *d_toff = nullptr, *d_ba = nullptr;
    uint32_t *d_plen = nullptr, *d_tlen = nullptr, *d_morder = nullptr, *d_aorder = nullptr;
    uint64_t *h_poff = nullptr, *h_toff = nullptr, *h_ba = nullptr;
    uint32_t *h_plen = nullptr, *h_tlen = nullptr, *h_morder = nullptr, *h_aorder = nullptr;
    PairDesc *d_mdesc = nullptr, *d_adesc = nullptr, *h_mdesc = nullptr, *h_adesc = nullptr;

    void views(uint32_t seg) {
        auto cut = [seg](uint64_t *b, uint64_t *&poff, uint64_t *&toff, uint64_t *&ba, uint32_t *&plen,
                         uint32_t *&tlen, uint32_t *&mo, uint32_t *&ao, PairDesc *&md, PairDesc *&ad) {
            poff = b, toff = b + seg, ba = b + 2 * (uint64_t)seg;
            plen = reinterpret_cast<uint32_t *>(b + 3 * (uint64_t)seg);
            tlen = plen + seg, mo = tlen + seg, ao = mo + seg;
            md = reinterpret_cast<PairDesc *>(b + 5 * (uint64_t)seg);
            ad = md + seg;
        };
        cut(d_stage, d_poff, d_toff, d_ba, d_plen, d_tlen, d_morder, d_aorder, d_mdesc, d_adesc);
        cut(h_stage, h_poff, h_toff, h_ba, h_plen, h_tlen, h_morder, h_aorder, h_mdesc, h_adesc);
    }
    std::vector<Req *> reqs;
    std::vector<uint32_t> seg0;
    std::vector<uint64_t> key; // launch-order sort keys (host work space)
    std::vector<uint32_t> tmp;
    uint32_t nseg = 0;
    bool busy = false;   // a batch is staged or on the device (the scheduler thread's own state)
    int fail = SF_OK;    // set when the batch could not be queued: it completes with this status
    int status = SF_OK;  // the finished batch's outcome, once `ended`
    bool ended = false;  // results are in the callers' arrays; not yet published under the lock
    uint64_t cells = 0;  // DP cells of the batch (both kernels), counted when it is published
    uint64_t launch_no = 0; // 1-based count of batches the server has taken in (never reset)
};

template <class T>
bool dev_alloc(T **p, uint64_t count) {
    return hipMalloc((void **)p, std::max<uint64_t>(count, 1) * sizeof(T)) == hipSuccess;
}
template <class T>
bool host_alloc(T **p, uint64_t count) {
    return hipHostMalloc((void **)p, std::max<uint64_t>(count, 1) * sizeof(T), hipHostMallocDefault) == hipSuccess;
}

constexpr uint32_t kPipeBinMax = 8192; // segments of a Myers lane bin that still run the 16-lane pipeline
constexpr uint32_t kWideAffMax = 2048; // segments of a batch whose affine runs 64 lanes a pair, in one launch

// Fills order[0, n) with 0..n-1 sorted by key, ties in index order: LSD radix over 11-bit digits,
// skipping a digit that is the same in every key (the high digits, mostly), so a batch of a few
// segments costs microseconds and one of 10^5 about a millisecond.
void key_order(const std::vector<uint64_t> &key, uint32_t *order, std::vector<uint32_t> &tmp) {
    constexpr int D = 11, NB = 1 << D, P = (64 + D - 1) / D;
    const uint32_t n = (uint32_t)key.size();
    std::iota(order, order + n, 0u);
    if (n < 2) return;
    std::vector<uint32_t> cnt((size_t)P * NB, 0u);
    for (uint32_t i = 0; i < n; ++i)
        for (int p = 0; p < P; ++p) ++cnt[(size_t)p * NB + ((key[i] >> (D * p)) & (NB - 1))];
    tmp.resize(n);
    uint32_t *src = order, *dst = tmp.data();
    for (int p = 0; p < P; ++p) {
        uint32_t *c = &cnt[(size_t)p * NB];
        if (c[(key[0] >> (D * p)) & (NB - 1)] == n) continue;
        uint32_t sum = 0;
        for (int d = 0; d < NB; ++d) {
            const uint32_t x = c[d];
            c[d] = sum;
            sum += x;
        }
        for (uint32_t q = 0; q < n; ++q) {
            const uint32_t i = src[q];
            dst[c[(key[i] >> (D * p)) & (NB - 1)]++] = i;
        }
        std::swap(src, dst);
    }
    if (src != order) std::memcpy(order, src, (size_t)n * 4);
}

// The q-quantile of v (linear between the two nearest ranks); reorders v.
double pct(std::vector<double> &v, double q) {
    if (v.empty()) return 0;
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    std::nth_element(v.begin(), v.begin() + lo, v.end());
    const double a = v[lo];
    const double b = lo + 1 < v.size() ? *std::min_element(v.begin() + lo + 1, v.end()) : a;
    return a + (b - a) * (pos - (double)lo);
}

} // namespace

struct sf_server {
    sf_server_config cfg{};
    // SF_SERVE_FAIL_BATCH=n (read at creation): the n-th batch (1-based) fails without a launch.
    // A fault injected from the host for tests of failure handling; 0 (unset) is off.
    uint64_t fail_batch = 0;
    uint64_t store_bases = 0, store_words = 0;
    uint32_t *d_seq = nullptr; // [store words | slot read areas]
    uint64_t cap_words = 0, cap_seg = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    // Held while a finished batch's results are copied into the callers' arrays, and by
    // sf_server_cancel while it marks a request in a batch as cancelled, so that no copy into a
    // cancelled request's outputs happens after the cancel returns.  Order: copy_mu before mu.
    std::mutex copy_mu;
    std::condition_variable cv_sched, cv_done;
    std::deque<Req *> pending;
    uint64_t pend_words = 0, pend_seg = 0; // the waiting requests' read words and segments
    // Recycled requests (pool_mu).  Filled at creation, so that the first submissions allocate
    // nothing either.
    std::mutex pool_mu;
    std::vector<Req *> pool;
    static constexpr size_t kPoolPrefill = 1024, kPoolMax = 1u << 16;
    std::unordered_map<uint64_t, Req *> reqs;
    uint64_t next_ticket = 1, launches = 0;
    bool stop = false;
    std::thread th;
    // stats
    std::vector<double> lat_ms;
    uint64_t reads_done = 0, reads_failed = 0, batches = 0, segments = 0, batch_reads = 0, max_batch_seg = 0;
    uint64_t cells = 0;
    Clock::time_point t_first, t_last;
    bool any = false;
    // SF_SERVE_TRACE set: the longest scheduler iteration (busy time: an idle wait for requests is
    // not counted), launch and gap between submissions (with when they ended, from creation),
    // reported on stderr at each stats reset and when the server is destroyed.  Guarded by
    // trace_mu (the scheduler and submitters write them, sf_server_get_stats reads them).
    bool trace = std::getenv("SF_SERVE_TRACE") != nullptr;
    std::mutex trace_mu;
    Clock::time_point t_created = Clock::now(), t_last_submit{};
    double max_iter_ms = 0, max_iter_at = 0, max_launch_ms = 0, max_launch_at = 0, max_submit_gap_ms = 0,
           max_submit_gap_at = 0, max_build_ms = 0, max_lock_ms = 0;

    Req *take_req() {
        {
            std::lock_guard<std::mutex> lk(pool_mu);
            if (!pool.empty()) {
                Req *r = pool.back();
                pool.pop_back();
                return r;
            }
        }
        return new Req();
    }
    void give_req(Req *r) {
        r->reset();
        {
            std::lock_guard<std::mutex> lk(pool_mu);
            if (pool.size() < kPoolMax) {
                pool.push_back(r);
                return;
            }
        }
        delete r;
    }
    void trace_max(double &field, double &at, double ms, Clock::time_point when) {
        std::lock_guard<std::mutex> lk(trace_mu);
        if (ms > field) field = ms, at = std::chrono::duration<double>(when - t_created).count();
    }

    void loop();
    void launch(Slot &s);
    void copy_out(Slot &s, int status);
    void publish(Slot &s);
};

// Results of a finished batch into the callers' arrays.  Runs without the server's lock: a
// request is not freed before it is marked done (sf_server_wait), and its fields do not change
// after submission; copy_mu keeps a cancel from overlapping the copy.
void sf_server::copy_out(Slot &s, int status) {
    const uint32_t ns = s.nseg;
    const uint32_t *rd = s.h_res, *re = s.h_res + ns, *rc = s.h_res + 3 * (uint64_t)ns;
    const int32_t *rs = reinterpret_cast<const int32_t *>(s.h_res + 2 * (uint64_t)ns);
    if (status == SF_OK) {
        std::lock_guard<std::mutex> lk(copy_mu);
        for (size_t q = 0; q < s.reqs.size(); ++q) {
            Req *r = s.reqs[q];
            if (r->cancelled) continue;
            const uint32_t k0 = s.seg0[q], n = (uint32_t)r->rlen.size();
            if (r->dist) std::memcpy(r->dist, rd + k0, n * 4);
            if (r->end) std::memcpy(r->end, re + k0, n * 4);
            if (r->score) std::memcpy(r->score, rs + k0, n * 4);
            if (r->col) std::memcpy(r->col, rc + k0, n * 4);
        }
    }
    s.status = status;
    s.ended = true;
}

// Marks a finished batch's requests done and counts it.  Called with the lock held.
void sf_server::publish(Slot &s) {
    const auto now = Clock::now();
    for (Req *r : s.reqs) {
        if (r->cancelled) {
            r->status = SF_ECANCELED;
        } else if (s.status == SF_OK) {
            ++reads_done;
            lat_ms.push_back(std::chrono::duration<double, std::milli>(now - r->t_submit).count());
            r->status = SF_OK;
        } else {
            ++reads_failed;
            r->status = s.status;
        }
        r->t_done = now;
        r->done = true;
    }
    ++batches;
    segments += s.nseg;
    max_batch_seg = std::max<uint64_t>(max_batch_seg, s.nseg);
    batch_reads += s.reqs.size();
    cells += s.cells;
    t_last = now;
    s.reqs.clear();
    s.seg0.clear();
    s.ended = false;
    s.busy = false;
}

// Stages the slot's batch (s.reqs), copies it over and queues the kernels and the copy of the
// results back on the slot's stream.  Runs without the lock.
void sf_server::launch(Slot &s) {
    // Stage: reads, then one descriptor per segment.
    uint64_t words = 0;
    uint32_t seg = 0;
    s.seg0.clear();
    s.cells = 0;
    uint32_t nseg = 0;
    for (Req *r : s.reqs) nseg += (uint32_t)r->rlen.size();
    s.views(nseg);
    for (Req *r : s.reqs) {
        std::memcpy(s.h_words + words, r->words.data(), r->words.size() * 4);
        const uint64_t rbase = (store_words + s.words_base + words) * 16;
        s.seg0.push_back(seg);
        for (size_t k = 0; k < r->rlen.size(); ++k, ++seg) {
            s.h_poff[seg] = rbase + r->roff[k];
            s.h_plen[seg] = r->rlen[k];
            s.h_toff[seg] = r->toff[k];
            s.h_tlen[seg] = r->tlen[k];
        }
        words += r->words.size();
    }
    s.nseg = seg;
    // Shape classes (as sf_batch_*): Myers by pattern blocks, affine by band width.
    std::vector<int> mcls(seg), acls(seg);
    uint32_t lane_bin[25] = {};
    // Up to kWideAffMax segments (a wave each: two a SIMD) the affine run is chain-bound: each
    // pair takes its 64-lane class (a row costs 3.5 kp + 30 VALU at kp slots a lane, against
    // 3.5 kp + 24 at 16 lanes with 4 times the slots), and the classes run side by side in one
    // mixed launch, so the run takes its longest pair's chain rather than the sum over classes.
    const bool wide = seg <= kWideAffMax;
    for (uint32_t i = 0; i < seg; ++i) {
        mcls[i] = sf_myers_wide_class(sf_myers_maxb_for((s.h_plen[i] + 63) / 64)); // (small batches)
        if (mcls[i] <= 24) ++lane_bin[mcls[i]];
        uint64_t below, above;
        band_geom(s.h_plen[i], s.h_tlen[i], cfg.affine_mode, cfg.band, &below, &above);
        s.h_ba[i] = below | (above << 32);
        acls[i] = sf_affine_k_for(below + above + 1);
        int c64 = wide ? sf_affine_class(below + above + 1, false) : 0;
        if (c64) {
            const int kp = sf_affine_pk16_class(aff_k(c64), cfg.sc.mat, cfg.sc.mis, cfg.sc.open, cfg.sc.ext,
                                                s.h_plen[i], (uint32_t)below, false, 64);
            if (kp) c64 = kp;
            if (!sf_affine_mixable(c64) || !scores_in_range(cfg.sc.mat, cfg.sc.mis, cfg.sc.open, cfg.sc.ext,
                                                            s.h_plen[i], s.h_tlen[i], aff_slots(c64)))
                c64 = 0;
        }
        if (c64) {
            acls[i] = c64;
        } else if (aff_lp(acls[i]) == 16) { // scores only: the 16-bit kernel where its range holds
            const int kp = sf_affine_pk16_class(aff_k(acls[i]), cfg.sc.mat, cfg.sc.mis, cfg.sc.open, cfg.sc.ext,
                                                s.h_plen[i], (uint32_t)below, false, 16);
            if (kp) acls[i] = kp;
        }
        s.cells += (uint64_t)s.h_plen[i] * s.h_tlen[i] + band_cells(s.h_plen[i], s.h_tlen[i], below, above);
    }
    // A serving batch is small, so a segment's own chain sets its time, not the chip's issue rate.
    // The lane kernel walks a pattern's blocks one after another in one lane (20 blocks: 20 block
    // steps a column); the 16-lane pipeline gives each lane 1 block (patterns of up to 16 blocks)
    // or 2 and moves 4 columns a step, a chain a fifth to a tenth as long for 3 times the
    // lane-instructions.  A lane bin of 5-24 blocks (classes 8-24) therefore runs the pipeline
    // while it is at most kPipeBinMax segments (2,048 waves: two a SIMD), beyond which the extra
    // instructions cost more than the chain saves.
    for (uint32_t i = 0; i < seg; ++i)
        if (mcls[i] >= 8 && mcls[i] <= 24 && lane_bin[mcls[i]] <= kPipeBinMax) mcls[i] = mcls[i] <= 16 ? 16001 : 16002;
    // Within a class, similar lengths side by side: the lanes (Myers, shorter texts first) and the
    // pairs sharing a wave (affine, longer patterns first) then run about as many columns and rows.
    s.key.resize(seg);
    for (uint32_t i = 0; i < seg; ++i) s.key[i] = ((uint64_t)(uint32_t)mcls[i] << 32) | s.h_tlen[i];
    key_order(s.key, s.h_morder, s.tmp);
    for (uint32_t i = 0; i < seg; ++i) s.key[i] = ((uint64_t)(uint32_t)acls[i] << 32) | (0xFFFFFFFFu - s.h_plen[i]);
    key_order(s.key, s.h_aorder, s.tmp);
    hipStream_t st = s.stream;
    int status = SF_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && status == SF_OK) status = SF_EDEVICE;
    };
    ok(hipMemcpyAsync(d_seq + store_words + s.words_base, s.h_words, words * 4, hipMemcpyHostToDevice, st));
    const PairsView pv{d_seq, s.d_poff, s.d_toff, s.d_plen, s.d_tlen};
    if (wide) {
        // The pair records in both launch orders (as desc_kernel makes them), sent with the rest.
        for (uint32_t w = 0; w < seg; ++w) {
            const uint32_t i = s.h_morder[w], k = s.h_aorder[w];
            s.h_mdesc[w] = PairDesc{s.h_poff[i], s.h_toff[i], s.h_plen[i], s.h_tlen[i], 0u, 0u};
            s.h_adesc[w] = PairDesc{s.h_poff[k], s.h_toff[k], s.h_plen[k], s.h_tlen[k], (uint32_t)s.h_ba[k],
                                    (uint32_t)(s.h_ba[k] >> 32)};
        }
        ok(hipMemcpyAsync(s.d_stage, s.h_stage, (uint64_t)seg * Slot::kStageWords * 8, hipMemcpyHostToDevice, st));
    } else {
        ok(hipMemcpyAsync(s.d_stage, s.h_stage, (uint64_t)seg * 40, hipMemcpyHostToDevice, st));
        ok(sf_launch_desc(seg, s.d_morder, pv, nullptr, s.d_mdesc, st));
        ok(sf_launch_desc(seg, s.d_aorder, pv, s.d_ba, s.d_adesc, st));
    }
    uint32_t *dist = s.d_res, *end = s.d_res + seg, *col = s.d_res + 3 * (uint64_t)seg;
    int32_t *score = reinterpret_cast<int32_t *>(s.d_res + 2 * (uint64_t)seg);
    const bool inject = fail_batch && s.launch_no == fail_batch;
    if (!inject) {
        MyersArgs ma{};
        ma.pv = pv;
        ma.mode = cfg.myers_mode;
        ma.k = UINT32_MAX;
        ma.dist = dist;
        ma.end = end;
        AffineArgs aa{};
        aa.pv = pv;
        aa.mode = cfg.affine_mode;
        aa.mat = cfg.sc.mat, aa.mis = cfg.sc.mis, aa.open = cfg.sc.open, aa.ext = cfg.sc.ext;
        aa.score = score;
        aa.score_col = col;
        // Class runs (a, e) of each launch order.  A small batch puts the runs the combined kernel
        // holds into its one launch (Myers and affine side by side, each taking its longest
        // chain) and launches the rest after it; a large one launches every run on its own.
        std::vector<std::pair<uint32_t, uint32_t>> mrest, arest;
        MixArgs mmx{}, amx{};
        auto add = [](MixArgs &mx, int c, uint32_t a, uint32_t e, uint32_t blocks) {
            mx.cls[mx.nseg] = c;
            mx.off[mx.nseg] = a;
            mx.count[mx.nseg] = e - a;
            mx.blk0[mx.nseg] = mx.blocks;
            mx.blocks += blocks;
            ++mx.nseg;
        };
        for (uint32_t a = 0; a < seg;) {
            uint32_t e = a;
            const int c = mcls[s.h_morder[a]];
            while (e < seg && mcls[s.h_morder[e]] == c) ++e;
            if (wide && sf_serve_mix_myers(c) && mmx.nseg < MixArgs::kMax) add(mmx, c, a, e, sf_myers_blocks_for(c, e - a));
            else mrest.push_back({a, e});
            a = e;
        }
        for (uint32_t a = 0; a < seg;) {
            uint32_t e = a;
            const int c = acls[s.h_aorder[a]];
            while (e < seg && acls[s.h_aorder[e]] == c) ++e;
            if (wide && sf_serve_mix_affine(c) && amx.nseg < MixArgs::kMax) add(amx, c, a, e, sf_affine_blocks_for(c, e - a));
            else arest.push_back({a, e});
            a = e;
        }
        if (mmx.nseg || amx.nseg) {
            MyersArgs xm = ma;
            xm.order = s.d_morder;
            xm.desc = s.d_mdesc;
            AffineArgs xa = aa;
            xa.order = s.d_aorder;
            xa.desc = s.d_adesc;
            ok(sf_launch_serve_mix(xa, amx, xm, mmx, st));
        }
        for (auto [a, e] : mrest) {
            MyersArgs x = ma;
            x.order = s.d_morder + a;
            x.desc = s.d_mdesc + a;
            x.count = e - a;
            ok(sf_launch_myers(x, mcls[s.h_morder[a]], st));
        }
        for (auto [a, e] : arest) {
            AffineArgs x = aa;
            x.order = s.d_aorder + a;
            x.desc = s.d_adesc + a;
            x.count = e - a;
            ok(sf_launch_affine(x, acls[s.h_aorder[a]], false, st));
        }
    } else {
        status = SF_EDEVICE; // fault injected from the host: the batch is failed without a launch
    }
    ok(hipMemcpyAsync(s.h_res, s.d_res, (uint64_t)seg * 16, hipMemcpyDeviceToHost, st));
    ok(hipEventRecord(s.ev, st));
    s.fail = status;
}

// The scheduler.  The lock guards only the hand-over: taking waiting requests into free slots
// and marking finished ones done.  Staging, launching and copying results out run without it,
// so submitters and waiters never wait behind a batch's host work.
void sf_server::loop() {
    (void)hipSetDevice(sf_runtime_device());
    // The loop sleeps 10 us between polls of the running batches; with the default 50 us of timer
    // slack such a sleep lasts up to 60 us, much of a small batch's latency.
    (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
    std::vector<Slot *> fresh;
    auto t_iter = Clock::now();
    for (;;) {
        if (trace) {
            const auto now = Clock::now();
            trace_max(max_iter_ms, max_iter_at, std::chrono::duration<double, std::milli>(now - t_iter).count(), now);
            t_iter = now;
        }
        // Batches that ended: results out to the callers (no lock).
        bool ended = false;
        for (Slot &s : slots) {
            if (!s.busy || s.ended || s.reqs.empty()) continue;
            int st = SF_OK;
            if (s.fail != SF_OK) {
                (void)hipStreamSynchronize(s.stream);
                st = s.fail;
            } else {
                const hipError_t q = hipEventQuery(s.ev);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) {
                    (void)hipGetLastError();
                    st = SF_EDEVICE;
                }
            }
            copy_out(s, st);
            ended = true;
        }
        std::unique_lock<std::mutex> lk(mu);
        if (ended) {
            for (Slot &s : slots)
                if (s.ended) publish(s);
            cv_done.notify_all();
        }
        // Free slots take the waiting requests, oldest first, while they fit the slot's share.
        // With a batch wait, they hold off until the waiting requests fill a share or the oldest
        // has waited that long.
        fresh.clear();
        bool hold = false;
        Clock::time_point hold_until{};
        if (cfg.batch_wait_us && !stop && !pending.empty()) {
            const bool fills = pend_words >= cap_words || pend_seg >= cap_seg ||
                               (cfg.max_reads_per_batch && pending.size() >= cfg.max_reads_per_batch);
            hold_until = pending.front()->t_submit + std::chrono::microseconds(cfg.batch_wait_us);
            hold = !fills && Clock::now() < hold_until;
        }
        for (Slot &s : slots) {
            if (s.busy || pending.empty() || hold) continue;
            uint64_t words = 0, seg = 0;
            while (!pending.empty()) {
                Req *r = pending.front();
                if (!s.reqs.empty() && (words + r->words.size() > cap_words || seg + r->rlen.size() > cap_seg ||
                                        (cfg.max_reads_per_batch && s.reqs.size() >= cfg.max_reads_per_batch)))
                    break;
                words += r->words.size();
                seg += r->rlen.size();
                r->queued = false;
                s.reqs.push_back(r);
                pending.pop_front();
                pend_words -= r->words.size();
                pend_seg -= r->rlen.size();
            }
            s.busy = true;
            s.launch_no = ++launches;
            fresh.push_back(&s);
        }
        bool busy = false;
        for (Slot &s : slots) busy |= s.busy;
        if (stop && pending.empty() && !busy) break;
        if (fresh.empty()) {
            if (!ended) {
                if (busy) {
                    cv_sched.wait_for(lk, std::chrono::microseconds(10));
                } else if (hold) {
                    cv_sched.wait_until(lk, hold_until); // (a submission wakes it to look again)
                } else {
                    cv_sched.wait(lk, [&] { return stop || !pending.empty(); });
                    if (trace) t_iter = Clock::now(); // idle, not an iteration's work
                }
            }
            continue;
        }
        lk.unlock();
        for (Slot *s : fresh) {
            const auto t0 = Clock::now();
            launch(*s);
            if (trace) {
                const auto t1 = Clock::now();
                trace_max(max_launch_ms, max_launch_at, std::chrono::duration<double, std::milli>(t1 - t0).count(), t1);
            }
        }
    }
}

extern "C" {

int sf_server_create(const uint32_t *store_words, uint64_t store_bases, const sf_server_config *cfg,
                     sf_server **out) {
    if (!out || !cfg) return sf_runtime_fail(SF_EINVAL, "null argument");
    *out = nullptr;
    if (const int st = sf_runtime_ensure()) return st;
    if (store_bases && !store_words) return sf_runtime_fail(SF_EINVAL, "store is null");
    // Every kernel the loop may launch is loaded now, so that no first use mid-run (a class that
    // turns up late, the first batch over the small-batch limit) holds up the scheduler while the
    // runtime loads its code object.
    if (sf_load_code_objects() != hipSuccess) return sf_runtime_fail(SF_EDEVICE, "loading the kernels failed");
    if (cfg->slots == 0 || cfg->slots > 16) return sf_runtime_fail(SF_EINVAL, "slots must be 1..16");
    if (cfg->myers_mode < SF_INFIX || cfg->myers_mode > SF_GLOBAL || cfg->affine_mode < SF_INFIX ||
        cfg->affine_mode > SF_GLOBAL)
        return sf_runtime_fail(SF_EINVAL, "mode");
    if (cfg->sc.mat < 0 || cfg->sc.mis < 0 || cfg->sc.open < 0 || cfg->sc.ext < 0)
        return sf_runtime_fail(SF_EINVAL, "scores must be non-negative");
    if (cfg->sc.mat > kScoreMax || cfg->sc.mis > kScoreMax || cfg->sc.open > kScoreMax || cfg->sc.ext > kScoreMax)
        return sf_runtime_fail(SF_ERANGE, "scoring values above 4,194,303 are outside this library's bound");
    // Per slot: half the share for read bases (4 B per 16 bases), half for segments (128 B each:
    // offsets, lengths, band, orders, two launch records and results).
    const uint64_t share = cfg->hbm_budget_bytes / cfg->slots;
    if (share < (1u << 16)) return sf_runtime_fail(SF_EINVAL, "hbm budget too small for the slots");
    auto *sv = new sf_server();
    sv->cfg = *cfg;
    if (const char *e = std::getenv("SF_SERVE_FAIL_BATCH")) sv->fail_batch = std::strtoull(e, nullptr, 10);
    sv->store_bases = store_bases;
    sv->store_words = (store_bases + 15) / 16;
    sv->cap_words = share / 2 / 4;
    sv->cap_seg = share / 2 / 128;
    // Sized up front: growing the ticket table rehashes it under the server's lock (the scheduler
    // then waits milliseconds to hand over a batch), and the latency record grows by copying.
    sv->reqs.reserve(1u << 18);
    sv->lat_ms.reserve(1u << 18);
    auto bail = [&](const char *m) {
        sf_server_destroy(sv); // frees what was made so far (no scheduler thread yet)
        return sf_runtime_fail(SF_ENOMEM, m);
    };
    if (!dev_alloc(&sv->d_seq, sv->store_words + cfg->slots * sv->cap_words)) return bail("device sequence buffer");
    if (sv->store_words && hipMemcpy(sv->d_seq, store_words, sv->store_words * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail("store upload");
    sv->slots.resize(cfg->slots);
    for (uint32_t q = 0; q < cfg->slots; ++q) {
        Slot &s = sv->slots[q];
        const uint64_t c = sv->cap_seg;
        s.words_base = q * sv->cap_words;
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess)
            return bail("stream");
        if (!dev_alloc(&s.d_stage, Slot::kStageWords * c) || !dev_alloc(&s.d_res, 4 * c))
            return bail("slot device buffers");
        if (!host_alloc(&s.h_words, sv->cap_words) || !host_alloc(&s.h_stage, Slot::kStageWords * c) ||
            !host_alloc(&s.h_res, 4 * c))
            return bail("slot pinned buffers");
    }
    // The request pool, each request sized for a long read (2,048 words: 32 kb) of 16 segments,
    // its memory touched now rather than during the first submissions.
    sv->pool.reserve(sf_server::kPoolMax);
    for (size_t q = 0; q < sf_server::kPoolPrefill; ++q) {
        Req *r = new Req();
        r->words.resize(2048);
        for (auto *v : {&r->roff, &r->rlen, &r->tlen}) v->resize(16);
        r->toff.resize(16);
        r->reset();
        sv->pool.push_back(r);
    }
    sv->th = std::thread([sv] { sv->loop(); });
    *out = sv;
    return SF_OK;
}

int sf_server_submit(sf_server *sv, const sf_read_request *rq, uint64_t *ticket) {
    if (!sv || !rq || !ticket) return sf_runtime_fail(SF_EINVAL, "null argument");
    if (rq->nseg && (!rq->seg_read_off || !rq->seg_read_len || !rq->seg_tx_off || !rq->seg_tx_len))
        return sf_runtime_fail(SF_EINVAL, "null segment arrays");
    if (rq->read_len && !rq->read) return sf_runtime_fail(SF_EINVAL, "read is null");
    const auto t_enter = Clock::now();
    Req *r = sv->take_req();
    r->roff.assign(rq->seg_read_off, rq->seg_read_off + rq->nseg);
    r->rlen.assign(rq->seg_read_len, rq->seg_read_len + rq->nseg);
    r->toff.assign(rq->seg_tx_off, rq->seg_tx_off + rq->nseg);
    r->tlen.assign(rq->seg_tx_len, rq->seg_tx_len + rq->nseg);
    for (uint32_t k = 0; k < rq->nseg; ++k) {
        const uint32_t m = r->rlen[k], n = r->tlen[k];
        bool bad = (uint64_t)r->roff[k] + m > rq->read_len || r->toff[k] + n > sv->store_bases;
        uint64_t below, above;
        band_geom(m, n, sv->cfg.affine_mode, sv->cfg.band, &below, &above);
        const int K = sf_affine_k_for(below + above + 1);
        const bool shape = sf_myers_maxb_for((m + 63) / 64) && K &&
                           scores_in_range(sv->cfg.sc.mat, sv->cfg.sc.mis, sv->cfg.sc.open, sv->cfg.sc.ext, m, n,
                                           aff_slots(K));
        if (bad || !shape) {
            sv->give_req(r);
            return sf_runtime_fail(bad ? SF_EINVAL : SF_ERANGE, bad ? "segment outside the read or the store"
                                                                    : "segment beyond this library's shape classes");
        }
    }
    r->words.resize((rq->read_len + 15) / 16);
    for (size_t w = 0; w < r->words.size(); ++w) {
        const uint8_t *b = rq->read + 16 * w;
        const uint32_t n = std::min<uint32_t>(16, rq->read_len - 16 * (uint32_t)w);
        uint32_t x = 0;
        for (uint32_t k = 0; k < n; ++k) x |= (uint32_t)(b[k] & 3u) << (2 * k);
        r->words[w] = x;
    }
    if (r->words.size() > sv->cap_words || r->rlen.size() > sv->cap_seg) {
        sv->give_req(r);
        return sf_runtime_fail(SF_ENOMEM, "request exceeds one slot's share of the HBM budget");
    }
    r->dist = rq->dist;
    r->end = rq->end;
    r->score = rq->score;
    r->col = rq->score_col;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        if (sv->trace) {
            const auto t_locked = Clock::now();
            std::lock_guard<std::mutex> tl(sv->trace_mu);
            sv->max_build_ms = std::max(sv->max_build_ms, std::chrono::duration<double, std::milli>(r->t_submit - t_enter).count());
            sv->max_lock_ms = std::max(sv->max_lock_ms, std::chrono::duration<double, std::milli>(t_locked - r->t_submit).count());
            if (sv->t_last_submit != Clock::time_point{}) {
                const double ms = std::chrono::duration<double, std::milli>(r->t_submit - sv->t_last_submit).count();
                if (ms > sv->max_submit_gap_ms)
                    sv->max_submit_gap_ms = ms,
                    sv->max_submit_gap_at = std::chrono::duration<double>(r->t_submit - sv->t_created).count();
            }
            sv->t_last_submit = r->t_submit;
        }
        if (sv->stop) {
            sv->give_req(r);
            return sf_runtime_fail(SF_EINVAL, "server is shutting down");
        }
        r->ticket = sv->next_ticket++;
        r->queued = true;
        sv->reqs[r->ticket] = r;
        sv->pending.push_back(r);
        sv->pend_words += r->words.size();
        sv->pend_seg += r->rlen.size();
        if (!sv->any) {
            sv->any = true;
            sv->t_first = r->t_submit;
        }
        *ticket = r->ticket;
    }
    sv->cv_sched.notify_one();
    return SF_OK;
}

int sf_server_wait(sf_server *sv, uint64_t ticket, double timeout_s) {
    if (!sv) return sf_runtime_fail(SF_EINVAL, "server is null");
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return sf_runtime_fail(SF_EINVAL, "unknown ticket");
    Req *r = it->second;
    auto pred = [&] { return r->done; };
    ++r->waiters;
    bool done = true;
    if (timeout_s < 0) sv->cv_done.wait(lk, pred);
    else done = sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
    --r->waiters;
    if (!done) return sf_runtime_fail(SF_ETIMEDOUT, "timed out; the request is still queued or running");
    const int st = r->status;
    if (r->waiters == 0) { // the last waiter retires the ticket
        sv->reqs.erase(ticket);
        lk.unlock();
        sv->give_req(r);
    }
    if (st == SF_ECANCELED) return sf_runtime_fail(st, "the request was cancelled");
    if (st != SF_OK) return sf_runtime_fail(st, "the request's batch failed");
    return SF_OK;
}

int sf_server_cancel(sf_server *sv, uint64_t ticket) {
    if (!sv) return sf_runtime_fail(SF_EINVAL, "server is null");
    std::lock_guard<std::mutex> ck(sv->copy_mu); // no batch's results are being copied out meanwhile
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return sf_runtime_fail(SF_EINVAL, "unknown ticket");
    Req *r = it->second;
    if (r->done) return r->status == SF_OK ? SF_OK : sf_runtime_fail(r->status, "the request completed before");
    if (r->queued) { // not in a batch yet: it completes now
        sv->pending.erase(std::find(sv->pending.begin(), sv->pending.end(), r));
        sv->pend_words -= r->words.size();
        sv->pend_seg -= r->rlen.size();
        r->queued = false;
        r->status = SF_ECANCELED;
        r->t_done = Clock::now();
        r->done = true;
        sv->cv_done.notify_all();
    } else { // in a batch: it finishes on the device, and its results are dropped
        r->cancelled = true;
    }
    return SF_OK;
}

int sf_server_get_stats(sf_server *sv, sf_server_stats *out, int reset) {
    if (!sv || !out) return sf_runtime_fail(SF_EINVAL, "null argument");
    std::vector<double> lat; // copied under the lock, ranked outside it
    std::unique_lock<std::mutex> lk(sv->mu);
    std::memset(out, 0, sizeof *out);
    out->reads_done = sv->reads_done;
    out->reads_failed = sv->reads_failed;
    out->batches = sv->batches;
    out->segments = sv->segments;
    out->cells = sv->cells;
    out->mean_reads_per_batch = sv->batches ? (double)sv->batch_reads / (double)sv->batches : 0;
    lat = sv->lat_ms;
    out->span_s = sv->any ? std::chrono::duration<double>(sv->t_last - sv->t_first).count() : 0;
    out->cap_segments_per_batch = sv->cap_seg;
    out->cap_read_bases_per_batch = sv->cap_words * 16;
    out->max_segments_per_batch = sv->max_batch_seg;
    if (reset) {
        sv->lat_ms.clear();
        sv->reads_done = sv->reads_failed = sv->batches = sv->segments = sv->batch_reads = sv->cells = 0;
        sv->max_batch_seg = 0;
        sv->any = false;
    }
    lk.unlock();
    out->p50_ms = pct(lat, 0.50);
    out->p99_ms = pct(lat, 0.99);
    out->max_ms = lat.empty() ? 0 : *std::max_element(lat.begin(), lat.end());
    if (reset) {
        if (sv->trace) {
            std::lock_guard<std::mutex> tl(sv->trace_mu);
            if (sv->max_iter_ms > 0)
                std::fprintf(stderr, "sf_server: longest scheduler iteration %.3f ms (ended %.3f s), launch %.3f ms "
                                     "(%.3f s), gap between submissions %.3f ms (%.3f s), request build %.3f ms, "
                                     "lock wait %.3f ms\n",
                             sv->max_iter_ms, sv->max_iter_at, sv->max_launch_ms, sv->max_launch_at,
                             sv->max_submit_gap_ms, sv->max_submit_gap_at, sv->max_build_ms, sv->max_lock_ms);
            sv->max_iter_ms = sv->max_launch_ms = sv->max_submit_gap_ms = 0;
            sv->max_build_ms = sv->max_lock_ms = 0;
            sv->t_last_submit = Clock::time_point{};
        }
    }
    return SF_OK;
}

void sf_server_destroy(sf_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_sched.notify_all();
    if (sv->th.joinable()) sv->th.join();
    if (sv->trace)
        std::fprintf(stderr,
                     "sf_server: longest scheduler iteration %.3f ms (ended %.3f s), launch %.3f ms (%.3f s), "
                     "gap between submissions %.3f ms (%.3f s)\n",
                     sv->max_iter_ms, sv->max_iter_at, sv->max_launch_ms, sv->max_launch_at, sv->max_submit_gap_ms,
                     sv->max_submit_gap_at);
    for (auto &kv : sv->reqs) delete kv.second;
    for (Req *r : sv->pool) delete r;
    for (Slot &s : sv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        (void)hipFree(s.d_stage), (void)hipFree(s.d_res);
        (void)hipHostFree(s.h_words), (void)hipHostFree(s.h_stage), (void)hipHostFree(s.h_res);
        if (s.ev) (void)hipEventDestroy(s.ev);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    (void)hipFree(sv->d_seq);
    delete sv;
}

// Replays a request stream: request i is submitted at t_arrive[i] seconds after the call starts
// (a submitter thread sleeps until each arrival), then every request is waited for.  Returns the
// wall time from the first arrival to the last completion in *wall_s and the number of requests
// refused at admission in *refused.  Latency statistics are the server's own (sf_server_get_stats).
int sf_server_replay(sf_server *sv, const sf_read_request *rq, const double *t_arrive, uint64_t n, double *wall_s,
                     uint64_t *refused) {
    if (!sv || (n && (!rq || !t_arrive))) return sf_runtime_fail(SF_EINVAL, "null argument");
    std::vector<uint64_t> tickets(n, 0);
    std::vector<int> sub(n, SF_OK);
    // The waits run beside the submissions (each request, once completed, goes back to the pool
    // for a later submission to reuse), on the calling thread.
    std::mutex m;
    std::condition_variable cv;
    uint64_t submitted = 0;
    const auto t0 = Clock::now();
    std::thread submitter([&] {
        for (uint64_t i = 0; i < n; ++i) {
            const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t_arrive[i]));
            std::this_thread::sleep_until(due);
            sub[i] = sf_server_submit(sv, &rq[i], &tickets[i]);
            {
                std::lock_guard<std::mutex> lk(m);
                submitted = i + 1;
            }
            cv.notify_one();
        }
    });
    int status = SF_OK;
    uint64_t nref = 0;
    for (uint64_t i = 0; i < n; ++i) {
        {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return submitted > i; });
        }
        if (sub[i] != SF_OK) {
            ++nref; // refused at admission: counted, not fatal
            continue;
        }
        const int st = sf_server_wait(sv, tickets[i], -1.0);
        if (st != SF_OK) status = st;
    }
    submitter.join();
    if (refused) *refused = nref;
    const auto t1 = Clock::now();
    if (wall_s) *wall_s = std::chrono::duration<double>(t1 - t0).count() - (n ? t_arrive[0] : 0.0);
    return status;
}

} // extern "C"
