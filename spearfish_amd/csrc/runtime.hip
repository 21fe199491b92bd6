// runtime.hip -- the C ABI of libsf_align.so (include/sf_align.h): device selection, batches
// resident in HBM, shape binning, traceback scratch, per-kernel event timing, errors.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/sf_align.h"
#include "internal.hpp"

using namespace sf;

// The autotuner's table (tune.hip).
bool sf_tune_cost(int cls, double *tput_ns, double *chain_ns);
int sf_tune_measure(const std::vector<int> &classes, int mode, const sf_scoring *sc);
void sf_tune_set(bool use, double chain_factor);
uint64_t sf_tune_generation();

// ------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------
namespace {
thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(SF_EDEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

std::mutex g_dev_mu;
int g_device = -1;

int ensure_device() {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_device >= 0) {
        HIP_TRY(hipSetDevice(g_device));
        return SF_OK;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(SF_ENODEV, "no HIP device visible");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, 0));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SF_ENODEV, "device 0 is %s, this library is built for gfx950", prop.gcnArchName);
    HIP_TRY(hipSetDevice(0));
    g_device = 0;
    return SF_OK;
}

// -1: not read yet (SF_PERSISTENT from the environment, default off).
std::atomic<int> g_persistent{-1};
std::atomic<int> g_myers_ukk{-1}; // -1: not read from the environment yet

unsigned host_threads() {
    unsigned t = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("SF_HOST_THREADS")) t = (unsigned)std::atoi(e);
    return std::max(1u, std::min(t ? t : 1u, 16u));
}

template <class F>
void parallel_for(uint64_t n, F f) {
    unsigned nt = host_threads();
    if (n < 4096 || nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    uint64_t chunk = (n + nt - 1) / nt;
    for (unsigned q = 0; q < nt; ++q) {
        uint64_t a = q * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([=] { f(a, b); });
    }
    for (auto &t : th) t.join();
}

template <class T>
int dalloc(T **p, uint64_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(SF_ENOMEM, "hipMalloc of %llu bytes failed: %s", (unsigned long long)(count * sizeof(T)),
                    hipGetErrorString(e));
    }
    return SF_OK;
}

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Streams are recycled across batches of this process's device: making one costs about 0.4 ms,
// and every one-shot call makes a batch.  A stream comes back with no work queued (its batch waited
// for it); beyond 64 idle ones they are destroyed.
std::mutex g_stream_mu;
std::vector<hipStream_t> g_streams;
hipError_t take_stream(hipStream_t *s) {
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        if (!g_streams.empty()) {
            *s = g_streams.back();
            g_streams.pop_back();
            return hipSuccess;
        }
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}
void give_stream(hipStream_t s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        if (g_streams.size() < 64) {
            g_streams.push_back(s);
            return;
        }
    }
    (void)hipStreamDestroy(s);
}

// Words of one traceback scratch buffer (two are used, alternately): SF_TB_BUDGET_MB, else a
// quarter of the device's memory, at most 64 GiB.  Larger chunks run more walks at once, and the
// walk (one dependent chain per pair) needs that concurrency.
// A launch plan's cache key over a per-pair array (null: every pair takes `absent`): eight
// independent multiply-xor chains, folded.  One serial chain costs a multiply's latency an
// element, which made the key of a 10^5-pair batch a third of a millisecond of host time a call.
uint64_t hash_words(uint64_t key, const uint64_t *v, uint64_t n, uint64_t absent) {
    constexpr uint64_t kP = 1099511628211ull;
    if (!v) return ((key ^ absent) * kP ^ n) * kP;
    uint64_t h[8];
    for (int k = 0; k < 8; ++k) h[k] = key + 0x9e3779b97f4a7c15ull * (uint64_t)(k + 1);
    uint64_t i = 0;
    for (; i + 8 <= n; i += 8)
        for (int k = 0; k < 8; ++k) h[k] = (h[k] ^ v[i + k]) * kP;
    for (; i < n; ++i) h[i & 7] = (h[i & 7] ^ v[i]) * kP;
    uint64_t r = key ^ n;
    for (int k = 0; k < 8; ++k) r = (r ^ h[k]) * kP;
    return r;
}

uint64_t tb_budget_words() {
    uint64_t mb = 0;
    if (const char *e = std::getenv("SF_TB_BUDGET_MB")) mb = std::strtoull(e, nullptr, 10);
    if (mb == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) tot = 16ull << 30;
        mb = std::min<uint64_t>(65536, tot / 4 / (1ull << 20));
    }
    return std::max<uint64_t>(mb, 16) * (1ull << 20) / 4;
}

// Exclusive prefix sum of n u32 into n + 1 u64 (out[n] = the total), in tiles of 4096 values (1024
// threads x 4): tile sums, their scan in one block, then each tile's own scan plus its offset.
constexpr int kScanTile = 4096;

// Exclusive scan of one u64 per thread over a block of 1024 threads; `total` gets the block's sum.
__device__ uint64_t block_excl_scan(uint64_t v, uint64_t &total) {
    __shared__ uint64_t wsum[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    uint64_t pre = 0, all = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const uint64_t t = wsum[k];
        pre += k < wv ? t : 0;
        all += t;
    }
    __syncthreads(); // wsum is reused by the next call
    total = all;
    return pre + x - v;
}

__device__ uint64_t tile_values(const uint32_t *in, uint64_t n, uint64_t base, uint32_t (&v)[4]) {
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t i = base + 4 * (uint64_t)threadIdx.x + k;
        v[k] = i < n ? in[i] : 0u;
        s += v[k];
    }
    return s;
}

__global__ __launch_bounds__(1024) void scan_sums_kernel(const uint32_t *in, uint64_t n, uint64_t *sums) {
    uint32_t v[4];
    uint64_t total;
    block_excl_scan(tile_values(in, n, (uint64_t)blockIdx.x * kScanTile, v), total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void scan_offsets_kernel(uint64_t *sums, uint64_t tiles, uint64_t *out_total) {
    uint64_t carry = 0;
    for (uint64_t c = 0; c < tiles; c += 1024) {
        const uint64_t i = c + threadIdx.x;
        const uint64_t v = i < tiles ? sums[i] : 0;
        uint64_t total;
        const uint64_t e = block_excl_scan(v, total);
        if (i < tiles) sums[i] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0) *out_total = carry;
}

__global__ __launch_bounds__(1024) void scan_apply_kernel(const uint32_t *in, uint64_t n, const uint64_t *offs,
                                                          uint64_t *out) {
    uint32_t v[4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint64_t total;
    uint64_t acc = offs[blockIdx.x] + block_excl_scan(tile_values(in, n, base, v), total);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint64_t i = base + 4 * (uint64_t)threadIdx.x + k;
        if (i < n) out[i] = acc;
        acc += v[k];
    }
}

hipError_t launch_scan(const uint32_t *in, uint64_t n, uint64_t *out, uint64_t *tmp, hipStream_t s) {
    const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles) hipLaunchKernelGGL(scan_sums_kernel, dim3((unsigned)tiles), dim3(1024), 0, s, in, n, tmp);
    hipLaunchKernelGGL(scan_offsets_kernel, dim3(1), dim3(1024), 0, s, tmp, tiles, out + n);
    if (tiles) hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)tiles), dim3(1024), 0, s, in, n, tmp, out);
    return hipGetLastError();
}

} // namespace

// ------------------------------------------------------------------------------------------
// batch
// ------------------------------------------------------------------------------------------
struct sf_batch {
    hipStream_t stream = nullptr;
    uint64_t n = 0;
    std::vector<uint32_t> plen, tlen;
    uint32_t *d_seq = nullptr;
    bool seq_borrowed = false; // d_seq belongs to the caller (sf_batch_create_resident)
    uint64_t *d_poff = nullptr, *d_toff = nullptr;
    uint32_t *d_plen = nullptr, *d_tlen = nullptr;
    // The texts as made, while the working texts may be restricted to located windows.
    uint64_t *d_toff0 = nullptr;
    uint32_t *d_tlen0 = nullptr;
    std::vector<uint32_t> tlen0;
    bool located_ready = false; // the last Myers run produced start and end
    // Myers
    struct Bin {
        int cls;
        uint32_t begin, end;
        // Estimated time: the longest single wave's chain, and the whole bin's work spread over
        // the chip (affine bins: ns, measured or estimated per class; Myers bins: block steps).
        // Orders the launches and places them on streams.
        double chain = 0, spread = 0;
        double est() const { return chain > spread ? chain : spread; }
    };
    std::vector<Bin> mbins;
    uint32_t *d_morder = nullptr;
    PairDesc *d_mdesc = nullptr; // pairs in Myers launch order
    uint32_t *d_dist = nullptr, *d_end = nullptr, *d_start = nullptr;
    std::string myers_unsupported;
    // Patterns over 1,536 blocks (kMyersStriped): two rows of column deltas per such pair, and the
    // most stripes any of them has.
    uint8_t *d_hcol = nullptr;
    uint64_t hcol_stride = 0, hcol_pairs = 0;
    uint32_t stripes_max = 0;
    // The start pass of patterns over 1,536 bases (the launch order's suffix from start_first):
    // its stripes' bottom-row deltas (2 bits a column, ping-pong), made on first use.
    std::vector<uint32_t> morder;     // host copy of the Myers launch order
    uint64_t start_first = 0;         // first launch slot of a pattern over 1,536 bases (n: none)
    uint64_t start_words = 0;         // words a row of start_hc
    uint32_t *d_start_hc = nullptr;
    // affine
    // Affine launch plans, cached per (mode, bands, shape tuning): one for score-only runs (which
    // may use the 16-bit kernel), one for runs with traceback.
    struct AffPlan {
        uint64_t key = 0;
        bool valid = false;
        std::vector<Bin> bins;
        std::vector<uint32_t> order; // host copy of the launch order
        uint64_t *d_ba = nullptr;    // per pair: below | (above << 32)
        uint32_t *d_order = nullptr;
        PairDesc *d_desc = nullptr;  // pairs in launch order
        uint64_t cells = 0;
        uint64_t wide_stride = 0;    // int32 a scratch row of the wide-band pairs (kAffWide), 0 if none
    };
    AffPlan aplan[4]; // scores only, with traceback, passes with bookkeeping, local bands
    DpOut *d_dp = nullptr;      // per pair: the last sf_batch_affine_dp run
    uint64_t *d_cut = nullptr;  // the cuts of the last sf_batch_affine_dp run that had them
    uint64_t cut_key = 0;       // their hash (0: none uploaded)
    int32_t *d_best = nullptr;  // per pair: the last sf_batch_local_band run
    uint32_t *d_pinv = nullptr; // bit (base offset) set where a pattern code is >= 4; null if none
    uint64_t pinv_words = 0;
    int32_t *d_score = nullptr;
    uint32_t *d_col = nullptr, *d_tstart = nullptr, *d_nruns = nullptr, *d_runs = nullptr, *d_cig = nullptr;
    uint64_t *d_runs_off = nullptr, *d_out_off = nullptr, *d_tboff = nullptr;
    uint64_t *d_scan = nullptr; // tile sums of the CIGAR offset scan
    uint32_t *d_tb[2] = {nullptr, nullptr}; // traceback scratch, alternating between chunks
    int32_t *d_wide = nullptr; // rows of the wide-band passes (kAffWide)
    uint64_t wide_words = 0;
    uint64_t tb_words = 0;
    hipStream_t wstream = nullptr;           // traceback walks, overlapping the next chunk's DP
    // Extra streams on which independent bins run beside the main stream's.
    // Two: with the main stream that is three, one a hardware queue (a process gets four and one is
    // taken); a third would share a queue with another and only serialise behind it.
    static constexpr int kAux = 2;
    hipStream_t aux[kAux] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, joins[kAux] = {nullptr, nullptr};
    // Extra streams whose work the main stream has not waited for (bit t: aux[t], its last work
    // marked by joins[t]).  A Myers run leaves them open so that a following independent run
    // (affine scores) starts beside its last bins; every other call joins them first (join_aux).
    uint32_t aux_open = 0;
    // Work queues of persistent mixed launches (MixArgs::queue): kQueues pairs of words, zero
    // between launches.  Slots [0, 8) serve a Myers run's launches, [8, 16) a score run's grouped
    // launches, 16 its mixed 64-lane launch: launches that may run at the same time never share one.
    static constexpr int kQueues = 17;
    uint32_t *d_queues = nullptr;
    // Graph capture (sf_batch_capture_begin/end) and the captured step, replayed by sf_batch_replay.
    bool capturing = false;
    int capture_error = SF_OK;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_myers_cells = 0, graph_affine_cells = 0, cap_myers_cells = 0, cap_affine_cells = 0;
    uint64_t graph_replays = 0;
    hipEvent_t dp_done[2] = {nullptr, nullptr}, walk_done[2] = {nullptr, nullptr}, walks_end = nullptr;
    bool cigar_valid = false;
    // timing
    struct Ev { // an interval from a to the latest of its ends (more: other streams' ends, or null)
        int cat;
        hipEvent_t a, b;
        hipEvent_t more[kAux];
    };
    std::vector<Ev> pending;
    std::vector<hipEvent_t> pool;
    double ms[5] = {0, 0, 0, 0, 0}; // Myers, start pass, affine, walk, graph replays
    uint64_t myers_cells = 0, affine_cells = 0, affine_launches = 0;
    uint64_t myers_cells_per_run = 0;

    hipEvent_t ev() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void collect() {
        for (auto &e : pending) {
            float t = 0, u = 0;
            if (hipEventElapsedTime(&t, e.a, e.b) != hipSuccess) t = 0;
            for (hipEvent_t m : e.more)
                if (m) {
                    if (hipEventElapsedTime(&u, e.a, m) == hipSuccess) t = std::max(t, u);
                    pool.push_back(m);
                }
            ms[e.cat] += t;
            pool.push_back(e.a);
            pool.push_back(e.b);
        }
        pending.clear();
    }
    ~sf_batch() {
        for (hipStream_t x : aux)
            if (x) (void)hipStreamSynchronize(x);
        if (wstream) (void)hipStreamSynchronize(wstream);
        if (stream) (void)hipStreamSynchronize(stream);
        collect();
        for (auto e : pool) (void)hipEventDestroy(e);
        if (seq_borrowed) d_seq = nullptr;
        dfree(d_seq), dfree(d_poff), dfree(d_toff), dfree(d_plen), dfree(d_tlen), dfree(d_morder);
        dfree(d_dist), dfree(d_end), dfree(d_start), dfree(aplan[0].d_ba), dfree(aplan[0].d_order), dfree(aplan[1].d_ba), dfree(aplan[1].d_order), dfree(d_score);
        dfree(aplan[0].d_desc), dfree(aplan[1].d_desc), dfree(d_mdesc);
        for (int q = 2; q < 4; ++q) dfree(aplan[q].d_ba), dfree(aplan[q].d_order), dfree(aplan[q].d_desc);
        dfree(d_dp), dfree(d_best), dfree(d_pinv), dfree(d_col), dfree(d_cut);
        dfree(d_tstart), dfree(d_nruns), dfree(d_runs), dfree(d_cig), dfree(d_runs_off), dfree(d_out_off);
        dfree(d_scan);
        dfree(d_tboff), dfree(d_tb[0]), dfree(d_tb[1]), dfree(d_toff0), dfree(d_tlen0), dfree(d_queues), dfree(d_wide), dfree(d_hcol),
            dfree(d_start_hc);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipStream_t x : aux)
            if (x) (void)hipStreamSynchronize(x);
        for (hipEvent_t e : {dp_done[0], dp_done[1], walk_done[0], walk_done[1], walks_end, fork, joins[0], joins[1]})
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t x : aux) give_stream(x);
        give_stream(wstream);
        give_stream(stream);
    }
};

namespace {

PairsView view(const sf_batch *b) { return PairsView{b->d_seq, b->d_poff, b->d_toff, b->d_plen, b->d_tlen}; }

int check_pairs(const sf_pairs *p) {
    if (!p) return fail(SF_EINVAL, "pairs is null");
    if (p->n == 0) return SF_OK;
    if (!p->pat || !p->txt || !p->pat_off || !p->pat_len || !p->txt_off || !p->txt_len)
        return fail(SF_EINVAL, "null sequence or offset pointer");
    if (p->packed != 0 && p->packed != 1) return fail(SF_EINVAL, "packed must be 0 or 1");
    for (uint64_t i = 0; i < p->n; ++i) {
        if (p->pat_off[i] + p->pat_len[i] > p->pat_bases || p->pat_off[i] + p->pat_len[i] < p->pat_off[i])
            return fail(SF_EINVAL, "pattern %llu lies outside the pattern buffer", (unsigned long long)i);
        if (p->txt_off[i] + p->txt_len[i] > p->txt_bases || p->txt_off[i] + p->txt_len[i] < p->txt_off[i])
            return fail(SF_EINVAL, "text %llu lies outside the text buffer", (unsigned long long)i);
    }
    return SF_OK;
}

inline uint32_t code_at(const void *buf, int packed, uint64_t k) {
    if (!packed) return static_cast<const uint8_t *>(buf)[k] & 3u;
    return (static_cast<const uint32_t *>(buf)[k >> 4] >> (2 * (k & 15))) & 3u;
}

// The low two bits of 8 bytes, compacted into 16 bits (byte k's bits at 2k).
inline uint32_t pack8(uint64_t x) {
    x &= 0x0303030303030303ull;
    x = (x | (x >> 6)) & 0x000F000F000F000Full;
    x = (x | (x >> 12)) & 0x000000FF000000FFull;
    return (uint32_t)((x | (x >> 24)) & 0xFFFFull);
}

void pack_into(uint32_t *dst, const void *buf, int packed, uint64_t off, uint32_t len) {
    const uint64_t nw = (len + 15) / 16;
    if (!packed) {
        // One code a byte: 16 bases a word, 8 bytes at a time.
        const uint8_t *src = static_cast<const uint8_t *>(buf) + off;
        const uint64_t full = len / 16;
        for (uint64_t w = 0; w < full; ++w) {
            uint64_t lo, hi;
            std::memcpy(&lo, src + 16 * w, 8);
            std::memcpy(&hi, src + 16 * w + 8, 8);
            dst[w] = pack8(lo) | (pack8(hi) << 16);
        }
        if (full < nw) {
            uint32_t v = 0;
            for (uint32_t k = 0; k < len - 16 * full; ++k) v |= (src[16 * full + k] & 3u) << (2 * k);
            dst[full] = v;
        }
        return;
    }
    for (uint64_t w = 0; w < nw; ++w) {
        uint32_t v = 0;
        const uint64_t k0 = w * 16;
        const uint32_t cnt = (uint32_t)std::min<uint64_t>(16, len - k0);
        for (uint32_t k = 0; k < cnt; ++k) v |= code_at(buf, packed, off + k0 + k) << (2 * k);
        dst[w] = v;
    }
}

// (While a step is being captured into a graph nothing is timed: the replays are.)
int timed_begin(sf_batch *b, int cat, hipEvent_t *a, hipStream_t s = nullptr) {
    if (b->capturing) {
        *a = nullptr;
        return SF_OK;
    }
    *a = b->ev();
    HIP_TRY(hipEventRecord(*a, s ? s : b->stream));
    (void)cat;
    return SF_OK;
}

int timed_end(sf_batch *b, int cat, hipEvent_t a, hipStream_t s = nullptr) {
    if (b->capturing) return SF_OK;
    hipEvent_t e = b->ev();
    HIP_TRY(hipEventRecord(e, s ? s : b->stream));
    b->pending.push_back({cat, a, e, {}});
    return SF_OK;
}

// Makes the main stream wait for the extra streams' open work.
int join_aux(sf_batch *b) {
    for (int t = 0; t < sf_batch::kAux; ++t)
        if (b->aux_open & (1u << t)) HIP_TRY(hipStreamWaitEvent(b->stream, b->joins[t], 0));
    b->aux_open = 0;
    return SF_OK;
}

// timed_end for a run that left extra streams open: an end on the main stream and one on each open
// stream, the interval running to the latest (no stream waits for another).
int timed_end_open(sf_batch *b, int cat, hipEvent_t a) {
    if (b->capturing) return SF_OK;
    if (const int st = timed_end(b, cat, a)) return st;
    sf_batch::Ev &ev = b->pending.back();
    for (int t = 0; t < sf_batch::kAux; ++t)
        if (b->aux_open & (1u << t)) {
            ev.more[t] = b->ev();
            HIP_TRY(hipEventRecord(ev.more[t], b->aux[t]));
        }
    return SF_OK;
}

__global__ void restrict_kernel(uint64_t n, const uint32_t *dist, const uint32_t *start, const uint32_t *end,
                                const uint64_t *toff0, const uint32_t *tlen0, uint64_t *toff, uint32_t *tlen) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (dist[i] == kNone) {
        toff[i] = toff0[i];
        tlen[i] = tlen0[i];
    } else {
        toff[i] = toff0[i] + start[i];
        tlen[i] = end[i] - start[i];
    }
}

// Runs the bins on the main stream and the extra streams.  A bin whose work fills the chip
// (spread >= chain) runs alone well, so those go one after another on the main stream; the bins
// bound by their longest chain (few waves, long sequences) go beside them on the extra streams,
// longest first, each to the stream with the least estimated time so far, so that none of them
// holds up a big bin.  With `mix`, the chain-bound bins whose classes `mixable` accepts share one
// launch (they then run side by side instead of one after another).  Without big bins every
// stream takes chain-bound ones.  The main stream then waits for every stream used.  Adds the
// number of launches to *launches.
using MixFn = std::function<hipError_t(const std::vector<size_t> &, hipStream_t)>;
// The extra streams (and their fork and join events), made on first use.
int ensure_aux(sf_batch *b) {
    if (b->fork) return SF_OK;
    if (b->capturing) return fail(SF_EINVAL, "the batch's streams are made by a run before the capture");
    HIP_TRY(hipMalloc((void **)&b->d_queues, sf_batch::kQueues * 2 * sizeof(uint32_t)));
    HIP_TRY(hipMemset(b->d_queues, 0, sf_batch::kQueues * 2 * sizeof(uint32_t)));
    HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (int i = 0; i < sf_batch::kAux; ++i) {
        HIP_TRY(take_stream(&b->aux[i]));
        HIP_TRY(hipEventCreateWithFlags(&b->joins[i], hipEventDisableTiming));
    }
    return SF_OK;
}

// The queue of a persistent mixed launch (slot `q` of the batch's queues: Myers below 8), or null
// for a launch of one workgroup a block index (the policy's bit for the run's kind clear, or no free
// slot).
uint32_t *queue_for(sf_batch *b, int q, int end) {
    const int bit = q < 8 ? 1 : 2;
    return ((sf_persistent_policy() & bit) && b->d_queues && q < end) ? b->d_queues + 2 * q : nullptr;
}

template <class F>
int run_bins(sf_batch *b, const std::vector<sf_batch::Bin> &bins, F launch, uint64_t *launches = nullptr,
             const MixFn &mix = nullptr, bool (*mixable)(int) = nullptr, bool spread_big = false) {
    constexpr int S = 1 + sf_batch::kAux;
    if (bins.size() <= 1) {
        for (auto &bin : bins) HIP_TRY(launch(bin, b->stream));
        if (launches) *launches += bins.size();
        return SF_OK;
    }
    if (const int st = ensure_aux(b)) return st;
    std::vector<size_t> big;
    std::vector<std::vector<size_t>> small; // launch units: one bin, or the mixed bins together
    std::vector<size_t> mixed;
    for (size_t i = 0; i < bins.size(); ++i) {
        if (bins[i].spread >= bins[i].chain) big.push_back(i);
        else if (mix && mixable && mixable(bins[i].cls) && mixed.size() < MixArgs::kMax) mixed.push_back(i);
        else small.push_back({i});
    }
    if (!mixed.empty()) small.push_back(mixed);
    auto est = [&](const std::vector<size_t> &u) {
        double c = 0, sp = 0;
        for (size_t i : u) {
            c = std::max(c, bins[i].chain);
            sp += bins[i].spread;
        }
        return std::max(c, sp);
    };
    std::stable_sort(small.begin(), small.end(),
                     [&](const std::vector<size_t> &x, const std::vector<size_t> &y) { return est(x) > est(y); });
    hipStream_t ss[S] = {b->stream, b->aux[0], b->aux[1]};
    double load[S] = {0, 0, 0};
    bool used[S] = {true, false, false};
    HIP_TRY(hipEventRecord(b->fork, b->stream));
    if (spread_big) {
        // Big bins too go to the least-loaded stream, the longest first, so that one bin's tail
        // overlaps the next one's start (passes whose classes each fill the chip a few times over).
        std::stable_sort(big.begin(), big.end(), [&](size_t x, size_t y) { return bins[x].est() > bins[y].est(); });
        for (size_t i : big) {
            int s = 0;
            for (int t = 1; t < S; ++t)
                if (load[t] < load[s]) s = t;
            if (!used[s]) {
                HIP_TRY(hipStreamWaitEvent(ss[s], b->fork, 0));
                used[s] = true;
            }
            HIP_TRY(launch(bins[i], ss[s]));
            load[s] += bins[i].est();
        }
        big.clear();
    }
    for (size_t i : big) {
        HIP_TRY(launch(bins[i], b->stream));
        load[0] += bins[i].est();
    }
    const int first = big.empty() ? 0 : 1;
    for (const auto &u : small) {
        int s = first;
        for (int t = first + 1; t < S; ++t)
            if (load[t] < load[s]) s = t;
        if (!used[s]) {
            HIP_TRY(hipStreamWaitEvent(ss[s], b->fork, 0));
            used[s] = true;
        }
        if (u.size() == 1) HIP_TRY(launch(bins[u[0]], ss[s]));
        else HIP_TRY(mix(u, ss[s]));
        load[s] += est(u);
    }
    if (launches) *launches += big.size() + small.size();
    for (int t = 1; t < S; ++t)
        if (used[t]) {
            HIP_TRY(hipEventRecord(b->joins[t - 1], ss[t]));
            HIP_TRY(hipStreamWaitEvent(b->stream, b->joins[t - 1], 0));
            b->aux_open &= ~(1u << (t - 1));
        }
    return SF_OK;
}

// Runs a run's bins grouped by register footprint.  The bins whose classes share a register group
// (group_of > 0; myers_mix.hip, affine_mix.hip, affine_mix1.hip) run as one launch per group, the
// longest chain's workgroups first, so the chip refills as waves end instead of draining at each
// bin's tail; the launches (and bins of classes outside the groups, or alone in theirs, on their
// own kernels) go to the main and the two extra streams side by side, the most work on the main
// stream.  A single bin is one plain launch.
template <class One, class Mix>
int run_grouped(sf_batch *b, const std::vector<sf_batch::Bin> &bins, One one, int (*group_of)(int),
                uint32_t (*blocks_for)(int, uint32_t), Mix mix, uint64_t *launches, bool leave_open = false,
                int queue0 = 0) {
    if (bins.size() <= 1) {
        for (const auto &bin : bins) HIP_TRY(one(bin, b->stream));
        if (launches) *launches += bins.size();
        return SF_OK;
    }
    struct Unit {
        int group;
        std::vector<size_t> ids;
        double est;
    };
    std::vector<Unit> units;
    std::vector<std::vector<size_t>> grp;
    for (size_t i = 0; i < bins.size(); ++i) {
        const int g = group_of(bins[i].cls);
        if ((size_t)g >= grp.size()) grp.resize(g + 1);
        grp[g].push_back(i);
    }
    for (size_t g = 1; g < grp.size(); ++g) {
        // A bin holding nearly all of its group's work (a batch of one dominant length class) that
        // fills the chip alone launches on its own kernel, and the group's other bins (its few
        // chain-bound stragglers) beside it: they then end on an extra stream, which a following
        // independent run does not wait for.
        if (grp[g].size() >= 2) {
            double total = 0;
            size_t top = grp[g][0];
            for (size_t i : grp[g]) {
                total += bins[i].spread;
                if (bins[i].spread > bins[top].spread) top = i;
            }
            if (bins[top].spread >= 0.9 * total && bins[top].spread >= bins[top].chain) {
                units.push_back(Unit{0, {top}, bins[top].est()});
                grp[g].erase(std::find(grp[g].begin(), grp[g].end(), top));
            }
        }
        // A group's launch holds at most MixArgs::kMax bins; more run as further launches.
        for (size_t q = 0; q < grp[g].size(); q += MixArgs::kMax) {
            Unit u{0, {}, 0};
            u.ids.assign(grp[g].begin() + q, grp[g].begin() + std::min(grp[g].size(), q + MixArgs::kMax));
            if (u.ids.size() > 1) u.group = (int)g;
            double chain = 0, spread = 0;
            for (size_t i : u.ids) chain = std::max(chain, bins[i].chain), spread += bins[i].spread;
            u.est = std::max(chain, spread);
            if (u.group == 0) {
                for (size_t i : u.ids) units.push_back(Unit{0, {i}, bins[i].est()});
            } else {
                units.push_back(std::move(u));
            }
        }
    }
    if (!grp.empty())
        for (size_t i : grp[0]) units.push_back(Unit{0, {i}, bins[i].est()});
    std::stable_sort(units.begin(), units.end(), [](const Unit &x, const Unit &y) { return x.est > y.est; });
    if (const int st = ensure_aux(b)) return st;
    constexpr int S = 1 + sf_batch::kAux;
    hipStream_t ss[S] = {b->stream, b->aux[0], b->aux[1]};
    double load[S] = {0, 0, 0};
    bool used[S] = {true, false, false};
    int queue = queue0;
    HIP_TRY(hipEventRecord(b->fork, b->stream));
    for (const Unit &u : units) {
        int s = 0;
        for (int t = 1; t < S; ++t)
            if (load[t] < load[s]) s = t;
        if (!used[s]) {
            HIP_TRY(hipStreamWaitEvent(ss[s], b->fork, 0));
            used[s] = true;
        }
        if (u.group == 0) {
            HIP_TRY(one(bins[u.ids[0]], ss[s]));
        } else {
            std::vector<size_t> o(u.ids);
            std::stable_sort(o.begin(), o.end(), [&](size_t x, size_t y) { return bins[x].chain > bins[y].chain; });
            MixArgs mx{};
            for (size_t i : o) {
                const sf_batch::Bin &bin = bins[i];
                mx.cls[mx.nseg] = bin.cls;
                mx.off[mx.nseg] = bin.begin;
                mx.count[mx.nseg] = bin.end - bin.begin;
                mx.blk0[mx.nseg] = mx.blocks;
                mx.blocks += blocks_for(bin.cls, bin.end - bin.begin);
                ++mx.nseg;
            }
            mx.queue = queue_for(b, queue++, queue0 + 8);
            HIP_TRY(mix(mx, u.group, ss[s]));
        }
        load[s] += u.est;
    }
    if (launches) *launches += units.size();
    for (int t = 1; t < S; ++t)
        if (used[t]) {
            HIP_TRY(hipEventRecord(b->joins[t - 1], ss[t]));
            if (leave_open) {
                b->aux_open |= 1u << (t - 1);
            } else {
                HIP_TRY(hipStreamWaitEvent(b->stream, b->joins[t - 1], 0));
                b->aux_open &= ~(1u << (t - 1));
            }
        }
    return SF_OK;
}

int myers_ukk_policy() {
    int v = g_myers_ukk.load();
    if (v < 0) {
        const char *e = std::getenv("SF_MYERS_UKK");
        v = e ? (std::atoi(e) != 0) : 1;
        int expected = -1;
        if (!g_myers_ukk.compare_exchange_strong(expected, v)) v = expected;
    }
    return v;
}

// Whether a lane class runs over the wave's Ukkonen band: the reversed pass (its threshold is each
// pair's distance) and passes whose k leaves at least one of the class's blocks out at the start.
bool myers_use_ukk(const MyersArgs &a, int cls) {
    if (!sf_myers_ukk_class(cls) || !myers_ukk_policy()) return false;
    return a.reverse != 0 || a.kvec || (uint64_t)a.k + 1 <= 64ull * (uint64_t)(cls - 1);
}

// The mix groups for passes whose lane classes run over the Ukkonen band: those launch alone.
int myers_group_ukk(int cls) { return sf_myers_ukk_class(cls) ? 0 : sf_myers_mix_group(cls); }

// Myers bins, grouped (myers_mix.hip).
int run_myers_bins(sf_batch *b, const MyersArgs &a, const std::vector<sf_batch::Bin> &bins) {
    auto one = [&](const sf_batch::Bin &bin, hipStream_t s) {
        MyersArgs x = a;
        x.order = b->d_morder + bin.begin;
        x.desc = b->d_mdesc + bin.begin;
        x.count = bin.end - bin.begin;
        if (myers_use_ukk(a, bin.cls)) return sf_launch_myers_ukk(x, bin.cls, s);
        if (bin.cls != kMyersStriped) return sf_launch_myers(x, bin.cls, s);
        // The stripes in order, on one stream: stripe t reads the deltas stripe t - 1 wrote.
        const uint64_t half = b->hcol_pairs * b->hcol_stride;
        for (uint32_t t = 0; t < b->stripes_max; ++t) {
            MyersArgs y = x;
            y.stripe = t;
            y.hcol_in = b->d_hcol + (t % 2) * half;
            y.hcol_out = b->d_hcol + ((t + 1) % 2) * half;
            y.hcol_stride = b->hcol_stride;
            if (const hipError_t e = sf_launch_myers_striped(y, s)) return e;
        }
        return hipSuccess;
    };
    auto mix = [&](const MixArgs &mx, int group, hipStream_t s) {
        MyersArgs x = a;
        x.order = b->d_morder;
        x.desc = b->d_mdesc;
        return sf_launch_myers_mix(x, mx, group, s);
    };
    // With the Ukkonen band the lane classes launch alone (a threshold that skips no block at the
    // start of the smallest lane class leaves the plain grouping).
    const bool ukk = myers_use_ukk(a, 24);
    return run_grouped(b, bins, one, ukk ? myers_group_ukk : sf_myers_mix_group, sf_myers_blocks_for, mix,
                       nullptr, true, 0);
}

// Every Myers bin.
int run_myers(sf_batch *b, const MyersArgs &a) { return run_myers_bins(b, a, b->mbins); }

// The start pass (a.reverse): lane classes as their bins run it (over Ukkonen's band), patterns
// over 1,536 bases stripe by stripe over the diagonal band (myers_start.hip), one launch on the
// batch's stream.  Under a capture without the stripe buffers made, every bin runs the pipelines'
// full pass.
int run_myers_start(sf_batch *b, const MyersArgs &a) {
    const uint64_t n = b->n, q0 = b->start_first;
    if (q0 >= n || !myers_ukk_policy() || (b->capturing && !b->d_start_hc)) return run_myers(b, a);
    int st;
    if (!b->d_start_hc) {
        // The stripe kernel keeps columns in 31 bits; its buffers are optional (without them, or
        // for texts of 2^30 bases or more, the pipelines run the pass).
        uint32_t tmax = 0;
        for (uint64_t q = q0; q < n; ++q) tmax = std::max(tmax, b->tlen[b->morder[q]]);
        if (tmax >= (1u << 30)) return run_myers(b, a);
        b->start_words = (uint64_t)tmax / 16 + 2;
        if (hipMalloc((void **)&b->d_start_hc, 2 * (n - q0) * b->start_words * sizeof(uint32_t)) != hipSuccess) {
            (void)hipGetLastError();
            b->d_start_hc = nullptr;
            return run_myers(b, a);
        }
    }
    std::vector<sf_batch::Bin> lane;
    for (const auto &bin : b->mbins)
        if (bin.end <= q0) lane.push_back(bin);
    if (!lane.empty() && (st = run_myers_bins(b, a, lane))) return st;
    MyersStartArgs s{};
    s.pv = a.pv;
    s.desc = b->d_mdesc;
    s.order = b->d_morder;
    s.first = (uint32_t)q0;
    s.count = (uint32_t)(n - q0);
    s.dist = a.dist;
    s.end = a.end;
    s.start = a.start;
    s.hc_words = b->start_words;
    s.hc0 = b->d_start_hc;
    s.hc1 = b->d_start_hc + (n - q0) * b->start_words;
    HIP_TRY(sf_launch_myers_start_stripe(s, b->stream));
    return SF_OK;
}

// Estimated VALU wave-instructions a row of one wave costs in shape class `cls`, shared by its
// 64 / LP pairs: about 9 a slot and 15 besides for the 32-bit kernels; 3.5 a slot (two slots a
// register) and 24 besides for the 16-bit packed ones, 27 on 32 lanes, 30 with the 64-lane scan.
double aff_row_cost(int cls) {
    if (cls == kAffWide) return 30.0 * 8 + 60.0; // (a 256-thread workgroup: about 30 a slot, 8 a thread)
    const double k = aff_k(cls);
    if (!aff_packed(cls)) return 9.0 * k + 15.0;
    return 3.5 * k + (aff_lp(cls) == 64 ? 30.0 : (aff_lp(cls) == 32 ? 27.0 : 24.0));
}

// The planner's two costs of a class, in ns: a pair-row's share of the chip's time when the class
// fills it (throughput), and one row of a wave running alone (chain).  Measured by the autotuner
// where its table holds the class (tune.hip), else estimated: aff_row_cost wave-instructions at one
// per SIMD every 4 cycles (2.4 GHz), spread over the chip's 1,024 SIMDs for the throughput, and
// 1.5 times that interval for a lone wave's dependent instructions.  The autotuner times the score
// kernels, so only score-only plans (`tuned`) take its measurements: a class code in a traceback
// plan or a pass names another kernel of the same shape.
constexpr double kNsPerIssue = 4.0 / 2.4;
double aff_tput_ns(int cls, bool tuned) {
    double t, c;
    if (tuned && sf_tune_cost(cls, &t, &c)) return t;
    return aff_row_cost(cls) * aff_lp(cls) / 64.0 * kNsPerIssue / 1024.0;
}
double aff_chain_ns(int cls, bool tuned) {
    double t, c;
    if (tuned && sf_tune_cost(cls, &t, &c)) return c;
    return 1.5 * aff_row_cost(cls) * kNsPerIssue;
}

// Per pair, the shape an affine run uses.  A shape's cost per pair-row is its wave-row cost over
// its pairs a wave; its chain, the time one pair's wave needs alone, is m wave-rows at about 1.5
// times the issue interval (dependent VALU, DPP and waits); the run's spread is every pair's rows
// at its cost per pair-row over the chip's 1024 SIMDs.  Every pair starts in its cheapest shape per
// pair-row (8 lanes for scores where the 16-bit range holds and the policy allows; 16 up to the
// 16-lane width limit; 32 for traceback runs of wider bands up to 1,024 slots where the 16-bit
// range holds; 16-bit packed where its range holds), and a pair whose chain would outlast
// the spread moves to the cheapest shape whose chain does not (else 64 lanes, the shortest).  A
// class left with too few waves to fill the chip (chain above its own spread) moves to 64 lanes
// as a whole where that adds little to the run (its 64-lane spread within 5% of the run's): its
// pairs then run at a quarter to an eighth of the chain, and score-only runs launch all such
// 64-lane classes together (run_bins).  Policy 1 keeps every pair in its fewest-lanes shape.
// A pair's candidate shapes: 4, 8, 16, 32 and 64 lanes (0 where not allowed), each at the smallest
// class that holds its band and, for 8, 16 and 64 lanes, at the next class edge up (a tile of more
// slots a lane: chosen only where the autotuner measured it cheaper).  Index kC64 is the 64-lane
// shape at its smallest class.  The fewest-lanes policies take the first candidate: 4 lanes under
// policy 2, 8 under policy 1 (which leaves 4 lanes out).
constexpr int kCands = 8, kC64 = 6;
using Cands = std::array<int, kCands>;
Cands pair_candidates(uint32_t m, uint64_t below, uint64_t width, const sf_scoring *sc, bool cigar, int lp8,
                      uint64_t lp16max) {
    if (width > 64 * 32) { // wider than any lane shape: the wide kernel, in the 64-lane fallback's place
        Cands c{};
        c[kC64] = cigar ? 0 : kAffWide;
        return c;
    }
    auto pk = [&](int c, int lanes) {
        if (!c) return 0;
        const int p = sf_affine_pk16_class(aff_k(c), sc->mat, sc->mis, sc->open, sc->ext, m, (uint32_t)below, cigar, lanes);
        return p ? p : c;
    };
    int c64 = pk(sf_affine_class(width, false), 64);
    int c64n = c64 ? pk(sf_affine_class(aff_slots(sf_affine_class(width, false)) + 1, false), 64) : 0;
    int c16 = 0, c16n = 0;
    if (width <= lp16max && width <= 16 * 32) {
        const int w16 = sf_affine_class(width, true);
        c16 = pk(w16, 16);
        const uint64_t nx = aff_slots(w16) + 1;
        if (nx <= lp16max && nx <= 16 * 32) c16n = pk(sf_affine_class(nx, true), 16);
    }
    const int c4 = (!cigar && (lp8 == 0 || lp8 == 2))
                       ? sf_affine_pk4_class(width, sc->mat, sc->mis, sc->open, sc->ext, m, (uint32_t)below)
                       : 0;
    int c8 = 0, c8n = 0;
    if (!cigar && lp8 >= 0) {
        c8 = sf_affine_pk8_class(width, sc->mat, sc->mis, sc->open, sc->ext, m, (uint32_t)below);
        if (c8) c8n = sf_affine_pk8_class(aff_slots(c8) + 1, sc->mat, sc->mis, sc->open, sc->ext, m, (uint32_t)below);
    }
    // 32 lanes: traceback runs of bands over 512 slots (2 * ceil(width / 64) slots a lane).
    const int c32 = (cigar && width > 16 * 32 && width <= 32 * 32)
                        ? sf_affine_pk16_class(2 * (int)((width + 63) / 64), sc->mat, sc->mis, sc->open, sc->ext, m,
                                               (uint32_t)below, true, 32)
                        : 0;
    // (A next-edge class the kernels lack, or outside the 16-bit range, comes back as 0 or as the
    // 32-bit class: drop it where it is not a different shape of the same kind.)
    if (cigar || c64n == c64 || (c64n && aff_packed(c64n) != aff_packed(c64))) c64n = 0;
    if (cigar || c16n == c16 || (c16n && aff_packed(c16n) != aff_packed(c16))) c16n = 0;
    if (c8n == c8) c8n = 0;
    return Cands{c4, c8, c16, c32, c8n, c16n, c64, c64n};
}

// The planner's costs of the classes one plan considers, each read once (the autotuner's lookups
// lock; the pair loops run on all host threads).  Class codes are below 65536.
struct ClassCosts {
    std::vector<uint16_t> slot = std::vector<uint16_t>(65536, 0xFFFF);
    std::vector<double> tput, chain;
    void add(int cls, bool tuned) {
        if (cls <= 0 || cls >= 65536 || slot[cls] != 0xFFFF) return;
        slot[cls] = (uint16_t)tput.size();
        tput.push_back(aff_tput_ns(cls, tuned));
        chain.push_back(aff_chain_ns(cls, tuned));
    }
};

void choose_shapes(const sf_batch *b, const sf_scoring *sc, const std::vector<uint64_t> &ba, bool cigar,
                   std::vector<int> &kcls) {
    const uint64_t n = b->n;
    const int lp8 = sf_affine_lp8_policy();
    const uint64_t lp16max = sf_affine_lp16_max();
    std::vector<Cands> cand(n);
    const bool tuned = !cigar;
    parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            const uint32_t m = b->plen[i], below = (uint32_t)ba[i];
            const uint64_t width = below + (ba[i] >> 32) + 1;
            cand[i] = pair_candidates(m, below, width, sc, cigar, lp8, lp16max);
        }
    });
    ClassCosts cc;
    for (uint64_t i = 0; i < n; ++i)
        for (int c : cand[i]) cc.add(c, tuned);
    auto slot = [&](int c) { return (c > 0 && c < 65536) ? cc.slot[c] : (uint16_t)0xFFFF; };
    auto tput = [&](int c) { const uint16_t q = slot(c); return q != 0xFFFF ? cc.tput[q] : aff_tput_ns(c, tuned); };
    auto chain = [&](uint64_t i, int c) {
        const uint16_t q = slot(c);
        return b->plen[i] * (q != 0xFFFF ? cc.chain[q] : aff_chain_ns(c, tuned));
    };
    parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            int pick = 0;
            for (int c : cand[i])
                if (c && (!pick || (lp8 <= 0 && tput(c) < tput(pick)))) pick = c;
            kcls[i] = pick;
        }
    });
    if (lp8 > 0) return;
    double spread = 0;
    for (uint64_t i = 0; i < n; ++i) spread += b->plen[i] * tput(kcls[i]);
    parallel_for(n, [&](uint64_t lo, uint64_t hi) {
        for (uint64_t i = lo; i < hi; ++i) {
            if (chain(i, kcls[i]) <= spread) continue;
            int pick = 0;
            for (int c : cand[i])
                if (c && chain(i, c) <= spread && (!pick || tput(c) < tput(pick))) pick = c;
            kcls[i] = pick ? pick : cand[i][kC64];
        }
    });
    // Per class: its chain, its spread, and its spread on 64 lanes.
    const size_t nc = cc.tput.size();
    std::vector<double> cchain(nc, 0.0), cspread(nc, 0.0), cspread64(nc, 0.0);
    for (uint64_t i = 0; i < n; ++i) {
        const int k = kcls[i];
        const uint16_t q = slot(k);
        if (q == 0xFFFF) continue;
        cchain[q] = std::max(cchain[q], chain(i, k));
        cspread[q] += b->plen[i] * tput(k);
        // (The move costs a class 4 to 8 times its spread: only where that stays small.)
        if (aff_lp(k) != 64) cspread64[q] += b->plen[i] * tput(cand[i][kC64]);
    }
    for (uint64_t i = 0; i < n; ++i) {
        const int k = kcls[i];
        const uint16_t q = slot(k);
        if (q != 0xFFFF && aff_lp(k) != 64 && cchain[q] > cspread[q] && cspread64[q] <= 0.05 * spread) kcls[i] = cand[i][kC64];
    }
}

// The indices 0..n-1 ordered by a 48-bit key, ties in index order (a stable LSD radix sort, three
// passes of 16 bits).
std::vector<uint32_t> radix_order(const std::vector<uint64_t> &key) {
    const uint64_t n = key.size();
    std::vector<uint32_t> order(n), tmp(n);
    std::iota(order.begin(), order.end(), 0u);
    std::vector<uint32_t> cnt(65536 + 1);
    for (int pass = 0; pass < 3; ++pass) {
        const int sh = 16 * pass;
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (uint64_t i = 0; i < n; ++i) ++cnt[((key[i] >> sh) & 0xFFFF) + 1];
        for (int d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
        for (uint64_t q = 0; q < n; ++q) {
            const uint32_t i = order[q];
            tmp[cnt[(key[i] >> sh) & 0xFFFF]++] = i;
        }
        order.swap(tmp);
    }
    return order;
}

// The pairs in launch order: by shape class, longer patterns first inside a class, ties in index
// order (key: class code, then 2^31 - 1 - pattern length).
std::vector<uint32_t> launch_order(const std::vector<int> &kcls, const std::vector<uint32_t> &plen) {
    const uint64_t n = kcls.size();
    std::vector<uint64_t> key(n);
    for (uint64_t i = 0; i < n; ++i) key[i] = ((uint64_t)(uint32_t)kcls[i] << 31) | (0x7FFFFFFFu - (plen[i] & 0x7FFFFFFFu));
    return radix_order(key);
}

// Groups the pairs of a batch into bins of one shape class each (longer patterns first inside a
// class), estimates each bin's time for the stream schedule, and uploads the launch order, the
// band geometry and the launch-order pair records.
// Waits for the batch's queued work.  Blocking uploads go on the null stream, which the batch's
// non-blocking streams do not order with, so one into a buffer that queued kernels may still read
// (a plan being replaced, the traceback offsets) waits first.  The main stream covers the aux and
// walk streams: every run joins them back before it returns.
int quiesce(sf_batch *b) {
    if (const int st = join_aux(b)) return st;
    HIP_TRY(hipStreamSynchronize(b->stream));
    return SF_OK;
}

// A call that cannot be part of a captured step (it waits for the device, copies between host and
// device or builds a launch plan) fails while the batch is capturing, and so does the capture.
int refuse_in_capture(sf_batch *b, const char *what) {
    if (!b->capturing) return SF_OK;
    b->capture_error = SF_EINVAL;
    return fail(SF_EINVAL, "%s cannot be captured (captured steps hold Myers runs and score-only affine runs "
                           "whose launch plans exist: run the step once before capturing it)", what);
}

int build_plan(sf_batch *b, sf_batch::AffPlan &pl, const std::vector<uint64_t> &ba, const std::vector<int> &kcls,
               uint64_t key, uint64_t cells, bool tuned = false) {
    const uint64_t n = b->n;
    int st;
    if (pl.d_ba && (st = quiesce(b))) return st; // the plan being replaced may still be in use
    std::vector<uint32_t> order = launch_order(kcls, b->plen);
    pl.bins.clear();
    for (uint64_t q = 0; q < n;) {
        uint64_t e = q;
        while (e < n && kcls[order[e]] == kcls[order[q]]) ++e;
        sf_batch::Bin bin{kcls[order[q]], (uint32_t)q, (uint32_t)e};
        const double tput = aff_tput_ns(bin.cls, tuned), lat = aff_chain_ns(bin.cls, tuned);
        for (uint64_t x = q; x < e; ++x) {
            bin.chain = std::max(bin.chain, b->plen[order[x]] * lat);
            bin.spread += b->plen[order[x]] * tput;
        }
        pl.bins.push_back(bin);
        q = e;
    }
    pl.wide_stride = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (kcls[i] == kAffWide) { // below and above as signed (a local pass's below is -dlo)
            const int64_t w = (int64_t)(int32_t)(uint32_t)ba[i] + (int64_t)(int32_t)(uint32_t)(ba[i] >> 32) + 1;
            pl.wide_stride = std::max<uint64_t>(pl.wide_stride, (uint64_t)std::max<int64_t>(w, 0) + 1);
        }
    pl.wide_stride = (pl.wide_stride + 63) & ~63ull;
    if (!pl.d_ba && ((st = dalloc(&pl.d_ba, n)) || (st = dalloc(&pl.d_order, n)) || (st = dalloc(&pl.d_desc, n))))
        return st;
    HIP_TRY(hipMemcpy(pl.d_ba, ba.data(), n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(pl.d_order, order.data(), n * 4, hipMemcpyHostToDevice));
    HIP_TRY(sf_launch_desc(n, pl.d_order, view(b), pl.d_ba, pl.d_desc, b->stream));
    pl.order = std::move(order);
    pl.key = key;
    pl.valid = true;
    pl.cells = cells;
    return SF_OK;
}

// Wide-band pairs (kAffWide) run as launches of as many pairs as 1 GiB of row scratch holds.
constexpr uint64_t kWideScratchWords = (1ull << 30) / 4;
uint64_t wide_pairs_per_launch(const sf_batch::AffPlan &pl) {
    return pl.wide_stride ? std::max<uint64_t>(1, kWideScratchWords / (4 * pl.wide_stride)) : 0;
}

// Scratch for a plan's wide-band bins, made before a run queues anything (waits for the batch's
// work if the scratch grows).
int ensure_wide_scratch(sf_batch *b, const sf_batch::AffPlan &pl) {
    if (!pl.wide_stride) return SF_OK;
    uint64_t count = 0;
    for (const auto &bin : pl.bins)
        if (bin.cls == kAffWide) count = std::max<uint64_t>(count, bin.end - bin.begin);
    const uint64_t words = std::min(count, wide_pairs_per_launch(pl)) * 4 * pl.wide_stride;
    if (words <= b->wide_words) return SF_OK;
    int st;
    if ((st = refuse_in_capture(b, "a run whose wide-band scratch must grow")) || (st = quiesce(b))) return st;
    dfree(b->d_wide);
    b->wide_words = 0;
    if ((st = dalloc(&b->d_wide, words))) return st;
    b->wide_words = words;
    return SF_OK;
}

// A wide-band bin's launches, one after another on stream s (they share the scratch).
hipError_t launch_wide(sf_batch *b, const sf_batch::AffPlan &pl, const sf_batch::Bin &bin, WideArgs x,
                       hipStream_t s) {
    const uint64_t per = wide_pairs_per_launch(pl);
    for (uint64_t q = bin.begin; q < bin.end; q += per) {
        WideArgs y = x;
        y.p.order = pl.d_order + q;
        y.p.desc = pl.d_desc + q;
        y.p.count = (uint32_t)std::min<uint64_t>(per, bin.end - q);
        y.rows = b->d_wide;
        y.stride = pl.wide_stride;
        if (const hipError_t e = sf_launch_affine_wide(y, s)) return e;
    }
    return hipSuccess;
}

// The rest of a batch's set-up once b->n, plen and tlen are set and its sequences are on the
// device (d_seq): pair i's pattern at base pbase[i], its text at tbase[i].  Streams, per-pair
// arrays, Myers bins and launch records.  On failure the caller deletes b.
int batch_init(sf_batch *b, const std::vector<uint64_t> &pbase, const std::vector<uint64_t> &tbase) {
    const uint64_t n = b->n;
    int st;
    if (take_stream(&b->stream) != hipSuccess) return fail(SF_EDEVICE, "hipStreamCreate failed");
    if ((st = dalloc(&b->d_poff, n)) || (st = dalloc(&b->d_toff, n)) ||
        (st = dalloc(&b->d_plen, n)) || (st = dalloc(&b->d_tlen, n)) || (st = dalloc(&b->d_dist, n)) ||
        (st = dalloc(&b->d_end, n)) || (st = dalloc(&b->d_start, n)) || (st = dalloc(&b->d_morder, n)) ||
        (st = dalloc(&b->d_mdesc, n)) ||
        (st = dalloc(&b->d_toff0, n)) || (st = dalloc(&b->d_tlen0, n)))
        return st;
    // Myers bins: by blocks per pattern (the lane kernel's register class), then text length.
    // Key: blocks (clamped to 2^17: patterns past 98,304 bases have no class anyway) over 30 bits
    // of text length, 47 bits.
    std::vector<uint64_t> mkey(n);
    for (uint64_t i = 0; i < n; ++i)
        mkey[i] = ((uint64_t)std::min<uint32_t>((b->plen[i] + 63) / 64, 1u << 17) << 30) | b->tlen[i];
    std::vector<uint32_t> order = radix_order(mkey);
    // A pattern past the largest class runs striped (kMyersStriped), up to kMyersStripedMax blocks.
    auto myers_class = [&](uint32_t m) {
        const uint32_t nb = (m + 63) / 64;
        const int c = sf_myers_maxb_for(nb);
        return (c == 0 && nb <= kMyersStripedMax) ? kMyersStriped : c;
    };
    for (uint64_t q = 0; q < n;) {
        int cls = myers_class(b->plen[order[q]]);
        uint64_t e = q;
        while (e < n && myers_class(b->plen[order[e]]) == cls) ++e;
        if (cls == 0 && b->myers_unsupported.empty()) {
            char msg[160];
            snprintf(msg, sizeof msg, "pattern of %u bases: Myers supports up to %llu", b->plen[order[q]],
                     (unsigned long long)kMyersStripedMax * 64);
            b->myers_unsupported = msg;
        }
        sf_batch::Bin bin{cls, (uint32_t)q, (uint32_t)e};
        // Per wave: lane classes take 64 pairs of `cls` blocks a column; pipelined ones take
        // 64 / G pairs, R blocks a lane a column, G - 1 columns of skew; a striped pattern runs
        // its stripes (64 x 24 blocks) one after another.
        auto estimate = [&](sf_batch::Bin &x) {
            const bool st = x.cls == kMyersStriped;
            const double G = st ? 64 : (x.cls >= 16000 ? x.cls / 1000 : 1),
                         R = st ? 24 : (x.cls >= 16000 ? x.cls % 1000 : x.cls);
            x.chain = x.spread = 0;
            for (uint64_t y = x.begin; y < x.end; ++y) {
                const double stripes = st ? (double)((b->plen[order[y]] + 63) / 64 + kMyersStripedBlocks - 1) /
                                                kMyersStripedBlocks
                                          : 1.0;
                const double cols = ((double)b->tlen[order[y]] + (G - 1)) * (st ? std::floor(stripes) : 1.0);
                x.chain = std::max(x.chain, cols * R);
                x.spread += cols * R * G / 64.0 / 1024.0;
            }
        };
        estimate(bin);
        // A 32-lane bin that cannot fill the chip is bound by its chain: 64 lanes halve it.  It
        // then joins a neighbouring bin of that class (one launch, not two in a row).
        if (bin.chain > bin.spread && sf_myers_wide_class(cls) != cls) {
            bin.cls = sf_myers_wide_class(cls);
            estimate(bin);
        }
        if (!b->mbins.empty() && b->mbins.back().cls == bin.cls && b->mbins.back().end == bin.begin) {
            b->mbins.back().end = bin.end;
            estimate(b->mbins.back());
        } else {
            b->mbins.push_back(bin);
        }
        q = e;
    }
    for (uint64_t i = 0; i < n; ++i) b->myers_cells_per_run += (uint64_t)b->plen[i] * b->tlen[i];
    // Striped patterns: two rows of a byte per text column (the stripes' deltas, ping-pong), each
    // row rounded to whole words.
    for (uint64_t i = 0; i < n; ++i)
        if (myers_class(b->plen[i]) == kMyersStriped) {
            ++b->hcol_pairs;
            b->hcol_stride = std::max<uint64_t>(b->hcol_stride, ((uint64_t)b->tlen[i] + 8) & ~3ull);
            b->stripes_max = std::max<uint32_t>(
                b->stripes_max, ((b->plen[i] + 63) / 64 + kMyersStripedBlocks - 1) / kMyersStripedBlocks);
        }
    if (b->hcol_pairs && (st = dalloc(&b->d_hcol, 2 * b->hcol_pairs * b->hcol_stride))) return st;
    // Patterns over 1,536 bases sort last in the launch order (by blocks).
    b->morder = order;
    b->start_first = n;
    for (uint64_t q = 0; q < n; ++q)
        if (b->plen[order[q]] > 64u * 24u) {
            b->start_first = q;
            break;
        }
    hipError_t e2 = hipMemcpy(b->d_poff, pbase.data(), n * 8, hipMemcpyHostToDevice);
    hipError_t e3 = hipMemcpy(b->d_toff, tbase.data(), n * 8, hipMemcpyHostToDevice);
    hipError_t e4 = hipMemcpy(b->d_plen, b->plen.data(), n * 4, hipMemcpyHostToDevice);
    hipError_t e5 = hipMemcpy(b->d_tlen, b->tlen.data(), n * 4, hipMemcpyHostToDevice);
    hipError_t e6 = hipMemcpy(b->d_morder, order.data(), n * 4, hipMemcpyHostToDevice);
    hipError_t e7 = hipMemcpy(b->d_toff0, tbase.data(), n * 8, hipMemcpyHostToDevice);
    hipError_t e8 = hipMemcpy(b->d_tlen0, b->tlen.data(), n * 4, hipMemcpyHostToDevice);
    if (e2 || e3 || e4 || e5 || e6 || e7 || e8) return fail(SF_EDEVICE, "upload failed");
    if (sf_launch_desc(n, b->d_morder, view(b), nullptr, b->d_mdesc, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return fail(SF_EDEVICE, "launch records failed");
    b->tlen0 = b->tlen;
    return SF_OK;
}

} // namespace

int sf_persistent_policy() {
    int v = g_persistent.load();
    if (v < 0) {
        const char *e = std::getenv("SF_PERSISTENT");
        v = e ? std::max(0, std::min(3, std::atoi(e))) : 0;
        int expected = -1;
        if (!g_persistent.compare_exchange_strong(expected, v)) v = expected;
    }
    return v;
}

uint32_t sf_persistent_grid(const void *kernel, uint32_t blocks) {
    static std::mutex mu;
    static std::map<const void *, uint32_t> cap;
    uint32_t c = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cap.find(kernel);
        if (it != cap.end()) c = it->second;
    }
    if (!c) {
        int per_cu = 0, dev = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu < 1)
            per_cu = 1;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
            cus = 256;
        c = (uint32_t)per_cu * (uint32_t)cus;
        std::lock_guard<std::mutex> lk(mu);
        cap[kernel] = c;
    }
    return std::min(c, blocks);
}

// Hooks for the server (serve.hip).
int sf_runtime_fail(int code, const char *msg) { return fail(code, "%s", msg); }
int sf_runtime_device() {
    (void)ensure_device();
    return g_device < 0 ? 0 : g_device;
}
int sf_runtime_ensure() { return ensure_device(); }

extern "C" {

const char *sf_strerror(int s) {
    switch (s) {
    case SF_OK: return "ok";
    case SF_EINVAL: return "invalid argument";
    case SF_ENOMEM: return "out of memory";
    case SF_EDEVICE: return "device error";
    case SF_ENODEV: return "no usable device";
    case SF_ERANGE: return "out of range";
    default: return "unknown status";
    }
}

const char *sf_last_error(void) { return g_err.c_str(); }
int sf_abi_version(void) { return SF_ABI_VERSION; }

int sf_init(int device) {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(SF_ENODEV, "no HIP device visible");
    if (device < 0 || device >= count) return fail(SF_EINVAL, "device %d out of range (%d devices)", device, count);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SF_ENODEV, "device %d is %s, this library is built for gfx950", device, prop.gcnArchName);
    HIP_TRY(hipSetDevice(device));
    g_device = device;
    return SF_OK;
}

uint64_t sf_batch_size(const sf_batch *b) { return b ? b->n : 0; }

void sf_batch_destroy(sf_batch *b) {
    if (b && b->capturing) { // the capture is abandoned
        hipGraph_t g = nullptr;
        (void)join_aux(b);
        (void)hipStreamEndCapture(b->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        b->capturing = false;
    }
    delete b;
}

int sf_batch_create(const sf_pairs *pairs, sf_batch **out) {
    if (!out) return fail(SF_EINVAL, "out is null");
    *out = nullptr;
    int st = check_pairs(pairs);
    if (st) return st;
    if ((st = ensure_device())) return st;
    for (uint64_t i = 0; i < pairs->n; ++i)
        if (pairs->pat_len[i] >= (1u << 30) || pairs->txt_len[i] >= (1u << 30))
            return fail(SF_ERANGE, "pair %llu: sequences of 2^30 bases or more", (unsigned long long)i);
    auto *b = new sf_batch();
    b->n = pairs->n;
    const uint64_t n = pairs->n;
    b->plen.assign(pairs->pat_len, pairs->pat_len + n);
    b->tlen.assign(pairs->txt_len, pairs->txt_len + n);
    std::vector<uint64_t> poff(n), toff(n);
    uint64_t words = 0;
    for (uint64_t i = 0; i < n; ++i) {
        poff[i] = words;
        words += (b->plen[i] + 15) / 16;
        toff[i] = words;
        words += (b->tlen[i] + 15) / 16;
    }
    // Every word belongs to one pair's range and is written by its packing (no zero fill).
    const uint64_t seq_words = std::max<uint64_t>(words, 1);
    std::unique_ptr<uint32_t[]> seqbuf(new uint32_t[seq_words]);
    uint32_t *seq = seqbuf.get();
    seq[seq_words - 1] = 0; // (the one word of a batch without bases)
    parallel_for(n, [&](uint64_t a, uint64_t e) {
        for (uint64_t i = a; i < e; ++i) {
            pack_into(&seq[poff[i]], pairs->pat, pairs->packed, pairs->pat_off[i], b->plen[i]);
            pack_into(&seq[toff[i]], pairs->txt, pairs->packed, pairs->txt_off[i], b->tlen[i]);
        }
    });
    auto bail = [&](int s) {
        delete b;
        return s;
    };
    if ((st = dalloc(&b->d_seq, seq_words))) return bail(st);
    if (hipMemcpy(b->d_seq, seq, seq_words * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SF_EDEVICE, "upload failed"));
    // The kernels address bases: word offsets * 16.
    std::vector<uint64_t> pbase(n), tbase(n);
    for (uint64_t i = 0; i < n; ++i) {
        pbase[i] = poff[i] * 16;
        tbase[i] = toff[i] * 16;
    }
    if ((st = batch_init(b, pbase, tbase))) return bail(st);
    // Pattern codes of 4 or more (unpacked input only) match nothing in local passes
    // (align.rs:731-734); the other functions mask codes with & 3 as the reference does.
    if (!pairs->packed) {
        const uint8_t *pb = static_cast<const uint8_t *>(pairs->pat);
        // First whether any pattern holds one: 8 codes at a time, on the host threads.
        std::atomic<bool> any{false};
        parallel_for(n, [&](uint64_t a, uint64_t e) {
            bool found = false;
            for (uint64_t i = a; i < e && !found; ++i) {
                const uint8_t *src = pb + pairs->pat_off[i];
                const uint32_t len = b->plen[i];
                uint32_t k = 0;
                for (; k + 8 <= len; k += 8) {
                    uint64_t x;
                    std::memcpy(&x, src + k, 8);
                    if (x & 0xFCFCFCFCFCFCFCFCull) found = true;
                }
                for (; k < len; ++k)
                    if (src[k] >= 4) found = true;
            }
            if (found) any.store(true, std::memory_order_relaxed);
        });
        std::vector<uint32_t> inv;
        for (uint64_t i = 0; any.load() && i < n; ++i)
            for (uint32_t k = 0; k < b->plen[i]; ++k)
                if (pb[pairs->pat_off[i] + k] >= 4) {
                    if (inv.empty()) inv.assign((seq_words * 16 + 64) / 32 + 2, 0u);
                    const uint64_t bit = pbase[i] + k;
                    inv[bit >> 5] |= 1u << (bit & 31);
                }
        if (!inv.empty()) {
            if ((st = dalloc(&b->d_pinv, inv.size()))) return bail(st);
            if (hipMemcpy(b->d_pinv, inv.data(), inv.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
                return bail(fail(SF_EDEVICE, "upload failed"));
            b->pinv_words = inv.size();
        }
    }
    *out = b;
    return SF_OK;
}


int sf_batch_myers(sf_batch *b, int mode, uint32_t k, int locate) {
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (mode < SF_INFIX || mode > SF_GLOBAL) return fail(SF_EINVAL, "mode %d", mode);
    if (locate < 0 || locate > 2) return fail(SF_EINVAL, "locate %d", locate);
    if (!b->myers_unsupported.empty()) return fail(SF_ERANGE, "%s", b->myers_unsupported.c_str());
    int st;
    if ((st = ensure_device())) return st;
    if ((st = join_aux(b))) return st; // the last Myers run may still write the outputs
    MyersArgs a{};
    a.pv = view(b);
    a.mode = mode;
    a.k = k;
    a.dist = b->d_dist;
    a.end = locate ? b->d_end : nullptr;
    a.start = locate == 2 ? b->d_start : nullptr;
    hipEvent_t e0;
    if ((st = timed_begin(b, 0, &e0))) return st;
    a.reverse = 0;
    if ((st = run_myers(b, a))) return st;
    if ((st = timed_end_open(b, 0, e0))) return st;
    (b->capturing ? b->cap_myers_cells : b->myers_cells) += b->myers_cells_per_run;
    b->located_ready = locate == 2;
    if (locate == 2 && mode == SF_INFIX) {
        if ((st = join_aux(b))) return st; // the start pass reads the forward pass's results
        if ((st = timed_begin(b, 1, &e0))) return st;
        a.reverse = 1;
        if ((st = run_myers_start(b, a))) return st;
        if ((st = timed_end_open(b, 1, e0))) return st;
    }
    return SF_OK;
}

int sf_batch_affine(sf_batch *b, int mode, const sf_scoring *sc, const uint64_t *band, int cigar) {
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (!sc) return fail(SF_EINVAL, "scoring is null");
    if (mode < SF_INFIX || mode > SF_GLOBAL) return fail(SF_EINVAL, "mode %d", mode);
    if (sc->open < 0 || sc->ext < 0 || sc->mat < 0 || sc->mis < 0)
        return fail(SF_EINVAL, "scores must be non-negative (align.rs:2511 requires open, ext >= 0)");
    if (sc->mat > kScoreMax || sc->mis > kScoreMax || sc->open > kScoreMax || sc->ext > kScoreMax)
        return fail(SF_ERANGE, "scoring values above %d are outside this library's bound", kScoreMax);
    int st;
    if ((st = ensure_device())) return st;
    if (cigar && (st = refuse_in_capture(b, "an affine run with traceback"))) return st;
    // A score run needs nothing from a Myers run: it starts beside that run's last bins.  The
    // traceback path (scratch, walks, uploads) starts from a joined state.
    if (cigar && (st = join_aux(b))) return st;
    const uint64_t n = b->n;
    // Shape classes, cached across runs with the same mode, bands, scoring and tuning.
    sf_batch::AffPlan &pl = b->aplan[cigar ? 1 : 0];
    uint64_t key = (1469598103934665603ull ^ (uint64_t)mode ^ (sf_affine_lp16_max() << 8) ^
                    ((uint64_t)(sf_affine_lp8_policy() + 1) << 4) ^ (cigar ? 0 : sf_tune_generation() << 24)) *
                   1099511628211ull;
    for (int32_t v : {sc->mat, sc->mis, sc->open, sc->ext}) key = (key ^ (uint64_t)(uint32_t)v) * 1099511628211ull;
    key = hash_words(key, band, n, SF_BAND_EXACT);
    if ((!pl.valid || pl.key != key) && (st = refuse_in_capture(b, "an affine run that needs a new launch plan")))
        return st;
    if (!pl.valid || pl.key != key) {
        std::vector<uint64_t> ba(n);
        std::vector<int> kcls(n);
        uint64_t cells = 0;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t below, above;
            const uint32_t m = b->plen[i], t = b->tlen[i];
            sf::band_geom(m, t, mode, band ? band[i] : SF_BAND_EXACT, &below, &above);
            const uint64_t width = below + above + 1;
            if (sf_affine_class(width, false) == 0 && (cigar || width > kAffWideMax))
                return fail(SF_ERANGE,
                            cigar ? "pair %llu: band of %llu cells per row; traceback handles up to 2048"
                                  : "pair %llu: band of %llu cells per row; this library handles up to 2^26",
                            (unsigned long long)i, (unsigned long long)width);
            ba[i] = below | (above << 32);
            cells += sf::band_cells(m, t, below, above); // DP_WORK (align.rs:2573-2578)
        }
        choose_shapes(b, sc, ba, cigar, kcls);
        for (uint64_t i = 0; i < n; ++i) {
            // Values stay far above NEG over all slots of the class (of the band, for the wide kernel).
            const uint64_t slots = kcls[i] == kAffWide ? (uint32_t)ba[i] + (ba[i] >> 32) + 1 : aff_slots(kcls[i]);
            if (!sf::scores_in_range(sc->mat, sc->mis, sc->open, sc->ext, b->plen[i], b->tlen[i], slots))
                return fail(SF_ERANGE, "pair %llu: scores could reach NEG", (unsigned long long)i);
        }
        if ((st = build_plan(b, pl, ba, kcls, key, cells, !cigar))) return st;
    }
    if (!b->d_score && ((st = dalloc(&b->d_score, n)) || (st = dalloc(&b->d_col, n)))) return st;
    if ((st = ensure_wide_scratch(b, pl))) return st;
    AffineArgs a{};
    a.pv = view(b);
    a.mode = mode;
    a.mat = sc->mat, a.mis = sc->mis, a.open = sc->open, a.ext = sc->ext;
    a.score = b->d_score;
    a.score_col = b->d_col;
    hipEvent_t e0;
    if (!cigar) {
        if ((st = timed_begin(b, 2, &e0))) return st;
        auto one = [&](const sf_batch::Bin &bin, hipStream_t s) {
            if (bin.cls == kAffWide) {
                WideArgs y{};
                y.p.pv = a.pv;
                y.p.mode = mode;
                y.p.mat = a.mat, y.p.mis = a.mis, y.p.open = a.open, y.p.ext = a.ext;
                y.score = a.score;
                y.score_col = a.score_col;
                return launch_wide(b, pl, bin, y, s);
            }
            AffineArgs x = a;
            x.order = pl.d_order + bin.begin;
            x.desc = pl.d_desc + bin.begin;
            x.count = bin.end - bin.begin;
            return sf_launch_affine(x, bin.cls, false, s);
        };
        auto mixed = [&](const std::vector<size_t> &ids, hipStream_t s) {
            MixArgs mx{};
            for (size_t t = 0; t < ids.size(); ++t) {
                const sf_batch::Bin &bin = pl.bins[ids[t]];
                mx.cls[t] = bin.cls;
                mx.off[t] = bin.begin;
                mx.count[t] = bin.end - bin.begin;
                mx.blk0[t] = mx.blocks;
                mx.blocks += sf_affine_blocks_for(bin.cls, mx.count[t]);
            }
            mx.nseg = (uint32_t)ids.size();
            mx.queue = queue_for(b, 16, 17);
            AffineArgs x = a;
            x.order = pl.d_order;
            x.desc = pl.d_desc;
            return sf_launch_affine_mix(x, mx, s);
        };
        // Many bins (a length-skewed batch): grouped launches, no tail between classes.  A few
        // (uniform lengths: one dominant class) keep a kernel each at that class's own occupancy.
        if (pl.bins.size() >= 4) {
            auto grouped = [&](const MixArgs &mx, int group, hipStream_t s) {
                AffineArgs x = a;
                x.order = pl.d_order;
                x.desc = pl.d_desc;
                return sf_launch_affine_mix_group(x, mx, group, s);
            };
            if ((st = run_grouped(b, pl.bins, one, sf_affine_mix_group, sf_affine_blocks_for, grouped,
                                  &b->affine_launches, false, 8)))
                return st;
        } else if ((st = run_bins(b, pl.bins, one, &b->affine_launches, mixed, sf_affine_mixable))) {
            return st;
        }
        if ((st = timed_end(b, 2, e0))) return st;
        (b->capturing ? b->cap_affine_cells : b->affine_cells) += pl.cells;
        b->cigar_valid = false;
        return SF_OK;
    }
    // Traceback: chunks of each bin whose bits fit the scratch budget.
    struct Chunk {
        int K;
        uint32_t begin, end;
    };
    std::vector<Chunk> chunks;
    std::vector<uint64_t> tboff(n);
    const uint64_t budget = tb_budget_words();
    uint64_t need_max = 0;
    for (auto &bin : pl.bins) {
        uint64_t used = 0;
        uint32_t cb = bin.begin;
        for (uint32_t q = bin.begin; q < bin.end; ++q) {
            // Whole 16-row blocks (the kernel writes and the walk reads a block's line at once).
            const uint64_t words = (((uint64_t)b->plen[pl.order[q]] + 15) & ~15ull) * aff_tb_words_per_row(bin.cls);
            if (used + words > budget && q > cb) {
                chunks.push_back({bin.cls, cb, q});
                need_max = std::max(need_max, used);
                used = 0;
                cb = q;
            }
            tboff[q] = used;
            used += words;
        }
        if (cb < bin.end) chunks.push_back({bin.cls, cb, bin.end});
        need_max = std::max(need_max, used);
    }
    // The last chunk's walk runs after every DP has ended, and a walk's time is set mostly by its
    // longest pattern (one dependent chain a pair), not by its chunk's DP time.  So the largest
    // bin's chunks run last, after the smaller bins (wider bands or longer patterns, whose walks
    // then overlap the large bin's DP), and its last chunk gives up about half of itself (its
    // shortest patterns, whose walks are the shortest too) as a piece that runs last: the walk
    // of the rest then overlaps the piece's DP, and only the piece's short walk is left after
    // the last DP.
    if (!chunks.empty()) {
        auto words = [&](int cls, uint32_t q) {
            return (((uint64_t)b->plen[pl.order[q]] + 15) & ~15ull) * aff_tb_words_per_row(cls);
        };
        std::map<int, uint64_t> bin_words;
        for (const Chunk &c : chunks)
            for (uint32_t q = c.begin; q < c.end; ++q) bin_words[c.K] += words(c.K, q);
        int big = chunks[0].K;
        for (auto &e : bin_words)
            if (e.second > bin_words[big]) big = e.first;
        std::stable_partition(chunks.begin(), chunks.end(), [&](const Chunk &c) { return c.K != big; });
        const size_t li = chunks.size() - 1;
        const Chunk last = chunks[li];
        uint64_t lw = 0;
        for (uint32_t q = last.begin; q < last.end; ++q) lw += words(last.K, q);
        if (last.end - last.begin >= 2 && lw > budget / 4) {
            uint32_t split = last.end;
            uint64_t tail = 0;
            while (split > last.begin + 1 && tail < lw / 2) tail += words(last.K, --split);
            uint64_t u = 0;
            for (uint32_t q = split; q < last.end; ++q) {
                tboff[q] = u;
                u += words(last.K, q);
            }
            chunks[li].end = split;
            chunks.push_back({last.K, split, last.end});
        }
    }
    if (need_max > b->tb_words) {
        dfree(b->d_tb[0]);
        dfree(b->d_tb[1]);
        if ((st = dalloc(&b->d_tb[0], need_max))) return st;
        if (chunks.size() > 1 && (st = dalloc(&b->d_tb[1], need_max))) return st;
        b->tb_words = need_max;
    }
    if (chunks.size() > 1 && !b->d_tb[1]) {
        if ((st = dalloc(&b->d_tb[1], b->tb_words))) return st;
    }
    if (!b->wstream) {
        HIP_TRY(take_stream(&b->wstream));
        for (hipEvent_t *e : {&b->dp_done[0], &b->dp_done[1], &b->walk_done[0], &b->walk_done[1], &b->walks_end})
            HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    if (b->d_tboff && (st = quiesce(b))) return st; // the last traceback run may still read the offsets
    if (!b->d_tboff) {
        if ((st = dalloc(&b->d_tboff, n)) || (st = dalloc(&b->d_tstart, n)) || (st = dalloc(&b->d_nruns, n)) ||
            (st = dalloc(&b->d_runs_off, n)) || (st = dalloc(&b->d_out_off, n + 1)) ||
            (st = dalloc(&b->d_scan, (n + kScanTile - 1) / kScanTile + 1)))
            return st;
        std::vector<uint64_t> roff(n);
        uint64_t tot = 0;
        for (uint64_t i = 0; i < n; ++i) {
            roff[i] = tot;
            tot += (uint64_t)b->plen[i] + b->tlen[i] + 1;
        }
        if ((st = dalloc(&b->d_runs, tot)) || (st = dalloc(&b->d_cig, tot))) return st;
        HIP_TRY(hipMemcpy(b->d_runs_off, roff.data(), n * 8, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(b->d_tboff, tboff.data(), n * 8, hipMemcpyHostToDevice));
    WalkArgs w{};
    w.pv = view(b);

    w.mode = mode;
    w.score = b->d_score;
    w.score_col = b->d_col;
    w.runs = b->d_runs;
    w.runs_off = b->d_runs_off;
    w.nruns = b->d_nruns;
    w.tstart = b->d_tstart;
    // Chunk c's DP (main stream) writes scratch c % 2; its walk (walk stream) reads it while the
    // DP of chunk c + 1 runs; the DP of chunk c + 2 waits for that walk before reusing the buffer.
    for (size_t c = 0; c < chunks.size(); ++c) {
        const Chunk &ch = chunks[c];
        const int buf = (int)(c & 1);
        if (c >= 2) HIP_TRY(hipStreamWaitEvent(b->stream, b->walk_done[buf], 0));
        a.tb = b->d_tb[buf];
        a.order = pl.d_order + ch.begin;
        a.desc = pl.d_desc + ch.begin;
        a.count = ch.end - ch.begin;
        a.tb_off = b->d_tboff + ch.begin;
        if ((st = timed_begin(b, 2, &e0))) return st;
        HIP_TRY(sf_launch_affine(a, ch.K, true, b->stream));
        if ((st = timed_end(b, 2, e0))) return st;
        ++b->affine_launches;
        HIP_TRY(hipEventRecord(b->dp_done[buf], b->stream));
        HIP_TRY(hipStreamWaitEvent(b->wstream, b->dp_done[buf], 0));
        w.tb = b->d_tb[buf];
        w.order = a.order;
        w.desc = a.desc;
        w.count = a.count;
        w.K = aff_k(ch.K);
        w.LP = aff_lp(ch.K);
        w.tb_off = a.tb_off;
        if ((st = timed_begin(b, 3, &e0, b->wstream))) return st;
        HIP_TRY(sf_launch_walk(w, b->wstream));
        if ((st = timed_end(b, 3, e0, b->wstream))) return st;
        HIP_TRY(hipEventRecord(b->walk_done[buf], b->wstream));
    }
    HIP_TRY(hipEventRecord(b->walks_end, b->wstream));
    HIP_TRY(hipStreamWaitEvent(b->stream, b->walks_end, 0));
    if ((st = timed_begin(b, 3, &e0))) return st;
    HIP_TRY(launch_scan(b->d_nruns, n, b->d_out_off, b->d_scan, b->stream));
    HIP_TRY(sf_launch_compact(pl.d_order, (uint32_t)n, b->d_runs, b->d_runs_off, b->d_nruns, b->d_out_off, b->d_cig,
                              b->stream));
    if ((st = timed_end(b, 3, e0))) return st;
    b->affine_cells += pl.cells;
    b->cigar_valid = true;
    return SF_OK;
}

int sf_batch_use_located(sf_batch *b, int on) {
    if (!b) return fail(SF_EINVAL, "batch is null");
    int st;
    if ((st = refuse_in_capture(b, "sf_batch_use_located"))) return st;
    if ((st = ensure_device())) return st;
    if ((st = join_aux(b))) return st; // reads the Myers results
    const uint64_t n = b->n;
    if (on) {
        if (!b->located_ready) return fail(SF_EINVAL, "no Myers run with start and end on this batch");
        if (n) {
            hipLaunchKernelGGL(restrict_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream, n,
                               b->d_dist, b->d_start, b->d_end, b->d_toff0, b->d_tlen0, b->d_toff, b->d_tlen);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(b->tlen.data(), b->d_tlen, n * 4, hipMemcpyDeviceToHost, b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
        }
    } else {
        HIP_TRY(hipMemcpyAsync(b->d_toff, b->d_toff0, n * 8, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync(b->d_tlen, b->d_tlen0, n * 4, hipMemcpyDeviceToDevice, b->stream));
        b->tlen = b->tlen0;
    }
    HIP_TRY(sf_launch_desc(n, b->d_morder, view(b), nullptr, b->d_mdesc, b->stream)); // the texts moved
    for (auto &pl : b->aplan) pl.valid = false; // the affine shape classes depend on the text lengths
    b->myers_cells_per_run = 0;
    for (uint64_t i = 0; i < n; ++i) b->myers_cells_per_run += (uint64_t)b->plen[i] * b->tlen[i];
    return SF_OK;
}

int sf_set_affine_pass_packed(int policy) {
    if (policy < 0 || policy > 2) return fail(SF_EINVAL, "policy must be 0, 1 or 2");
    sf_affine_set_pass_packed(policy);
    return SF_OK;
}

int sf_set_affine_lp8(int policy) {
    if (policy < -1 || policy > 2) return fail(SF_EINVAL, "policy must be -1, 0, 1 or 2");
    sf_affine_set_lp8_policy(policy);
    return SF_OK;
}

int sf_batch_autotune(sf_batch *b, int mode, const sf_scoring *sc, const uint64_t *band) {
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (!sc) return fail(SF_EINVAL, "scoring is null");
    if (mode < SF_INFIX || mode > SF_GLOBAL) return fail(SF_EINVAL, "mode %d", mode);
    if (sc->open < 0 || sc->ext < 0 || sc->mat < 0 || sc->mis < 0) return fail(SF_EINVAL, "scores must be non-negative");
    if (sc->mat > kScoreMax || sc->mis > kScoreMax || sc->open > kScoreMax || sc->ext > kScoreMax)
        return fail(SF_ERANGE, "scoring values above %d are outside this library's bound", kScoreMax);
    int st;
    if ((st = refuse_in_capture(b, "sf_batch_autotune")) || (st = ensure_device()) || (st = quiesce(b))) return st;
    const uint64_t lp16max = sf_affine_lp16_max();
    std::vector<char> seen(65536, 0);
    std::vector<int> classes;
    for (uint64_t i = 0; i < b->n; ++i) {
        uint64_t below, above;
        sf::band_geom(b->plen[i], b->tlen[i], mode, band ? band[i] : SF_BAND_EXACT, &below, &above);
        const uint64_t width = below + above + 1;
        if (sf_affine_class(width, false) == 0) continue; // sf_batch_affine refuses it
        for (int c : pair_candidates(b->plen[i], below, width, sc, false, 0, lp16max))
            if (c && !seen[c]) {
                seen[c] = 1;
                classes.push_back(c);
            }
    }
    if ((st = sf_tune_measure(classes, mode, sc))) return st;
    // The setting: the batch's score run timed under the estimates and under the measurements with
    // the chain scaled by 1, 1.5, 2 and 3 (median of three runs each, after one that plans it).
    struct Setting {
        bool use;
        double factor;
    };
    const Setting settings[] = {{false, 1.0}, {true, 1.0}, {true, 1.5}, {true, 2.0}, {true, 3.0}};
    // (The timing runs are not the batch's runs: its timing totals are restored afterwards.)
    b->collect();
    double ms_saved[5];
    std::copy(b->ms, b->ms + 5, ms_saved);
    const uint64_t cells_saved = b->affine_cells, launches_saved = b->affine_launches;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double best = 1e30;
    Setting pick = settings[0];
    for (const Setting &x : settings) {
        sf_tune_set(x.use, x.factor);
        double t[3];
        for (int r = -1; r < 3 && !st; ++r) {
            if ((st = sf_batch_affine(b, mode, sc, band, 0)) || (st = quiesce(b))) break;
            if (r < 0) continue;
            float ms = 0;
            if (hipEventRecord(e0, b->stream) != hipSuccess || (st = sf_batch_affine(b, mode, sc, band, 0)) ||
                hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                if (!st) st = fail(SF_EDEVICE, "autotune: timing the score run");
                break;
            }
            t[r] = ms;
        }
        if (st) break;
        std::sort(t, t + 3);
        if (t[1] < best) best = t[1], pick = x;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    sf_tune_set(pick.use, pick.factor);
    b->aplan[0].valid = false; // the next score run plans with the setting kept
    if (const int q = quiesce(b)) return st ? st : q;
    b->collect();
    std::copy(ms_saved, ms_saved + 5, b->ms);
    b->affine_cells = cells_saved;
    b->affine_launches = launches_saved;
    return st;
}

int sf_set_myers_ukk(int on) {
    if (on < 0 || on > 1) return fail(SF_EINVAL, "on must be 0 or 1");
    g_myers_ukk.store(on);
    return SF_OK;
}

int sf_set_persistent(int on) {
    if (on < 0 || on > 3) return fail(SF_EINVAL, "on must be 0..3 (bit 0: Myers runs, bit 1: score runs)");
    g_persistent.store(on);
    return SF_OK;
}

int sf_batch_capture_begin(sf_batch *b) {
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (b->capturing) return fail(SF_EINVAL, "the batch is capturing already");
    int st;
    if ((st = ensure_device()) || (st = ensure_aux(b))) return st;
    // Buffers a captured run writes are made now (no allocation while capturing).
    if (!b->d_score && ((st = dalloc(&b->d_score, b->n)) || (st = dalloc(&b->d_col, b->n)))) return st;
    if ((st = quiesce(b))) return st; // nothing queued before the capture is part of it
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
    b->capturing = true;
    b->capture_error = SF_OK;
    b->cap_myers_cells = b->cap_affine_cells = 0;
    return SF_OK;
}

int sf_batch_capture_end(sf_batch *b) {
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (!b->capturing) return fail(SF_EINVAL, "the batch is not capturing");
    const int joined = join_aux(b); // every stream the captured runs forked joins the main stream
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    b->capturing = false;
    if (joined || b->capture_error || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        if (joined) return joined;
        if (b->capture_error) return fail(b->capture_error, "a call during the capture could not be captured");
        return fail(SF_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(SF_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    }
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph = g;
    b->graph_exec = x;
    b->graph_myers_cells = b->cap_myers_cells;
    b->graph_affine_cells = b->cap_affine_cells;
    return SF_OK;
}

int sf_batch_replay(sf_batch *b, uint32_t times) {
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (b->capturing) return fail(SF_EINVAL, "the batch is capturing");
    if (!b->graph_exec) return fail(SF_EINVAL, "no captured step on this batch");
    int st;
    if ((st = ensure_device()) || (st = join_aux(b))) return st;
    hipEvent_t e0;
    if ((st = timed_begin(b, 4, &e0))) return st;
    for (uint32_t t = 0; t < times; ++t) HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
    if ((st = timed_end(b, 4, e0))) return st;
    b->myers_cells += b->graph_myers_cells * times;
    b->affine_cells += b->graph_affine_cells * times;
    b->graph_replays += times;
    return SF_OK;
}

int sf_batch_sync(sf_batch *b) {
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (const int st = refuse_in_capture(b, "sf_batch_sync")) return st;
    return quiesce(b);
}

int sf_batch_myers_results(sf_batch *b, uint32_t *dist, uint32_t *end, uint32_t *start) {
    if (b && b->capturing) return refuse_in_capture(b, "reading results");
    int st;
    if ((st = sf_batch_sync(b))) return st;
    const uint64_t n = b->n;
    if (dist) HIP_TRY(hipMemcpy(dist, b->d_dist, n * 4, hipMemcpyDeviceToHost));
    if (end) HIP_TRY(hipMemcpy(end, b->d_end, n * 4, hipMemcpyDeviceToHost));
    if (start) HIP_TRY(hipMemcpy(start, b->d_start, n * 4, hipMemcpyDeviceToHost));
    return SF_OK;
}

int sf_batch_affine_results(sf_batch *b, int32_t *score, uint32_t *score_col, uint32_t *tstart, uint32_t *cigar,
                            uint64_t cigar_cap, uint64_t *cigar_off) {
    int st;
    if ((st = sf_batch_sync(b))) return st;
    const uint64_t n = b->n;
    if (!b->d_score) return fail(SF_EINVAL, "no affine run on this batch");
    if (score) HIP_TRY(hipMemcpy(score, b->d_score, n * 4, hipMemcpyDeviceToHost));
    if (score_col) HIP_TRY(hipMemcpy(score_col, b->d_col, n * 4, hipMemcpyDeviceToHost));
    if (tstart || cigar || cigar_off) {
        if (!b->cigar_valid) return fail(SF_EINVAL, "the last affine run had no traceback");
        std::vector<uint32_t> ts(n);
        HIP_TRY(hipMemcpy(ts.data(), b->d_tstart, n * 4, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < n; ++i)
            if (ts[i] == kNone) return fail(SF_EDEVICE, "pair %llu: traceback left the band", (unsigned long long)i);
        if (tstart) std::memcpy(tstart, ts.data(), n * 4);
        std::vector<uint64_t> off(n + 1);
        HIP_TRY(hipMemcpy(off.data(), b->d_out_off, (n + 1) * 8, hipMemcpyDeviceToHost));
        if (cigar_off) std::memcpy(cigar_off, off.data(), (n + 1) * 8);
        if (cigar) {
            if (off[n] > cigar_cap)
                return fail(SF_ERANGE, "cigar buffer holds %llu ops, %llu needed", (unsigned long long)cigar_cap,
                            (unsigned long long)off[n]);
            HIP_TRY(hipMemcpy(cigar, b->d_cig, off[n] * 4, hipMemcpyDeviceToHost));
        }
    }
    return SF_OK;
}

int sf_batch_timing(sf_batch *b, sf_timing *t) {
    int st;
    if (!t) return fail(SF_EINVAL, "timing is null");
    if ((st = sf_batch_sync(b))) return st;
    b->collect();
    t->myers_ms = b->ms[0];
    t->locate_ms = b->ms[1];
    t->affine_ms = b->ms[2];
    t->walk_ms = b->ms[3];
    t->myers_cells = b->myers_cells;
    t->affine_cells = b->affine_cells;
    t->affine_launches = b->affine_launches;
    t->graph_ms = b->ms[4];
    t->graph_replays = b->graph_replays;
    return SF_OK;
}

void sf_batch_reset_timing(sf_batch *b) {
    if (!b || refuse_in_capture(b, "sf_batch_reset_timing")) return;
    (void)quiesce(b);
    b->collect();
    for (double &x : b->ms) x = 0;
    b->myers_cells = b->affine_cells = b->affine_launches = b->graph_replays = 0;
}

// ---- one-shot calls ----

// ---- passes with per-row bookkeeping (affine_dp with clip/cut/geom; local bands) ----

int sf_batch_affine_dp(sf_batch *b, int mode, const sf_scoring *sc, const uint64_t *band, const uint64_t *geom,
                       int has_clip, int32_t xdrop, int clip_cell, const uint64_t *cut) {
    if (b && b->capturing) return refuse_in_capture(b, "sf_batch_affine_dp");
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (!sc) return fail(SF_EINVAL, "scoring is null");
    if (mode < SF_INFIX || mode > SF_GLOBAL) return fail(SF_EINVAL, "mode %d", mode);
    if (sc->open < 0 || sc->ext < 0 || sc->mat < 0 || sc->mis < 0)
        return fail(SF_EINVAL, "scores must be non-negative (align.rs:2511 requires open, ext >= 0)");
    if (sc->mat > kScoreMax || sc->mis > kScoreMax || sc->open > kScoreMax || sc->ext > kScoreMax)
        return fail(SF_ERANGE, "scoring values above %d are outside this library's bound", kScoreMax);
    int st;
    if ((st = ensure_device())) return st;
    if ((st = join_aux(b))) return st;
    const uint64_t n = b->n;
    sf_batch::AffPlan &pl = b->aplan[2];
    const int packed = sf_affine_pass_packed();
    uint64_t key = (0x51ed270b27f1a3c5ull ^ (uint64_t)mode ^ (sf_affine_lp16_max() << 8) ^ ((uint64_t)packed << 61)) *
                   1099511628211ull;
    // (The scoring is in the key: the plan's range check and the packed classes depend on it.)
    for (int32_t v : {sc->mat, sc->mis, sc->open, sc->ext}) key = (key ^ (uint64_t)(uint32_t)v) * 1099511628211ull;
    key = hash_words(hash_words(key, band, n, SF_BAND_EXACT), geom, n, 0x9e3779b97f4a7c15ull);
    if (!pl.valid || pl.key != key) {
        std::vector<uint64_t> ba(n);
        std::vector<int> kcls(n);
        uint64_t cells = 0;
        for (uint64_t i = 0; i < n; ++i) {
            uint64_t below, above;
            const uint32_t m = b->plen[i], t = b->tlen[i];
            // align.rs:2464-2474: the band laid out for a text of geom bases.
            sf::band_geom_for(m, t, geom ? geom[i] : t, mode, band ? band[i] : SF_BAND_EXACT, &below, &above);
            const uint64_t width = below + above + 1;
            kcls[i] = sf_affine_pass_class(width, width <= sf_affine_lp16_max());
            // The 16-bit packed pass where its range holds: 8 lanes a pair up to 256 slots (the
            // per-row reductions shared by twice the slots), 16 up to 512.
            if (packed && kcls[i] != 0 && width <= sf_affine_lp16_max()) {
                int pc = 0;
                for (int lanes : {8, 16})
                    if (!pc && (lanes == 16 || packed == 1))
                        pc = sf_affine_pass_pk_class(width, lanes, sc->mat, sc->mis, sc->open, sc->ext, m, (uint32_t)below);
                if (pc) kcls[i] = pc;
            }
            if (kcls[i] == 0 && width <= kAffWideMax) kcls[i] = kAffWide;
            if (kcls[i] == 0)
                return fail(SF_ERANGE, "pair %llu: band of %llu cells per row; this library handles up to 2^26",
                            (unsigned long long)i, (unsigned long long)width);
            if (!sf::scores_in_range(sc->mat, sc->mis, sc->open, sc->ext, m, t,
                                     kcls[i] == kAffWide ? width : aff_slots(kcls[i])))
                return fail(SF_ERANGE, "pair %llu: scores could reach NEG", (unsigned long long)i);
            ba[i] = below | (above << 32);
            cells += sf::band_cells(m, t, below, above);
        }
        if ((st = build_plan(b, pl, ba, kcls, key, cells))) return st;
    }
    if (!b->d_dp && (st = dalloc(&b->d_dp, n))) return st;
    if ((st = ensure_wide_scratch(b, pl))) return st;
    // The cuts stay on the device for the next run with the same ones (an upload waits for the
    // batch's work, which may still read the previous ones).
    if (cut) {
        const uint64_t ck = hash_words(0x243f6a8885a308d3ull, cut, n, 0) | 1;
        if (ck != b->cut_key) {
            if ((st = quiesce(b))) return st;
            b->cut_key = 0;
            if (!b->d_cut && (st = dalloc(&b->d_cut, n))) return st;
            HIP_TRY(hipMemcpy(b->d_cut, cut, n * 8, hipMemcpyHostToDevice));
            b->cut_key = ck;
        }
    }
    uint64_t *d_cut = cut ? b->d_cut : nullptr;
    PassArgs a{};
    a.pv = view(b);
    a.mode = mode;
    a.mat = sc->mat, a.mis = sc->mis, a.open = sc->open, a.ext = sc->ext;
    a.has_clip = has_clip != 0;
    a.clip_cell = clip_cell != 0;
    a.xdrop = xdrop;
    a.cut = d_cut;
    a.out = b->d_dp;
    hipEvent_t e0;
    if ((st = timed_begin(b, 2, &e0))) return st;
    st = run_bins(b, pl.bins, [&](const sf_batch::Bin &bin, hipStream_t s) {
        if (bin.cls == kAffWide) {
            WideArgs y{};
            y.p = a;
            return launch_wide(b, pl, bin, y, s);
        }
        PassArgs x = a;
        x.order = pl.d_order + bin.begin;
        x.desc = pl.d_desc + bin.begin;
        x.count = bin.end - bin.begin;
        return sf_launch_pass(x, bin.cls, false, s);
    }, nullptr, nullptr, nullptr, true);
    if (!st) st = timed_end(b, 2, e0);
    if (st) return st;
    b->affine_cells += pl.cells;
    b->affine_launches += pl.bins.size();
    return SF_OK;
}

int sf_batch_dp_results(sf_batch *b, sf_dp *out) {
    if (b && b->capturing) return refuse_in_capture(b, "reading results");
    int st;
    if ((st = sf_batch_sync(b))) return st;
    if (!out) return fail(SF_EINVAL, "out is null");
    if (!b->d_dp) return fail(SF_EINVAL, "no sf_batch_affine_dp run on this batch");
    static_assert(sizeof(sf_dp) == sizeof(DpOut), "sf_dp and the device record agree");
    HIP_TRY(hipMemcpy(out, b->d_dp, b->n * sizeof(sf_dp), hipMemcpyDeviceToHost));
    return SF_OK;
}

int sf_batch_local_band(sf_batch *b, const sf_scoring *sc, const int64_t *dlo, const int64_t *dhi, int32_t bonus_start,
                        int32_t bonus_end) {
    if (b && b->capturing) return refuse_in_capture(b, "sf_batch_local_band");
    if (!b) return fail(SF_EINVAL, "batch is null");
    if (!sc || !dlo || !dhi) return fail(SF_EINVAL, "scoring and diagonals are required");
    if (sc->open < 0 || sc->ext < 0 || sc->mat < 0 || sc->mis < 0)
        return fail(SF_EINVAL, "scores must be non-negative (align.rs:2511 requires open, ext >= 0)");
    if (sc->mat > kScoreMax || sc->mis > kScoreMax || sc->open > kScoreMax || sc->ext > kScoreMax)
        return fail(SF_ERANGE, "scoring values above %d are outside this library's bound", kScoreMax);
    if (bonus_start < -(1 << 24) || bonus_start > (1 << 24) || bonus_end < -(1 << 24) || bonus_end > (1 << 24))
        return fail(SF_ERANGE, "end bonuses beyond +-2^24 are outside this library's bound");
    int st;
    if ((st = ensure_device())) return st;
    if ((st = join_aux(b))) return st;
    const uint64_t n = b->n;
    sf_batch::AffPlan &pl = b->aplan[3];
    uint64_t key = (0x2545f4914f6cdd1dull ^ (sf_affine_lp16_max() << 8)) * 1099511628211ull;
    key = hash_words(hash_words(key, reinterpret_cast<const uint64_t *>(dlo), n, 0),
                     reinterpret_cast<const uint64_t *>(dhi), n, 0);
    if (!pl.valid || pl.key != key) {
        std::vector<uint64_t> ba(n);
        std::vector<int> kcls(n);
        uint64_t cells = 0;
        for (uint64_t i = 0; i < n; ++i) {
            const uint32_t m = b->plen[i], t = b->tlen[i];
            // Diagonals dlo..=dhi (j - i): row i covers [i + dlo, i + dhi] within [0, n]
            // (align.rs:790-795) -- the band below = -dlo, above = dhi.  An empty range scores 0.
            const int64_t lo = dlo[i], hi = dhi[i] < lo ? lo : dhi[i];
            if (lo < -(int64_t)(1u << 30) || hi > (int64_t)(1u << 30))
                return fail(SF_ERANGE, "pair %llu: diagonals beyond +-2^30", (unsigned long long)i);
            const uint64_t width = (uint64_t)(hi - lo + 1);
            kcls[i] = sf_affine_pass_class(width, width <= sf_affine_lp16_max());
            if (kcls[i] == 0 && width <= kAffWideMax) kcls[i] = kAffWide;
            if (kcls[i] == 0)
                return fail(SF_ERANGE, "pair %llu: band of %llu diagonals; this library handles up to 2^26",
                            (unsigned long long)i, (unsigned long long)width);
            if (!sf::scores_in_range(sc->mat, sc->mis, sc->open, sc->ext, m, t,
                                     kcls[i] == kAffWide ? width : aff_slots(kcls[i])))
                return fail(SF_ERANGE, "pair %llu: scores could reach NEG", (unsigned long long)i);
            const uint32_t below = (uint32_t)(int32_t)(-lo), above = (uint32_t)(int32_t)(dhi[i] < lo ? lo - 1 : dhi[i]);
            ba[i] = (uint64_t)below | ((uint64_t)above << 32);
            for (uint64_t r = 1; r <= m && dhi[i] >= lo; ++r) {
                const int64_t a0 = (int64_t)r + lo < 0 ? 0 : (int64_t)r + lo;
                const int64_t a1 = (int64_t)r + dhi[i] > (int64_t)t ? (int64_t)t : (int64_t)r + dhi[i];
                if (a1 >= a0) cells += (uint64_t)(a1 - a0 + 1);
            }
        }
        if ((st = build_plan(b, pl, ba, kcls, key, cells))) return st;
    }
    if (!b->d_best && (st = dalloc(&b->d_best, n))) return st;
    if ((st = ensure_wide_scratch(b, pl))) return st;
    PassArgs a{};
    a.pv = view(b);
    a.mode = SF_PREFIX;
    a.mat = sc->mat, a.mis = sc->mis, a.open = sc->open, a.ext = sc->ext;
    a.bonus0 = bonus_start;
    a.bonus1 = bonus_end;
    a.pinv = b->d_pinv;
    a.best = b->d_best;
    hipEvent_t e0;
    if ((st = timed_begin(b, 2, &e0))) return st;
    if ((st = run_bins(b, pl.bins, [&](const sf_batch::Bin &bin, hipStream_t s) {
             if (bin.cls == kAffWide) {
                 WideArgs y{};
                 y.p = a;
                 y.local = 1;
                 return launch_wide(b, pl, bin, y, s);
             }
             PassArgs x = a;
             x.order = pl.d_order + bin.begin;
             x.desc = pl.d_desc + bin.begin;
             x.count = bin.end - bin.begin;
             return sf_launch_pass(x, bin.cls, true, s);
         })))
        return st;
    if ((st = timed_end(b, 2, e0))) return st;
    b->affine_cells += pl.cells;
    b->affine_launches += pl.bins.size();
    return SF_OK;
}

int sf_batch_local_results(sf_batch *b, int32_t *best) {
    if (b && b->capturing) return refuse_in_capture(b, "reading results");
    int st;
    if ((st = sf_batch_sync(b))) return st;
    if (!best) return fail(SF_EINVAL, "best is null");
    if (!b->d_best) return fail(SF_EINVAL, "no sf_batch_local_band run on this batch");
    HIP_TRY(hipMemcpy(best, b->d_best, b->n * 4, hipMemcpyDeviceToHost));
    return SF_OK;
}

namespace {
int32_t sat_add(int32_t a, int32_t b) {
    const int64_t r = (int64_t)a + b;
    return r > INT32_MAX ? INT32_MAX : (r < INT32_MIN ? INT32_MIN : (int32_t)r);
}

// One pass per pair on a temporary batch: the batch forms of the reference's per-call wrappers.
int run_dp(const sf_pairs *pairs, int mode, const sf_scoring *sc, const uint64_t *band, const uint64_t *geom,
           int has_clip, int32_t xdrop, int clip_cell, const uint64_t *cut, std::vector<sf_dp> &out) {
    out.assign(pairs ? pairs->n : 0, sf_dp{});
    if (pairs && pairs->n == 0) return check_pairs(pairs);
    sf_batch *b;
    int st = sf_batch_create(pairs, &b);
    if (st) return st;
    st = sf_batch_affine_dp(b, mode, sc, band, geom, has_clip, xdrop, clip_cell, cut);
    if (!st) st = sf_batch_dp_results(b, out.data());
    sf_batch_destroy(b);
    return st;
}
} // namespace

int sf_affine_dp(const sf_pairs *pairs, int mode, const sf_scoring *sc, const uint64_t *band, const uint64_t *geom,
                 int has_clip, int32_t xdrop, int clip_cell, const uint64_t *cut, sf_dp *out) {
    if (!out) return fail(SF_EINVAL, "out is null");
    std::vector<sf_dp> r;
    const int st = run_dp(pairs, mode, sc, band, geom, has_clip, xdrop, clip_cell, cut, r);
    if (!st) std::copy(r.begin(), r.end(), out);
    return st;
}

int sf_affine_extend(const sf_pairs *pairs, const sf_scoring *sc, const uint64_t *band, int32_t end_bonus,
                     int32_t xdrop, int32_t *score) {
    if (!score) return fail(SF_EINVAL, "score is null");
    std::vector<sf_dp> r;
    const int st = run_dp(pairs, SF_PREFIX, sc, band, nullptr, 1, xdrop, 0, nullptr, r);
    if (st) return st;
    for (size_t i = 0; i < r.size(); ++i) score[i] = std::max(r[i].clip_best, sat_add(r[i].score, end_bonus));
    return SF_OK;
}

int sf_affine_extend_window(const sf_pairs *pairs, const uint64_t *window, const sf_scoring *sc, const uint64_t *band,
                            int32_t end_bonus, int32_t xdrop, sf_extend_window *out) {
    if (!out || !window) return fail(SF_EINVAL, "window and out are required");
    const uint64_t n = pairs ? pairs->n : 0;
    std::vector<uint64_t> geom(n);
    for (uint64_t i = 0; i < n; ++i) geom[i] = std::max<uint64_t>(window[i], pairs->txt_len[i]);
    std::vector<sf_dp> r;
    const int st = run_dp(pairs, SF_PREFIX, sc, band, geom.data(), 1, xdrop, 1, nullptr, r);
    if (st) return st;
    for (uint64_t i = 0; i < n; ++i) {
        const int32_t whole = sat_add(r[i].score, end_bonus);
        if (r[i].clip_best >= whole)
            out[i] = sf_extend_window{r[i].clip_best, r[i].clip_i, r[i].clip_j, r[i].text_read};
        else
            out[i] = sf_extend_window{whole, pairs->pat_len[i], r[i].score_col, r[i].text_read};
    }
    return SF_OK;
}

int sf_affine_extend_cut_bonus(const sf_pairs *pairs, const sf_scoring *sc, const uint64_t *band, int32_t end_bonus,
                               int32_t xdrop, int diagonal, const uint64_t *window, const uint64_t *cut, int track_cut,
                               int bonus_within_cut, sf_extend_cut *out) {
    if (!out || !window || !cut) return fail(SF_EINVAL, "window, cut and out are required");
    const uint64_t n = pairs ? pairs->n : 0;
    std::vector<uint64_t> geom(n);
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t m = pairs->pat_len[i], t = pairs->txt_len[i];
        geom[i] = diagonal ? std::min(m, t) : std::max(t, window[i]);
    }
    // align.rs:924-933: the best cell is tracked only with track_cut and a cut.
    // align.rs:924-933: the best cell is tracked only with track_cut and a cut; untracked, it stays
    // (0, 0).  Tracking changes no score, so one pass tracking it for all pairs serves both.
    std::vector<sf_dp> r(n);
    const int st = n ? run_dp(pairs, SF_PREFIX, sc, band, geom.data(), 1, xdrop, track_cut ? 1 : 0, cut, r) : SF_OK;
    if (st) return st;
    for (uint64_t i = 0; i < n; ++i) {
        sf_dp &d = r[i];
        if (!(track_cut && cut[i] != SF_BAND_EXACT)) d.clip_i = d.clip_j = 0;
        const int32_t end = bonus_within_cut && cut[i] != SF_BAND_EXACT ? std::max(d.score, sat_add(d.cut_score, end_bonus))
                                                                       : sat_add(d.score, end_bonus);
        const int32_t whole = std::max(d.clip_best, end);
        const uint32_t cols = d.clip_best >= end ? d.clip_j : d.score_col;
        const int32_t cutv = cut[i] == SF_BAND_EXACT ? whole : std::max(d.cut_best, sat_add(d.cut_score, end_bonus));
        out[i] = sf_extend_cut{whole, cutv, cols};
    }
    return SF_OK;
}

int sf_local_band_ends(const sf_pairs *pairs, const sf_scoring *sc, const int64_t *dlo, const int64_t *dhi,
                       int32_t bonus_start, int32_t bonus_end, int32_t *best) {
    if (!best) return fail(SF_EINVAL, "best is null");
    if (pairs && pairs->n == 0) return check_pairs(pairs);
    sf_batch *b;
    int st = sf_batch_create(pairs, &b);
    if (st) return st;
    st = sf_batch_local_band(b, sc, dlo, dhi, bonus_start, bonus_end);
    if (!st) st = sf_batch_local_results(b, best);
    sf_batch_destroy(b);
    return st;
}

// sf_myers with a threshold per pair (distances only): the pairs' lane classes run Ukkonen's band.
static int myers_thresholds_run(sf_batch *b, int mode, const uint32_t *k, uint32_t *dk, uint32_t *dist) {
    if (!b->myers_unsupported.empty()) return fail(SF_ERANGE, "%s", b->myers_unsupported.c_str());
    int st;
    if ((st = ensure_device()) || (st = join_aux(b))) return st;
    HIP_TRY(hipMemcpyAsync(dk, k, b->n * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    MyersArgs a{};
    a.pv = view(b);
    a.mode = mode;
    a.k = UINT32_MAX;
    a.kvec = dk;
    a.dist = b->d_dist;
    if ((st = run_myers(b, a))) return st;
    return sf_batch_myers_results(b, dist, nullptr, nullptr);
}

static int myers_thresholds(const sf_pairs *pairs, int mode, const uint32_t *k, uint32_t *dist) {
    sf_batch *b;
    int st = sf_batch_create(pairs, &b);
    if (st) return st;
    uint32_t *dk = nullptr;
    if (!(st = dalloc(&dk, pairs->n))) st = myers_thresholds_run(b, mode, k, dk, dist);
    sf_batch_destroy(b); // (waits for the batch's work)
    if (dk) (void)hipFree(dk);
    return st;
}

int sf_myers_banded(const sf_pairs *pairs, const uint64_t *group_off, uint64_t ngroups, int mode, uint32_t cap,
                    int distinct, uint32_t *out) {
    int st = check_pairs(pairs);
    if (st) return st;
    if (!group_off || (pairs->n && !out)) return fail(SF_EINVAL, "null argument");
    if (mode < SF_INFIX || mode > SF_GLOBAL) return fail(SF_EINVAL, "mode %d", mode);
    const uint64_t n = pairs->n;
    if (group_off[0] != 0 || group_off[ngroups] != n) return fail(SF_EINVAL, "groups must cover the pairs");
    for (uint64_t g = 0; g < ngroups; ++g)
        if (group_off[g + 1] < group_off[g]) return fail(SF_EINVAL, "group offsets must not decrease");
    if (n == 0) return SF_OK;
    // Distinct texts of each group, in first-occurrence order (lib.rs:862-879).  Texts compare by
    // their codes masked & 3: the reference compares bytes, but texts whose bytes differ only
    // above bit 1 have the same distance, and a repeat and a distinct text of equal distance take
    // the same value in order (k + 1 > best either way), so the results are the same.
    std::vector<uint64_t> first(n); // pair -> the pair holding its group's first occurrence
    auto hash = [&](uint64_t i) {
        uint64_t h = 1469598103934665603ull ^ pairs->txt_len[i];
        for (uint32_t q = 0; q < pairs->txt_len[i]; ++q)
            h = (h ^ code_at(pairs->txt, pairs->packed, pairs->txt_off[i] + q)) * 1099511628211ull;
        return h;
    };
    auto same = [&](uint64_t i, uint64_t j) {
        if (pairs->txt_len[i] != pairs->txt_len[j]) return false;
        for (uint32_t q = 0; q < pairs->txt_len[i]; ++q)
            if (code_at(pairs->txt, pairs->packed, pairs->txt_off[i] + q) !=
                code_at(pairs->txt, pairs->packed, pairs->txt_off[j] + q))
                return false;
        return true;
    };
    std::vector<uint64_t> uniq, group_head(n); // group_head: the pair holding its group's first text
    auto first_of_group = [&](uint64_t i) { return group_head[i] == i; };
    for (uint64_t g = 0; g < ngroups; ++g) {
        std::unordered_multimap<uint64_t, uint64_t> seen;
        for (uint64_t i = group_off[g]; i < group_off[g + 1]; ++i) {
            group_head[i] = group_off[g];
            first[i] = i;
            if (!distinct) {
                const uint64_t h = hash(i);
                auto r = seen.equal_range(h);
                for (auto it = r.first; it != r.second; ++it)
                    if (same(it->second, i)) {
                        first[i] = it->second;
                        break;
                    }
                if (first[i] == i) seen.emplace(h, i);
            }
            if (first[i] == i) uniq.push_back(i);
        }
    }
    // The distinct texts' distances on the device, in two passes.  A group's first text runs
    // unbounded (its threshold is u32::MAX + cap, saturated); every later one at that distance +
    // cap, which no later threshold of the group exceeds (the best only falls).  distance_with is
    // exact wherever it is within its k, so each result is exact or known to lie above every
    // threshold the text meets in order, and the second pass computes only Ukkonen's bands.
    auto sat = [](uint32_t a, uint32_t b) { return a > UINT32_MAX - b ? UINT32_MAX : a + b; };
    std::vector<uint32_t> exact(n, 0);
    auto pass = [&](const std::vector<uint64_t> &ids, const std::vector<uint32_t> *kv) {
        if (ids.empty()) return SF_OK;
        std::vector<uint64_t> poff(ids.size()), toff(ids.size());
        std::vector<uint32_t> plen(ids.size()), tlen(ids.size()), d(ids.size());
        for (size_t q = 0; q < ids.size(); ++q) {
            const uint64_t i = ids[q];
            poff[q] = pairs->pat_off[i], plen[q] = pairs->pat_len[i];
            toff[q] = pairs->txt_off[i], tlen[q] = pairs->txt_len[i];
        }
        sf_pairs sub = *pairs;
        sub.pat_off = poff.data(), sub.pat_len = plen.data(), sub.txt_off = toff.data(), sub.txt_len = tlen.data();
        sub.n = ids.size();
        int r = kv ? myers_thresholds(&sub, mode, kv->data(), d.data()) : sf_myers(&sub, mode, UINT32_MAX, d.data(),
                                                                                    nullptr, nullptr);
        if (r) return r;
        for (size_t q = 0; q < ids.size(); ++q) exact[ids[q]] = d[q];
        return SF_OK;
    };
    std::vector<uint64_t> heads, rest;
    std::vector<uint32_t> kv;
    for (uint64_t i : uniq) (first_of_group(i) ? heads : rest).push_back(i);
    if ((st = pass(heads, nullptr))) return st;
    for (uint64_t i : rest) kv.push_back(sat(exact[group_head[i]], cap));
    if ((st = pass(rest, &kv))) return st;
    // The sequential values (lib.rs:3016-3034, equal to lib.rs:880-910 by its own argument).  A text
    // above its pass's threshold (SF_NONE) is above every threshold it meets here.
    for (uint64_t g = 0; g < ngroups; ++g) {
        uint32_t best = UINT32_MAX;
        for (uint64_t i = group_off[g]; i < group_off[g + 1]; ++i) {
            const uint32_t k = sat(best, cap);
            if (first[i] != i) {
                const uint32_t v = out[first[i]];
                out[i] = v <= k ? v : sat(k, 1);
                continue;
            }
            const uint32_t v = exact[i] <= k ? exact[i] : sat(k, 1);
            best = std::min(best, v);
            out[i] = v;
        }
    }
    return SF_OK;
}

int sf_myers(const sf_pairs *pairs, int mode, uint32_t k, uint32_t *dist, uint32_t *end, uint32_t *start) {
    if (!dist) return fail(SF_EINVAL, "dist is null");
    if (start && !end) return fail(SF_EINVAL, "start needs end");
    if (pairs && pairs->n == 0) return check_pairs(pairs);
    sf_batch *b;
    int st = sf_batch_create(pairs, &b);
    if (st) return st;
    st = sf_batch_myers(b, mode, k, start ? 2 : end ? 1 : 0);
    if (!st) st = sf_batch_myers_results(b, dist, end, start);
    sf_batch_destroy(b);
    return st;
}

int sf_affine(const sf_pairs *pairs, int mode, const sf_scoring *sc, const uint64_t *band, int32_t *score,
              uint32_t *score_col) {
    if (!score) return fail(SF_EINVAL, "score is null");
    if (pairs && pairs->n == 0) return check_pairs(pairs);
    sf_batch *b;
    int st = sf_batch_create(pairs, &b);
    if (st) return st;
    st = sf_batch_affine(b, mode, sc, band, 0);
    if (!st) st = sf_batch_affine_results(b, score, score_col, nullptr, nullptr, 0, nullptr);
    sf_batch_destroy(b);
    return st;
}

int sf_affine_cigar(const sf_pairs *pairs, int mode, const sf_scoring *sc, const uint64_t *band, int32_t *score,
                    uint32_t *score_col, uint32_t *tstart, uint32_t *cigar, uint64_t cigar_cap, uint64_t *cigar_off) {
    if (!score || !cigar_off) return fail(SF_EINVAL, "score and cigar_off are required");
    if (pairs && pairs->n == 0) {
        if (cigar_off) cigar_off[0] = 0;
        return check_pairs(pairs);
    }
    sf_batch *b;
    int st = sf_batch_create(pairs, &b);
    if (st) return st;
    st = sf_batch_affine(b, mode, sc, band, 1);
    if (!st) st = sf_batch_affine_results(b, score, score_col, tstart, cigar, cigar_cap, cigar_off);
    sf_batch_destroy(b);
    return st;
}

} // extern "C"

// A batch over sequences already on the device (declared in internal.hpp).
int sf_batch_create_resident(const uint32_t *seq, uint64_t n, const uint32_t *plen, const uint32_t *tlen,
                             const uint64_t *pbase, const uint64_t *tbase, sf_batch **out) {
    *out = nullptr;
    if (int st = ensure_device()) return st;
    auto *b = new sf_batch();
    b->n = n;
    b->plen.assign(plen, plen + n);
    b->tlen.assign(tlen, tlen + n);
    b->d_seq = const_cast<uint32_t *>(seq);
    b->seq_borrowed = true;
    if (int st = batch_init(b, std::vector<uint64_t>(pbase, pbase + n), std::vector<uint64_t>(tbase, tbase + n))) {
        delete b;
        return st;
    }
    *out = b;
    return SF_OK;
}
