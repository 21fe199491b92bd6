// noalloc.hip -- sf_myers_noalloc: sf_myers without any allocation, for callers that cannot allocate
// (harnesses that forbid it, tight serving loops).  No counterpart in the reference.
//
// The device memory a call needs beyond the caller's arrays -- the codes as given, the 2-bit packed
// words, the pair records, the launch order, the class cursors and the results before they are
// copied out -- lives in `scratch`, one file-scope device array of kScratchBytes.  Each GPU holds
// its own copy of it from the first call until the process ends.  Host code never uses the array's
// name as an address (on the host it only names the symbol): the first call on a device takes the
// array's device address with hipGetSymbolAddress, keeps it per device, and every kernel argument,
// copy and offset after that is derived from that address.
//
// A call: the caller's codes and per-pair arrays are copied into the array as they are; one kernel
// packs the codes into words (masked & 3, as align.rs:78, 120); the host counts the pairs of each
// Myers shape class (the classes of sf_batch_myers, by blocks per pattern) into arrays on its stack;
// one kernel sets each class's cursor to its first slot and another scatters the pair indices into
// their class's range with atomic cursors; then the pair records, one launch per class (the same
// kernels as sf_myers), the reversed pass for start columns when asked, and the copies back.  Calls
// on one GPU share its array, so overlapping calls would race on it: the library serialises them
// with a lock per device (a call waits until the one before it has copied its results out).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/sf_align.h"
#include "internal.hpp"

using namespace sf;

int sf_runtime_fail(int code, const char *msg); // runtime.hip
int sf_runtime_ensure();                        // runtime.hip
int sf_runtime_device();                        // runtime.hip

namespace {

constexpr uint64_t kScratchBytes = 32ull << 20;
constexpr int kMaxDevices = 64;
constexpr int kMaxClasses = 32;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

// Per device: the array's device address (set once, under g_mu) and the lock serialising calls.
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};
std::mutex g_call_mu[kMaxDevices];

// The Myers shape classes as ranges of blocks per pattern: class q holds patterns of
// (maxblk[q - 1], maxblk[q]] blocks; its pairs take order slots [begin[q], begin[q] + count[q]).
struct ClassTable {
    uint32_t ncls;
    uint32_t maxblk[kMaxClasses];
    uint32_t begin[kMaxClasses];
};

// Words of the packed sequences from codes one a byte: word w of a buffer is codes 16w..16w+15.
__global__ __launch_bounds__(256) void pack_kernel(const uint8_t *codes, uint64_t nbases, uint32_t *words,
                                                   uint64_t nwords) {
    const uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (w >= nwords) return;
    uint32_t x = 0;
    const uint64_t k0 = w * 16;
    for (uint32_t k = 0; k < 16 && k0 + k < nbases; ++k) x |= (uint32_t)(codes[k0 + k] & 3u) << (2 * k);
    words[w] = x;
}

__global__ void cursor_kernel(ClassTable t, uint32_t *cursor) {
    const uint32_t q = threadIdx.x;
    if (q < t.ncls) cursor[q] = t.begin[q];
}

// Pair i's text offset moved past the pattern words, and its index into its class's range.
__global__ __launch_bounds__(256) void scatter_kernel(uint64_t n, ClassTable t, const uint32_t *plen, uint64_t *toff,
                                                      uint64_t text_base, uint32_t *cursor, uint32_t *order) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    toff[i] += text_base;
    const uint32_t blocks = (plen[i] + 63) / 64;
    uint32_t q = 0;
    while (q + 1 < t.ncls && blocks > t.maxblk[q]) ++q;
    order[atomicAdd(&cursor[q], 1u)] = (uint32_t)i;
}

// The class table of sf_myers_maxb_for: every class with its largest block count.
void class_table(ClassTable &t, int *code) {
    t.ncls = 0;
    int cur = sf_myers_maxb_for(1);
    for (uint32_t b = 2;; ++b) {
        const int c = sf_myers_maxb_for(b);
        if (c != cur) {
            t.maxblk[t.ncls] = b - 1;
            code[t.ncls++] = cur;
            if (c == 0) break;
            cur = c;
        }
    }
}

inline uint64_t up256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

int dev_fail(const char *what, hipError_t e) {
    char msg[200];
    std::snprintf(msg, sizeof msg, "sf_myers_noalloc: %s: %s", what, hipGetErrorString(e));
    return sf_runtime_fail(SF_EDEVICE, msg);
}

} // namespace

extern "C" int sf_myers_noalloc(const sf_pairs *p, int mode, uint32_t k, uint32_t *dist, uint32_t *end,
                                uint32_t *start) {
    if (!p) return sf_runtime_fail(SF_EINVAL, "pairs is null");
    if (mode < SF_INFIX || mode > SF_GLOBAL) return sf_runtime_fail(SF_EINVAL, "mode");
    if (start && !end) return sf_runtime_fail(SF_EINVAL, "start needs end");
    const uint64_t n = p->n;
    if (n == 0) return SF_OK;
    if (!dist) return sf_runtime_fail(SF_EINVAL, "dist is null");
    if (!p->pat || !p->txt || !p->pat_off || !p->pat_len || !p->txt_off || !p->txt_len)
        return sf_runtime_fail(SF_EINVAL, "null sequence or offset pointer");
    if (p->packed != 0 && p->packed != 1) return sf_runtime_fail(SF_EINVAL, "packed must be 0 or 1");
    if (const int st = sf_runtime_ensure()) return st;
    const int dev = sf_runtime_device();
    if (dev < 0 || dev >= kMaxDevices) return sf_runtime_fail(SF_ENODEV, "device index");

    // Sizes, classes and the layout of the array, on the host's stack.
    ClassTable t{};
    int code[kMaxClasses];
    class_table(t, code);
    uint32_t count[kMaxClasses] = {};
    for (uint64_t i = 0; i < n; ++i) {
        if (p->pat_off[i] + p->pat_len[i] > p->pat_bases || p->pat_off[i] + p->pat_len[i] < p->pat_off[i] ||
            p->txt_off[i] + p->txt_len[i] > p->txt_bases || p->txt_off[i] + p->txt_len[i] < p->txt_off[i])
            return sf_runtime_fail(SF_EINVAL, "a sequence lies outside its buffer");
        if (p->txt_len[i] >= (1u << 30)) return sf_runtime_fail(SF_ERANGE, "sequences of 2^30 bases or more");
        const uint32_t blocks = (p->pat_len[i] + 63) / 64;
        if (blocks > t.maxblk[t.ncls - 1])
            return sf_runtime_fail(SF_ERANGE, "sf_myers_noalloc: a pattern over the largest Myers class");
        uint32_t q = 0;
        while (blocks > t.maxblk[q]) ++q;
        ++count[q];
    }
    for (uint32_t q = 0, b = 0; q < t.ncls; b += count[q], ++q) t.begin[q] = b;
    const uint64_t pw = (p->pat_bases + 15) / 16, tw = (p->txt_bases + 15) / 16;
    const uint64_t raw = p->packed ? 0 : up256(p->pat_bases) + up256(p->txt_bases);
    uint64_t at = 0;
    auto region = [&](uint64_t bytes) {
        const uint64_t r = at;
        at += up256(bytes);
        return r;
    };
    const uint64_t o_raw = region(raw), o_seq = region((pw + tw) * 4), o_poff = region(n * 8), o_toff = region(n * 8),
                   o_plen = region(n * 4), o_tlen = region(n * 4), o_order = region(n * 4),
                   o_desc = region(n * sizeof(PairDesc)), o_dist = region(n * 4), o_end = region(n * 4),
                   o_start = region(n * 4), o_cur = region(kMaxClasses * 4);
    if (at > kScratchBytes) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "sf_myers_noalloc: the call needs %llu bytes of device scratch; its limit is %llu (32 MiB)",
                      (unsigned long long)at, (unsigned long long)kScratchBytes);
        return sf_runtime_fail(SF_ERANGE, msg);
    }

    std::lock_guard<std::mutex> call(g_call_mu[dev]);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_scratch[dev]) {
            void *ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) return dev_fail("hipGetSymbolAddress(scratch)", e);
            g_scratch[dev] = ptr;
        }
    }
    uint8_t *base = static_cast<uint8_t *>(g_scratch[dev]);
    uint32_t *seq = reinterpret_cast<uint32_t *>(base + o_seq);
    uint64_t *poff = reinterpret_cast<uint64_t *>(base + o_poff), *toff = reinterpret_cast<uint64_t *>(base + o_toff);
    uint32_t *plen = reinterpret_cast<uint32_t *>(base + o_plen), *tlen = reinterpret_cast<uint32_t *>(base + o_tlen);
    uint32_t *order = reinterpret_cast<uint32_t *>(base + o_order), *cur = reinterpret_cast<uint32_t *>(base + o_cur);
    PairDesc *desc = reinterpret_cast<PairDesc *>(base + o_desc);
    uint32_t *d_dist = reinterpret_cast<uint32_t *>(base + o_dist), *d_end = reinterpret_cast<uint32_t *>(base + o_end),
             *d_start = reinterpret_cast<uint32_t *>(base + o_start);

    // Everything runs on the null stream, in order; the copies from the caller's pageable memory
    // and back are synchronous.
    hipStream_t s = nullptr;
    hipError_t e = hipSuccess;
    auto run = [&](const char *what, hipError_t x) {
        if (e == hipSuccess && x != hipSuccess) {
            e = x;
            return dev_fail(what, x);
        }
        return SF_OK;
    };
    int st = SF_OK;
    if (p->packed) {
        st = run("copy patterns", hipMemcpy(seq, p->pat, pw * 4, hipMemcpyHostToDevice));
        if (!st) st = run("copy texts", hipMemcpy(seq + pw, p->txt, tw * 4, hipMemcpyHostToDevice));
    } else {
        uint8_t *rp = base + o_raw, *rt = rp + up256(p->pat_bases);
        st = run("copy patterns", hipMemcpy(rp, p->pat, p->pat_bases, hipMemcpyHostToDevice));
        if (!st) st = run("copy texts", hipMemcpy(rt, p->txt, p->txt_bases, hipMemcpyHostToDevice));
        if (!st && pw) {
            hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((pw + 255) / 256)), dim3(256), 0, s, rp, p->pat_bases, seq,
                               pw);
            st = run("pack", hipGetLastError());
        }
        if (!st && tw) {
            hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((tw + 255) / 256)), dim3(256), 0, s, rt, p->txt_bases,
                               seq + pw, tw);
            st = run("pack", hipGetLastError());
        }
    }
    if (!st) st = run("copy offsets", hipMemcpy(poff, p->pat_off, n * 8, hipMemcpyHostToDevice));
    if (!st) st = run("copy offsets", hipMemcpy(toff, p->txt_off, n * 8, hipMemcpyHostToDevice));
    if (!st) st = run("copy lengths", hipMemcpy(plen, p->pat_len, n * 4, hipMemcpyHostToDevice));
    if (!st) st = run("copy lengths", hipMemcpy(tlen, p->txt_len, n * 4, hipMemcpyHostToDevice));
    if (!st) {
        hipLaunchKernelGGL(cursor_kernel, dim3(1), dim3(kMaxClasses), 0, s, t, cur);
        st = run("cursors", hipGetLastError());
    }
    if (!st) {
        hipLaunchKernelGGL(scatter_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, t, plen, toff,
                           pw * 16, cur, order);
        st = run("scatter", hipGetLastError());
    }
    const PairsView pv{seq, poff, toff, plen, tlen};
    if (!st) st = run("pair records", sf_launch_desc(n, order, pv, nullptr, desc, s));
    MyersArgs a{};
    a.pv = pv;
    a.mode = mode;
    a.k = k;
    a.dist = d_dist;
    a.end = end ? d_end : nullptr;
    a.start = start ? d_start : nullptr;
    for (int pass = 0; pass < (start && mode == SF_INFIX ? 2 : 1) && !st; ++pass) {
        a.reverse = pass;
        for (uint32_t q = 0; q < t.ncls && !st; ++q) {
            if (!count[q]) continue;
            MyersArgs x = a;
            x.order = order + t.begin[q];
            x.desc = desc + t.begin[q];
            x.count = count[q];
            st = run("Myers launch", sf_launch_myers(x, code[q], s));
        }
    }
    if (!st) st = run("results", hipMemcpy(dist, d_dist, n * 4, hipMemcpyDeviceToHost));
    if (!st && end) st = run("results", hipMemcpy(end, d_end, n * 4, hipMemcpyDeviceToHost));
    // (The forward pass writes the start of Prefix and Global alignments, and of pairs over k.)
    if (!st && start) st = run("results", hipMemcpy(start, d_start, n * 4, hipMemcpyDeviceToHost));
    return st;
}
