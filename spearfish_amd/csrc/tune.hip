// tune.hip -- the autotuner behind the affine launch planner (sf_batch_autotune, sf_tune_*).
//
// The planner (choose_shapes, runtime.hip) chooses per pair a shape class -- lanes per pair (8,
// 16, 32 or 64), slots per lane (the tile: which class edge a band falls under), 16-bit packed or
// 32-bit -- from two costs per class: its throughput, ns per pair-row when the class fills the chip,
// and its chain, ns per row of one wave running alone (the latency that bounds a class with few
// pairs).  Without measurements those come from an instruction-count estimate; this file measures
// them on the device and keeps them per class, for every later plan in the process, and saves and
// loads them as text so that a process can start from a previous one's measurements.
//
// The planner's thresholds compare chains with spreads, and a lone wave's measured chain is shorter
// than the same wave's rows among others on a loaded SIMD; so the tuner also times the batch's own
// score run (sf_batch_autotune, runtime.hip) under the estimates and under the measurements with
// the chain scaled by 1, 1.5, 2 and 3, and keeps the fastest of those settings (the estimates among
// them, so tuning never keeps a setting measured slower than the untuned one).
//
// A measurement runs the class's score kernel (the same kernels runs use) over synthetic pairs of
// 512 random bases against 512 random bases: a kernel's time per row depends on its class (slots
// per lane, lanes per pair, packing), not on the bases or the band inside the class, so one set of
// pairs serves every class.  Throughput: 8,192 waves' worth of pairs (several times what the chip
// holds at once), best of two launches after a warm-up.  Chain: one wave's pairs, best of three.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sf_align.h"
#include "internal.hpp"

using namespace sf;

int sf_runtime_fail(int code, const char *msg); // runtime.hip
int sf_runtime_ensure();                        // runtime.hip

namespace {

struct Cost {
    double tput_ns;  // ns per pair-row with the chip full
    double chain_ns; // ns per row of one wave alone
};

std::mutex g_mu;
std::map<int, Cost> g_table;
bool g_env_loaded = false;
// The setting the planner uses: measurements on (chain x factor) or off (the estimates), and a
// generation count that changes with the table or the setting (part of a plan's cache key).
bool g_use = true;
double g_chain_factor = 1.0;
uint64_t g_generation = 1;

constexpr uint32_t kRows = 512;
constexpr uint32_t kPairs = 65536; // 8,192 waves of the 8-lane classes

std::string device_key() {
    int dev = 0, cus = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "unknown";
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    char arch[64];
    std::snprintf(arch, sizeof arch, "%.*s", 6, prop.gcnArchName);
    return std::string(arch) + "-" + std::to_string(cus) + "cu";
}

int load_file(const char *path, bool must_exist) {
    FILE *f = std::fopen(path, "r");
    if (!f) {
        if (!must_exist) return SF_OK;
        char msg[300];
        std::snprintf(msg, sizeof msg, "sf_tune_load: cannot open %s", path);
        return sf_runtime_fail(SF_EINVAL, msg);
    }
    const std::string want = device_key();
    char key[64], tag[32];
    int cls = 0;
    double tput = 0, chain = 0;
    std::map<int, Cost> got;
    // Lines: "<device> <class> <tput ns> <chain ns>", and "<device> setting <use 0|1> <factor>".
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        if (std::sscanf(line, "%63s %d %lf %lf", key, &cls, &tput, &chain) == 4) {
            if (want == key && tput > 0 && chain > 0) got[cls] = Cost{tput, chain};
        } else if (std::sscanf(line, "%63s %31s %d %lf", key, tag, &cls, &tput) == 4 && want == key &&
                   std::string(tag) == "setting" && tput > 0) {
            g_use = cls != 0;
            g_chain_factor = tput;
        }
    }
    std::fclose(f);
    for (auto &e : got) g_table[e.first] = e.second;
    ++g_generation;
    return SF_OK;
}

// SF_TUNE_FILE, read on the planner's first look at the table.
void load_env_once() {
    if (g_env_loaded) return;
    g_env_loaded = true;
    if (const char *p = std::getenv("SF_TUNE_FILE")) (void)load_file(p, false);
}

struct Synth {
    uint32_t *seq = nullptr, *plen = nullptr, *tlen = nullptr, *order = nullptr, *col = nullptr;
    uint64_t *poff = nullptr, *toff = nullptr;
    int32_t *score = nullptr;
    PairDesc *desc = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Synth() {
        if (s) (void)hipStreamSynchronize(s);
        for (void *p : {(void *)seq, (void *)plen, (void *)tlen, (void *)order, (void *)col, (void *)poff, (void *)toff,
                        (void *)score, (void *)desc})
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (s) (void)hipStreamDestroy(s);
    }
};

template <class T>
bool alloc(T **p, uint64_t n) {
    return hipMalloc((void **)p, n * sizeof(T)) == hipSuccess;
}

// Random 2-bit sequences: pair i's pattern in words [64 i, 64 i + 32), its text in the next 32.
int make_synth(Synth &y) {
    const uint64_t words = (uint64_t)kPairs * 2 * (kRows / 16);
    if (!alloc(&y.seq, words) || !alloc(&y.plen, kPairs) || !alloc(&y.tlen, kPairs) || !alloc(&y.order, kPairs) ||
        !alloc(&y.col, kPairs) || !alloc(&y.poff, kPairs) || !alloc(&y.toff, kPairs) || !alloc(&y.score, kPairs) ||
        !alloc(&y.desc, kPairs))
        return sf_runtime_fail(SF_ENOMEM, "autotune: device memory for the timing pairs");
    if (hipStreamCreateWithFlags(&y.s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&y.e0) != hipSuccess ||
        hipEventCreate(&y.e1) != hipSuccess)
        return sf_runtime_fail(SF_EDEVICE, "autotune: stream or events");
    std::vector<uint32_t> w(words);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (auto &v : w) { // xorshift64*
        x ^= x >> 12, x ^= x << 25, x ^= x >> 27;
        v = (uint32_t)((x * 0x2545F4914F6CDD1Dull) >> 32);
    }
    std::vector<uint64_t> po(kPairs), to(kPairs);
    std::vector<uint32_t> len(kPairs, kRows), ord(kPairs);
    std::vector<PairDesc> d(kPairs);
    for (uint32_t i = 0; i < kPairs; ++i) {
        po[i] = (uint64_t)i * 2 * kRows;
        to[i] = po[i] + kRows;
        ord[i] = i;
        d[i] = PairDesc{po[i], to[i], kRows, kRows, 0u, 0u}; // band 0: inside every class
    }
    if (hipMemcpy(y.seq, w.data(), words * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(y.poff, po.data(), kPairs * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(y.toff, to.data(), kPairs * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(y.plen, len.data(), kPairs * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(y.tlen, len.data(), kPairs * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(y.order, ord.data(), kPairs * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(y.desc, d.data(), kPairs * sizeof(PairDesc), hipMemcpyHostToDevice) != hipSuccess)
        return sf_runtime_fail(SF_EDEVICE, "autotune: upload");
    return SF_OK;
}

// Best time (ms) of `reps` launches of class `cls` over the first `count` pairs, after a warm-up.
int time_class(Synth &y, const AffineArgs &a0, int cls, uint32_t count, int reps, double *best) {
    AffineArgs a = a0;
    a.count = count;
    *best = 1e30;
    for (int r = -1; r < reps; ++r) {
        if (hipEventRecord(y.e0, y.s) != hipSuccess || sf_launch_affine(a, cls, false, y.s) != hipSuccess ||
            hipEventRecord(y.e1, y.s) != hipSuccess || hipEventSynchronize(y.e1) != hipSuccess)
            return sf_runtime_fail(SF_EDEVICE, "autotune: a timing launch failed");
        float ms = 0;
        if (hipEventElapsedTime(&ms, y.e0, y.e1) != hipSuccess) return sf_runtime_fail(SF_EDEVICE, "autotune: events");
        if (r >= 0) *best = std::min(*best, (double)ms);
    }
    return SF_OK;
}

} // namespace

// The measured costs of a class under the planner's setting (chain scaled by its factor), if the
// table holds them and the setting uses them (SF_TUNE_FILE is read on the first call).
bool sf_tune_cost(int cls, double *tput_ns, double *chain_ns) {
    std::lock_guard<std::mutex> lk(g_mu);
    load_env_once();
    if (!g_use) return false;
    auto it = g_table.find(cls);
    if (it == g_table.end()) return false;
    *tput_ns = it->second.tput_ns;
    *chain_ns = it->second.chain_ns * g_chain_factor;
    return true;
}

void sf_tune_set(bool use, double chain_factor) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_use = use;
    g_chain_factor = chain_factor;
    ++g_generation;
}

uint64_t sf_tune_generation() {
    std::lock_guard<std::mutex> lk(g_mu);
    load_env_once();
    return g_generation;
}

// Measures the score classes in `classes` that the table does not hold yet (mode and scoring as
// the runs use them).
int sf_tune_measure(const std::vector<int> &classes, int mode, const sf_scoring *sc) {
    std::vector<int> todo;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        load_env_once();
        for (int c : classes)
            if (c && !g_table.count(c) && std::find(todo.begin(), todo.end(), c) == todo.end()) todo.push_back(c);
    }
    if (todo.empty()) return SF_OK;
    Synth y;
    if (const int st = make_synth(y)) return st;
    AffineArgs a{};
    a.pv = PairsView{y.seq, y.poff, y.toff, y.plen, y.tlen};
    a.desc = y.desc;
    a.order = y.order;
    a.mode = mode;
    a.mat = sc->mat, a.mis = sc->mis, a.open = sc->open, a.ext = sc->ext;
    a.score = y.score;
    a.score_col = y.col;
    for (int c : todo) {
        const uint32_t per_wave = 64u / (uint32_t)aff_lp(c);
        const uint32_t fill = std::min<uint32_t>(kPairs, 8192u * per_wave);
        double t_fill = 0, t_one = 0;
        if (const int st = time_class(y, a, c, fill, 2, &t_fill)) return st;
        if (const int st = time_class(y, a, c, per_wave, 3, &t_one)) return st;
        const Cost cost{t_fill * 1e6 / ((double)fill * kRows), t_one * 1e6 / kRows};
        std::lock_guard<std::mutex> lk(g_mu);
        g_table[c] = cost;
        ++g_generation;
    }
    return SF_OK;
}

extern "C" {

int sf_tune_save(const char *path) {
    if (!path) return sf_runtime_fail(SF_EINVAL, "path is null");
    if (const int st = sf_runtime_ensure()) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    FILE *f = std::fopen(path, "w");
    if (!f) return sf_runtime_fail(SF_EINVAL, "sf_tune_save: cannot write the file");
    const std::string key = device_key();
    for (auto &e : g_table) std::fprintf(f, "%s %d %.6g %.6g\n", key.c_str(), e.first, e.second.tput_ns, e.second.chain_ns);
    std::fprintf(f, "%s setting %d %.6g\n", key.c_str(), g_use ? 1 : 0, g_chain_factor);
    const bool ok = std::fclose(f) == 0;
    return ok ? SF_OK : sf_runtime_fail(SF_EINVAL, "sf_tune_save: write failed");
}

int sf_tune_load(const char *path) {
    if (!path) return sf_runtime_fail(SF_EINVAL, "path is null");
    if (const int st = sf_runtime_ensure()) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    g_env_loaded = true; // an explicit load takes the place of SF_TUNE_FILE
    return load_file(path, true);
}

int sf_tune_reset(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
    g_env_loaded = true;
    g_use = true;
    g_chain_factor = 1.0;
    ++g_generation;
    return SF_OK;
}

uint64_t sf_tune_classes(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_table.size();
}

} // extern "C"
