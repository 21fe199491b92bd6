// workload.cpp -- synthetic (pattern, text) pairs for the benchmark workloads (libsf_workload.so).
//
// Each pair draws from its own SplitMix64 stream, seeded per record as spearfish's simulator does
// (crates/spearfish-cli/src/rng.rs:1-40: state = seed ^ index * 0xd1b54a32d192ed03, one draw
// discarded), so a pair does not depend on how many threads generate the batch.
//
// Workloads (error rates are per target base; indel lengths geometric with p = 0.7,
// P(L = l) = 0.7 * 0.3^(l - 1)):
//   1  parity set: target U{200..1000}, 8% errors (3% substitution, 2.5% insertion, 2.5% deletion),
//      band 64.
//   2  uniform: target 1000, 6% errors (3% / 1.5% / 1.5%), band 128.
//   3  uneven: target LogNormal(ln 800, 0.8) clipped to [100, 20000], error U[5%, 15%] split half
//      substitutions, a quarter each insertions and deletions; band min(512, max(32,
//      ceil(1.2 |m - n|) + 32)).
//   4  long: query length L = U{2000..10000}, target ceil(1.05 L) random bases, the query a window
//      of L target bases at a random start with 12% errors (6% / 3% / 3%); band 256.
// A substitution draws one of the three other bases; an insertion keeps the base and then adds L
// random bases; a deletion drops L target bases starting at this one.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    static SplitMix64 for_record(uint64_t seed, uint64_t index) {
        SplitMix64 r{seed ^ (index * 0xd1b54a32d192ed03ull)};
        r.next();
        return r;
    }
    uint64_t next() {
        s += 0x9e3779b97f4a7c15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint64_t below(uint64_t n) { return (uint64_t)(((__uint128_t)next() * n) >> 64); }
    double unit() { return (double)(next() >> 11) / (double)(1ull << 53); }
    double normal() { // Box-Muller
        double u1 = unit(), u2 = unit();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    uint32_t geometric() {
        uint32_t l = 1;
        while (unit() >= 0.7) ++l;
        return l;
    }
};

struct Rec {
    std::vector<uint8_t> p, t;
    uint64_t band;
};

void mutate(SplitMix64 &r, const uint8_t *t, size_t n, double sub, double ins, double del, std::vector<uint8_t> &out) {
    out.clear();
    out.reserve(n + n / 8 + 16);
    for (size_t i = 0; i < n;) {
        const double u = r.unit();
        const uint8_t b = t[i];
        if (u < sub) {
            out.push_back((uint8_t)((b + 1 + r.below(3)) & 3));
            ++i;
        } else if (u < sub + ins) {
            out.push_back(b);
            const uint32_t l = r.geometric();
            for (uint32_t q = 0; q < l; ++q) out.push_back((uint8_t)r.below(4));
            ++i;
        } else if (u < sub + ins + del) {
            i += r.geometric();
        } else {
            out.push_back(b);
            ++i;
        }
    }
}

void make(int cfg, uint64_t seed, uint64_t idx, Rec &rec) {
    SplitMix64 r = SplitMix64::for_record(seed, idx);
    std::vector<uint8_t> &t = rec.t, &p = rec.p;
    switch (cfg) {
    case 1: {
        const size_t n = 200 + r.below(801);
        t.resize(n);
        for (auto &x : t) x = (uint8_t)r.below(4);
        mutate(r, t.data(), n, 0.03, 0.025, 0.025, p);
        rec.band = 64;
        break;
    }
    case 2: {
        t.resize(1000);
        for (auto &x : t) x = (uint8_t)r.below(4);
        mutate(r, t.data(), 1000, 0.03, 0.015, 0.015, p);
        rec.band = 128;
        break;
    }
    case 3: {
        double len = std::exp(std::log(800.0) + 0.8 * r.normal());
        len = std::fmin(20000.0, std::fmax(100.0, std::round(len)));
        const size_t n = (size_t)len;
        const double e = 0.05 + 0.10 * r.unit();
        t.resize(n);
        for (auto &x : t) x = (uint8_t)r.below(4);
        mutate(r, t.data(), n, e / 2, e / 4, e / 4, p);
        const double d = std::fabs((double)p.size() - (double)n);
        const double b = std::ceil(1.2 * d) + 32;
        rec.band = (uint64_t)std::fmin(512.0, std::fmax(32.0, b));
        break;
    }
    case 4: {
        const size_t L = 2000 + r.below(8001);
        const size_t n = (size_t)std::ceil(1.05 * (double)L);
        t.resize(n);
        for (auto &x : t) x = (uint8_t)r.below(4);
        const size_t w = r.below(n - L + 1);
        mutate(r, t.data() + w, L, 0.06, 0.03, 0.03, p);
        rec.band = 256;
        break;
    }
    default:
        t.clear();
        p.clear();
        rec.band = 0;
    }
}

} // namespace

struct sf_workload {
    std::vector<Rec> recs;
};

extern "C" {

// Generates n pairs of workload `cfg` (1..4) from `seed` with `threads` workers.
int sf_workload_generate(int cfg, uint64_t seed, uint64_t n, int threads, sf_workload **out) {
    if (!out || cfg < 1 || cfg > 4) return 1;
    auto *w = new sf_workload();
    w->recs.resize(n);
    if (threads < 1) threads = 1;
    std::vector<std::thread> th;
    const uint64_t chunk = (n + threads - 1) / threads;
    for (int q = 0; q < threads; ++q) {
        const uint64_t a = q * chunk, b = a + chunk < n ? a + chunk : n;
        if (a >= b) break;
        th.emplace_back([=] {
            for (uint64_t i = a; i < b; ++i) make(cfg, seed, i, w->recs[i]);
        });
    }
    for (auto &t : th) t.join();
    *out = w;
    return 0;
}

uint64_t sf_workload_size(const sf_workload *w) { return w->recs.size(); }

uint64_t sf_workload_pat_bases(const sf_workload *w) {
    uint64_t s = 0;
    for (auto &r : w->recs) s += r.p.size();
    return s;
}

uint64_t sf_workload_txt_bases(const sf_workload *w) {
    uint64_t s = 0;
    for (auto &r : w->recs) s += r.t.size();
    return s;
}

// Copies the pairs out as one code per byte (buffers sized by the two counts above).
void sf_workload_export(const sf_workload *w, uint8_t *pat, uint64_t *pat_off, uint32_t *pat_len, uint8_t *txt,
                        uint64_t *txt_off, uint32_t *txt_len, uint64_t *band) {
    uint64_t po = 0, to = 0;
    for (size_t i = 0; i < w->recs.size(); ++i) {
        const Rec &r = w->recs[i];
        pat_off[i] = po;
        pat_len[i] = (uint32_t)r.p.size();
        txt_off[i] = to;
        txt_len[i] = (uint32_t)r.t.size();
        band[i] = r.band;
        if (!r.p.empty()) std::memcpy(pat + po, r.p.data(), r.p.size());
        if (!r.t.empty()) std::memcpy(txt + to, r.t.data(), r.t.size());
        po += r.p.size();
        to += r.t.size();
    }
}

void sf_workload_free(sf_workload *w) { delete w; }

} // extern "C"
