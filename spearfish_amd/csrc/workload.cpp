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
//   serve: a synthetic transcriptome and reads against it (sf_txome_*, sf_reads_*, below).
// A substitution draws one of the three other bases; an insertion keeps the base and then adds L
// random bases; a deletion drops L target bases starting at this one.
#include <cmath>
#include <cstdint>
#include <algorithm>
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

// ---- serve workload -------------------------------------------------------------------------
//
// Transcriptome: `genes` genes, each with 5..16 exons of LogNormal(ln 150, 0.6) bases clipped
// to [30, 2000] and 1..8 isoforms (the counts then trimmed from the largest genes, or grown at the
// smallest, to exactly `transcripts`).  Isoform 0 has every exon; the others skip each interior
// exon with probability 0.25 and, with probability 0.5, start one non-first exon 3..60 bases late
// (an alternative 3' splice site), the exon-skipping and alternative-3'-site patterns of
// spearfish's toy fixture (fixtures/toy/make_toy.py:35-92).
//
// Reads: a gene and one of its isoforms X uniformly, length LogNormal(ln 1200, 0.5) clipped to
// [500, 10000] (and to X), a uniform start, 8% errors (4% substitutions, 2% insertions, 2%
// deletions, geometric indels).  Segments: X's window is cut into ceil(L / 600) equal parts;
// each part is projected onto every isoform Y of the gene through the exons (a position inside an
// exon Y lacks moves to the start of Y's next exon), and each non-empty projection gives one
// segment (read slice, Y's window, the window cut to at most 1,024 bases longer than the slice)
// -- at most 16 per read, as the tie-break aligns a read's isoform-differing regions against
// each candidate.
namespace {
struct Exon {
    std::vector<uint8_t> seq;
};
struct Iso {
    std::vector<uint32_t> exon, skip0; // included exons and their start offsets (alt 3' site)
    std::vector<uint32_t> start;       // position of each included exon in the transcript
    uint32_t len = 0;
    uint64_t off = 0; // in the store
};
struct Gene {
    std::vector<Exon> exons;
    std::vector<Iso> iso;
};
} // namespace

struct sf_txome {
    std::vector<Gene> genes;
    std::vector<uint8_t> store; // one code per base, transcripts back to back
};

struct sf_readset {
    std::vector<uint8_t> codes;
    std::vector<uint64_t> roff;
    std::vector<uint32_t> rlen, nseg, sro, srl, stl;
    std::vector<uint64_t> sto;
};

namespace {

void make_iso(SplitMix64 &r, Gene &g, bool full) {
    Iso iso;
    const uint32_t E = (uint32_t)g.exons.size();
    for (uint32_t e = 0; e < E; ++e) {
        if (!full && e > 0 && e + 1 < E && r.unit() < 0.25) continue;
        iso.exon.push_back(e);
        iso.skip0.push_back(0);
    }
    if (!full && iso.exon.size() > 1 && r.unit() < 0.5) {
        const uint32_t k = 1 + (uint32_t)r.below(iso.exon.size() - 1);
        const uint32_t d = 3 + (uint32_t)r.below(58);
        if (g.exons[iso.exon[k]].seq.size() > d + 20) iso.skip0[k] = d;
    }
    for (size_t k = 0; k < iso.exon.size(); ++k) {
        iso.start.push_back(iso.len);
        iso.len += (uint32_t)g.exons[iso.exon[k]].seq.size() - iso.skip0[k];
    }
    g.iso.push_back(std::move(iso));
}

// Position p of isoform X projected onto isoform Y (see the header comment).
uint32_t project(const Gene &g, const Iso &X, const Iso &Y, uint32_t p) {
    if (p >= X.len) return Y.len;
    size_t k = 0;
    while (k + 1 < X.start.size() && X.start[k + 1] <= p) ++k;
    const uint32_t e = X.exon[k];
    const uint32_t within = p - X.start[k] + X.skip0[k]; // offset inside the exon
    for (size_t q = 0; q < Y.exon.size(); ++q) {
        if (Y.exon[q] == e) {
            const uint32_t w = within < Y.skip0[q] ? Y.skip0[q] : within;
            return Y.start[q] + (w - Y.skip0[q]);
        }
        if (Y.exon[q] > e) return Y.start[q];
    }
    return Y.len;
}

} // namespace

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

int sf_txome_generate(uint64_t seed, uint32_t genes, uint32_t transcripts, sf_txome **out) {
    if (!out || genes == 0 || transcripts < genes || transcripts > 8ull * genes) return 1;
    auto *t = new sf_txome();
    t->genes.resize(genes);
    std::vector<uint32_t> want(genes);
    uint64_t total = 0;
    for (uint32_t g = 0; g < genes; ++g) {
        SplitMix64 r = SplitMix64::for_record(seed ^ 0x7a5c0e11ull, g);
        const uint32_t E = 5 + (uint32_t)r.below(12);
        t->genes[g].exons.resize(E);
        for (auto &ex : t->genes[g].exons) {
            double l = std::exp(std::log(150.0) + 0.6 * r.normal());
            l = std::fmin(2000.0, std::fmax(30.0, std::round(l)));
            ex.seq.resize((size_t)l);
            for (auto &b : ex.seq) b = (uint8_t)r.below(4);
        }
        want[g] = 1 + (uint32_t)r.below(8);
        total += want[g];
    }
    // Trim or grow the isoform counts (largest genes first / smallest first) to `transcripts`.
    for (uint32_t g = 0; total > transcripts; g = (g + 1) % genes)
        if (want[g] > 1) --want[g], --total;
    for (uint32_t g = 0; total < transcripts; g = (g + 1) % genes)
        if (want[g] < 8) ++want[g], ++total;
    for (uint32_t g = 0; g < genes; ++g) {
        SplitMix64 r = SplitMix64::for_record(seed ^ 0x150f0e5ull, g);
        for (uint32_t k = 0; k < want[g]; ++k) make_iso(r, t->genes[g], k == 0);
        for (Iso &iso : t->genes[g].iso) {
            iso.off = t->store.size();
            for (size_t q = 0; q < iso.exon.size(); ++q) {
                const auto &ex = t->genes[g].exons[iso.exon[q]].seq;
                t->store.insert(t->store.end(), ex.begin() + iso.skip0[q], ex.end());
            }
        }
    }
    *out = t;
    return 0;
}

uint64_t sf_txome_bases(const sf_txome *t) { return t->store.size(); }

uint32_t sf_txome_count(const sf_txome *t) {
    uint32_t c = 0;
    for (auto &g : t->genes) c += (uint32_t)g.iso.size();
    return c;
}

// words: (bases + 15) / 16 u32, 2-bit packed; tx_off/tx_len/tx_gene: one per transcript.
void sf_txome_export(const sf_txome *t, uint32_t *words, uint64_t *tx_off, uint32_t *tx_len, uint32_t *tx_gene) {
    const uint64_t nb = t->store.size();
    std::memset(words, 0, ((nb + 15) / 16) * 4);
    for (uint64_t k = 0; k < nb; ++k) words[k >> 4] |= (uint32_t)t->store[k] << (2 * (k & 15));
    uint32_t i = 0;
    for (uint32_t g = 0; g < t->genes.size(); ++g)
        for (auto &iso : t->genes[g].iso) {
            tx_off[i] = iso.off;
            tx_len[i] = iso.len;
            tx_gene[i] = g;
            ++i;
        }
}

void sf_txome_free(sf_txome *t) { delete t; }

int sf_reads_generate(const sf_txome *t, uint64_t seed, uint64_t n, sf_readset **out) {
    if (!t || !out) return 1;
    auto *rs = new sf_readset();
    const uint32_t G = (uint32_t)t->genes.size();
    std::vector<uint8_t> read;
    for (uint64_t i = 0; i < n; ++i) {
        SplitMix64 r = SplitMix64::for_record(seed ^ 0x4eadull, i);
        const Gene &g = t->genes[r.below(G)];
        const Iso &X = g.iso[r.below(g.iso.size())];
        double l = std::exp(std::log(1200.0) + 0.5 * r.normal());
        l = std::fmin(10000.0, std::fmax(500.0, std::round(l)));
        const uint32_t L = std::min<uint32_t>((uint32_t)l, X.len);
        const uint32_t a = (uint32_t)r.below(X.len - L + 1);
        mutate(r, t->store.data() + X.off + a, L, 0.04, 0.02, 0.02, read);
        rs->roff.push_back(rs->codes.size());
        rs->rlen.push_back((uint32_t)read.size());
        rs->codes.insert(rs->codes.end(), read.begin(), read.end());
        const uint32_t parts = std::max<uint32_t>(1, (L + 599) / 600);
        uint32_t segs = 0;
        for (uint32_t c = 0; c < parts && segs < 16; ++c) {
            const uint32_t x0 = a + (uint32_t)((uint64_t)L * c / parts), x1 = a + (uint32_t)((uint64_t)L * (c + 1) / parts);
            const uint32_t r0 = (uint32_t)((uint64_t)read.size() * c / parts);
            const uint32_t r1 = (uint32_t)((uint64_t)read.size() * (c + 1) / parts);
            for (const Iso &Y : g.iso) {
                if (segs >= 16) break;
                const uint32_t y0 = project(g, X, Y, x0);
                uint32_t y1 = project(g, X, Y, x1);
                if (y1 <= y0 || r1 <= r0) continue;
                y1 = std::min<uint32_t>(y1, y0 + (r1 - r0) + 1024); // window at most 1 kb longer than the slice
                rs->sro.push_back(r0);
                rs->srl.push_back(r1 - r0);
                rs->sto.push_back(Y.off + y0);
                rs->stl.push_back(y1 - y0);
                ++segs;
            }
        }
        rs->nseg.push_back(segs);
    }
    *out = rs;
    return 0;
}

uint64_t sf_reads_bases(const sf_readset *r) { return r->codes.size(); }
uint64_t sf_reads_segments(const sf_readset *r) { return r->sro.size(); }

void sf_reads_export(const sf_readset *r, uint8_t *codes, uint64_t *read_off, uint32_t *read_len, uint32_t *nseg,
                     uint32_t *seg_read_off, uint32_t *seg_read_len, uint64_t *seg_tx_off, uint32_t *seg_tx_len) {
    std::memcpy(codes, r->codes.data(), r->codes.size());
    std::memcpy(read_off, r->roff.data(), r->roff.size() * 8);
    std::memcpy(read_len, r->rlen.data(), r->rlen.size() * 4);
    std::memcpy(nseg, r->nseg.data(), r->nseg.size() * 4);
    std::memcpy(seg_read_off, r->sro.data(), r->sro.size() * 4);
    std::memcpy(seg_read_len, r->srl.data(), r->srl.size() * 4);
    std::memcpy(seg_tx_off, r->sto.data(), r->sto.size() * 8);
    std::memcpy(seg_tx_len, r->stl.data(), r->stl.size() * 4);
}

void sf_reads_free(sf_readset *r) { delete r; }

} // extern "C"
