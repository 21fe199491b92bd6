/*
 * sf_align.h -- C ABI of the MI355X (gfx950) alignment library, libsf_align.so.
 *
 * A drop-in for the base-level alignment step of spearfish (crates/map/src/align.rs, v0.3.2):
 * Myers bit-parallel edit distance and banded affine-gap (Gotoh) alignment, computed by HIP
 * kernels with results bit-identical to the reference's CPU kernels.  Every entry point is a
 * batch form of one or more of the reference's functions; the comment on each cites the
 * function(s) it replaces.  The reference binds it from Rust through a thin `-sys` crate of
 * `extern "C"` declarations mirroring the ones below.
 *
 * Conventions
 *   - Plain C types only; the caller owns every buffer it passes, the library keeps no pointer
 *     past the call's return (sf_batch objects own their device copies).
 *   - Sequences are base codes A=0 C=1 G=2 T=3 (crates/lcp-core/src/encode.rs), one per byte and
 *     masked with `& 3` as the reference does (align.rs:78, 120, 1054, 2557), or 2-bit packed:
 *     16 bases per little-endian u32 word, base k at bits 2*(k % 16) of word k / 16 -- the same
 *     bit order as the index's u64 packing (crates/index/src/lib.rs:45-151).
 *   - Every call returns an int status: SF_OK or an SF_E* code; sf_last_error() gives the
 *     calling thread's last message.  No call aborts the process.
 *   - Thread-safe: calls from several host threads may overlap; each has its own stream.
 */
#ifndef SF_ALIGN_H
#define SF_ALIGN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SF_ABI_VERSION 2

/* Status codes. */
#define SF_OK 0
#define SF_EINVAL 1  /* invalid argument */
#define SF_ENOMEM 2  /* host or device memory exhausted */
#define SF_EDEVICE 3 /* HIP runtime or kernel failure */
#define SF_ENODEV 4  /* no gfx950 device, or the library was built without device code */
#define SF_ERANGE 5  /* an input exceeds this library's limits, or an output buffer is too small */
/* Limits (beyond them SF_ERANGE): sequences under 2^30 bases; Myers patterns up to 2^24 blocks of
 * 64 bases (patterns over 98,304 bases run as stripes); affine bands up to 2^26 cells a row (over
 * 2,048 on the wide-band kernel; tracebacks up to 2,048); scoring values 0..4,194,303 and scores that
 * stay far from NEG (|score| under about 2^28); the serving loop's shape classes (patterns up to
 * 98,304 bases, bands up to 2,048 cells). */
#define SF_ETIMEDOUT 6 /* sf_server_wait: the timeout passed before the request completed */
#define SF_ECANCELED 7 /* the request was cancelled (sf_server_cancel) */

/* align.rs:40-45 `Mode`. */
#define SF_INFIX 0
#define SF_PREFIX 1
#define SF_GLOBAL 2

#define SF_NEG (INT32_MIN / 4) /* align.rs:693 `NEG`: unreachable / stopped pass */
#define SF_NONE UINT32_MAX     /* Option<u32>::None (Myers distance above k) */
#define SF_BAND_EXACT UINT64_MAX /* `band = usize::MAX`: exact (unbanded) pass */

/* align.rs:660-691 `Affine`: +mat per match, -mis per mismatch, -(open + ext * l) per gap of l. */
typedef struct {
    int32_t mat, mis, open, ext;
} sf_scoring;

/* A batch of (pattern, text) pairs.  Pattern i is bases [pat_off[i], pat_off[i] + pat_len[i])
 * of `pat`; text i likewise of `txt`.  `packed` = 0: one code per byte; 1: 2-bit words (offsets
 * are still in bases).  pat_bases / txt_bases bound the buffers (in bases) for checking. */
typedef struct {
    const void *pat;
    const uint64_t *pat_off;
    const uint32_t *pat_len;
    const void *txt;
    const uint64_t *txt_off;
    const uint32_t *txt_len;
    uint64_t n;
    int32_t packed;
    uint64_t pat_bases;
    uint64_t txt_bases;
} sf_pairs;

/* Per-kernel device time of the last run on a batch (HIP events on the batch's stream). */
typedef struct {
    double myers_ms;     /* Myers distance (+ end) kernels */
    double locate_ms;    /* reversed Myers pass for start coordinates */
    double affine_ms;    /* banded affine DP kernels (with traceback bits when asked) */
    double walk_ms;      /* traceback walk + CIGAR compaction */
    uint64_t myers_cells;  /* sum of m * n over the pairs run (the GCUPS numerator) */
    uint64_t affine_cells; /* DP cells as the reference's DP_WORK counts them (align.rs:2573-2578) */
    uint64_t affine_launches;
    double graph_ms;        /* replays of the captured step (sf_batch_replay); the cells above count them */
    uint64_t graph_replays;
} sf_timing;

/* ---- library ---- */
const char *sf_strerror(int status);
const char *sf_last_error(void);
int sf_abi_version(void);
/* Selects the device for this process (default: device 0 on first use).  SF_ENODEV if it is
 * not a gfx950. */
int sf_init(int device);

/* ---- one-shot batch calls (host buffers in, host buffers out) ---- */

/* Replaces align.rs:87-168 `distance` / `distance_with`, align.rs:170-200 `batch_distances`
 * (equal to one distance_with per text, align.rs:170-172, 3373-3432) and align.rs:655-658
 * `infix_distance`: dist[i] = edit distance of pair i in `mode`, or SF_NONE if above k.
 * `end` (nullable): smallest text end column j with D[m][j] = dist (n for Global).
 * `start` (nullable, needs end): start column of the shortest text substring at that end with the
 * same distance (0 for Prefix and Global).  end/start have no reference counterpart. */
int sf_myers(const sf_pairs *pairs, int mode, uint32_t k, uint32_t *dist, uint32_t *end, uint32_t *start);

/* Replaces crates/map/src/lib.rs:851-911 `Mapper::banded` (equal to its definition `sequential`,
 * lib.rs:3016-3034) for many reads at once.  Group g -- pairs [group_off[g], group_off[g + 1]) --
 * is one read (every pair's pattern) against its texts in the caller's order; group_off has
 * ngroups + 1 entries, from 0 to pairs->n.  out[i] is the value the reference gives text i: in its
 * group, in order, with k = (best so far) + cap (saturating, best starting at u32::MAX), the edit
 * distance in `mode` where it is <= k, else k + 1, best falling to it; a text equal to an earlier
 * one of its group takes that one's value re-capped at the current k.  `distinct` != 0: the
 * caller promises no text repeats within a group (no comparison is made).  Texts compare by codes
 * masked & 3 (the results are those of comparing bytes, see the implementation). */
int sf_myers_banded(const sf_pairs *pairs, const uint64_t *group_off, uint64_t ngroups, int mode, uint32_t cap,
                    int distinct, uint32_t *out);

/* No counterpart in the reference: sf_myers for callers that cannot allocate (harnesses that forbid
 * device allocation, tight serving loops).  The same arguments and results as sf_myers, computed
 * by the same kernels, with no hipMalloc, hipFree, stream creation or host heap allocation: the
 * device memory it needs lives in one fixed 32 MiB device array per GPU, held from the first call
 * until the process ends.  A call needs about 1.25 B per base of the two buffers (0.25 B if packed)
 * plus 72 B per pair; SF_ERANGE if that exceeds 32 MiB.  Calls on one GPU share the array and are
 * serialised by the library (a call waits for the one before it); each call leaves its data in the
 * array and the next call overwrites it. */
int sf_myers_noalloc(const sf_pairs *pairs, int mode, uint32_t k, uint32_t *dist, uint32_t *end, uint32_t *start);

/* Replaces align.rs:699-701 `affine` and align.rs:854-861 `affine_sorted` (equal to separate
 * calls, align.rs:2210-2213): score[i] of pair i in `mode` with band[i] (SF_BAND_EXACT for
 * exact; band == NULL: all exact), SF_NEG where the pass stops (align.rs:2543-2556).
 * score_col (nullable): the text column where the score ends (align.rs:2643-2650). */
int sf_affine(const sf_pairs *pairs, int mode, const sf_scoring *sc, const uint64_t *band, int32_t *score,
              uint32_t *score_col);

/* sf_affine plus a traceback (no reference counterpart: the reference returns scores only,
 * align.rs:2111-2129).  CIGARs are BAM-encoded u32 (len << 4 | op; '=' 7, 'X' 8, 'I' 1 pattern
 * base only, 'D' 2 text base only), pair i's at cigar[cigar_off[i] .. cigar_off[i + 1]);
 * cigar_off has n + 1 entries.  tstart[i] (nullable): the text column where the alignment
 * starts (0 unless Infix).  SF_ERANGE if cigar_cap is too small (sum of m + n + 1 always fits).
 * A pair whose pass stopped gets an empty CIGAR. */
int sf_affine_cigar(const sf_pairs *pairs, int mode, const sf_scoring *sc, const uint64_t *band, int32_t *score,
                    uint32_t *score_col, uint32_t *tstart, uint32_t *cigar, uint64_t cigar_cap, uint64_t *cigar_off);

/* ---- device-resident batches ---- */
/* A batch packs its pairs 2-bit into HBM once, bins them by shape, and can then be run many
 * times; runs are asynchronous on the batch's own HIP streams (a main one and two more) until
 * sf_batch_sync.  A run that needs another's results, or rewrites its outputs, waits for it; a
 * score-only affine run may overlap the Myers run queued before it. */
typedef struct sf_batch sf_batch;

int sf_batch_create(const sf_pairs *pairs, sf_batch **out);
void sf_batch_destroy(sf_batch *b);
uint64_t sf_batch_size(const sf_batch *b);
/* Enqueues Myers over every pair (results stay on the device).  locate: 0 distance only,
 * 1 + end, 2 + end + start. */
int sf_batch_myers(sf_batch *b, int mode, uint32_t k, int locate);
/* Enqueues the banded affine pass over every pair; band as in sf_affine; cigar != 0 adds the
 * traceback. */
int sf_batch_affine(sf_batch *b, int mode, const sf_scoring *sc, const uint64_t *band, int cigar);
/* Restricts each pair's text to the window [start, end) located by the last sf_batch_myers run
 * with locate = 2 (pairs with no distance keep their text), so that the next sf_batch_affine
 * aligns each pattern to its best-matching substring; on = 0 restores the texts.  No reference
 * counterpart: it chains T2 (end/start) into the affine pass for long segments. */
int sf_batch_use_located(sf_batch *b, int on);
int sf_batch_sync(sf_batch *b);
/* Copies results of the last runs out (implies sync).  Nullable outputs are skipped. */
int sf_batch_myers_results(sf_batch *b, uint32_t *dist, uint32_t *end, uint32_t *start);
int sf_batch_affine_results(sf_batch *b, int32_t *score, uint32_t *score_col, uint32_t *tstart, uint32_t *cigar,
                            uint64_t cigar_cap, uint64_t *cigar_off);
/* Graph capture (no reference counterpart): the runs queued between sf_batch_capture_begin and
 * sf_batch_capture_end -- sf_batch_myers (any mode, k, locate) and score-only sf_batch_affine whose
 * launch plan exists (the same mode, bands and scoring ran on this batch before) -- are recorded
 * as one HIP graph instead of running; sf_batch_replay launches that graph `times` times on the
 * batch's main stream, asynchronously, with the results of running those calls eagerly.  Calls that
 * wait for the device, copy results or build a plan fail during a capture with SF_EINVAL, and so
 * does the capture's end.  A new capture replaces the batch's graph. */
int sf_batch_capture_begin(sf_batch *b);
int sf_batch_capture_end(sf_batch *b);
int sf_batch_replay(sf_batch *b, uint32_t times);
/* Device time of the runs since the last sf_batch_reset_timing (implies sync). */
int sf_batch_timing(sf_batch *b, sf_timing *t);
void sf_batch_reset_timing(sf_batch *b);

/* ---- passes with per-row bookkeeping: X-drop extension, cut, band layout, local bands ---- */

/* align.rs:2111-2129 `Dp`: the usize fields as u32 (sequences are under 2^30 bases). */
typedef struct {
    int32_t score;
    uint32_t score_col;
    int32_t clip_best;
    uint32_t clip_i, clip_j; /* clip_cell (pattern bases, text bases); (0, 0) when not tracked */
    uint32_t text_read, last_row;
    int32_t cut_best, cut_score;
} sf_dp;

/* align.rs:2183-2208 `affine_dp` per pair: `geom` (nullable: the text length) lays the band out as
 * for a text of that many bases; has_clip with `xdrop` and `clip_cell` is `clip: Some((xdrop,
 * clip_cell))`; `cut` (nullable: usize::MAX for every pair) tracks scores over the first cut text
 * bases.  Results are read with sf_batch_dp_results. */
int sf_batch_affine_dp(sf_batch *b, int mode, const sf_scoring *sc, const uint64_t *band, const uint64_t *geom,
                       int has_clip, int32_t xdrop, int clip_cell, const uint64_t *cut);
int sf_batch_dp_results(sf_batch *b, sf_dp *out);
/* align.rs:741-760 `local_band_ends` per pair: local affine over the diagonals dlo[i]..=dhi[i]
 * (text position minus read position), bonus_start / bonus_end for reaching the read's ends.
 * Pattern codes of 4 or more (unpacked input) match nothing.  Results: sf_batch_local_results. */
int sf_batch_local_band(sf_batch *b, const sf_scoring *sc, const int64_t *dlo, const int64_t *dhi, int32_t bonus_start,
                        int32_t bonus_end);
int sf_batch_local_results(sf_batch *b, int32_t *best);

/* Host-buffer forms (one device pass over all pairs; the same results as per-pair calls). */
int sf_affine_dp(const sf_pairs *pairs, int mode, const sf_scoring *sc, const uint64_t *band, const uint64_t *geom,
                 int has_clip, int32_t xdrop, int clip_cell, const uint64_t *cut, sf_dp *out);
/* align.rs:710-729 `affine_extend`: max(clip_best, score + end_bonus) of a Prefix pass with X-drop. */
int sf_affine_extend(const sf_pairs *pairs, const sf_scoring *sc, const uint64_t *band, int32_t end_bonus,
                     int32_t xdrop, int32_t *score);
/* align.rs:957-982 `affine_extend_window`: (score, pattern bases, text bases, text bases read)
 * with the band laid out for max(window[i], text length). */
typedef struct {
    int32_t score;
    uint32_t pat_bases, txt_bases, txt_read;
} sf_extend_window;
int sf_affine_extend_window(const sf_pairs *pairs, const uint64_t *window, const sf_scoring *sc, const uint64_t *band,
                            int32_t end_bonus, int32_t xdrop, sf_extend_window *out);
/* align.rs:903-949 `affine_extend_sorted_cut_bonus` (and `affine_extend_sorted_cut` =
 * track_cut 1, bonus_within_cut 0), one text per pair: (whole, cut, text bases covered).  Pairs
 * sharing a pattern give the same results as the reference's sorted call over their texts
 * (align.rs:2210-2213). */
typedef struct {
    int32_t whole, cut;
    uint32_t cols;
} sf_extend_cut;
int sf_affine_extend_cut_bonus(const sf_pairs *pairs, const sf_scoring *sc, const uint64_t *band, int32_t end_bonus,
                               int32_t xdrop, int diagonal, const uint64_t *window, const uint64_t *cut, int track_cut,
                               int bonus_within_cut, sf_extend_cut *out);
int sf_local_band_ends(const sf_pairs *pairs, const sf_scoring *sc, const int64_t *dlo, const int64_t *dhi,
                       int32_t bonus_start, int32_t bonus_end, int32_t *best);

/* ---- the tie-break's segment alignments over a device-resident transcript store ---- */

/* The index's packed transcript sequence (crates/index/src/lib.rs:45-151): `bases` bases in u64
 * words, base i at bits 2*(i % 32) of words[i / 32], copied to the device once.  Candidate texts
 * are ranges of it. */
typedef struct sf_store sf_store;
int sf_store_create(const uint64_t *words, uint64_t bases, sf_store **out);
void sf_store_destroy(sf_store *s);
/* crates/index/src/lib.rs:135-137 `PackedRange::sequence_eq` for ranges of equal length (lib.rs:104-114
 * `packed_eq`, 32 bases a compare), on the device: eq[i] = bases [a_off[i], +len[i]) equal bases
 * [b_off[i], +len[i]). */
int sf_store_codes_eq(sf_store *s, const uint64_t *a_off, const uint64_t *b_off, const uint32_t *len, uint64_t n,
                      uint8_t *eq);

/* One read segment of crates/map/src/lib.rs:1307-1492 (`seg`): read bases [read_off, +read_len)
 * of `reads` (one code a byte) against candidates cand0 .. cand0 + ncand - 1. */
#define SF_SEG_REVERSE 1          /* `rev`: a left read end, read and texts aligned reversed */
#define SF_SEG_EXTEND 2           /* Prefix with `clip_ends`: align.rs:903-949 (else align.rs:854-861) */
#define SF_SEG_DIAGONAL 4         /* `end_diagonal` */
#define SF_SEG_TRACK_CUT 8        /* `flank_min > 0` */
#define SF_SEG_BONUS_WITHIN_CUT 16 /* `end_bonus_cut` */
typedef struct {
    uint64_t read_off;
    uint32_t read_len;
    uint32_t cand0, ncand;
    int32_t mode;    /* SF_INFIX / SF_PREFIX / SF_GLOBAL (SF_PREFIX with SF_SEG_EXTEND) */
    uint32_t flags;  /* SF_SEG_* */
    uint64_t window; /* `tb.window(read length)`; the band is laid out for the larger of it and the
                        segment's longest text (lib.rs:1446-1451) */
    uint64_t cut;    /* `cut`, UINT64_MAX for None */
} sf_tb_segment;
typedef struct {
    sf_scoring sc;     /* `tb.affine` */
    uint64_t band;     /* `tb.band` */
    int32_t end_bonus; /* `tb.end_bonus` */
    int32_t xdrop;     /* `tb.xdrop` */
} sf_tb_params;
/* Per candidate c, the scores the reference's `seg` gives it with `tb.affine` set and
 * `affine_within` unset (a caller with `affine_within` runs sf_myers_banded first and passes the
 * texts within reach): out[c] = (whole, cut, text bases) of align.rs:903-949 for an SF_SEG_EXTEND
 * segment, else (x, x, 0) with x of align.rs:854-861 (lib.rs:1442-1445).  Candidate c's text is
 * store bases [cand_off[c], +cand_len[c]).  Equal texts within a segment are found on the device
 * (packed equality) and aligned once; rep (nullable) gets each candidate's first equal candidate,
 * counted from its segment's cand0 (lib.rs:1337-1345 `rep0`).  Every candidate must belong to
 * exactly one segment. */
int sf_tiebreak_segments(sf_store *st, const uint8_t *reads, uint64_t reads_len, const sf_tb_segment *segs,
                         uint64_t nseg, const uint64_t *cand_off, const uint32_t *cand_len, uint64_t ncand,
                         const sf_tb_params *p, sf_extend_cut *out, uint32_t *rep);
/* sf_tiebreak_segments in steps, for callers that keep a call's inputs on the device: create
 * uploads the reads and candidates, finds the distinct texts and plans the passes (it copies the
 * store's words into the batch, which may outlive the store); run queues the passes with `p` and
 * returns; results waits and fills out (and rep, nullable) as sf_tiebreak_segments does.  Runs may
 * be repeated (with other parameters too); results reads the last. */
typedef struct sf_tb_batch sf_tb_batch;
int sf_tb_batch_create(sf_store *st, const uint8_t *reads, uint64_t reads_len, const sf_tb_segment *segs,
                       uint64_t nseg, const uint64_t *cand_off, const uint32_t *cand_len, uint64_t ncand,
                       sf_tb_batch **out);
int sf_tb_batch_run(sf_tb_batch *b, const sf_tb_params *p);
int sf_tb_batch_results(sf_tb_batch *b, sf_extend_cut *out, uint32_t *rep);
void sf_tb_batch_destroy(sf_tb_batch *b);
/* Device time of the runs since the last reset (HIP events around each pass): the extension pass
 * (all its launches) and the sorted-score passes (summed over modes), with their cells counted over
 * the whole band (an upper bound of DP_WORK where an X-drop stops a pass early). */
typedef struct {
    double ext_ms, sorted_ms;
    uint64_t ext_cells, sorted_cells;
} sf_tb_timing;
int sf_tb_batch_timing(sf_tb_batch *b, sf_tb_timing *t, int reset);

/* ---- serving: continuous batching over a device-resident transcript store ----
 * No reference counterpart as an object: it serves the calls the tie-break makes per read
 * (crates/map/src/lib.rs:1307-1492, `seg`), many reads at a time.  Each request is a read and
 * its segments (read slices against windows of the store); per segment the server computes the
 * Myers distance and end column (`myers_mode`, k unbounded) and the banded affine score and end
 * column (`affine_mode`, `sc`, `band`).  Up to `slots` batches run at once; a free slot admits the
 * oldest waiting requests while they fit its share of `hbm_budget_bytes`; a request larger than
 * one share is refused (SF_ENOMEM) at submit. */
typedef struct sf_server sf_server;

typedef struct {
    uint64_t hbm_budget_bytes;    /* device memory for batches in flight, split over the slots */
    uint32_t slots;               /* batches in flight at once (1..16), one HIP stream each */
    uint32_t max_reads_per_batch; /* 0: limited by memory only; 1: one read per launch */
    int32_t myers_mode;
    int32_t affine_mode;
    sf_scoring sc;
    uint64_t band;
    /* 0: a free slot takes whatever is waiting at once.  Otherwise a free slot waits until the
     * waiting requests fill its share (or max_reads_per_batch), or until the oldest of them has
     * waited this many microseconds: fuller batches for a bounded extra latency. */
    uint32_t batch_wait_us;
} sf_server_config;

typedef struct {
    const uint8_t *read; /* one code per byte */
    uint32_t read_len;
    uint32_t nseg;
    const uint32_t *seg_read_off, *seg_read_len; /* read slice of each segment */
    const uint64_t *seg_tx_off;                  /* text of each segment: store bases */
    const uint32_t *seg_tx_len;
    /* outputs, nseg each (nullable), written before the request completes; caller-owned */
    uint32_t *dist, *end;
    int32_t *score;
    uint32_t *score_col;
} sf_read_request;

typedef struct {
    uint64_t reads_done, reads_failed, batches, segments, cells;
    double mean_reads_per_batch;
    double p50_ms, p99_ms, max_ms; /* submit-to-completion latency of the completed reads */
    double span_s;                 /* first submit to last completion */
    uint64_t cap_segments_per_batch, cap_read_bases_per_batch;
    uint64_t max_segments_per_batch; /* the largest batch taken in since the last reset */
} sf_server_stats;

/* store_words: the transcripts, 2-bit packed as in sf_pairs, store_bases long; copied to HBM. */
int sf_server_create(const uint32_t *store_words, uint64_t store_bases, const sf_server_config *cfg,
                     sf_server **out);
/* Queues one read; the request's input arrays may be reused after the call.  Its outputs must stay
 * valid until the request completes: until a sf_server_wait on its ticket returns anything other
 * than SF_ETIMEDOUT, or a sf_server_cancel of it returns SF_OK. */
int sf_server_submit(sf_server *s, const sf_read_request *req, uint64_t *ticket);
/* Waits for a ticket (timeout_s < 0: forever).  Returns the request's status once it has completed
 * (SF_OK, SF_ECANCELED, or the failure of its batch), after which the ticket is retired; or
 * SF_ETIMEDOUT, after which the request is still queued or running and the ticket stays valid (wait
 * again, or cancel it).  Several threads may wait on one ticket: each gets the status, and the
 * ticket is retired when the last of them returns. */
int sf_server_wait(sf_server *s, uint64_t ticket, double timeout_s);
/* Cancels a request that has not completed.  SF_OK: it will never be computed or have its outputs
 * written again (one already in a batch on the device finishes there, and its results are
 * dropped); it completes with SF_ECANCELED, which a wait on its ticket returns.  SF_EINVAL for an
 * unknown or retired ticket; a request that completed before the call keeps its status, and the
 * call returns it. */
int sf_server_cancel(sf_server *s, uint64_t ticket);
int sf_server_get_stats(sf_server *s, sf_server_stats *out, int reset);
/* Drains the queue (every request already submitted finishes), then frees the server.  With
 * SF_SERVE_TRACE set in the environment the server reports on stderr, at each stats reset and
 * here, its longest scheduler iteration and launch and the longest gap between submissions
 * (a diagnostic). */
void sf_server_destroy(sf_server *s);
/* Load generator for measurements: submits req[i] at t_arrive[i] seconds after the call, waits
 * for all; *wall_s = first arrival to last completion, *refused = requests refused at admission
 * (nullable).  Returns the first failure status of an admitted request, else SF_OK. */
int sf_server_replay(sf_server *s, const sf_read_request *req, const double *t_arrive, uint64_t n, double *wall_s,
                     uint64_t *refused);

/* Tuning (no reference counterpart; results never depend on it): bands of up to `slots` cells
 * per row may run with 16 lanes per pair, four pairs per wave (the launch planner moves a pair
 * whose rows would outlast the rest of the run to 64 lanes); wider bands, or every band when
 * `slots` is 0, run with 64 lanes per pair.  Default 512 (or SF_AFFINE_LP16_MAX from the
 * environment).  Applies to affine runs started afterwards in this process. */
int sf_set_affine_lp16_max(uint64_t slots);

/* Autotuner (no reference counterpart; results never depend on it).  The launch planner picks each
 * pair's shape -- lanes per pair (8, 16, 32 or 64), slots per lane (the class edge its band falls
 * under, or the next one up), 16-bit packed or 32-bit -- from two costs per shape class: ns per
 * pair-row with the chip full, and ns per row of one wave alone.  Untuned they are estimates from
 * instruction counts.  sf_batch_autotune times, on this device, the score kernel of every class a
 * score run of `b` with this mode, scoring and bands could pick (those not measured yet in this
 * process), then times b's score run planned with the estimates and with the measurements (a lone
 * wave's chain scaled by 1, 1.5, 2 or 3 for the load it runs under) and keeps the fastest setting
 * for later score runs' plans, of every batch in the process.  It runs b's score run 20 times (the
 * results are those of any run).  sf_tune_save / sf_tune_load write and read the measurements and
 * the setting as text (per device model; lines of another device are skipped); SF_TUNE_FILE in the
 * environment names a file read on first use.  sf_tune_reset forgets them; sf_tune_classes counts
 * the classes measured. */
int sf_batch_autotune(sf_batch *b, int mode, const sf_scoring *sc, const uint64_t *band);
int sf_tune_save(const char *path);
int sf_tune_load(const char *path);
int sf_tune_reset(void);
uint64_t sf_tune_classes(void);

/* Tuning (no reference counterpart; results never depend on it): makes the launches that run several
 * shape classes together (a length-skewed batch's Myers runs: bit 0 of `on`; its score runs: bit 1)
 * persistent: as many workgroups as the GPU holds at once, each taking the next block of pairs from
 * a device-side work queue (the classes' bins in the launch's order, longest chains first) until
 * none is left; a clear bit launches one workgroup per block.  Default 0 (or SF_PERSISTENT from the
 * environment).  Applies to runs started (or captured) afterwards in this process. */
int sf_set_persistent(int on);

/* Tuning (results never depend on it): Myers passes of patterns of 2-24 blocks (65-1,536 bases)
 * with a threshold below the pattern (k + 1 <= 64 * (blocks - 1)), and the reversed pass that finds
 * `start`, compute only the blocks of Ukkonen's band (align.rs:113-149), over the union of a
 * wave's 64 pairs' bands as align.rs:194-350 runs eight texts; 0 computes every block.  Default 1
 * (or SF_MYERS_UKK from the environment).  Applies to passes started afterwards in this process. */
int sf_set_myers_ukk(int on);

/* Tuning (no reference counterpart; results never depend on it): lanes per pair of affine runs.
 * policy 0 (default): per pair, by the launch planner's estimate (score-only runs may use 4 or 8
 * lanes per pair, sixteen or eight pairs per wave, where the 16-bit range holds); 1: every pair in
 * its 8-lane shape wherever the 8-lane kernel's range holds, else 16 up to the 16-lane width limit,
 * else 64; 2: as 1 with the 4-lane shape first; -1: never 4 or 8, else as 0.  Default 0 (or
 * SF_AFFINE_LP8 from the environment).  Applies to affine runs started afterwards in this process. */
int sf_set_affine_lp8(int policy);

/* Tuning (no reference counterpart; results never depend on it): the passes with per-row
 * bookkeeping (sf_affine_dp, sf_batch_affine_dp and the calls built on them) run in the 16-bit
 * packed layout where its range holds -- as the reference runs them in 16 bits where `narrow`
 * holds (align.rs:2381-2426) -- on 8 lanes a pair where the band fits 256 slots, else 16
 * (policy 1, the default, or SF_AFFINE_PASS_PACKED from the environment), on 16 lanes only (2),
 * or in 32 bits throughout (0).  Applies to passes started afterwards in this process. */
int sf_set_affine_pass_packed(int policy);

#ifdef __cplusplus
}
#endif
#endif
