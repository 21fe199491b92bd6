/*
 * sf_oracle.h -- CPU oracle for the alignment hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load this library, and only to check the product's results against (or, for the
 * baseline, as the timed CPU restatement).  The product (libsf_align.so) never links, loads or calls it.
 *
 * A scalar C restatement of spearfish's crates/map/src/align.rs (Rust, v0.3.2), which cannot be
 * built here (no Rust toolchain).  Every function cites the reference lines it follows.  Two
 * parts have no reference counterpart and are this project's own definitions:
 *   - sfo_align():  affine traceback / CIGAR (tie rules in sf_oracle.c, "Traceback").
 *   - sfo_locate(): Myers end and start coordinates.
 * Both are pinned by invariants and brute force in tests/, not by the reference.
 *
 * Sequences are one code per byte; the kernels use `code & 3` exactly as the reference does
 * (align.rs:78, 120, 1054, 2557).  Sizes are uint64_t (Rust usize); SFO_UMAX is usize::MAX.
 */
#ifndef SF_ORACLE_H
#define SF_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFO_INFIX 0
#define SFO_PREFIX 1
#define SFO_GLOBAL 2

#define SFO_NEG (INT32_MIN / 4) /* align.rs:693 */
#define SFO_NONE UINT32_MAX     /* Option<u32>::None */
#define SFO_UMAX UINT64_MAX     /* usize::MAX */

typedef struct {
    int32_t mat, mis, open, ext; /* align.rs:662-668 */
} sfo_scoring;

/* align.rs:2111-2129 `Dp`. */
typedef struct {
    int32_t score;
    uint64_t score_col;
    int32_t clip_best;
    uint64_t clip_i, clip_j;
    uint64_t text_read;
    uint64_t last_row;
    int32_t cut_best;
    int32_t cut_score;
} sfo_dp;

/* ---- Myers (align.rs:40-658) ---- */

/* distance / distance_with (align.rs:87-168): SFO_NONE for None. */
uint32_t sfo_distance(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode,
                      uint32_t k);

/* batch_lanes::<8> (align.rs:552-653): one pattern against `nt` texts, union band.  Results in
 * `out` (SFO_NONE for None). */
void sfo_batch_distances(const uint8_t *pat, uint64_t m, const uint8_t *const *txts,
                         const uint64_t *lens, uint64_t nt, int mode, uint32_t k, uint32_t *out);

/* crates/map/src/lib.rs:851-911 `Mapper::banded` (batches of eight distinct texts, re-capped in
 * order) and lib.rs:3016-3034 `sequential` (its test definition): one u32 per text. */
void sfo_banded(const uint8_t *read, uint64_t m, const uint8_t *const *texts, const uint64_t *lens, uint64_t nt,
                int mode, uint32_t cap, int distinct, uint32_t *out);
void sfo_sequential(const uint8_t *read, uint64_t m, const uint8_t *const *texts, const uint64_t *lens, uint64_t nt,
                    int mode, uint32_t cap, uint32_t *out);

/* Own definition (no reference counterpart): distance as sfo_distance, `end` = smallest j with
 * D[m][j] == distance (exclusive text end; n for Global), `start` = end - L where L is the
 * smallest prefix of reverse(text[..end]) whose Prefix-mode distance to reverse(pattern) is the
 * distance (0 for Prefix and Global).  All three SFO_NONE when the distance exceeds k. */
void sfo_locate(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode,
                uint32_t k, uint32_t *dist, uint32_t *end, uint32_t *start);

/* ---- Affine (align.rs:660-2666) ---- */

/* affine_dp (align.rs:2183-2208) -> run_core with the scalar rows (align.rs:2446-2666,
 * 1018-1072).  `has_clip`/`xdrop`/`clip_cell` mirror `clip: Option<(i32, bool)>`; `cut` is
 * SFO_UMAX for none. */
void sfo_affine_dp(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode,
                   const sfo_scoring *sc, uint64_t band, uint64_t geom, int has_clip, int32_t xdrop,
                   int clip_cell, uint64_t cut, sfo_dp *out);

/* affine (align.rs:699-701). */
int32_t sfo_affine(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode,
                   const sfo_scoring *sc, uint64_t band);

/* affine_extend (align.rs:710-729). */
int32_t sfo_affine_extend(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n,
                          const sfo_scoring *sc, uint64_t band, int32_t end_bonus, int32_t xdrop);

/* affine_extend_window (align.rs:957-982): out4 = (score, pattern bases, text bases, read). */
void sfo_affine_extend_window(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n,
                              uint64_t window, const sfo_scoring *sc, uint64_t band,
                              int32_t end_bonus, int32_t xdrop, int64_t out4[4]);

/* One text of affine_extend_sorted_cut_bonus (align.rs:903-949); the sorted batch equals
 * separate passes (align.rs:2210-2213, test sorted_batch_matches_single align.rs:3243-3310).
 * out3 = (whole, cut, cols). */
void sfo_affine_extend_cut_bonus(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n,
                                 const sfo_scoring *sc, uint64_t band, int32_t end_bonus,
                                 int32_t xdrop, int diagonal, uint64_t window, uint64_t cut,
                                 int track_cut, int bonus_within_cut, int64_t out3[3]);

/* local_band_ends (align.rs:741-760) -> local_band_core (align.rs:783-852). */
int32_t sfo_local_band_ends(const uint8_t *read, uint64_t m, const uint8_t *txt, uint64_t n,
                            const sfo_scoring *sc, int64_t dlo, int64_t dhi, int32_t bonus0,
                            int32_t bonus1);

/* ---- Traceback (own definition; no reference counterpart) ----
 * The pass of sfo_affine_dp (no clip, no cut, geom = n) with every row kept, then a walk from
 * (m, score_col) under the tie rules in sf_oracle.c.  CIGAR ops are BAM-encoded u32
 * (len << 4 | op) with op '=' 7, 'X' 8, 'I' 1 (pattern only), 'D' 2 (text only).
 * Returns the number of ops written (<= cap), or -1 if cap is too small, or -2 if the pass
 * stopped (no alignment; score NEG).  *score, *score_col as sfo_affine_dp; *tstart is the text
 * column where the alignment starts (0 except in Infix mode). */
int64_t sfo_align(const uint8_t *pat, uint64_t m, const uint8_t *txt, uint64_t n, int mode,
                  const sfo_scoring *sc, uint64_t band, uint32_t *cigar, uint64_t cap,
                  int32_t *score, uint64_t *score_col, uint64_t *tstart);

/* Naive definitions from the reference's own tests (align.rs:2801-2827, 2928-2957, 2959-2965,
 * 2969-2983, 2987-3014), used to pin the restatement.  Naive DP over whole matrices. */
uint32_t sfo_naive_edit(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, int mode);
/* Every D[m][j], j = 0..n (n + 1 values), by the naive DP. */
void sfo_naive_edit_lastrow(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, int mode,
                            uint32_t *row);
int32_t sfo_naive_affine(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n, int mode,
                         const sfo_scoring *sc);
int32_t sfo_naive_extend(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n,
                         const sfo_scoring *sc);
int32_t sfo_naive_local(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n,
                        const sfo_scoring *sc, int32_t b0, int32_t b1);
int32_t sfo_naive_local_band(const uint8_t *p, uint64_t m, const uint8_t *t, uint64_t n,
                             const sfo_scoring *sc, int64_t dlo, int64_t dhi, int32_t b0,
                             int32_t b1);

/* Cells a pass computes, as the reference's DP_WORK counts them (align.rs:2573-2578):
 * sum over rows i = 1..m of (hi + 1 - lo) for a banded pass with geom = n, until a stop. */
uint64_t sfo_affine_cells(uint64_t m, uint64_t n, int mode, uint64_t band);

/* Multi-threaded batch drivers for the CPU baseline: pairs i in [0, npairs), sequences given
 * by offsets into one code buffer each.  `threads` workers pull pairs dynamically. */
void sfo_batch_myers(const uint8_t *pat, const uint64_t *poff, const uint32_t *plen,
                     const uint8_t *txt, const uint64_t *toff, const uint32_t *tlen,
                     uint64_t npairs, int mode, uint32_t k, uint32_t *dist, int threads);
void sfo_batch_affine(const uint8_t *pat, const uint64_t *poff, const uint32_t *plen,
                      const uint8_t *txt, const uint64_t *toff, const uint32_t *tlen,
                      uint64_t npairs, int mode, const sfo_scoring *sc, const uint64_t *band,
                      int32_t *score, uint64_t *score_col, int threads);
/* The same with AVX-512 rows (the reference's run_core::<Avx512Rows>, align.rs:1150-1267):
 * equal results; returns -1 (and computes nothing) on a host without AVX-512 F/BW/VL. */
int sfo_avx512_available(void);
int sfo_batch_affine_avx512(const uint8_t *pat, const uint64_t *poff, const uint32_t *plen,
                            const uint8_t *txt, const uint64_t *toff, const uint32_t *tlen,
                            uint64_t npairs, int mode, const sfo_scoring *sc, const uint64_t *band,
                            int32_t *score, uint64_t *score_col, int threads);

/* crates/index/src/lib.rs:104-114 `packed_eq` on u64 words (base i at bits 2 (i % 32) of word i / 32). */
int sfo_packed_eq(const uint64_t *words, uint64_t nwords, uint64_t a, uint64_t b, uint32_t len);

/* crates/map/src/lib.rs:1307-1492 `seg` (with `tb.affine` set, `affine_within` unset) over many
 * segments: segment s aligns reads[read_off, +read_len) against candidates cand0 .. cand0+ncand-1,
 * each a range of the packed store `words`; per candidate out[3c..3c+2] = (whole, cut, text bases)
 * and rep[c] = its first equal candidate counted from cand0.  Flags as sf_align.h's SF_SEG_*.
 * Threads over segments; avx512 runs the reference's x86 rows (32-bit AVX-512 rows for the
 * extension passes, 16-bit rows where `narrow` holds for the others); -1 if the host lacks them. */
#define SFO_SEG_REVERSE 1
#define SFO_SEG_EXTEND 2
#define SFO_SEG_DIAGONAL 4
#define SFO_SEG_TRACK_CUT 8
#define SFO_SEG_BONUS_WITHIN_CUT 16
typedef struct {
    uint64_t read_off;
    uint32_t read_len, cand0, ncand;
    int32_t mode;
    uint32_t flags;
    uint64_t window, cut;
} sfo_tb_segment;
int sfo_tiebreak_segments(const uint64_t *words, uint64_t bases, const uint8_t *reads, const sfo_tb_segment *segs,
                          uint64_t nseg, const uint64_t *cand_off, const uint32_t *cand_len, const sfo_scoring *sc,
                          uint64_t band, int32_t end_bonus, int32_t xdrop, int64_t *out, uint32_t *rep, int threads,
                          int avx512);

#ifdef __cplusplus
}
#endif
#endif
