"""Synthetic tie-break calls for the tests: an isoform store (genes of exons, transcripts that keep
some of them, packed back to back) and segments of the reference's three kinds (crates/map/src/lib.rs
1307-1492 `seg` as `segmented` makes its calls): right read ends (Prefix, windows of
`tb.window(read length)` bases, cut at the shortest window that stops at its transcript's end), left
read ends (the same, reversed), interior stretches (Global), plus Infix."""
from oracle import sf_oracle as O
from spearfish_amd import tiebreak as T
from spearfish_amd.align import Affine, Mode


def _seq(rng, n):
    return bytes(rng.randrange(4) for _ in range(n))


def _mutate(rng, s, rate=0.08):
    out = bytearray()
    for b in s:
        r = rng.random()
        if r < rate / 3:  # deletion
            continue
        if r < 2 * rate / 3:  # insertion before the base
            out.append(rng.randrange(4))
        out.append(rng.randrange(4) if r >= 2 * rate / 3 and r < rate else b)  # substitution
    return bytes(out)


def make_store(rng, genes=30):
    codes = bytearray()
    out = []
    for _ in range(genes):
        exons = [_seq(rng, rng.randrange(6, 140)) for _ in range(rng.randrange(2, 7))]
        isos = []
        for _ in range(rng.randrange(2, 8)):
            keep = [e for e in range(len(exons)) if rng.random() < 0.7] or [rng.randrange(len(exons))]
            codes += _seq(rng, rng.randrange(0, 40))  # intergenic bases: offsets at any word position
            off, pos, p = len(codes), {}, 0
            for e in keep:
                pos[e] = p
                p += len(exons[e])
                codes += exons[e]
            isos.append((off, p, pos))
        out.append((exons, isos))
    return bytes(codes), out


def make_segments(rng, codes, genes, tb):
    """Segments of the three kinds; returns (reads buffer, [Segment])."""
    reads, segs = bytearray(), []

    def add_read(r):
        reads.extend(r)
        return len(reads) - len(r), len(r)

    for exons, isos in genes:
        for e in range(len(exons)):
            has = [iso for iso in isos if e in iso[2]]
            if not has:
                continue
            kind = rng.randrange(4)
            if kind == 0:  # right end from exon e's start
                base = codes[has[0][0] + has[0][2][e]: has[0][0] + has[0][1]]
                read = _mutate(rng, base[: rng.randrange(0, 90)])
                w = tb.window(len(read))
                cands = [(off + pos[e], min(w, ln - pos[e])) for off, ln, pos in has]
                short = [ln - pos[e] < w for off, ln, pos in has]
                lens = [c[1] for c in cands]
                cut = min((l for l, s in zip(lens, short) if s), default=None)
                if cut is None or not (tb.flank_min > 0 or tb.end_bonus_cut) or not any(l > cut for l in lens):
                    cut = None
                ro, rl = add_read(read)
                segs.append(T.Segment(ro, rl, cands, Mode.Prefix, False, cut))
            elif kind == 1:  # left end up to exon e's end, aligned reversed
                ends = [(off, pos[e] + len(exons[e])) for off, ln, pos in has]
                base = codes[ends[0][0]: ends[0][0] + ends[0][1]]
                read = _mutate(rng, base[max(0, len(base) - rng.randrange(0, 90)):])
                w = tb.window(len(read))
                cands = [(off + max(0, end - w), min(w, end)) for off, end in ends]
                lens = [c[1] for c in cands]
                short = [end < w for off, end in ends]
                cut = min((l for l, s in zip(lens, short) if s), default=None)
                if cut is None or not (tb.flank_min > 0 or tb.end_bonus_cut) or not any(l > cut for l in lens):
                    cut = None
                ro, rl = add_read(read)
                segs.append(T.Segment(ro, rl, cands, Mode.Prefix, True, cut))
            else:  # an interior stretch from exon e's start, some running into the next exon
                extra = rng.randrange(0, 12)
                cands = [(off + pos[e], min(len(exons[e]) + extra, ln - pos[e])) for off, ln, pos in has]
                read = _mutate(rng, exons[e])
                ro, rl = add_read(read)
                segs.append(T.Segment(ro, rl, cands, Mode.Global if kind == 2 else Mode.Infix))
    return bytes(reads), segs


def flags_of(s, tb):
    f = O.SEG_REVERSE if s.rev else 0
    if s.mode == Mode.Prefix and tb.clip_ends:
        f |= O.SEG_EXTEND
    f |= (O.SEG_DIAGONAL if tb.end_diagonal else 0) | (O.SEG_TRACK_CUT if tb.flank_min > 0 else 0)
    return f | (O.SEG_BONUS_WITHIN_CUT if tb.end_bonus_cut else 0)


def oracle(words, reads, segs, tb):
    res, reps = [], []
    for s in segs:
        r, rep = O.tiebreak_segment(words, reads[s.read_off: s.read_off + s.read_len], list(s.cands), int(s.mode),
                                    flags_of(s, tb), tb.window(s.read_len), O.UMAX if s.cut is None else s.cut,
                                    (tb.affine.mat, tb.affine.mis, tb.affine.open, tb.affine.ext),
                                    tb.band, tb.end_bonus, tb.xdrop)
        res += [tuple(x) for x in r]
        reps += rep
    return res, reps


TBS = {
    "default": T.TieBreak(),
    "hifi_band64": T.TieBreak(affine=Affine.HIFI, band=64),
    "diagonal": T.TieBreak(end_diagonal=True),
    "flank_min": T.TieBreak(flank_min=12, end_bonus_cut=False),
    "flank_min_bonus_cut": T.TieBreak(flank_min=12),
    "no_clip_ends": T.TieBreak(clip_ends=False, band=8),
    "xdrop_small_band_wide": T.TieBreak(xdrop=8, band=300, end_bonus=9),
}
