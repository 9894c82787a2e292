"""GPU parity of the pairwise functions: align_pair, align_pairs, align_pair_position_max and
score2 must equal the pinned oracle bit for bit.

Mirrors the reference's unit tests of src/kernel.rs (position-max hand cases and proptest,
align_pairs against per-pair align_pair across modes, search types and widths with empty
geometries, scratch reuse, error propagation, score2 cases), then widens to every short pair in
every mode, random schemes, protein and DNA batches large enough for the one-pair-per-lane kernel,
long and multi-pass pairs, each of the three batch kernels forced in turn (one pair per lane, per
wavefront, per multi-wavefront band), and the band kernel at every wavefront count it launches.
"""
import random

import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from common import (ALL_MODES, BLOSUM62, MODE_SW, all_sequences, dna_scheme, identity_matrix)

pytestmark = pytest.mark.gpu

ST = (hy.SearchType.Score, hy.SearchType.ScoreEnd)


def scoring(scheme):
    al, mat, go, ge = scheme
    return hy.Scoring(al, mat, go, ge)


def want_hit(q, t, scheme, mode, st, index=0):
    al, mat, go, ge = scheme
    s, qe, te = orc.align_core(q, t, al, mat, go, ge, mode)
    if st == hy.SearchType.Score:
        qe = te = None
    return hy.BestHit(s, index, qe, te)


@pytest.fixture(params=["lanes", "waves", "band"])
def kernel(request, monkeypatch):
    monkeypatch.setenv("HYALITE_HIP_PAIRS_KERNEL", request.param)
    return request.param


def check_pairs(pairs, scheme, mode, st, scratch=None):
    got = hy.align_pairs(pairs, scoring(scheme), hy.Mode(mode), st, scratch=scratch)
    assert len(got) == len(pairs)
    for k, (q, t) in enumerate(pairs):
        assert got[k] == want_hit(list(q), list(t), scheme, mode, st, k), \
            (hy.Mode(mode), st, k, list(q)[:20], list(t)[:20])


# ---- src/kernel.rs unit tests ---------------------------------------------------------------------

def mk(seed, n):
    return [((seed * (i + 1)) >> 3) % 4 for i in range(n)]


def test_align_pairs_matches_per_pair_and_reuses_out(kernel):
    """src/kernel.rs:1104-1153: lengths {0, 1, 4, 9, 20} crossed, three widths, every mode and
    search type; a reused `out` is cleared; an out-of-range symbol in any pair surfaces."""
    s_i8 = (4, identity_matrix(4, 2, -1), 2, 1)
    s_i16 = (4, identity_matrix(4, 20, -5), 8, 2)
    s_i32 = (4, identity_matrix(4, 3000, -1000), 500, 100)
    lens = [0, 1, 4, 9, 20]
    pairs = [(mk(0x1000 + a, ql), mk(0x2000 + b, tl))
             for a, ql in enumerate(lens) for b, tl in enumerate(lens)]
    sc = hy.PairScratch()
    out = []
    for scheme in (s_i8, s_i16, s_i32):
        for mode in ALL_MODES:
            for st in ST:
                hy.align_pairs(pairs, scoring(scheme), hy.Mode(mode), st, out=out, scratch=sc)
                assert len(out) == len(pairs)
                for k, (q, t) in enumerate(pairs):
                    assert out[k] == want_hit(q, t, scheme, mode, st, k), (mode, st, k)
    hy.align_pairs(pairs[:3], scoring(s_i8), hy.Mode.Sw, hy.SearchType.Score, out=out,
                   scratch=sc)
    assert len(out) == 3
    bad = [([0, 1], [0]), ([9], [0])]
    with pytest.raises(hy.SymbolOutOfRange):
        hy.align_pairs(bad, scoring(s_i8), hy.Mode.Nw, hy.SearchType.Score, out=out, scratch=sc)
    assert hy.align_pairs([], scoring(s_i8), hy.Mode.Nw, hy.SearchType.Score) == []


def test_pair_scratch_reuse_matches_one_shot():
    """src/kernel.rs:1156-1204: grow-then-shrink lengths through one scratch."""
    s_i8 = (4, identity_matrix(4, 2, -1), 2, 1)
    s_i16 = (4, identity_matrix(4, 20, -5), 8, 2)
    lens = [40, 3, 55, 1, 30, 0, 12, 48, 2]
    sc = hy.PairScratch()
    for k, ql in enumerate(lens):
        tl = lens[(k + 3) % len(lens)]
        for scheme in (s_i8, s_i16):
            q, t = mk(0x9E3779B1 + k, ql), mk(0x12345678 + k, tl)
            for mode in ALL_MODES:
                for st in ST:
                    got = hy.align_pair_with(sc, q, t, scoring(scheme), hy.Mode(mode), st)
                    assert got == want_hit(q, t, scheme, mode, st), (mode, st, ql, tl)
                    assert got == hy.align_pair(q, t, scoring(scheme), hy.Mode(mode), st)
            out_r, out_o = [], []
            a = hy.align_pair_position_max_with(sc, q, t, scoring(scheme), out_r)
            b = hy.align_pair_position_max(q, t, scoring(scheme), out_o)
            assert a == b and out_r == out_o
            col, s, te = orc.position_max(q, t, *scheme)
            assert out_r == col and (a.score, a.target_end, a.query_end) == (s, te, None)


def test_position_max_hand_computed_cases():
    """src/kernel.rs:1028-1059."""
    s = scoring(dna_scheme())
    out = []
    hit = hy.align_pair_position_max([0, 1, 2, 3], [0, 1, 2, 3], s, out)
    assert out == [2, 4, 6, 8] and hit.score == 8 and (hit.query_end, hit.target_end) == (None, 3)
    hit = hy.align_pair_position_max([0], [1, 1, 0], s, out)
    assert out == [0, 0, 2] and hit.score == 2 and hit.target_end == 2
    hit = hy.align_pair_position_max([0, 0], [1, 1], s, out)
    assert out == [0, 0] and hit.score == 0 and hit.target_end is None
    hit = hy.align_pair_position_max([0, 1], [], s, out)
    assert out == [] and hit.score == 0 and hit.target_end is None
    hit = hy.align_pair_position_max([], [0, 1, 2], s, out)
    assert out == [0, 0, 0] and hit.target_end is None


def test_position_max_out_of_range_symbol_is_reported():
    """src/kernel.rs:1062-1075: the error carries the symbol; `out` is cleared."""
    out = [1, 2, 3]
    with pytest.raises(hy.SymbolOutOfRange) as e:
        hy.align_pair_position_max([0, 5], [0], scoring(dna_scheme()), out)
    assert (e.value.symbol, e.value.alphabet_len) == (5, 4)
    assert out == []


def test_position_max_random_schemes():
    """src/kernel.rs:1076-1102: array, score and target end equal the oracle's."""
    rng = random.Random(3)
    sc = hy.PairScratch()
    for _ in range(150):
        al = rng.randint(2, 5)
        mat = [rng.randint(-6, 6) for _ in range(al * al)]
        go = rng.randint(0, 7)
        ge = rng.randint(0, go)
        q = [rng.randrange(al) for _ in range(rng.randrange(0, 30))]
        t = [rng.randrange(al) for _ in range(rng.randrange(0, 30))]
        out = []
        hit = hy.align_pair_position_max_with(sc, q, t, hy.Scoring(al, mat, go, ge), out)
        col, s, te = orc.position_max(q, t, al, mat, go, ge)
        assert (out, hit.score, hit.target_end) == (col, s, te), (q, t, al, mat, go, ge)


@pytest.mark.parametrize("qlen,tlen", [(100, 3000), (300, 700), (600, 1200), (1100, 5000)])
def test_position_max_protein_shapes(qlen, tlen):
    """Every rows-per-lane shape and multi-pass queries (beyond 512 rows)."""
    rng = np.random.default_rng(qlen)
    q = rng.integers(0, 20, size=qlen, dtype=np.uint8)
    t = np.concatenate([rng.integers(0, 20, size=tlen // 2, dtype=np.uint8), q[: qlen // 2],
                        rng.integers(0, 20, size=tlen // 2, dtype=np.uint8)])
    out = []
    hit = hy.align_pair_position_max(q, t, hy.Scoring(20, BLOSUM62, 11, 1), out)
    col, s, te = orc.position_max(q, t, 20, BLOSUM62, 11, 1)
    assert out == col and (hit.score, hit.target_end) == (s, te)


def test_score2_hand_computed_cases():
    """src/kernel.rs:1245-1276."""
    assert hy.score2([0, 5, 0, 0, 0, 0, 0, 0, 4, 0], 5, 1, 1, 1) == (4, 8)
    assert hy.score2([0, 3, 5, 4, 0], 5, 2, 5, 1) is None
    assert hy.score2([3, 5, 4, 6], 1, 20, 1, 1) == (6, 3)
    assert hy.score2([7, 0, 7], 1, 10, 1, 1) == (7, 0)
    assert hy.score2([0, 0, 0], 0, 0, 1, 1) is None
    assert hy.score2([5, 5], 5, 0, 0, 1) is None
    assert hy.score2([2, 0, 0, 0, 0, 0, 9], 9, 6, 9, 5) is None
    assert hy.score2([2, 0, 0, 0, 0, 0, 9], 9, 6, 9, 1) == (2, 0)


def test_score2_matches_the_oracle_on_random_arrays():
    rng = random.Random(8)
    for _ in range(2000):
        colmax = [rng.randint(0, 20) for _ in range(rng.randint(0, 40))]
        args = (rng.randint(0, 40), rng.randint(0, 40), rng.randint(-2, 8), rng.randint(1, 6))
        assert hy.score2(colmax, *args) == orc.score2(colmax, *args)


def test_score2_on_a_device_position_max():
    """The mate-rescue flow: position maxima of a read against a window holding two copies."""
    rng = np.random.default_rng(4)
    read = rng.integers(0, 4, size=120, dtype=np.uint8)
    win = np.concatenate([rng.integers(0, 4, size=300, dtype=np.uint8), read,
                          rng.integers(0, 4, size=200, dtype=np.uint8), read[:100],
                          rng.integers(0, 4, size=100, dtype=np.uint8)])
    s = scoring(dna_scheme())
    out = []
    hit = hy.align_pair_position_max(read, win, s, out)
    got = hy.score2(out, hit.score, hit.target_end, 2, 20)
    col, bs, te = orc.position_max(read, win, *dna_scheme())
    assert got == orc.score2(col, bs, te, 2, 20) and got is not None and got[0] >= 190


# ---- widened parity -------------------------------------------------------------------------------

SCHEMES = [
    dna_scheme(),
    (4, identity_matrix(4, 1, -3), 10, 5),
    (4, identity_matrix(4, 3, -2), 4, 0),
    (4, identity_matrix(4, 5, -1), 1, 1),
]


@pytest.mark.parametrize("scheme_ix", range(4))
def test_every_short_pair_in_one_batch(scheme_ix, kernel):
    scheme = SCHEMES[scheme_ix]
    seqs = all_sequences(3, 3)
    pairs = [(q, t) for q in seqs for t in seqs]
    for mode in ALL_MODES:
        for st in ST:
            check_pairs(pairs, scheme, mode, st)


def test_random_schemes_and_lengths(kernel):
    rng = random.Random(21)
    for _ in range(25):
        al = rng.randint(2, 6)
        mat = [rng.randint(-6, 6) for _ in range(al * al)]
        go = rng.randint(0, 8)
        ge = rng.randint(0, go)
        pairs = [([rng.randrange(al) for _ in range(rng.randrange(0, 70))],
                  [rng.randrange(al) for _ in range(rng.randrange(0, 90))]) for _ in range(150)]
        for mode in ALL_MODES:
            check_pairs(pairs, (al, mat, go, ge), mode, hy.SearchType.ScoreEnd)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_protein_pairs_across_tiles_and_long_pairs(mode, kernel):
    """Queries past several 16/32-row tiles, targets of 0-2000 residues, related pairs (ties and
    high scores), and pairs long enough to leave the lane kernel in the default split."""
    rng = np.random.default_rng(40 + mode)
    pairs = []
    for _ in range(90):
        q = rng.integers(0, 20, size=int(rng.integers(0, 400)), dtype=np.uint8)
        if rng.random() < 0.5 and q.size:
            t = np.concatenate([rng.integers(0, 20, size=int(rng.integers(0, 300)),
                                             dtype=np.uint8), q,
                                rng.integers(0, 20, size=int(rng.integers(0, 300)),
                                             dtype=np.uint8)])
        else:
            t = rng.integers(0, 20, size=int(rng.integers(0, 2000)), dtype=np.uint8)
        pairs.append((q, t))
    pairs.append((rng.integers(0, 20, size=700, dtype=np.uint8),
                  rng.integers(0, 20, size=3000, dtype=np.uint8)))
    q = rng.integers(0, 20, size=2300, dtype=np.uint8)
    pairs.append((q, np.concatenate([rng.integers(0, 20, size=200, dtype=np.uint8), q[400:1500],
                                     rng.integers(0, 20, size=150, dtype=np.uint8)])))
    for st in ST:
        check_pairs(pairs, (20, BLOSUM62, 11, 1), mode, st)


def test_default_split_on_a_large_read_batch():
    """40,000 DNA read pairs (reads 100-150 nt, windows 200-600 nt) take the one-pair-per-lane
    kernel by default; a few long pairs run one per wavefront beside it. Checked against the
    oracle on a sample spread over the batch, and on every pair against the forced one-pair-per-
    wavefront run."""
    rng = np.random.default_rng(77)
    n = 40000
    pairs = []
    for k in range(n):
        r = rng.integers(0, 4, size=int(rng.integers(100, 151)), dtype=np.uint8)
        w = rng.integers(0, 4, size=int(rng.integers(200, 601)), dtype=np.uint8)
        if k % 3 == 0:
            at = int(rng.integers(0, w.size - 60))
            w[at:at + 60] = r[:60]
        pairs.append((r, w))
    pairs[5] = (rng.integers(0, 4, size=5000, dtype=np.uint8),
                rng.integers(0, 4, size=9000, dtype=np.uint8))
    s = scoring(dna_scheme())
    for mode in (MODE_SW, 2):
        got = hy.align_pairs(pairs, s, hy.Mode(mode), hy.SearchType.ScoreEnd)
        for k in list(range(0, n, 97)) + [5]:
            q, t = pairs[k]
            assert got[k] == want_hit(list(q), list(t), dna_scheme(), mode,
                                      hy.SearchType.ScoreEnd, k), (mode, k)
        import os
        os.environ["HYALITE_HIP_PAIRS_KERNEL"] = "waves"
        try:
            ref = hy.align_pairs(pairs, s, hy.Mode(mode), hy.SearchType.ScoreEnd)
        finally:
            del os.environ["HYALITE_HIP_PAIRS_KERNEL"]
        assert got == ref


def test_pair_errors_follow_the_reference_order():
    s = scoring(dna_scheme())
    with pytest.raises(hy.SymbolOutOfRange) as e:
        hy.align_pair([0, 7], [9], s, hy.Mode.Nw, hy.SearchType.Score)
    assert e.value.symbol == 7          # the query is checked first
    with pytest.raises(hy.SymbolOutOfRange) as e:
        hy.align_pair([0, 1], [0, 8], s, hy.Mode.Nw, hy.SearchType.Score)
    assert e.value.symbol == 8
    huge = hy.Scoring(2, [0, 0, 0, 0], 2_000_000_000, 1)
    with pytest.raises(hy.ScoreRangeTooWide) as e:
        hy.align_pair([0], [1], huge, hy.Mode.Nw, hy.SearchType.Score)
    assert e.value.bound == 2_000_000_001
    with pytest.raises(hy.SymbolOutOfRange):
        hy.align_pairs([([0], [1]), ([1], [0, 4]), ([5], [0])], s, hy.Mode.Sw,
                       hy.SearchType.Score)


def test_flat_batch_equals_the_pair_list():
    """align_pairs_flat over one buffer per side (offsets not starting at 0, as a slice of a
    larger buffer would give) equals align_pairs over the same pairs."""
    rng = np.random.default_rng(12)
    lens_q = rng.integers(0, 200, size=300)
    lens_t = rng.integers(0, 500, size=300)
    qf = rng.integers(0, 4, size=int(lens_q.sum()) + 7, dtype=np.uint8)
    tf = rng.integers(0, 4, size=int(lens_t.sum()) + 5, dtype=np.uint8)
    qo = np.concatenate([[7], 7 + np.cumsum(lens_q)]).astype(np.uint64)
    to = np.concatenate([[5], 5 + np.cumsum(lens_t)]).astype(np.uint64)
    pairs = [(qf[qo[k]:qo[k + 1]], tf[to[k]:to[k + 1]]) for k in range(300)]
    s = scoring(dna_scheme())
    sc = hy.PairScratch()
    for mode in ALL_MODES:
        sco, qe, te = hy.align_pairs_flat(qf, qo, tf, to, s, hy.Mode(mode),
                                          hy.SearchType.ScoreEnd, scratch=sc)
        hits = hy.align_pairs(pairs, s, hy.Mode(mode), hy.SearchType.ScoreEnd, scratch=sc)
        for k, h in enumerate(hits):
            assert (h.score, -1 if h.query_end is None else h.query_end,
                    -1 if h.target_end is None else h.target_end) == (sco[k], qe[k], te[k])
            assert h == want_hit(list(pairs[k][0]), list(pairs[k][1]), dna_scheme(), mode,
                                 hy.SearchType.ScoreEnd, k)
    ms, calls = sc.profile_read(reset=True)
    assert calls == 10 and ms > 0
    with pytest.raises(ValueError):
        hy.align_pairs_flat(qf, qo[:-1], tf, to, s, hy.Mode.Sw, hy.SearchType.Score)


# ---- the band kernel (one pair per multi-wavefront workgroup) -------------------------------------

# Query lengths on both sides of every rows-per-pass class of the default launch shapes (64,
# 128, 256, 512, 1,024, 2,048 and 4,096 rows) and of multi-pass queries.
BAND_QLENS = [1, 63, 64, 65, 128, 129, 256, 257, 512, 513, 1024, 1025, 2048, 2049, 4100]
# Every launch shape (rows per lane, wavefronts) the band launcher instantiates.
BAND_SHAPES = [(1, 1), (1, 2), (1, 4), (2, 1), (2, 2), (2, 4), (2, 8), (2, 16), (4, 1), (4, 2),
               (4, 4), (4, 8), (4, 16)]


def band_pair(rng, m, n, al):
    """A query of m and a target of n symbols; the target holds a mutated slice of the query so
    scores are high and ends sit inside the band boundaries."""
    q = rng.integers(0, al, size=m, dtype=np.uint8)
    t = rng.integers(0, al, size=n, dtype=np.uint8)
    if m and n > 8:
        k = min(m, n) // 2
        at = int(rng.integers(0, n - k + 1))
        src = int(rng.integers(0, m - k + 1))
        t[at:at + k] = q[src:src + k]
        t[at:at + k:7] = (t[at:at + k:7] + 1) % al
    return q, t


@pytest.mark.parametrize("mode", ALL_MODES)
def test_band_single_pairs_at_every_wavefront_count(mode, monkeypatch):
    """align_pair (band kernel by default) equals the oracle at every rows-per-pass class and
    across pass boundaries, for targets shorter than, equal to and longer than one wavefront's
    64-step skew; the forced one-wavefront kernel gives the same hits."""
    rng = np.random.default_rng(500 + mode)
    sc = hy.PairScratch()
    al, mat, go, ge = (4, identity_matrix(4, 3, -2), 5, 1)
    s = hy.Scoring(al, mat, go, ge)
    for m in BAND_QLENS:
        for n in (1, 63, 64, 65, 300):
            q, t = band_pair(rng, m, n, al)
            want = orc.align_core(q, t, al, mat, go, ge, mode)
            for st in ST:
                got = hy.align_pair_with(sc, q, t, s, hy.Mode(mode), st)
                exp = want_hit(q, t, (al, mat, go, ge), mode, st)
                assert got == exp, (hy.Mode(mode), st, m, n, got, want)
    for m in (65, 1025, 2049):
        q, t = band_pair(rng, m, 700, al)
        monkeypatch.setenv("HYALITE_HIP_PAIRS_KERNEL", "waves")
        one = hy.align_pair_with(sc, q, t, s, hy.Mode(mode), hy.SearchType.ScoreEnd)
        monkeypatch.delenv("HYALITE_HIP_PAIRS_KERNEL")
        assert hy.align_pair_with(sc, q, t, s, hy.Mode(mode), hy.SearchType.ScoreEnd) == one


def test_band_position_max_at_every_wavefront_count():
    """align_pair_position_max (band kernel, column maxima folded by every wavefront) equals the
    oracle's array, best and first column at every class and across pass boundaries."""
    rng = np.random.default_rng(600)
    sc = hy.PairScratch()
    al, mat, go, ge = dna_scheme()
    s = hy.Scoring(al, mat, go, ge)
    for m in BAND_QLENS:
        for n in (1, 64, 65, 900):
            q, t = band_pair(rng, m, n, al)
            out = []
            hit = hy.align_pair_position_max_with(sc, q, t, s, out)
            col, bs, te = orc.position_max(q, t, al, mat, go, ge)
            assert out == col and (hit.score, hit.target_end) == (bs, te), (m, n)


def test_band_batch_classes_and_multipass_chunks(monkeypatch):
    """A batch forced onto the band kernel: pairs of every rows-per-pass class in one call, in
    shuffled order (one launch per class), multi-pass queries beside single-pass ones; every hit
    equals the oracle and the one-wavefront kernel's."""
    rng = np.random.default_rng(700)
    al, mat, go, ge = (20, BLOSUM62, 11, 1)
    pairs = []
    for m in BAND_QLENS + [0, 30, 700, 3000]:
        for n in (0, 50, 400):
            pairs.append(band_pair(rng, m, n, al))
    order = rng.permutation(len(pairs))
    pairs = [pairs[k] for k in order]
    s = hy.Scoring(al, mat, go, ge)
    for mode in ALL_MODES:
        monkeypatch.setenv("HYALITE_HIP_PAIRS_KERNEL", "band")
        got = hy.align_pairs(pairs, s, hy.Mode(mode), hy.SearchType.ScoreEnd)
        for k, (q, t) in enumerate(pairs):
            assert got[k] == want_hit(q, t, (al, mat, go, ge), mode, hy.SearchType.ScoreEnd, k), \
                (hy.Mode(mode), k, len(q), len(t))
        monkeypatch.setenv("HYALITE_HIP_PAIRS_KERNEL", "waves")
        assert hy.align_pairs(pairs, s, hy.Mode(mode), hy.SearchType.ScoreEnd) == got


@pytest.mark.parametrize("shape", BAND_SHAPES)
def test_band_every_launch_shape_over_several_passes(shape, monkeypatch):
    """Each launch shape forced in turn on queries of two passes and a ragged third, targets
    shorter and longer than the 256-slot rings: scores and ends in local and global mode, and the
    column maxima, equal the oracle's."""
    r, w = shape
    monkeypatch.setenv("HYALITE_HIP_BAND_SHAPE", f"{r},{w}")
    rng = np.random.default_rng(800 + 100 * r + w)
    al, mat, go, ge = (4, identity_matrix(4, 3, -2), 5, 1)
    s = hy.Scoring(al, mat, go, ge)
    sc = hy.PairScratch()
    rows = 64 * r * w
    for m in (rows - 1, 2 * rows + 3):
        for n in (70, 300):
            q, t = band_pair(rng, m, n, al)
            for mode in (MODE_SW, 1):
                for st in ST:
                    got = hy.align_pair_with(sc, q, t, s, hy.Mode(mode), st)
                    assert got == want_hit(q, t, (al, mat, go, ge), mode, st), (shape, m, n, mode)
            out = []
            hit = hy.align_pair_position_max_with(sc, q, t, s, out)
            col, bs, te = orc.position_max(q, t, al, mat, go, ge)
            assert out == col and (hit.score, hit.target_end) == (bs, te), (shape, m, n)


def test_band_batches_on_both_upload_paths(monkeypatch):
    """Batches forced onto the band kernel through the single-upload path for small batches
    (under 1 MB of sequences) and the general upload path (a batch of about 2 MB), in local and
    global mode with ends: every hit of a sample equals the oracle, and the whole batch equals
    the one-wavefront kernel's."""
    rng = np.random.default_rng(900)
    al, mat, go, ge = dna_scheme()
    s = hy.Scoring(al, mat, go, ge)
    for npairs in (300, 4000):
        pairs = [band_pair(rng, int(rng.integers(0, 300)), int(rng.integers(0, 420)), al)
                 for _ in range(npairs)]
        for mode in (MODE_SW, 2):
            monkeypatch.setenv("HYALITE_HIP_PAIRS_KERNEL", "band")
            got = hy.align_pairs(pairs, s, hy.Mode(mode), hy.SearchType.ScoreEnd)
            for k in range(0, npairs, 37):
                q, t = pairs[k]
                assert got[k] == want_hit(q, t, (al, mat, go, ge), mode, hy.SearchType.ScoreEnd,
                                          k), (npairs, mode, k)
            monkeypatch.setenv("HYALITE_HIP_PAIRS_KERNEL", "waves")
            assert hy.align_pairs(pairs, s, hy.Mode(mode), hy.SearchType.ScoreEnd) == got


def test_large_batch_uploads_through_pinned_slots():
    """A batch whose query and target buffers both exceed 8 MB: both go up through the pinned
    slots in one call, each worker refilling its slots, the target upload reusing slots whose
    query pieces may still be in flight, and the next call reusing them again. Every pair equals
    the same pair aligned in sub-batches small enough for plain copies (a different kernel), and
    a sample equals the oracle."""
    rng = np.random.default_rng(1100)
    n = 100000
    ql = rng.integers(100, 141, size=n)
    tl = rng.integers(200, 301, size=n)
    q = rng.integers(0, 4, size=int(ql.sum()), dtype=np.uint8)
    t = rng.integers(0, 4, size=int(tl.sum()), dtype=np.uint8)
    qo = np.concatenate([[0], np.cumsum(ql)]).astype(np.uint64)
    to = np.concatenate([[0], np.cumsum(tl)]).astype(np.uint64)
    for k in range(0, n, 2):          # every other window holds its read
        at = int(to[k]) + int(rng.integers(0, tl[k] - ql[k] + 1))
        t[at:at + ql[k]] = q[qo[k]:qo[k + 1]]
    assert q.size > (8 << 20) and t.size > (8 << 20)
    scheme = dna_scheme()
    s = scoring(scheme)
    sc = hy.PairScratch()
    st = hy.SearchType.ScoreEnd
    whole = hy.align_pairs_flat(q, qo, t, to, s, hy.Mode.Sw, st, scratch=sc)
    again = hy.align_pairs_flat(q, qo, t, to, s, hy.Mode.Sw, st, scratch=sc)
    for a, b in zip(whole, again):
        assert np.array_equal(a, b)
    step = 20000
    for k0 in range(0, n, step):
        k1 = k0 + step
        qa, qb, ta, tb = int(qo[k0]), int(qo[k1]), int(to[k0]), int(to[k1])
        part = hy.align_pairs_flat(q[qa:qb], qo[k0:k1 + 1] - qo[k0], t[ta:tb],
                                   to[k0:k1 + 1] - to[k0], s, hy.Mode.Sw, st, scratch=sc)
        for a, b in zip(whole, part):
            bad = np.flatnonzero(a[k0:k1] != b)
            assert bad.size == 0, ("pair", k0 + int(bad[0]) if bad.size else None)
    for k in range(0, n, 4999):
        want = want_hit(q[qo[k]:qo[k + 1]], t[to[k]:to[k + 1]], scheme, MODE_SW, st)
        assert (whole[0][k], whole[1][k], whole[2][k]) == (
            want.score, -1 if want.query_end is None else want.query_end,
            -1 if want.target_end is None else want.target_end), k
    # Symbols are checked on the host while the buffers are copied up: a bad one late in either
    # buffer is still reported, with the reference's error.
    bad_t = t.copy()
    bad_t[int(to[n - 2]) + 3] = 7
    with pytest.raises(hy.SymbolOutOfRange) as e:
        hy.align_pairs_flat(q, qo, bad_t, to, s, hy.Mode.Sw, st, scratch=sc)
    assert e.value.symbol == 7
    bad_q = q.copy()
    bad_q[int(qo[n // 2]) + 1] = 200
    with pytest.raises(hy.SymbolOutOfRange) as e:
        hy.align_pairs_flat(bad_q, qo, t, to, s, hy.Mode.Sw, st, scratch=sc)
    assert e.value.symbol == 200
    assert hy.align_pairs_flat(q, qo, t, to, s, hy.Mode.Sw, st, scratch=sc)[0].tolist() == \
        whole[0].tolist()


@pytest.mark.parametrize("al", [2, 4, 7, 8, 20, 24, 25])
def test_two_pairs_per_lane_equal_one_per_lane(al, monkeypatch):
    """Local-mode lane pairs of alphabets up to 24 symbols run two per lane in 16-bit halves
    (k_pairs16: one pair-table lookup per cell pair up to 7 symbols, two lookups in a replicated
    score table from 8 to 24; 25 stays one per lane): scores and ends equal the one-pair-per-lane
    kernel and the oracle, with ties (identity and all-equal matrices), queries across 16- and
    32-row tiles (both tile sizes), an odd last group (only low halves live; with 257 pairs, the
    last four-group workgroup of the replicated table has one live group), empty sides, and
    pairs whose width proof needs 32 bits mixed in (they stay one per lane in the same call)."""
    rng = np.random.default_rng(1200 + al)
    schemes = [(al, identity_matrix(al, 2, -1), 2, 1),
               (al, [int(v) for v in rng.integers(-3, 5, size=al * al)], 3, 1),
               (al, [1] * (al * al), 1, 0)]
    monkeypatch.setenv("HYALITE_HIP_PAIRS_KERNEL", "lanes")
    for scheme in schemes:
        s = scoring(scheme)
        pairs = []
        for k in range(257):
            m = int(rng.choice([0, 1, 15, 16, 17, 31, 32, 33, 64, 100]))
            n = int(rng.choice([0, 1, 3, 4, 5, 63, 64, 65, 300]))
            q = rng.integers(0, al, size=m, dtype=np.uint8)
            t = rng.integers(0, al, size=n, dtype=np.uint8)
            if m and n >= m and k % 3 == 0:
                at = int(rng.integers(0, n - m + 1))
                t[at:at + m] = q
            pairs.append((q, t))
        for st in ST:
            monkeypatch.setenv("HYALITE_HIP_PAIRS16", "0")
            one = hy.align_pairs(pairs, s, hy.Mode.Sw, st)
            monkeypatch.setenv("HYALITE_HIP_PAIRS16", "2")  # two per lane at any batch size
            for tile in ("32", "16"):
                monkeypatch.setenv("HYALITE_HIP_PAIRS16_TILE", tile)
                two = hy.align_pairs(pairs, s, hy.Mode.Sw, st)
                assert two == one, (al, scheme[1][:4], st, tile)
            monkeypatch.delenv("HYALITE_HIP_PAIRS16_TILE", raising=False)
            for k, (q, t) in enumerate(pairs):
                assert one[k] == want_hit(list(q), list(t), scheme, MODE_SW, st, k), (al, k)
    # 32-bit pairs beside 16-bit ones: a large match score pushes long pairs past 32,767.
    big = (al, identity_matrix(al, 300, -1), 5, 1)
    s = scoring(big)
    pairs = [(rng.integers(0, al, size=m, dtype=np.uint8), rng.integers(0, al, size=n,
                                                                        dtype=np.uint8))
             for m, n in [(20, 40), (150, 200), (10, 10), (120, 130), (3, 500)] * 30]
    monkeypatch.setenv("HYALITE_HIP_PAIRS16", "2")
    got = hy.align_pairs(pairs, s, hy.Mode.Sw, hy.SearchType.ScoreEnd)
    for k, (q, t) in enumerate(pairs):
        assert got[k] == want_hit(list(q), list(t), big, MODE_SW, hy.SearchType.ScoreEnd, k), k


def test_two_per_lane_protein_batch(monkeypatch):
    """A realistic protein batch (20,000 pairs of 250-350 x 300-400, 24 symbols; 157 groups of
    128, the last workgroup of four partly live) two per lane with two lookups, at 16- and 32-row
    tiles: every score and end equals one pair per lane, and a sample equals the oracle."""
    from hyalite_amd import workloads as wl
    b = wl.protein_pairs(n=20_000, seed=31)
    s = hy.Scoring(b.al, b.matrix, b.go, b.ge)
    scheme = (b.al, [int(v) for v in np.asarray(b.matrix).ravel()], b.go, b.ge)
    sc = hy.PairScratch()
    monkeypatch.setenv("HYALITE_HIP_PAIRS_KERNEL", "lanes")  # below the lane kernels' batch size
    for st in ST:
        monkeypatch.setenv("HYALITE_HIP_PAIRS16", "0")
        one = hy.align_pairs_flat(b.queries, b.qoff, b.targets, b.toff, s, hy.Mode.Sw, st,
                                  scratch=sc)
        monkeypatch.setenv("HYALITE_HIP_PAIRS16", "2")
        for tile in ("16", "32"):
            monkeypatch.setenv("HYALITE_HIP_PAIRS16_TILE", tile)
            two = hy.align_pairs_flat(b.queries, b.qoff, b.targets, b.toff, s, hy.Mode.Sw, st,
                                      scratch=sc)
            for x, y in zip(two, one):
                assert np.array_equal(x, y), (st, tile)
        monkeypatch.delenv("HYALITE_HIP_PAIRS16_TILE", raising=False)
        for k in range(0, b.n, 997):
            q, t = b.pair(k)
            w = want_hit(list(q), list(t), scheme, MODE_SW, st, k)
            got = (int(one[0][k]), None if st == hy.SearchType.Score or one[1][k] < 0
                   else int(one[1][k]), None if st == hy.SearchType.Score or one[2][k] < 0
                   else int(one[2][k]))
            assert got == (w.score, w.query_end, w.target_end), k
