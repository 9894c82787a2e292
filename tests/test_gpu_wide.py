"""Device differential at the 32-bit width, up to the width proof's ceiling.

The 32-bit kernels (k_inter32, k_sys32, k_band32 in winner mode, k_pairs32, k_trace_dp) add and
subtract with wrapping 32-bit arithmetic around the sentinel i32::MIN/4, exact for every cell the
width proof admits (DETERMINISM.md §1, src/inter.rs:80-98). These tests run them where that
matters, against the scalar oracle, bit for bit, through every entry point of the C ABI:

* the four magnitude regimes of the reference's differential fuzz target
  (fuzz/fuzz_targets/differential.rs:40-46: alphabets 2-6, lengths 0-24, 1-5 targets, entries
  within +-8, +-1e3, +-1e5, +-6e8, gap_open up to the regime's bound, gap_ext up to gap_open);
  schemes the proof rejects must be rejected with the same bound;
* the i32 proptest (tests/properties.rs:400-444: alphabets 2-4, entries +-5000, gaps 0-10, up to
  6 targets of length <= 12, max_query_len 12, every mode whose proof gives i32);
* the deterministic i32 guard (tests/properties.rs:451-500);
* per-sequence width escalation (tests/properties.rs:708-758: entries -40..70, 2-8 targets of
  length <= 40, queries <= 20, max_query_len 40);
* schemes scaled so that the proof's bound sits just below 2^29 - 1, with gap_ext equal or close
  to gap_open, over databases large enough for several 32-bit groups, several query tiles,
  long (systolic) targets and multi-pass winner passes;
* schemes scaled to the edge of the packed 16-bit kernel (bound + gap_open around 32767).

Each database is checked with scan, scan_all, scan_scores and scan_batch; the pairs with
align_pairs; a share of them with scan_aligned and scan_all_aligned. Non-vacuity: every test
asserts that the 32-bit partition (or the packed edge) was actually reached.
"""
from __future__ import annotations

import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from common import ALL_MODES, identity_matrix

pytestmark = pytest.mark.gpu

CEIL = 536870911  # the largest bound the proof accepts (i32 width, src/width.rs:48-55)
ST = (hy.SearchType.Score, hy.SearchType.ScoreEnd)
ALIGN_ANY = hy.SearchType.Alignment((1 << 64) - 1)


def _builder(seqs, scheme, mode, st, mql):
    al, mat, go, ge = scheme
    return (hy.Database.builder().sequences(seqs).scoring(hy.Scoring(al, mat, go, ge))
            .mode(hy.Mode(mode)).search_type(st).max_query_len(mql)
            .backend(hy.BackendChoice.force(hy.Backend.Hip)))


class Tally:
    def __init__(self):
        self.built = self.rejected = self.i32 = self.i16 = self.i8 = self.mixed = 0

    def add(self, db):
        p = db.partition_sizes()
        self.built += 1
        self.i32 += p[hy.ScoreWidth.I32] > 0
        self.i16 += p[hy.ScoreWidth.I16] > 0
        self.i8 += p[hy.ScoreWidth.I8] > 0
        self.mixed += sum(v > 0 for v in p.values()) > 1


def check_db(seqs, scheme, mode, st, queries, mql, tally, pairs=True, aligned=False):
    """One database: the proof's verdict, then every scan entry point against the oracle."""
    al, mat, go, ge = scheme
    mtl = max(len(s) for s in seqs)
    _, bound, bits = orc.required_width(mode, min(mat), max(mat), go, ge, mql, mtl)
    if bits == 0:
        with pytest.raises(hy.ScoreRangeTooWide) as ei:
            _builder(seqs, scheme, mode, st, mql).build()
        assert ei.value.bound == bound, (mode, bound)
        tally.rejected += 1
        return None
    db = _builder(seqs, scheme, mode, st, mql).build()
    tally.add(db)
    sc = hy.Scratch(db)
    flat, off = orc.flatten(seqs)
    tracks = st == hy.SearchType.ScoreEnd
    wants = []
    ctx = (mode, int(st), scheme[2], scheme[3], bound)
    for q in queries:
        want = orc.scan_flat(flat, off, q, al, mat, go, ge, mode, tracks, threads=4)
        wants.append(want)
        got = db.scan(sc, q)
        assert (got.score, got.db_index, got.query_end, got.target_end) == want, (ctx, list(q))
        ws, wqe, wte = orc.scan_all_flat(flat, off, q, al, mat, go, ge, mode, threads=4)
        s, qe, te = db.scan_all_arrays(sc, q)
        assert np.array_equal(s, ws), (ctx, list(q))
        if tracks:
            assert np.array_equal(qe, wqe) and np.array_equal(te, wte), (ctx, list(q))
        else:
            assert (qe == -1).all() and (te == -1).all()
        assert np.array_equal(db.scan_scores_array(sc, q), ws), (ctx, list(q))
        if pairs and len(seqs) <= 64:
            hits = hy.align_pairs([(q, t) for t in seqs], hy.Scoring(al, mat, go, ge),
                                  hy.Mode(mode), st)
            for k, h in enumerate(hits):
                w = (int(ws[k]), int(wqe[k]) if tracks and wqe[k] >= 0 else None,
                     int(wte[k]) if tracks and wte[k] >= 0 else None)
                assert (h.score, h.query_end, h.target_end) == w, (ctx, list(q), k)
    if len(queries) > 1:
        got = db.scan_batch(sc, queries)
        for g, w, q in zip(got, wants, queries):
            assert (g.score, g.db_index, g.query_end, g.target_end) == w, (ctx, "batch", list(q))
    if aligned:
        adb = _builder(seqs, scheme, mode, ALIGN_ANY, mql).build()
        asc = hy.Scratch(adb)
        for q in queries:
            tbs = [orc.traceback(q, t, al, mat, go, ge, mode) for t in seqs]
            got = adb.scan_all_aligned(asc, q)
            for k, (g, w) in enumerate(zip(got, tbs)):
                assert g.as_tuple() == w, (ctx, "aligned", list(q), k)
            hit = adb.scan_aligned(asc, q)
            best = max(w[0] for w in tbs)
            k = next(i for i, w in enumerate(tbs) if w[0] == best)
            assert hit.db_index == k and hit.alignment.as_tuple() == tbs[k], (ctx, list(q))
    return db


def _seq(rng, al, lo, hi):
    return rng.integers(0, al, size=int(rng.integers(lo, hi + 1)), dtype=np.uint8)


def _extra_queries(rng, al, q, mql):
    """The fuzz case's query plus two more within max_query_len, for the batched form."""
    return [q] + [_seq(rng, al, 0, mql) for _ in range(2)]


REGIMES = {0: (-8, 8), 1: (-1_000, 1_000), 2: (-100_000, 100_000),
           3: (-600_000_000, 600_000_000)}


@pytest.mark.parametrize("regime", sorted(REGIMES))
def test_fuzz_magnitude_regime(regime):
    """fuzz/fuzz_targets/differential.rs:36-82 decoded from a seeded generator instead of fuzzer
    bytes: each case draws its scheme, mode, search type, 1-5 targets and a query, then runs
    every scan entry point and align_pairs against the oracle. Regime 3 straddles the i32
    ceiling: most of its schemes fail the proof and must fail the device build identically."""
    lo, hi = REGIMES[regime]
    rng = np.random.default_rng(7100 + regime)
    tally = Tally()
    ncases = 300
    for case in range(ncases):
        # Regime 3 at the fuzz target's lengths (0-24) almost never passes the proof (2 of 300
        # here); every other case of it draws lengths 0-2, as the fuzzer's coverage guidance
        # would, so that accepted schemes with entries and gaps near 2^29 are common.
        short = regime == 3 and case % 2 == 1
        al = int(rng.integers(2, 4 if short else 7))
        mat = [int(v) for v in rng.integers(lo, hi + 1, size=al * al)]
        go = int(rng.integers(0, max(hi, 0) + 1))
        ge = int(rng.integers(0, go + 1))
        mode = int(rng.integers(0, 5))
        st = ST[int(rng.integers(0, 2))]
        lmax = 2 if short else 24
        seqs = [_seq(rng, al, 0, lmax) for _ in range(int(rng.integers(1, 6)))]
        q = _seq(rng, al, 0, lmax)
        mql = max(len(q), 1)
        check_db(seqs, (al, mat, go, ge), mode, st, _extra_queries(rng, al, q, mql), mql, tally,
                 aligned=case % 10 == 0)
    assert tally.built + tally.rejected == ncases
    if regime >= 1:
        assert tally.i32 >= 20, vars(tally)  # the 32-bit partition is reached
    if regime == 3:
        assert tally.rejected >= 50 and tally.built >= 20, vars(tally)


def test_verywide_proptest_i32():
    """tests/properties.rs:400-444 (scheme_db_query_verywide): every mode whose proof gives i32
    for (max_query_len 12, longest target), both search types, against the oracle."""
    rng = np.random.default_rng(7200)
    tally = Tally()
    modes_hit = set()
    for _ in range(150):
        al = int(rng.integers(2, 5))
        mat = [int(v) for v in rng.integers(-5000, 5001, size=al * al)]
        go = int(rng.integers(0, 11))
        ge = int(rng.integers(0, go + 1))
        seqs = [_seq(rng, al, 0, 12) for _ in range(int(rng.integers(1, 7)))]
        q = _seq(rng, al, 0, 12)
        mtl = max(len(s) for s in seqs)
        for mode in ALL_MODES:
            if orc.required_width(mode, min(mat), max(mat), go, ge, 12, mtl)[2] != 32:
                continue
            modes_hit.add(mode)
            for st in ST:
                check_db(seqs, (al, mat, go, ge), mode, st, [q, _seq(rng, al, 0, 12)], 12, tally,
                         pairs=st == hy.SearchType.ScoreEnd)
    assert modes_hit == set(ALL_MODES)
    assert tally.i32 == tally.built and tally.built >= 500, vars(tally)


def test_i32_guard_every_mode():
    """tests/properties.rs:451-500: identity 5000/-3000, gaps 600/100, six rotations of a
    12-symbol cycle, in every mode (each proves i32) and both search types; plus the aligned
    scans."""
    scheme = (4, identity_matrix(4, 5000, -3000), 600, 100)
    seqs = [[(i + k) % 4 for i in range(12)] for k in range(6)]
    q = [(i * 3 + 1) % 4 for i in range(12)]
    tally = Tally()
    for mode in ALL_MODES:
        assert orc.required_width(mode, -3000, 5000, 600, 100, 12, 12)[2] == 32
        for st in ST:
            check_db(seqs, scheme, mode, st, [q, q[:7], []], 12, tally, aligned=True)
    assert tally.i32 == tally.built == 10


def test_escalation_proptest():
    """tests/properties.rs:708-758 (scheme_mixed_length_db): sequences of length 0-40 span
    several proven widths, so the database splits into partitions; every mode and search type
    against the oracle."""
    rng = np.random.default_rng(7300)
    tally = Tally()
    for _ in range(80):
        al = int(rng.integers(2, 5))
        mat = [int(v) for v in rng.integers(-40, 71, size=al * al)]
        go = int(rng.integers(0, 11))
        ge = int(rng.integers(0, go + 1))
        seqs = [_seq(rng, al, 0, 40) for _ in range(int(rng.integers(2, 9)))]
        qs = [_seq(rng, al, 0, 20) for _ in range(2)]
        for mode in ALL_MODES:
            for st in ST:
                check_db(seqs, (al, mat, go, ge), mode, st, qs, 40, tally,
                         pairs=st == hy.SearchType.ScoreEnd)
    # Entries up to 70 at lengths up to 40 stay within i16, as in the reference's generator; the
    # databases that hold both an i8 and an i16 partition are the escalated ones.
    assert tally.built >= 700 and tally.mixed >= 50, vars(tally)


def _scaled(rng, mode, al, mql, mtl, target, ge_close, emax=1000):
    """A scheme whose proof bound for (mode, mql, mtl) lands in (target - base, target]: a base
    scheme with entries in +-emax scaled by an integer factor (the bound is linear in the
    entries and gaps). ge_close draws gap_ext equal or next to gap_open."""
    for _ in range(1000):
        mat = [int(v) for v in rng.integers(-emax, emax + 1, size=al * al)]
        go = int(rng.integers(1, emax + 1))
        ge = go - int(rng.integers(0, 2)) if ge_close else int(rng.integers(0, go + 1))
        _, base, _ = orc.required_width(mode, min(mat), max(mat), go, ge, mql, mtl)
        if base <= 0:
            continue
        k = target // base
        if k < 1:
            continue
        # Entries the bound does not depend on (local mode's negative ones, any entry when a
        # length is 0) can outgrow i32 when scaled: they are held at the i32 limits, where the
        # device's own entry clamp then applies.
        lim = (1 << 31) - 1
        mat = [max(-lim, min(lim, v * k)) for v in mat]
        go, ge = go * k, ge * k
        _, b, bits = orc.required_width(mode, min(mat), max(mat), go, ge, mql, mtl)
        assert b <= target and bits != 0 and b == base * k
        return (al, mat, go, ge), b
    raise AssertionError("no base scheme fits under the target bound")


@pytest.mark.parametrize("mode", ALL_MODES)
def test_bound_at_the_i32_ceiling(mode):
    """Schemes whose proof bound lies just below 2^29 - 1, gap_ext equal or next to gap_open in
    half the cases, on small databases (every entry point, aligned scans on a share) and on
    databases of several 32-bit groups against queries of several 32-row tiles."""
    rng = np.random.default_rng(7400 + mode)
    tally = Tally()
    for case in range(40):
        al = int(rng.integers(2, 7))
        seqs = [_seq(rng, al, 0, 30) for _ in range(int(rng.integers(1, 8)))]
        mql = int(rng.integers(1, 40))
        mtl = max(len(s) for s in seqs)
        scheme, b = _scaled(rng, mode, al, mql, mtl, CEIL, case % 2 == 0)
        assert b > CEIL - 1_200_000
        qs = [_seq(rng, al, 0, mql) for _ in range(3)]
        for st in ST:
            check_db(seqs, scheme, mode, st, qs, mql, tally, aligned=(case % 8 == 0 and
                                                                      st == ST[1]))
    for case in range(4):
        al = 4 if case % 2 else 20
        seqs = [_seq(rng, al, 30, 300) for _ in range(int(rng.integers(150, 400)))]
        mql = 150
        mtl = max(len(s) for s in seqs)
        scheme, _ = _scaled(rng, mode, al, mql, mtl, CEIL, case >= 2)
        qs = [_seq(rng, al, 1, mql), _seq(rng, al, 100, mql), _seq(rng, al, 1, 40)]
        for st in ST:
            check_db(seqs, scheme, mode, st, qs, mql, tally, pairs=False)
    assert tally.i32 == tally.built, vars(tally)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_long_targets_at_the_i32_ceiling(mode):
    """Long targets (the systolic kernel beside the inter-sequence kernels) and queries past one
    systolic pass (multi-pass winner ends, boundary rows in HBM) with the bound just below the
    ceiling."""
    rng = np.random.default_rng(7500 + mode)
    tally = Tally()
    al = 5
    seqs = [_seq(rng, al, 20, 200) for _ in range(300)]
    seqs += [_seq(rng, al, 1100, 1800) for _ in range(6)]
    mql = 700
    mtl = max(len(s) for s in seqs)
    q_long = _seq(rng, al, 600, 700)
    # A copy of a long target's middle makes a long winner in every mode.
    q_copy = seqs[-2][200:200 + 640].copy()
    for k, ge_close in enumerate((False, True)):
        scheme, _ = _scaled(rng, mode, al, mql, mtl, CEIL, ge_close)
        for st in ST:
            check_db(seqs, scheme, mode, st, [q_long, q_copy, _seq(rng, al, 1, 60)], mql, tally,
                     pairs=False)
    assert tally.i32 == tally.built == 4, vars(tally)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_packed_edge_scaled(mode):
    """Schemes scaled so that the proof bound plus gap_open falls on either side of 32767, the
    edge between the packed 16-bit kernel and the 32-bit kernels, with lengths spread so that
    one database holds targets on both sides."""
    rng = np.random.default_rng(7600 + mode)
    tally = Tally()
    for case in range(24):
        al = int(rng.integers(2, 7))
        seqs = [_seq(rng, al, 0, 200) for _ in range(int(rng.integers(40, 300)))]
        mql = int(rng.integers(20, 80))
        mtl = max(len(s) for s in seqs)
        target = 32767 - (0 if case % 3 else int(rng.integers(0, 400)))
        # Scaled at half the reach of the longest target (the SW and OV bounds stop growing at
        # the query length), so the targets on either side of the middle length fall on either
        # side of the edge.
        half = (mql if mode in (0, 3) else mtl) // 2 + 1
        scheme, b = _scaled(rng, mode, al, mql, half, target, case % 2 == 0, emax=3)
        qs = [_seq(rng, al, 0, mql) for _ in range(3)]
        for st in ST:
            check_db(seqs, scheme, mode, st, qs, mql, tally, pairs=False)
    assert tally.i16 > 0 and tally.i32 > 0, vars(tally)
