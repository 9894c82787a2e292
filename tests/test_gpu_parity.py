"""GPU parity: every scan through the C ABI must equal the CPU oracle bit for bit.

Mirrors the reference's differential tests (tests/database.rs, tests/sweeps.rs,
tests/properties.rs, tests/opal_regressions.rs, tests/real_data.rs): the Hip backend is forced and
its (score, db_index, query_end, target_end) compared with the oracle's scalar scan, per target
(scan_all, scan_scores) and single best (scan), on the same seeded inputs.
"""
import dataclasses
import threading

import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from common import (ALL_MODES, BLOSUM62, MODE_NW, MODE_OV, MODE_SW, all_sequences,
                    cr4_adapters, cr4_scheme, dna_scheme, identity_matrix, lcg_next, lcg_seq,
                    phix)

pytestmark = pytest.mark.gpu

ST = (hy.SearchType.Score, hy.SearchType.ScoreEnd)


def build(seqs, scheme, mode, st, mql):
    al, mat, go, ge = scheme
    return (hy.Database.builder().sequences(seqs).scoring(hy.Scoring(al, mat, go, ge))
            .mode(hy.Mode(mode)).search_type(st).max_query_len(mql)
            .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())


def check(seqs, scheme, mode, st, queries, mql=None, all_targets=True):
    al, mat, go, ge = scheme
    mql = max([len(q) for q in queries] + [1]) if mql is None else mql
    db = build(seqs, scheme, mode, st, mql)
    sc = hy.Scratch(db)
    flat, off = orc.flatten(seqs)
    tracks = st == hy.SearchType.ScoreEnd
    for q in queries:
        want = orc.scan_flat(flat, off, q, al, mat, go, ge, mode, tracks, threads=4)
        got = db.scan(sc, q)
        assert (got.score, got.db_index, got.query_end, got.target_end) == want, \
            (mode, st, list(q)[:20], len(seqs))
        if not all_targets:
            continue
        ws, wqe, wte = orc.scan_all_flat(flat, off, q, al, mat, go, ge, mode, threads=4)
        s, qe, te = db.scan_all_arrays(sc, q)
        assert np.array_equal(s, ws), (mode, st)
        if tracks:
            assert np.array_equal(qe, wqe) and np.array_equal(te, wte), (mode, st)
        else:
            assert (qe == -1).all() and (te == -1).all()
        assert np.array_equal(db.scan_scores_array(sc, q), ws)


def test_backend_is_available():
    assert hy.Backend.Hip.is_available()


@pytest.mark.parametrize("mode", ALL_MODES)
def test_all_short_sequences_against_all_short_queries(mode):
    """tests/database.rs:67-100 at database scale: all 21 sequences of length <=2 (plus a
    reversed copy so equal scores tie across indices) against all 85 queries of length <=3."""
    seqs = all_sequences(4, 2)
    seqs = seqs + seqs[::-1]
    for st in ST:
        check(seqs, dna_scheme(), mode, st, all_sequences(4, 3), mql=3)


def test_two_sequence_databases_tie_break_both_orders():
    rng = np.random.default_rng(5)
    seqs = all_sequences(4, 2)
    queries = all_sequences(4, 3)
    for _ in range(12):
        a, b = rng.integers(0, len(seqs), 2)
        for mode in ALL_MODES:
            check([seqs[a], seqs[b]], dna_scheme(), mode, hy.SearchType.ScoreEnd,
                  queries[::7], mql=3)


def test_database_sizes_around_lane_boundaries():
    """tests/sweeps.rs:130-158 with the wavefront boundaries added."""
    state = [0x5EED1234]
    for n in (1, 2, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 191, 192, 193, 257):
        seqs = [lcg_seq(state, 3 + lcg_next(state) % 8, 4) for _ in range(n)]
        queries = [lcg_seq(state, 4 + lcg_next(state) % 6, 4) for _ in range(3)]
        for mode in ALL_MODES:
            for st in ST:
                check(seqs, dna_scheme(), mode, st, queries)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_winner_in_every_wavefront_of_a_workgroup(mode):
    """Regression: the single-best key must collect every wavefront of a multi-wavefront
    workgroup. The winner (an exact query copy) is planted among the shortest targets, so after
    the length sort it lands in the last group, for each position of that group in its block."""
    rng = np.random.default_rng(400 + mode)
    q = rng.integers(0, 4, size=12, dtype=np.uint8)
    for ngroups in range(1, 10):
        n = 128 * ngroups
        seqs = [rng.integers(0, 4, size=int(k), dtype=np.uint8) for k in rng.integers(30, 60, n)]
        seqs[n // 3] = q.copy()
        check(seqs, dna_scheme(), mode, hy.SearchType.Score, [q], all_targets=False)


def test_degenerate_sequences_and_queries():
    """tests/sweeps.rs:160-194."""
    cases = [
        [[]] * 20,
        [[] if i % 2 == 0 else [0, 1, 2] for i in range(33)],
        [[0] * 10] * 40,
        [[i % 4] for i in range(17)],
        [[0, 1, 2, 3]] * 32 + [[0] * 20],
        [[]] * 70 + [[1, 2, 3]],
    ]
    queries = [[], [0], [0, 1, 2, 3], [2, 2, 2, 2, 2]]
    for seqs in cases:
        for mode in ALL_MODES:
            for st in ST:
                check(seqs, dna_scheme(), mode, st, queries, mql=5)


def test_scoring_parameter_grid():
    """tests/sweeps.rs:196-236, including free gaps and gap_open at the i8 edge."""
    state = [0xA11CE]
    seqs = [lcg_seq(state, lcg_next(state) % 9, 4) for _ in range(20)] + [[]]
    queries = [lcg_seq(state, lcg_next(state) % 9, 4) for _ in range(5)]
    for m, x in ((1, -1), (2, -1), (3, -2), (1, 0)):
        for go, ge in ((1, 1), (2, 1), (3, 0), (5, 5), (4, 2), (0, 0), (100, 1), (127, 1)):
            scheme = (4, identity_matrix(4, m, x), go, ge)
            for mode in ALL_MODES:
                check(seqs, scheme, mode, hy.SearchType.ScoreEnd, queries, mql=9)


@pytest.mark.parametrize("al", [8, 16, 17, 24, 64, 255, 256])
def test_alphabet_sizes(al):
    """tests/sweeps.rs:238-262 plus protein-sized, large and byte-exhausting alphabets (256 has
    no free padding symbol and runs entirely in the systolic kernel)."""
    mat = [-1] * (al * al)
    for i in range(al):
        mat[i * al + i] = 1 + (i % 3)
    state = [0xF00D + al]
    seqs = [lcg_seq(state, 4 + lcg_next(state) % 40, al) for _ in range(90)]
    queries = [lcg_seq(state, 6 + lcg_next(state) % 30, al) for _ in range(3)]
    for mode in ALL_MODES:
        check(seqs, (al, mat, 2, 1), mode, hy.SearchType.ScoreEnd, queries)


@pytest.mark.parametrize("mode", [MODE_SW, MODE_NW, MODE_OV])
def test_small_alphabet_profiles_every_tile_height(mode, monkeypatch):
    """One-wavefront tiles hold a residue-pair profile when it fits 16 KB of LDS
    ((al + 1)^2 x (rows + 4) words) and the per-symbol profile otherwise. Alphabets 4, 9, 10 and
    12 put 32-, 24- and 16-row tiles on either side of that limit (10: pair profile at 16 and 24
    rows only; 12: at 16 rows only); every tile height of the one-query kernel (one wavefront
    per group, and pipelined), the batch kernel and the end-tracking scan equals the oracle."""
    rng = np.random.default_rng(1500 + mode)
    for al in (4, 9, 10, 12):
        mat = [int(rng.integers(-3, 4)) for _ in range(al * al)]
        for i in range(al):
            mat[i * al + i] = 2 + i % 3
        seqs = [rng.integers(0, al, size=int(n), dtype=np.uint8)
                for n in rng.integers(0, 200, 400)]
        qs = [rng.integers(0, al, size=int(m), dtype=np.uint8) for m in (7, 33, 70)]
        flat, off = orc.flatten(seqs)
        db = build(seqs, (al, mat, 3, 1), mode, hy.SearchType.ScoreEnd, 70)
        sc = hy.Scratch(db)
        want = [orc.scan_flat(flat, off, q, al, mat, 3, 1, mode, True, threads=4) for q in qs]
        ends = [orc.scan_all_flat(flat, off, q, al, mat, 3, 1, mode, threads=4) for q in qs]
        monkeypatch.setenv("HYALITE_HIP_CHAIN", "0")
        for rows in (16, 24, 32):
            sc.set_tuning(hy.Tuning(tile_rows=rows, batch_tile_rows=rows, batch_pairs=0))
            monkeypatch.setenv("HYALITE_HIP_PIPE_ROWS", str(rows))
            for pipe in ("0", "1"):
                monkeypatch.setenv("HYALITE_HIP_PIPE", pipe)
                for q, w, (ws, wqe, wte) in zip(qs, want, ends):
                    h = db.scan(sc, q)
                    assert (h.score, h.db_index, h.query_end, h.target_end) == w, (al, rows, pipe)
                    assert np.array_equal(db.scan_scores_array(sc, q), ws), (al, rows, pipe)
                    s, qe, te = db.scan_all_arrays(sc, q)
                    assert (np.array_equal(s, ws) and np.array_equal(qe, wqe) and
                            np.array_equal(te, wte)), (al, rows, pipe)
            got = db.scan_batch(sc, qs)
            assert [(h.score, h.db_index, h.query_end, h.target_end) for h in got] == want, \
                (al, rows)


def test_score_width_boundaries():
    """tests/sweeps.rs:264-285."""
    scheme = (2, [1, -1, -1, 1], 2, 1)
    for n in (126, 127, 128, 129):
        check([[0] * n], scheme, MODE_SW, hy.SearchType.ScoreEnd, [[0] * n])


def test_overlap_negative_reach_at_the_i8_edge():
    """tests/sweeps.rs:287-316."""
    scheme = (2, [1, -1, -1, 1], 1, 1)
    check([[1] * 126], scheme, MODE_OV, hy.SearchType.ScoreEnd, [[0] * 126])
    check([[1] * 6], scheme, MODE_OV, hy.SearchType.ScoreEnd, [[0] * 120 + [1] * 6])


@pytest.mark.parametrize("mode", ALL_MODES)
def test_packed_kernel_eligibility_edge(mode):
    """Targets whose width-proof bound sits just inside (packed 16-bit kernel) and just outside
    (32-bit kernel) bound + gap_open <= 32767, with mismatch-heavy content so that real cells run
    deep negative and non-real intermediates saturate."""
    scheme = (2, [1, -8, -8, 1], 6, 2)
    # NW/HW/SHW bound = 10 (m + n) + 4; OV/SW bounds are far smaller at these lengths.
    m = 1000
    rng = np.random.default_rng(300 + mode)
    lens = [2270, 2275, 2276, 2277, 2290, 5, 0, 1500]
    seqs = [np.where(rng.random(n) < 0.9, 1, 0).astype(np.uint8) for n in lens]
    q = np.where(rng.random(m) < 0.9, 0, 1).astype(np.uint8)
    check(seqs, scheme, mode, hy.SearchType.ScoreEnd, [q, q[:400]], mql=m)


def test_opal_regressions_and_proptest_seeds():
    """tests/opal_regressions.rs:68-82 and the proptest regression seeds."""
    check([[0, 0, 1, 1, 2]], (3, identity_matrix(3, 2, -2), 1, 1), MODE_NW,
          hy.SearchType.ScoreEnd, [[0, 0, 1, 1]])
    for mode in ALL_MODES:
        check([[]], (2, [0, 0, 0, 0], 0, 0), mode, hy.SearchType.ScoreEnd, [[0]])
        check([[0, 0, 1, 0, 1, 1, 0, 1, 1, 0, 0, 1]], (2, [0, 0, 0, -5], 8, 0), mode,
              hy.SearchType.ScoreEnd, [[1, 1, 0, 1, 1]])
        check([[0] * 12], (2, [0, 0, 11, 0], 1, 0), mode, hy.SearchType.ScoreEnd, [[1]])


def test_protein_scale_scoring():
    """tests/opal_regressions.rs:140-216 (24-letter matrix, 13/2 gaps)."""
    AL = 24
    mat = [4 + (i % 3) if i == j else -1 - ((i + j) % 4) for i in range(AL) for j in range(AL)]
    state = [0xB105]
    seqs = [lcg_seq(state, 30 + lcg_next(state) % 40, AL) for _ in range(12)]
    queries = [lcg_seq(state, 40 + lcg_next(state) % 30, AL) for _ in range(4)]
    for mode in ALL_MODES:
        check(seqs, (AL, mat, 13, 2), mode, hy.SearchType.ScoreEnd, queries)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_random_differential_multi_tile(mode):
    """Random databases with lengths 0..700 and queries long enough to need several query tiles
    (inter kernel) and several passes (systolic kernel, > 512 rows)."""
    rng = np.random.default_rng(100 + mode)
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(0, 700, 400)]
    queries = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in (0, 1, 31, 33, 600)]
    check(seqs, (20, BLOSUM62, 11, 1), mode, hy.SearchType.ScoreEnd, queries, mql=700)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_long_targets_take_the_systolic_path(mode):
    """A few targets far above the long-target threshold next to many short ones, plus
    mutated copies of the query so high scores tie across indices."""
    rng = np.random.default_rng(200 + mode)
    q = rng.integers(0, 4, size=150, dtype=np.uint8)
    seqs = [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in rng.integers(20, 80, 600)]
    for k in (3, 77, 301):
        seqs[k] = np.concatenate([rng.integers(0, 4, 2500, dtype=np.uint8), q,
                                  rng.integers(0, 4, 900, dtype=np.uint8)])
    seqs[450] = q.copy()
    seqs[451] = q.copy()
    check(seqs, dna_scheme(), mode, hy.SearchType.ScoreEnd, [q, q[:40]], mql=150)


def test_real_phix_windows_and_cr4():
    """tests/real_data.rs:63-103 and 145-201."""
    genome = np.array(phix(), dtype=np.uint8)
    windows = [genome[o:o + 60] for o in (0, 500, 1500, 2500, 4000)]
    queries = [genome[500:560], genome[510:570], genome[3000:3060], genome[1500:1530]]
    for mode in ALL_MODES:
        for st in ST:
            check(windows, dna_scheme(), mode, st, queries, mql=64)
    seqs = [np.array(s, dtype=np.uint8) for _, s in cr4_adapters()]
    db = build(seqs, cr4_scheme(), MODE_OV, hy.SearchType.ScoreEnd, 128)
    sc = hy.Scratch(db)
    tso_read = np.concatenate([seqs[0], genome[800:840]])
    hit = db.scan(sc, tso_read)
    assert (hit.db_index, hit.score) == (0, 30)
    polya = np.concatenate([genome[1500:1530], np.zeros(25, dtype=np.uint8)])
    hit = db.scan(sc, polya)
    assert hit.db_index == 1 and hit.score >= 25


def test_concurrent_scans_from_threads():
    """tests/database.rs:136-176: one database, a Scratch per thread, identical results."""
    rng = np.random.default_rng(9)
    seqs = [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in rng.integers(0, 200, 2000)]
    db = build(seqs, dna_scheme(), MODE_SW, hy.SearchType.ScoreEnd, 100)
    q = rng.integers(0, 4, size=100, dtype=np.uint8)
    expected = db.scan(hy.Scratch(db), q)
    results, errors = [], []

    def work():
        try:
            sc = hy.Scratch(db)
            for _ in range(20):
                results.append(db.scan(sc, q))
        except Exception as exc:  # pragma: no cover - surfaced below
            errors.append(exc)

    threads = [threading.Thread(target=work) for _ in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors
    assert len(results) == 80 and all(r == expected for r in results)


def test_scratch_reuse_across_query_sizes():
    """src/database.rs scratch_reuse_across_many_scans_is_consistent."""
    db = build([[0, 1], [0, 1, 2, 3, 3, 2, 1, 0], [2]], dna_scheme(), MODE_NW,
               hy.SearchType.ScoreEnd, 8)
    sc = hy.Scratch(db)
    qs = [[0, 1, 2, 3], [2], [0, 1, 2, 3, 3, 2, 1, 0], [0, 1, 2, 3]]
    hits = [db.scan(sc, q) for q in qs]
    assert hits[0] == hits[3]
    for q, h in zip(qs, hits):
        assert db.scan(hy.Scratch(db), q) == h


def test_scan_contract_violations_are_errors():
    db = build([[0, 1, 2]], dna_scheme(), MODE_SW, hy.SearchType.ScoreEnd, 4)
    sc = hy.Scratch(db)
    with pytest.raises(hy.ScanContractViolation):
        db.scan(sc, [0, 1, 2, 3, 0])
    with pytest.raises(hy.ScanContractViolation):
        db.scan(sc, [0, 9])


def test_batch_equals_individual_scans():
    rng = np.random.default_rng(21)
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(10, 300, 3000)]
    db = build(seqs, (20, BLOSUM62, 11, 1), MODE_SW, hy.SearchType.ScoreEnd, 400)
    sc = hy.Scratch(db)
    qs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in (5, 50, 399, 0, 120)]
    batch = db.scan_batch(sc, qs)
    assert batch == [db.scan(sc, q) for q in qs]


# ---- batched form -------------------------------------------------------------------------------

@pytest.mark.parametrize("mode", ALL_MODES)
def test_batch_matches_oracle_all_modes(mode):
    """The persistent batch kernel + batched winner pass: every hit equals the oracle's scan, for
    both search types, with long targets (packed and, at 600-row queries, multi-pass)."""
    rng = np.random.default_rng(500 + mode)
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(5, 200, 900)]
    for k in (7, 300, 899):
        seqs[k] = rng.integers(0, 20, size=int(rng.integers(2500, 3200)), dtype=np.uint8)
    qs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in (0, 1, 17, 200, 600, 33, 90)]
    flat, off = orc.flatten(seqs)
    for st in ST:
        db = build(seqs, (20, BLOSUM62, 11, 1), mode, st, 600)
        sc = hy.Scratch(db)
        got = db.scan_batch(sc, qs)
        for q, h in zip(qs, got):
            want = orc.scan_flat(flat, off, q, 20, BLOSUM62, 11, 1, mode,
                                 st == hy.SearchType.ScoreEnd, threads=4)
            assert (h.score, h.db_index, h.query_end, h.target_end) == want, (mode, st, len(q))


def test_batch_graph_replays_and_chunking():
    """Several batches through one scratch (the first captures the graph, later ones replay it),
    a batch larger than the reservation (several submissions), each hit == the one-query scan."""
    rng = np.random.default_rng(77)
    seqs = [rng.integers(0, 24, size=int(n), dtype=np.uint8) for n in rng.integers(10, 400, 4000)]
    mat = [4 + (i % 8) if i == j else -4 + ((i * 7 + j * 3) % 8) for i in range(24)
           for j in range(24)]
    db = build(seqs, (24, mat, 11, 1), MODE_SW, hy.SearchType.ScoreEnd, 300)
    sc = hy.Scratch(db)
    sc.reserve_batch(8)
    one = hy.Scratch(db)
    for rep in range(3):
        qs = [rng.integers(0, 24, size=int(n), dtype=np.uint8)
              for n in rng.integers(1, 300, 5 + 4 * rep)]
        got = db.scan_batch(sc, qs)
        assert got == [db.scan(one, q) for q in qs], rep


def test_batch_with_unpackable_long_targets():
    """Targets whose proof needs 32 bits (long, and short ones of the 32-bit partition) go through
    the batched systolic list."""
    rng = np.random.default_rng(88)
    scheme = (4, [40 if a == b else -3 for a in range(4) for b in range(4)], 6, 2)
    seqs = [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in rng.integers(20, 120, 700)]
    q = rng.integers(0, 4, size=900, dtype=np.uint8)
    seqs[5] = np.concatenate([q[:800], rng.integers(0, 4, 400, dtype=np.uint8)])  # i32, long
    seqs[6] = rng.integers(0, 4, size=1500, dtype=np.uint8)
    # Short (below the long-target threshold) but beyond the packed bound: the 32-bit partition.
    seqs[7] = np.concatenate([rng.integers(0, 4, 60, dtype=np.uint8), q[100:900]])
    seqs[8] = rng.integers(0, 4, size=850, dtype=np.uint8)
    db = build(seqs, scheme, MODE_SW, hy.SearchType.ScoreEnd, 900)
    assert db.partition_sizes()[hy.ScoreWidth.I32] >= 1
    sc = hy.Scratch(db)
    qs = [q, q[:100], q[300:], rng.integers(0, 4, size=40, dtype=np.uint8)]
    flat, off = orc.flatten(seqs)
    got = db.scan_batch(sc, qs)
    for qq, h in zip(qs, got):
        want = orc.scan_flat(flat, off, qq, *scheme, MODE_SW, True, threads=4)
        assert (h.score, h.db_index, h.query_end, h.target_end) == want


@pytest.mark.parametrize("mode", ALL_MODES)
def test_batch_slots_skip_packed_groups_wider_than_a_slot(mode):
    """A packable target wider than a batched-form slot (8,192 columns) runs in the batched
    form's systolic list, so one long target does not widen every resident workgroup's boundary
    slot: batched hits (one of them inside that target) equal the oracle and the one-query scans,
    and a serving lane stays far below the 2,048 x 9,000-column x 512-byte slots it would need."""
    rng = np.random.default_rng(1900 + mode)
    scheme = dna_scheme()
    seqs = [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in rng.integers(0, 300, 1500)]
    q = rng.integers(0, 4, size=150, dtype=np.uint8)
    wide = rng.integers(0, 4, size=9_000, dtype=np.uint8)
    wide[8_700:8_850] = q  # the best hit of q sits past column 8,192 of the wide target
    seqs[700] = wide
    seqs[701] = rng.integers(0, 4, size=8_300, dtype=np.uint8)
    db = build(seqs, scheme, mode, hy.SearchType.ScoreEnd, 200)
    assert db.partition_sizes()[hy.ScoreWidth.I32] == 0  # every target packable
    sc = hy.Scratch(db)
    qs = [q, q[:60], rng.integers(0, 4, size=200, dtype=np.uint8), q[20:], q[:9]]
    flat, off = orc.flatten(seqs)
    want = [orc.scan_flat(flat, off, qq, *scheme, mode, True, threads=8) for qq in qs]
    assert want[0][1] == 700 or mode != MODE_SW
    got = db.scan_batch(sc, qs)
    assert [(h.score, h.db_index, h.query_end, h.target_end) for h in got] == want
    for qq, w in zip(qs, want):
        h = db.scan(sc, qq)
        assert (h.score, h.db_index, h.query_end, h.target_end) == w
    with hy.Server(db, max_lanes=1, max_batch=8, prealloc=True) as srv:
        assert srv.stats().lane_bytes < 2 * 2**30, srv.stats().lane_bytes
        tickets = [srv.submit(qq) for qq in qs]
        assert [_tuple(srv.wait(t).hit) for t in tickets] == want


@pytest.mark.parametrize("mode", ALL_MODES)
def test_unpackable_groups_beside_the_packed_kernel(mode):
    """Targets that fail the packed width proof but stay below the long-target threshold run in
    the 32-bit kernel on the scratch's second stream while the packed kernel runs on the first,
    each with boundary rows of its own: three 32-bit groups (some targets holding a piece of the
    query, so the winner falls there for some queries) beside 16 packed groups, queries of one to
    many register tiles."""
    rng = np.random.default_rng(1400 + mode)
    scheme = (4, [40 if a == b else -3 for a in range(4) for b in range(4)], 6, 2)
    q = rng.integers(0, 4, size=900, dtype=np.uint8)
    seqs = [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in rng.integers(20, 120, 2000)]
    for k in range(0, 2000, 13):  # 154 targets of 850..1000 residues: beyond the packed bound
        t = rng.integers(0, 4, size=int(rng.integers(850, 1001)), dtype=np.uint8)
        if k % 5 == 0:
            a, at = int(rng.integers(0, 600)), int(rng.integers(0, 500))
            t[at:at + 300] = q[a:a + 300]
        seqs[k] = t
    queries = [q, q[:300], q[600:], q[:65], q[:33], rng.integers(0, 4, size=40, dtype=np.uint8)]
    for st in ST:
        check(seqs, scheme, mode, st, queries, mql=900)
    db = build(seqs, scheme, mode, hy.SearchType.Score, 900)
    assert db.partition_sizes()[hy.ScoreWidth.I32] == 154


@pytest.mark.parametrize("mode", ALL_MODES)
def test_every_remainder_tile_size(mode):
    """Query lengths 1..73 take every remainder tile (8, 16 and full rows) after 0, 1 and 2 full
    tiles; per-target scores and ends equal the oracle's."""
    rng = np.random.default_rng(600 + mode)
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(0, 90, 300)]
    base = rng.integers(0, 20, size=73, dtype=np.uint8)
    queries = [base[:m] for m in range(1, 74)]
    check(seqs, (20, BLOSUM62, 11, 1), mode, hy.SearchType.ScoreEnd, queries, mql=73)
    check(seqs, (20, BLOSUM62, 11, 1), mode, hy.SearchType.Score, queries[::5], mql=73)


# ---- launch shapes and the autotuner ------------------------------------------------------------

def _tuning_db(mode, st):
    rng = np.random.default_rng(700 + mode)
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(1, 300, 1500)]
    seqs[11] = rng.integers(0, 20, size=2600, dtype=np.uint8)  # beyond the long-target threshold
    qs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in (5, 23, 64, 150)]
    return seqs, qs, build(seqs, (20, BLOSUM62, 11, 1), mode, st, 150)


def default_shapes(sc, db):
    """The scratch starts at the default shapes; its batched submissions hold 256 queries, or
    more (a power of two up to 4,096) for a database of few packed groups."""
    t = sc.tuning()
    return (dataclasses.replace(t, batch_chunk=256) == hy.Tuning() and
            t.batch_chunk in (256, 512, 1024, 2048, 4096))


@pytest.mark.parametrize("mode", [MODE_SW, MODE_NW, MODE_OV])
def test_every_launch_shape_gives_the_same_hits(mode):
    """Tile rows x workgroup shape of the one-query kernel and tile rows x resident workgroups x
    submission size of the batch kernel change speed only: every shape equals the oracle."""
    st = hy.SearchType.ScoreEnd
    seqs, qs, db = _tuning_db(mode, st)
    flat, off = orc.flatten(seqs)
    want = [orc.scan_flat(flat, off, q, 20, BLOSUM62, 11, 1, mode, True, threads=4) for q in qs]
    want_scores = [orc.scan_all_flat(flat, off, q, 20, BLOSUM62, 11, 1, mode, threads=4)[0]
                   for q in qs[:2]]
    sc = hy.Scratch(db)
    assert default_shapes(sc, db)
    for rows in (16, 24, 32):
        for waves in (1, 4, 8):
            sc.set_tuning(hy.Tuning(tile_rows=rows, waves=waves))
            for q, w in zip(qs, want):
                h = db.scan(sc, q)
                assert (h.score, h.db_index, h.query_end, h.target_end) == w, (rows, waves)
            for q, w in zip(qs, want_scores):
                assert np.array_equal(db.scan_scores_array(sc, q), w), (rows, waves)
    for rows in (16, 24, 32):
        for wpc in (0, 1):
            for chunk in (3, 256):
                for pairs in (0, 1):
                    sc.set_tuning(hy.Tuning(batch_tile_rows=rows, batch_wgs_per_cu=wpc,
                                            batch_chunk=chunk, batch_pairs=pairs))
                    got = db.scan_batch(sc, qs)
                    assert [(h.score, h.db_index, h.query_end, h.target_end)
                            for h in got] == want, (rows, wpc, chunk, pairs)


def test_invalid_launch_shapes_are_rejected():
    seqs, qs, db = _tuning_db(MODE_SW, hy.SearchType.Score)
    sc = hy.Scratch(db)
    for bad in (hy.Tuning(tile_rows=20), hy.Tuning(waves=2), hy.Tuning(batch_tile_rows=8),
                hy.Tuning(batch_chunk=0), hy.Tuning(batch_chunk=65537),
                hy.Tuning(batch_wgs_per_cu=-1), hy.Tuning(batch_pairs=2)):
        with pytest.raises(ValueError):
            sc.set_tuning(bad)
    assert default_shapes(sc, db)


def test_autotune_keeps_results_and_uses_its_cache(tmp_path, monkeypatch):
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("HYALITE_HIP_TUNE_CACHE", str(cache))
    seqs, qs, db = _tuning_db(MODE_SW, hy.SearchType.ScoreEnd)
    flat, off = orc.flatten(seqs)
    want = [orc.scan_flat(flat, off, q, 20, BLOSUM62, 11, 1, MODE_SW, True, threads=4) for q in qs]
    sc = hy.Scratch(db)
    t = sc.autotune(qs)
    assert sc.tuning() == t
    assert t.tile_rows in (16, 24, 32) and t.waves in (1, 4, 8)
    assert t.batch_tile_rows in (16, 24, 32) and 1 <= t.batch_chunk <= 65536
    assert [(h.score, h.db_index, h.query_end, h.target_end) for h in db.scan_batch(sc, qs)] == want
    for q, w in zip(qs, want):
        h = db.scan(sc, q)
        assert (h.score, h.db_index, h.query_end, h.target_end) == w
    lines = cache.read_text().splitlines()
    assert len(lines) == 1
    other = hy.Scratch(db)
    assert other.autotune(qs) == t  # served from the cache
    assert len(cache.read_text().splitlines()) == 1


# ---- serving form -------------------------------------------------------------------------------

def _serving_db(st=hy.SearchType.ScoreEnd):
    rng = np.random.default_rng(900)
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(1, 400, 3000)]
    seqs[17] = rng.integers(0, 20, size=3500, dtype=np.uint8)
    qs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(1, 300, 40)]
    qs[5] = seqs[1234][10:200].copy()  # a strong hit
    db = build(seqs, (20, BLOSUM62, 11, 1), MODE_SW, st, 300)
    flat, off = orc.flatten(seqs)
    want = [orc.scan_flat(flat, off, q, 20, BLOSUM62, 11, 1, MODE_SW,
                          st == hy.SearchType.ScoreEnd, threads=4) for q in qs]
    return db, qs, want


def _tuple(h):
    return (h.score, h.db_index, h.query_end, h.target_end)


@pytest.mark.parametrize("st", ST)
def test_served_hit_is_the_same_alone_and_in_the_mix(st):
    """Each served hit equals the oracle's scan of that query, whether it ran alone, in a burst
    where it shares batches with others, or under Poisson arrivals while batches overlap."""
    from hyalite_amd import loadgen
    db, qs, want = _serving_db(st)
    with hy.Server(db, max_lanes=2, max_batch=8) as srv:
        alone = [srv.wait(srv.submit(q)) for q in qs[:6]]
        assert all(a.batch_size == 1 for a in alone)
        assert [_tuple(a.hit) for a in alone] == want[:6]
        tickets = [srv.submit(q) for q in qs]
        burst = [srv.wait(t) for t in tickets]
        assert [_tuple(b.hit) for b in burst] == want
        assert max(b.batch_size for b in burst) > 1
        res = loadgen.run_poisson(srv, qs, qps=400.0, seed=1)
        assert [_tuple(s.hit) for s in res.served] == want
        assert all(s.latency_s >= s.queued_s >= 0 for s in res.served)
        stt = srv.stats()
        assert stt.completed == stt.submitted == 6 + 2 * len(qs)
        assert 1 <= stt.lanes <= 2 and stt.peak_queries > 1 and stt.inflight_bytes == 0


def test_server_memory_budget_sets_the_lanes():
    """Lanes are built on demand within the byte budget; prealloc builds them all at once."""
    db, qs, want = _serving_db()
    with hy.Server(db, max_lanes=4, max_batch=8, prealloc=True) as srv:
        lane = srv.stats().lane_bytes
        assert srv.stats().lanes == 4 and lane > 0
    with hy.Server(db, max_lanes=4, max_batch=8) as srv:
        assert srv.stats().lanes == 1  # on demand: one lane until requests wait
    budget = int(lane * 2.5)
    with hy.Server(db, scratch_budget_bytes=budget, max_lanes=4, max_batch=2) as srv:
        assert srv.stats().lanes == 1
        tickets = [srv.submit(q) for q in qs]  # a burst: every lane busy, requests waiting
        assert [_tuple(srv.wait(t).hit) for t in tickets] == want
        st = srv.stats()
        assert 1 <= st.lanes <= 2 and st.peak_bytes <= budget, st
    with hy.Server(db, scratch_budget_bytes=budget, max_lanes=4, max_batch=8,
                   prealloc=True) as srv:
        assert srv.stats().lanes == 2
        assert [_tuple(srv.wait(srv.submit(q)).hit) for q in qs[:3]] == want[:3]
    with pytest.raises(hy.DeviceError):
        hy.Server(db, scratch_budget_bytes=lane // 2, max_lanes=4, max_batch=8)
    for bad in (dict(max_lanes=0), dict(max_batch=0), dict(max_lanes=65)):
        with pytest.raises(ValueError):
            hy.Server(db, **bad)


def test_server_admits_requests_by_their_bytes():
    """A budget that holds one lane plus about two long requests: requests beyond 512 rows carry
    a winner boundary row each, so at most that many are in flight at once; every hit still
    equals the oracle."""
    rng = np.random.default_rng(901)
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(1, 400, 500)]
    seqs[3] = rng.integers(0, 20, size=60_000, dtype=np.uint8)  # max_target_len 60,000
    db = build(seqs, (20, BLOSUM62, 11, 1), MODE_SW, hy.SearchType.ScoreEnd, 700)
    qs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(520, 700, 8)]
    flat, off = orc.flatten(seqs)
    want = [orc.scan_flat(flat, off, q, 20, BLOSUM62, 11, 1, MODE_SW, True, threads=4)
            for q in qs]
    with hy.Server(db, max_lanes=1, max_batch=8) as probe:
        lane = probe.stats().lane_bytes
    row = (60_000 + 1) * 8  # one winner boundary row
    budget = lane + 2 * (row + 1024) + row // 2
    with hy.Server(db, scratch_budget_bytes=budget, max_lanes=1, max_batch=8) as srv:
        tickets = [srv.submit(q) for q in qs]
        served = [srv.wait(t) for t in tickets]
        assert [_tuple(s.hit) for s in served] == want
        st = srv.stats()
        assert st.peak_queries == 2 and max(s.batch_size for s in served) <= 2, st
        assert st.peak_bytes <= budget and st.inflight_bytes == 0


def test_scratches_built_beside_first_batch_captures():
    """Scratches built, reserved and freed on one thread while fresh scratches on another thread
    capture their first batched submission as a graph: every batch equals the oracle, and no
    scratch fails to build (the library's allocations never fall inside a capture)."""
    db, qs, want = _serving_db()
    stop = threading.Event()
    errors = []

    def churn():
        while not stop.is_set():
            try:
                sc = hy.Scratch(db)
                sc.reserve_batch(8)
                del sc
            except Exception as e:  # reported below
                errors.append(e)
                return

    t = threading.Thread(target=churn)
    t.start()
    try:
        for k in range(8):
            sc = hy.Scratch(db)
            part = slice(4 * k % 36, 4 * k % 36 + 4)
            got = [_tuple(h) for h in db.scan_batch(sc, qs[part])]
            assert got == want[part], k
            del sc
    finally:
        stop.set()
        t.join()
    assert not errors, errors


def test_server_contract_errors():
    db, qs, want = _serving_db()
    with hy.Server(db, max_lanes=1, max_batch=4) as srv:
        with pytest.raises(hy.ScanContractViolation):
            srv.submit(np.zeros(301, dtype=np.uint8))
        with pytest.raises(hy.ScanContractViolation):
            srv.submit(np.array([0, 20], dtype=np.uint8))
        with pytest.raises(KeyError):
            srv.wait(12345)
        t = srv.submit(qs[0])
        try:
            got = srv.wait(t, timeout=0.0)
        except hy.ServerTimeout:
            got = srv.wait(t)
        assert _tuple(got.hit) == want[0]
        with pytest.raises(KeyError):
            srv.wait(t)  # already collected


@pytest.mark.parametrize("mode", ALL_MODES)
def test_winner_ends_across_band_passes(mode, monkeypatch):
    """The one-query scan's winner ends come from the band kernel in winner mode. Forced to its
    smallest shape (64 rows per pass), a 300-row query takes five passes with boundary rows in
    the scratch (max_query_len 600 gives it a target's worth); with max_query_len 300 the scratch
    has none and the winner pass falls back to the one-wavefront kernel. Hits equal the oracle."""
    monkeypatch.setenv("HYALITE_HIP_BAND_SHAPE", "1,1")
    rng = np.random.default_rng(1500 + mode)
    q = rng.integers(0, 20, size=300, dtype=np.uint8)
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(1, 700, 400)]
    for k in range(0, 400, 37):  # pieces of the query, so winners sit on long alignments
        t = seqs[k]
        if t.size > 120:
            a = int(rng.integers(0, 180))
            t[10:110] = q[a:a + 100]
    queries = [q, q[:200], q[150:], q[:64], q[:65]]
    for mql in (600, 300):
        check(seqs, (20, BLOSUM62, 11, 1), mode, hy.SearchType.ScoreEnd, queries, mql=mql,
              all_targets=False)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_pipelined_kernel_every_tile_count(mode, monkeypatch):
    """The pipelined packed kernel (one query tile per wavefront of a workgroup, bottom rows
    through LDS) for every tile count 2..8 at each tile height (16, 24, 32 rows) and every
    remainder size, against the oracle and against the one-wavefront-per-group kernel; two
    alphabets, targets of every padding class."""
    rng = np.random.default_rng(1200 + mode)
    monkeypatch.setenv("HYALITE_HIP_CHAIN", "0")
    for al, mat, go, ge in ((4, identity_matrix(4, 2, -1), 2, 1), (20, BLOSUM62, 11, 1)):
        seqs = [rng.integers(0, al, size=int(n), dtype=np.uint8)
                for n in rng.integers(0, 300, 700)]
        flat, off = orc.flatten(seqs)
        db = build(seqs, (al, mat, go, ge), mode, hy.SearchType.ScoreEnd, 256)
        sc = hy.Scratch(db)
        for m in (17, 24, 31, 40, 56, 97, 100, 128, 129, 150, 177, 192, 193, 230, 256):
            q = rng.integers(0, al, size=m, dtype=np.uint8)
            ws = orc.scan_all_flat(flat, off, q, al, mat, go, ge, mode, ends=False, threads=8)
            want = orc.scan_flat(flat, off, q, al, mat, go, ge, mode, True, threads=8)
            for pipe, kernel in (("1", "pipelined"), ("0", "packed")):
                monkeypatch.setenv("HYALITE_HIP_PIPE", pipe)
                assert np.array_equal(db.scan_scores_array(sc, q), ws), (al, m, pipe)
                assert sc.last_kernel() == kernel, (al, m, pipe)
                h = db.scan(sc, q)
                assert (h.score, h.db_index, h.query_end, h.target_end) == want, (al, m, pipe)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_pipelined_kernel_tracks_ends(mode, monkeypatch):
    """scan_all with ends on a database of few groups: the pipelined kernel, each wavefront
    tracking its tile's best cells (local mode) or answer cells (the other modes) and wavefront 0
    folding them with the full tie-break, gives every target's (score, query end, target end) as
    the oracle and the one-wavefront end-tracking kernel do, over tile counts 2..8 at each tile
    height the mode allows (local mode at most 24 rows), two alphabets, targets of every padding
    class (empty ones included). The chained kernel, which would take these databases first, is
    turned off."""
    monkeypatch.setenv("HYALITE_HIP_CHAIN", "0")
    rng = np.random.default_rng(1250 + mode)
    for al, mat, go, ge in ((4, identity_matrix(4, 2, -1), 2, 1), (20, BLOSUM62, 11, 1)):
        seqs = [rng.integers(0, al, size=int(n), dtype=np.uint8)
                for n in rng.integers(0, 300, 700)]
        flat, off = orc.flatten(seqs)
        db = build(seqs, (al, mat, go, ge), mode, hy.SearchType.ScoreEnd, 256)
        sc = hy.Scratch(db)
        for m in (17, 24, 40, 63, 100, 129, 177, 192, 193, 256):
            q = rng.integers(0, al, size=m, dtype=np.uint8)
            want = orc.scan_all_flat(flat, off, q, al, mat, go, ge, mode, ends=True, threads=8)
            for pipe in ("1", "0"):
                monkeypatch.setenv("HYALITE_HIP_PIPE", pipe)
                got = db.scan_all_arrays(sc, q)
                for g, w_ in zip(got, want):
                    assert np.array_equal(np.asarray(g), np.asarray(w_)), (al, m, pipe)
                kernel = sc.last_kernel()
                if pipe == "0":
                    assert kernel == "ends", (al, m)
                elif m in (40, 100):  # shapes whose LDS fits for both alphabets
                    assert kernel == "pipelined", (al, m)
                else:
                    assert kernel in ("pipelined", "ends"), (al, m)


def chain_chunks(m, rows, waves):
    tiles = -(-m // rows)
    return -(-tiles // min(waves, tiles))


@pytest.mark.parametrize("mode", ALL_MODES)
def test_chained_kernel_every_shape(mode, monkeypatch):
    """The chained pipelined kernel (a group's query tiles spread over 2-8 workgroups, each
    reading the previous one's bottom rows and folded maxima through the group's boundary buffer
    under published progress counts) at every tile height (8, 16, 24, 32 rows), 1-4 and 8
    wavefronts per workgroup (short last chunks included), reader slack 1 and 4 and writers
    publishing every block and every third block, for query
    lengths giving every chunk count and remainder tile, against the oracle: scan_scores and
    scan, on 23 packed groups with targets of every padding class, two alphabets. Repeated
    scans check that every launch's epoch and tickets start afresh."""
    rng = np.random.default_rng(1700 + mode)
    monkeypatch.setenv("HYALITE_HIP_CHAIN", "1")
    ran = 0
    for al, mat, go, ge in ((4, identity_matrix(4, 2, -1), 2, 1), (20, BLOSUM62, 11, 1)):
        seqs = [rng.integers(0, al, size=int(n), dtype=np.uint8)
                for n in rng.integers(0, 300, 2900)]
        flat, off = orc.flatten(seqs)
        db = build(seqs, (al, mat, go, ge), mode, hy.SearchType.ScoreEnd, 256)
        sc = hy.Scratch(db)
        for m in (9, 17, 40, 65, 100, 129, 200, 256):
            q = rng.integers(0, al, size=m, dtype=np.uint8)
            ws = orc.scan_all_flat(flat, off, q, al, mat, go, ge, mode, ends=False, threads=8)
            want = orc.scan_flat(flat, off, q, al, mat, go, ge, mode, True, threads=8)
            for rows in (8, 16, 24, 32):
                for waves in (1, 2, 3, 4, 8):
                    if not 2 <= chain_chunks(m, rows, waves) <= 8:
                        continue
                    monkeypatch.setenv("HYALITE_HIP_CHAIN_ROWS", str(rows))
                    monkeypatch.setenv("HYALITE_HIP_CHAIN_WAVES", str(waves))
                    for slack, every in ((1, 1), (4, 3)):
                        monkeypatch.setenv("HYALITE_HIP_CHAIN_SLACK", str(slack))
                        monkeypatch.setenv("HYALITE_HIP_CHAIN_EVERY", str(every))
                        case = (al, m, rows, waves, slack, every)
                        assert np.array_equal(db.scan_scores_array(sc, q), ws), case
                        assert sc.last_kernel() == "chained", case
                        h = db.scan(sc, q)
                        assert (h.score, h.db_index, h.query_end, h.target_end) == want, case
                        ran += 1
            # The default rule on a database this small: the chained kernel once 8-row tiles make
            # two chunks of four (33 rows on), the pipelined kernel below, the one-wavefront
            # kernel for one tile; at 400 rows (below) no pipelined shape exists.
            for var in ("HYALITE_HIP_CHAIN", "HYALITE_HIP_CHAIN_ROWS", "HYALITE_HIP_CHAIN_WAVES",
                        "HYALITE_HIP_CHAIN_SLACK", "HYALITE_HIP_CHAIN_EVERY"):
                monkeypatch.delenv(var, raising=False)
            assert np.array_equal(db.scan_scores_array(sc, q), ws), (al, m)
            assert sc.last_kernel() == ("chained" if m > 32 else
                                        "pipelined" if m > 16 else "packed"), (al, m)
            monkeypatch.setenv("HYALITE_HIP_CHAIN", "1")
        q = rng.integers(0, al, size=400, dtype=np.uint8)
        db = build(seqs, (al, mat, go, ge), mode, hy.SearchType.ScoreEnd, 400)
        sc = hy.Scratch(db)
        monkeypatch.delenv("HYALITE_HIP_CHAIN")
        ws = orc.scan_all_flat(flat, off, q, al, mat, go, ge, mode, ends=False, threads=8)
        assert np.array_equal(db.scan_scores_array(sc, q), ws), al
        assert sc.last_kernel() == "chained", al
        monkeypatch.setenv("HYALITE_HIP_CHAIN", "1")
    assert ran > 100


@pytest.mark.parametrize("mode", ALL_MODES)
def test_chained_kernel_tracks_ends(mode, monkeypatch):
    """The chained kernel with end tracking: each chunk folds its wavefronts' best cells, joins
    those the chunk before it passed on and passes the result to the next; the last chunk writes
    every target's (score, query end, target end). scan_all with ends against the oracle at every
    tile height the mode allows (local mode at most 16 rows) and 1, 2, 4 and 8 wavefronts per
    workgroup, two alphabets, targets of every padding class, and the default rule's choice."""
    rng = np.random.default_rng(1800 + mode)
    rows_ok = (8, 16) if mode == MODE_SW else (8, 16, 24, 32)
    ran = 0
    for al, mat, go, ge in ((4, identity_matrix(4, 2, -1), 2, 1), (20, BLOSUM62, 11, 1)):
        seqs = [rng.integers(0, al, size=int(n), dtype=np.uint8)
                for n in rng.integers(0, 300, 2900)]
        flat, off = orc.flatten(seqs)
        db = build(seqs, (al, mat, go, ge), mode, hy.SearchType.ScoreEnd, 256)
        sc = hy.Scratch(db)
        for m in (17, 40, 100, 129, 256):
            q = rng.integers(0, al, size=m, dtype=np.uint8)
            want_all = orc.scan_all_flat(flat, off, q, al, mat, go, ge, mode, ends=True,
                                         threads=8)
            monkeypatch.setenv("HYALITE_HIP_CHAIN", "1")
            for rows in rows_ok:
                for waves in (1, 2, 4, 8):
                    if not 2 <= chain_chunks(m, rows, waves) <= 8:
                        continue
                    monkeypatch.setenv("HYALITE_HIP_CHAIN_ROWS", str(rows))
                    monkeypatch.setenv("HYALITE_HIP_CHAIN_WAVES", str(waves))
                    case = (al, m, rows, waves)
                    got = db.scan_all_arrays(sc, q)
                    for g, w_ in zip(got, want_all):
                        assert np.array_equal(np.asarray(g), np.asarray(w_)), case
                    assert sc.last_kernel() == "chained", case
                    ran += 1
            for var in ("HYALITE_HIP_CHAIN", "HYALITE_HIP_CHAIN_ROWS",
                        "HYALITE_HIP_CHAIN_WAVES"):
                monkeypatch.delenv(var, raising=False)
            # The default rule: the chained kernel once 8-row tiles make two chunks of four.
            got = db.scan_all_arrays(sc, q)
            for g, w_ in zip(got, want_all):
                assert np.array_equal(np.asarray(g), np.asarray(w_)), (al, m)
            if m > 32:
                assert sc.last_kernel() == "chained", (al, m)
    assert ran > 40


def test_chained_wait_that_gives_up_is_reported(monkeypatch):
    """The chained kernel's cross-workgroup waits are bounded: with the bound forced to 0 polls
    a reader gives up at its first unmet poll, the kernel still runs to its end, and the scan
    reports an internal error instead of results (it may also find every wait already met and
    return exact results). The same scratch's next scans, with the default bound, are exact."""
    rng = np.random.default_rng(1750)
    seqs = [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in rng.integers(50, 300, 3000)]
    flat, off = orc.flatten(seqs)
    al, mat, go, ge = 4, identity_matrix(4, 2, -1), 2, 1
    db = build(seqs, (al, mat, go, ge), MODE_SW, hy.SearchType.ScoreEnd, 200)
    sc = hy.Scratch(db)
    q = rng.integers(0, al, size=200, dtype=np.uint8)
    ws = orc.scan_all_flat(flat, off, q, al, mat, go, ge, MODE_SW, ends=False, threads=8)
    want = orc.scan_flat(flat, off, q, al, mat, go, ge, MODE_SW, True, threads=8)
    for var, v in (("HYALITE_HIP_CHAIN", "1"), ("HYALITE_HIP_CHAIN_ROWS", "8"),
                   ("HYALITE_HIP_CHAIN_WAVES", "4"), ("HYALITE_HIP_CHAIN_SPIN_LIMIT", "0")):
        monkeypatch.setenv(var, v)
    errors = 0
    for _ in range(4):
        try:
            got = db.scan_scores_array(sc, q)
        except hy.DeviceError as e:
            assert "status 106" in str(e), e
            errors += 1
        else:
            assert np.array_equal(got, ws)
        assert sc.last_kernel() == "chained"
    assert errors >= 1
    monkeypatch.delenv("HYALITE_HIP_CHAIN_SPIN_LIMIT")
    for _ in range(2):
        assert np.array_equal(db.scan_scores_array(sc, q), ws)
        assert sc.last_kernel() == "chained"
        h = db.scan(sc, q)
        assert (h.score, h.db_index, h.query_end, h.target_end) == want


@pytest.mark.parametrize("mode", [MODE_SW, MODE_NW, MODE_OV])
def test_packed_groups_wider_than_16bit_columns(mode, monkeypatch):
    """Targets of 66,000-70,000 residues kept out of the systolic kernel (the long-target
    threshold raised) sit in packed groups wider than 65,535 columns: local-mode end tracking
    keeps columns in 16 bits, so its scan_all then runs every non-long target at 32 bits; the
    other modes keep their ends in 32 bits inside the packed kernel. Both against the oracle."""
    monkeypatch.setenv("HYALITE_HIP_LONG_THRESHOLD", "200000")
    rng = np.random.default_rng(1300 + mode)
    seqs = [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in (66_000, 70_000, 67_500)]
    seqs += [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in rng.integers(0, 300, 200)]
    q = rng.integers(0, 4, size=40, dtype=np.uint8)
    seqs[1][30_000:30_040] = q  # a full-length copy far beyond column 65,535 in one case
    seqs[2][66_000:66_040] = q
    check(seqs, dna_scheme(), mode, hy.SearchType.ScoreEnd, [q, q[:20]], mql=40)
