"""GPU parity of the aligned scans (Database::scan_aligned / scan_all_aligned): every alignment the
device returns must equal the pinned oracle's canonical traceback of the same pair, op for op.

Mirrors tests/database_alignment.rs (per-target alignments in database order, the winner of the
score pass with its alignment, the budget check at build) and the hand vectors of
src/align/tests.rs, then widens to every short pair, protein pairs across the kernel's
rows-per-lane shapes and multi-pass queries, long targets, and chunked runs.
"""
import os
import random

import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from common import (ALL_MODES, BLOSUM62, MODE_HW, MODE_NW, MODE_OV, MODE_SHW, MODE_SW,
                    all_sequences, dna_scheme, identity_matrix)

pytestmark = pytest.mark.gpu

M, X, I, D = 0, 1, 2, 3


def build(seqs, scheme, mode, mql, budget=(1 << 64) - 1):
    al, mat, go, ge = scheme
    return (hy.Database.builder().sequences(seqs).scoring(hy.Scoring(al, mat, go, ge))
            .mode(hy.Mode(mode)).search_type(hy.SearchType.Alignment(budget))
            .max_query_len(mql).backend(hy.BackendChoice.force(hy.Backend.Hip)).build())


def want_all(seqs, q, scheme, mode):
    al, mat, go, ge = scheme
    return [orc.traceback(q, t, al, mat, go, ge, mode) for t in seqs]


def winner(wants):
    best = max(w[0] for w in wants)
    return next(k for k, w in enumerate(wants) if w[0] == best)


def check_db(seqs, scheme, mode, queries, mql=None, aligned=True):
    mql = max([len(q) for q in queries] + [1]) if mql is None else mql
    db = build(seqs, scheme, mode, mql)
    sc = hy.Scratch(db)
    for q in queries:
        wants = want_all(seqs, q, scheme, mode)
        got = db.scan_all_aligned(sc, q)
        assert len(got) == len(seqs)
        for k, (g, w) in enumerate(zip(got, wants)):
            assert g.as_tuple() == w, (hy.Mode(mode), k, list(q), list(seqs[k]))
        if aligned:
            hit = db.scan_aligned(sc, q)
            k = winner(wants)
            assert hit.db_index == k, (hy.Mode(mode), list(q))
            assert hit.alignment.as_tuple() == wants[k], (hy.Mode(mode), list(q))
            # The winner is the plain scan's winner, and its ends are the scan's ends.
            best = db.scan(sc, q)
            assert (best.db_index, best.score) == (k, wants[k][0])
    return db, sc


# ---- src/align/tests.rs and tests/database_alignment.rs known answers ---------------------------

def test_reference_hand_vectors_through_the_database():
    dna = dna_scheme()
    cases = [
        (MODE_NW, [0, 1, 2, 3], [0, 1, 2, 3], (8, 0, 4, 0, 4, [M] * 4)),
        (MODE_NW, [0, 1, 2, 3], [0, 1, 1, 3], (5, 0, 4, 0, 4, [M, M, X, M])),
        (MODE_SW, [2, 2, 0, 1, 2, 3, 2, 2], [0, 1, 2, 3], (8, 2, 6, 0, 4, [M] * 4)),
        (MODE_HW, [1, 2], [0, 1, 2, 3], (4, 0, 2, 1, 3, [M, M])),
        (MODE_SHW, [0, 1, 2, 3], [1, 2], (4, 1, 3, 0, 2, [M, M])),
        (MODE_NW, [], [0, 1, 2], (-4, 0, 0, 0, 3, [D, D, D])),
        (MODE_SW, [0, 0, 0], [1, 1, 1], (0, 0, 0, 0, 0, [])),
    ]
    for mode, q, t, want in cases:
        db = build([t], dna, mode, max(len(q), 1))
        sc = hy.Scratch(db)
        hit = db.scan_aligned(sc, q)
        assert hit.db_index == 0 and hit.alignment.as_tuple() == want, (mode, q, t)
        assert db.scan_all_aligned(sc, q)[0].as_tuple() == want
    # OV tie: the smaller target end wins.
    db = build([[2, 3, 2, 3]], dna, MODE_OV, 2)
    a = db.scan_aligned(hy.Scratch(db), [2, 3]).alignment
    assert (a.score, a.target_start, a.target_end) == (4, 0, 2)
    # database_alignment.rs:84-118: eight matches, "8M".
    seq = [0, 1, 2, 3, 0, 1, 2, 3]
    db = build([seq], dna, MODE_NW, 8)
    hit = db.scan_aligned(hy.Scratch(db), seq)
    assert hit.db_index == 0 and hit.alignment.score == 16 and hit.alignment.cigar() == "8M"
    # database_alignment.rs:120-156: per-target in order; the winner is target 1.
    seqs = [[0, 0, 0, 0], [0, 1, 2, 3], [3, 3, 3, 3]]
    db = build(seqs, dna, MODE_SW, 4)
    sc = hy.Scratch(db)
    out = []
    db.scan_all_aligned(sc, [0, 1, 2, 3], out)
    assert len(out) == 3 and out[1].score == 8 and out[1].cigar() == "4M"
    hit = db.scan_aligned(sc, [0, 1, 2, 3])
    assert hit.db_index == 1 and hit.alignment == out[1]


def test_unknown_symbol_at_the_start_traces_back():
    """src/align/tests.rs:235-256: an in-alphabet 'N' (index 4) leading both sequences."""
    mat = identity_matrix(5, 2, -1)
    mat[24] = -1
    for mode in ALL_MODES:
        check_db([[4, 0, 1]], (5, mat, 2, 1), mode, [[4, 0, 1]])


def test_aligned_scan_requires_an_alignment_database():
    dna = dna_scheme()
    al, mat, go, ge = dna
    db = (hy.Database.builder().sequences([[0, 1]]).scoring(hy.Scoring(al, mat, go, ge))
          .mode(hy.Mode.Sw).search_type(hy.SearchType.ScoreEnd).max_query_len(4)
          .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())
    sc = hy.Scratch(db)
    with pytest.raises(hy.ScanContractViolation):
        db.scan_aligned(sc, [0, 1])
    with pytest.raises(hy.ScanContractViolation):
        db.scan_all_aligned(sc, [0, 1])
    adb = build([[0, 1]], dna, MODE_SW, 4)
    with pytest.raises(hy.ScanContractViolation):
        adb.scan_aligned(hy.Scratch(adb), [0, 1, 2, 3, 0])   # longer than max_query_len
    with pytest.raises(hy.ScanContractViolation):
        adb.scan_aligned(hy.Scratch(adb), [0, 9])            # symbol out of range


def test_budget_is_validated_at_build_on_the_device_path():
    with pytest.raises(hy.TracebackBudgetExceeded):
        build([[0, 1, 2, 3, 0, 1, 2, 3]], dna_scheme(), MODE_NW, 8, budget=1)
    db = build([[0, 1, 2, 3, 0, 1, 2, 3]], dna_scheme(), MODE_NW, 8, budget=684)
    want = orc.traceback([0, 1, 2, 3], [0, 1, 2, 3, 0, 1, 2, 3], *dna_scheme(), MODE_NW)
    assert db.scan_aligned(hy.Scratch(db), [0, 1, 2, 3]).alignment.as_tuple() == want


# ---- every short pair, every mode, the reference's four scorings --------------------------------

SCORINGS = [
    dna_scheme(),
    (4, identity_matrix(4, 1, -3), 10, 5),
    (4, identity_matrix(4, 3, -2), 4, 0),
    (4, identity_matrix(4, 5, -1), 1, 1),
]


@pytest.mark.parametrize("scheme_ix", range(4))
def test_every_short_pair_in_every_mode(scheme_ix):
    scheme = SCORINGS[scheme_ix]
    seqs = all_sequences(3, 3)          # 40 sequences, the empty one included
    for mode in ALL_MODES:
        check_db(seqs, scheme, mode, seqs[::3], mql=3)


def test_random_schemes_like_the_reference_proptest():
    """tests/database_alignment.rs:18-81: alphabets 2-4, entries in [-5, 5], 0 <= ge <= go <= 8,
    one to five targets and queries of 0-7 symbols."""
    rng = random.Random(7)
    for _ in range(120):
        al = rng.randint(2, 4)
        mat = [rng.randint(-5, 5) for _ in range(al * al)]
        go = rng.randint(0, 8)
        ge = rng.randint(0, go)
        seqs = [[rng.randrange(al) for _ in range(rng.randrange(8))]
                for _ in range(rng.randint(1, 5))]
        q = [rng.randrange(al) for _ in range(rng.randrange(8))]
        for mode in ALL_MODES:
            check_db(seqs, (al, mat, go, ge), mode, [q], mql=12)


# ---- protein pairs across the DP shapes ---------------------------------------------------------

def protein_pairs(rng, n, lo, hi):
    return [rng.integers(0, 20, size=int(rng.integers(lo, hi + 1)), dtype=np.uint8)
            for _ in range(n)]


def mutate(rng, seq, rate=0.15):
    out = []
    for c in seq:
        r = rng.random()
        if r < rate / 3:
            continue
        if r < 2 * rate / 3:
            out.append(int(rng.integers(0, 20)))
        elif r < rate:
            out.extend([int(c), int(rng.integers(0, 20))])
        else:
            out.append(int(c))
    return np.array(out, dtype=np.uint8)


@pytest.mark.parametrize("qlen", [1, 63, 100, 129, 250, 300, 384, 385, 600])
def test_protein_pairs_every_rows_per_lane_shape(qlen):
    """Query lengths on both sides of each rows-per-lane step (128, 256, 384) and past one
    pass (512 rows), against related and unrelated targets of 0-900 residues."""
    rng = np.random.default_rng(qlen)
    q = rng.integers(0, 20, size=qlen, dtype=np.uint8)
    seqs = protein_pairs(rng, 10, 0, 900) + [mutate(rng, q) for _ in range(6)] + [
        np.zeros(0, dtype=np.uint8), q.copy()]
    scheme = (20, BLOSUM62, 11, 1)
    for mode in ALL_MODES:
        check_db(seqs, scheme, mode, [q, mutate(rng, q, 0.3)[:qlen]], mql=qlen)


def test_long_target_multi_pass_query():
    rng = np.random.default_rng(11)
    q = rng.integers(0, 20, size=700, dtype=np.uint8)
    core = mutate(rng, q, 0.2)
    t = np.concatenate([rng.integers(0, 20, size=4000, dtype=np.uint8), core,
                        rng.integers(0, 20, size=4000, dtype=np.uint8)])
    scheme = (20, BLOSUM62, 11, 1)
    for mode in (MODE_SW, MODE_HW, MODE_NW):
        check_db([t, core], scheme, mode, [q], mql=700)


def test_chunked_runs_match_one_chunk(monkeypatch):
    """A traceback pool smaller than the database's move codes splits the run into chunks
    (down to one target per chunk); the alignments do not change."""
    rng = np.random.default_rng(5)
    q = rng.integers(0, 20, size=200, dtype=np.uint8)
    seqs = protein_pairs(rng, 40, 0, 600) + [mutate(rng, q) for _ in range(8)]
    scheme = (20, BLOSUM62, 11, 1)
    db = build(seqs, scheme, MODE_SW, 200)
    sc = hy.Scratch(db)
    whole = [a.as_tuple() for a in db.scan_all_aligned(sc, q)]
    assert whole == want_all(seqs, q, scheme, MODE_SW)
    for pool in ("1", "300000"):
        monkeypatch.setenv("HYALITE_HIP_TRACE_POOL_BYTES", pool)
        sc2 = hy.Scratch(db)
        assert [a.as_tuple() for a in db.scan_all_aligned(sc2, q)] == whole, pool
        # The pool of a live scratch grows but never shrinks; a second call reuses it.
        assert [a.as_tuple() for a in db.scan_all_aligned(sc2, q[:150])] == want_all(
            seqs, q[:150], scheme, MODE_SW)


def test_repeated_calls_and_query_lengths_reuse_the_scratch():
    rng = np.random.default_rng(9)
    seqs = protein_pairs(rng, 30, 1, 500)
    scheme = (20, BLOSUM62, 11, 1)
    db = build(seqs, scheme, MODE_OV, 520)
    sc = hy.Scratch(db)
    for qlen in (520, 3, 130, 520, 0, 64):
        q = rng.integers(0, 20, size=qlen, dtype=np.uint8)
        wants = want_all(seqs, q, scheme, MODE_OV)
        assert [a.as_tuple() for a in db.scan_all_aligned(sc, q)] == wants
        hit = db.scan_aligned(sc, q)
        assert hit.alignment.as_tuple() == wants[winner(wants)]


def test_scan_paths_track_ends_on_an_alignment_database():
    """SearchType::Alignment tracks ends (search.rs:33-35): scan and scan_all report them."""
    rng = np.random.default_rng(2)
    seqs = protein_pairs(rng, 50, 1, 300)
    q = rng.integers(0, 20, size=120, dtype=np.uint8)
    al, mat, go, ge = 20, BLOSUM62, 11, 1
    db = build(seqs, (al, mat, go, ge), MODE_SW, 120)
    sc = hy.Scratch(db)
    flat, off = orc.flatten(seqs)
    want = orc.scan_flat(flat, off, q, al, mat, go, ge, MODE_SW, True)
    got = db.scan(sc, q)
    assert (got.score, got.db_index, got.query_end, got.target_end) == tuple(want)
    s, qe, te = db.scan_all_arrays(sc, q)
    for k, t in enumerate(seqs):
        ws, wq, wt = orc.align_core(q, t, al, mat, go, ge, MODE_SW)
        assert (s[k], qe[k], te[k]) == (ws, -1 if wq is None else wq, -1 if wt is None else wt)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_packed_traceback_beside_the_systolic_one(mode, monkeypatch):
    """scan_all_aligned runs the targets of the packed 16-bit layout in the packed traceback
    (k_trace16: tiles of 32 rows, codes per tile, column and 4-row block) and the others (here:
    targets of 2,980-3,300 residues whose proof needs 32 bits at max_query_len 3,000) in the
    systolic one, in the same call: every alignment equals the all-systolic run's and the
    oracle's, with long packable targets (past the long-target threshold) and empty ones among
    them, with a code pool small enough to split the run into chunks of single groups."""
    rng = np.random.default_rng(60 + mode)
    q = rng.integers(0, 20, size=300, dtype=np.uint8)
    seqs = protein_pairs(rng, 300, 0, 700) + [mutate(rng, q) for _ in range(10)]
    seqs += protein_pairs(rng, 6, 1200, 2500)           # long, packable
    seqs += protein_pairs(rng, 3, 2980, 3300)           # 32-bit
    seqs += [np.zeros(0, dtype=np.uint8)] * 3
    seqs[5] = np.concatenate([rng.integers(0, 20, 2990, dtype=np.uint8), q[:200]])
    scheme = (20, BLOSUM62, 11, 1)
    db = build(seqs, scheme, mode, 3000)
    assert db.partition_sizes()[hy.ScoreWidth.I32] >= 3
    sc = hy.Scratch(db)
    got = [a.as_tuple() for a in db.scan_all_aligned(sc, q)]
    # The ops come back packed in database order.
    recs, ops = db.scan_all_aligned_arrays(sc, q)
    n_ops = recs["n_ops"].astype(np.uint64)
    assert np.array_equal(recs["ops_offset"],
                          np.concatenate([[0], np.cumsum(n_ops)[:-1]]).astype(np.uint64))
    assert int(n_ops.sum()) <= ops.size
    monkeypatch.setenv("HYALITE_HIP_TRACE16", "0")
    systolic = [a.as_tuple() for a in db.scan_all_aligned(sc, q)]
    assert got == systolic
    monkeypatch.delenv("HYALITE_HIP_TRACE16")
    idx = list(range(0, len(seqs), 7)) + list(range(len(seqs) - 20, len(seqs)))
    al, mat, go, ge = scheme
    for k in idx:
        assert got[k] == orc.traceback(q, seqs[k], al, mat, go, ge, mode), k
    monkeypatch.setenv("HYALITE_HIP_TRACE_POOL_BYTES", "1")
    sc2 = hy.Scratch(db)
    assert [a.as_tuple() for a in db.scan_all_aligned(sc2, q[:150])] == [
        a.as_tuple() for a in db.scan_all_aligned(sc, q[:150])]


@pytest.mark.parametrize("mode", [MODE_SW, MODE_NW, MODE_OV])
def test_packed_traceback_large_alphabet_fits_its_lds(mode, monkeypatch):
    """An alphabet of 120 symbols: each wavefront of the packed traceback stages 121 symbols of
    its tile profile in LDS, so a 300-row query (10 tiles) runs with as many wavefronts as fit
    the workgroup's 64 KB (5), not one per tile; every alignment equals the all-systolic run's
    and the oracle's."""
    rng = np.random.default_rng(80 + mode)
    al = 120
    mat = rng.integers(-4, 4, size=(al, al))
    mat = np.triu(mat) + np.triu(mat, 1).T
    np.fill_diagonal(mat, rng.integers(4, 12, size=al))
    scheme = (al, [int(x) for x in mat.ravel()], 11, 1)
    q = rng.integers(0, al, size=300, dtype=np.uint8)
    seqs = [rng.integers(0, al, size=int(rng.integers(40, 400)), dtype=np.uint8)
            for _ in range(250)]
    for k in range(0, 250, 25):
        at = int(rng.integers(0, 100))
        seqs[k] = np.concatenate([seqs[k][:at], q[20:220], seqs[k][at:]])
    db = build(seqs, scheme, mode, 300)
    assert db.partition_sizes()[hy.ScoreWidth.I32] == 0  # every target in the packed layout
    sc = hy.Scratch(db)
    got = [a.as_tuple() for a in db.scan_all_aligned(sc, q)]
    monkeypatch.setenv("HYALITE_HIP_TRACE16", "0")
    assert got == [a.as_tuple() for a in db.scan_all_aligned(sc, q)]
    monkeypatch.delenv("HYALITE_HIP_TRACE16")
    al_, mat_, go, ge = scheme
    for k in list(range(0, len(seqs), 9)):
        assert got[k] == orc.traceback(q, seqs[k], al_, mat_, go, ge, mode), k
