"""The benchmark configurations at full size on the GPU.

Small configurations are compared with the oracle target by target. At full size the oracle would
take minutes, so the checks are size-independent properties plus oracle samples:
  * scan == the smallest-index argmax of scan_scores (tests/properties.rs:506-571),
  * scan_all scores == scan_scores, and the ScoreEnd winner's ends == the oracle's align_core of
    that pair,
  * a seeded random sample of targets scored by the oracle equals the GPU's scores,
  * a mutated copy of the query (the DNA configuration plants 5% of them) is among the best hits.
"""
import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from hyalite_amd import workloads as wl

pytestmark = pytest.mark.gpu


def build(w, mode=None, st=None):
    return (hy.Database.builder().sequences_flat(w.flat, w.off)
            .scoring(hy.Scoring(w.al, w.matrix, w.go, w.ge))
            .mode(hy.Mode(w.mode if mode is None else mode))
            .search_type(hy.SearchType(w.search_type if st is None else st))
            .max_query_len(w.mql).backend(hy.BackendChoice.force(hy.Backend.Hip)).build())


def sample_check(w, q, scores, mode, k=1500, seed=0):
    sub_idx = np.sort(np.random.default_rng(seed).choice(w.n, size=min(k, w.n), replace=False))
    seqs = [w.flat[int(w.off[i]):int(w.off[i + 1])] for i in sub_idx]
    flat, off = orc.flatten(seqs)
    want = orc.scan_all_flat(flat, off, q, w.al, w.matrix, w.go, w.ge, mode, ends=False,
                             threads=8)
    assert np.array_equal(scores[sub_idx], want)


def winner_check(w, q, hit, mode):
    t = w.flat[int(w.off[hit.db_index]):int(w.off[hit.db_index + 1])]
    s, qe, te = orc.align_core(q, t, w.al, w.matrix, w.go, w.ge, mode)
    assert (hit.score, hit.query_end, hit.target_end) == (s, qe, te)


def test_dna_sw_100x10k_every_target():
    """Plain DNA SW ScoreEnd, 100-nt query x 10,000 targets U{50..300} (i8 and i16 widths)."""
    w = wl.c1_dna()
    db = build(w)
    assert db.partition_sizes()[hy.ScoreWidth.I8] > 0
    sc = hy.Scratch(db)
    q = w.queries[0]
    s, qe, te = db.scan_all_arrays(sc, q)
    ws, wqe, wte = orc.scan_all_flat(w.flat, w.off, q, w.al, w.matrix, w.go, w.ge, w.mode,
                                     threads=8)
    assert np.array_equal(s, ws) and np.array_equal(qe, wqe) and np.array_equal(te, wte)
    hit = db.scan(sc, q)
    want = orc.scan_flat(w.flat, w.off, q, w.al, w.matrix, w.go, w.ge, w.mode, True, threads=8)
    assert (hit.score, hit.db_index, hit.query_end, hit.target_end) == want


@pytest.fixture(scope="module")
def dna_modes():
    return wl.c3_dna_modes()


@pytest.mark.parametrize("mode", [1, 2, 4, 3])
def test_dna_modes_150x2M_scan_scores(dna_modes, mode):
    """Mode sweep: 150-nt query x 2,000,000 targets U{100..1000}, 5% mutated query copies."""
    w = dna_modes
    q = w.queries[0]
    db = build(w, mode=mode, st=hy.SearchType.ScoreEnd)
    sc = hy.Scratch(db)
    scores = db.scan_scores_array(sc, q)
    assert scores.shape == (w.n,)
    hit = db.scan(sc, q)
    best = int(scores.max())
    assert hit.score == best and hit.db_index == int(np.flatnonzero(scores == best)[0])
    winner_check(w, q, hit, mode)
    sample_check(w, q, scores, mode, seed=mode)
    # Ties are common here (many mutated copies): scan_all agrees with scan_scores everywhere.
    s_all, _, _ = db.scan_all_arrays(sc, q)
    assert np.array_equal(s_all, scores)


def test_protein_sw_360x1M_properties():
    """Protein SW ScoreEnd, 360-aa query x 1,000,000 log-normal targets."""
    w = wl.c2_protein()
    q = w.queries[0]
    db = build(w)
    sc = hy.Scratch(db)
    scores = db.scan_scores_array(sc, q)
    hit = db.scan(sc, q)
    best = int(scores.max())
    assert hit.score == best and hit.db_index == int(np.flatnonzero(scores == best)[0])
    winner_check(w, q, hit, w.mode)
    sample_check(w, q, scores, w.mode, k=1000, seed=2)
    # The batched form on a few queries equals the one-query scans.
    qs = [x[:w.mql] for x in wl.c4_batch_queries(nq=6)] + [q]
    sc.set_tuning(hy.Tuning(batch_chunk=4))
    assert db.scan_batch(sc, qs) == [db.scan(sc, x) for x in qs]


@pytest.mark.parametrize("mode", [1, 2, 4])
def test_dna_modes_scaled_150x2M_two_widths(dna_modes, mode):
    """The mode-sweep database with the scheme scaled by 25: targets up to 504 nt prove i16
    (packed kernel), longer ones i32 (32-bit kernel beside it); scan_scores, scan and scan_all
    with ends over all 2,000,000 targets."""
    w = wl.c3_scaled(dna_modes, mode)
    q = w.queries[0]
    db = build(w)
    parts = db.partition_sizes()
    lens = w.lens()
    assert parts[hy.ScoreWidth.I32] == int((lens > 504).sum()) > 500_000
    assert parts[hy.ScoreWidth.I16] == int((lens <= 504).sum())
    sc = hy.Scratch(db)
    scores = db.scan_scores_array(sc, q)
    hit = db.scan(sc, q)
    best = int(scores.max())
    assert hit.score == best and hit.db_index == int(np.flatnonzero(scores == best)[0])
    winner_check(w, q, hit, mode)
    sample_check(w, q, scores, mode, seed=10 + mode)
    s_all, qe, te = db.scan_all_arrays(sc, q)
    assert np.array_equal(s_all, scores)
    # Ends of a sample against the oracle's align_core, both widths represented.
    rng = np.random.default_rng(20 + mode)
    for k in list(rng.choice(np.flatnonzero(lens <= 504), 40)) + \
            list(rng.choice(np.flatnonzero(lens > 504), 40)):
        t = w.flat[int(w.off[k]):int(w.off[k + 1])]
        s, e1, e2 = orc.align_core(q, t, w.al, w.matrix, w.go, w.ge, mode)
        assert (int(s_all[k]), int(qe[k]), int(te[k])) == (s, -1 if e1 is None else e1,
                                                            -1 if e2 is None else e2)
