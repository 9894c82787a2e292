"""The batched and serving configurations at full size, checked against the oracle.

* Batched: 256 protein queries (log-uniform 20-2000 aa) against the 1,000,000-target protein
  database built with max_query_len 2000, in one hy_scan_batch call (the persistent queue, query
  pairs per lane, the graph-captured submission, the winner pass over queries past one pass).
* Serving: the 2,000,000-target protein database at max_query_len 3000 (the targets of 2,979
  residues or more prove i32), queries with log-normal lengths (median 350, 20-3000) served under
  Poisson arrivals, one query planted inside an i32 target so that an i32 winner occurs.

The oracle cannot scan 4e13 cells, so each hit is pinned by what it must satisfy (the reference's
contract is bit-identity to the scalar kernel, DETERMINISM.md:1-16):
  * score and ends == the oracle's align_core of the query against the hit's target
    (src/kernel.rs:267-277, the role align_core plays for the winner in src/database.rs:183-197);
  * hit == the one-query scan of that query (Database::scan, src/database.rs:83);
  * for queries spread over the length range: scan_scores on a seeded 1,000-target sample ==
    the oracle, and scan == the smallest-index argmax of scan_scores over every target
    (tests/properties.rs:506-571).
"""
import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from hyalite_amd import loadgen
from hyalite_amd import workloads as wl

pytestmark = pytest.mark.gpu


def build(w):
    return (hy.Database.builder().sequences_flat(w.flat, w.off)
            .scoring(hy.Scoring(w.al, w.matrix, w.go, w.ge)).mode(hy.Mode(w.mode))
            .search_type(hy.SearchType(w.search_type)).max_query_len(w.mql)
            .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())


def target(w, k):
    return w.flat[int(w.off[k]):int(w.off[k + 1])]


def check_winner(w, q, hit):
    s, qe, te = orc.align_core(q, target(w, hit.db_index), w.al, w.matrix, w.go, w.ge, w.mode)
    assert (hit.score, hit.query_end, hit.target_end) == (s, qe, te), (len(q), hit)


def check_scores(w, db, sc, q, seed):
    """scan_scores over every target: a seeded sample equals the oracle, and scan is its
    smallest-index argmax."""
    scores = db.scan_scores_array(sc, q)
    idx = np.sort(np.random.default_rng(seed).choice(w.n, size=1000, replace=False))
    flat, off = orc.flatten([target(w, int(i)) for i in idx])
    want = orc.scan_all_flat(flat, off, q, w.al, w.matrix, w.go, w.ge, w.mode, ends=False,
                             threads=8)
    assert np.array_equal(scores[idx], want), len(q)
    hit = db.scan(sc, q)
    best = int(scores.max())
    assert (hit.score, hit.db_index) == (best, int(np.flatnonzero(scores == best)[0]))
    return hit


def spread(lengths, k):
    """Indices of k queries spread evenly over the sorted lengths."""
    order = np.argsort(lengths, kind="stable")
    return [int(order[int(round(p))]) for p in np.linspace(0, len(order) - 1, k)]


def test_batched_256_queries_x_1M_protein_targets():
    w = wl.c4_batch()
    assert w.n == 1_000_000 and w.mql == 2000 and len(w.queries) == 256
    db = build(w)
    sc = hy.Scratch(db)
    hits = db.scan_batch(sc, w.queries)
    assert len(hits) == 256
    for q, h in zip(w.queries, hits):
        check_winner(w, q, h)
        assert db.scan(sc, q) == h, len(q)
    lens = np.array([len(q) for q in w.queries])
    assert lens.min() < 30 and lens.max() > 1700  # the sample spans the range
    for k, i in enumerate(spread(lens, 8)):
        assert check_scores(w, db, sc, w.queries[i], seed=40 + k) == hits[i]


def test_served_queries_x_2M_protein_targets_with_i32_winner():
    w = wl.c5_serving(nq=48)
    assert w.n == 2_000_000 and w.mql == 3000
    lens = w.lens()
    # Plant the longest query that fits inside an i32 target (length >= 2979 here), so that
    # the best hit of that query is a 32-bit target.
    qlens = np.array([len(q) for q in w.queries])
    i32_targets = np.flatnonzero(lens >= 2979)
    assert i32_targets.size > 100
    host = int(i32_targets[np.argmax(lens[i32_targets])])
    pq = int(np.argmax(np.where(qlens <= lens[host] - 100, qlens, 0)))
    q = w.queries[pq]
    at = int(w.off[host]) + 50
    w.flat[at:at + len(q)] = q
    db = build(w)
    parts = db.partition_sizes()
    assert parts[hy.ScoreWidth.I32] == i32_targets.size and parts[hy.ScoreWidth.I16] > 0
    with hy.Server(db, max_lanes=3, max_batch=16) as srv:
        res = loadgen.run_poisson(srv, w.queries, qps=40.0, seed=5)
    served = [s.hit for s in res.served]
    assert served[pq].db_index == host
    sc = hy.Scratch(db)
    for q, h in zip(w.queries, served):
        check_winner(w, q, h)
        assert db.scan(sc, q) == h, len(q)
    assert max(s.batch_size for s in res.served) > 1  # the load made batches
    for k, i in enumerate(spread(qlens, 6) + [pq]):
        assert check_scores(w, db, sc, w.queries[i], seed=50 + k) == served[i]
