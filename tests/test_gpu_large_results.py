"""Large result copies: per-target results of 8 MB or more come back to the caller's (pageable)
buffers through the pinned slots of the transfer workers instead of the runtime's staged copy.
A database of 450,000 short DNA targets puts scan_all (scores and both ends, 9 MB) and
scan_all_aligned (22 MB of records and the local alignments' ops) on that path, and the same
targets in global mode (NW) scan_all_aligned with 22 MB of records and about 80 MB of ops, every
worker refilling its slots; their results must equal the small-result paths and the oracle. The
ops come back packed in database order: target k's start where target k - 1's end.
"""
import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from common import MODE_NW, MODE_SW, dna_scheme

pytestmark = pytest.mark.gpu


def big_dna_database(rng, n, lo, hi, q):
    seqs = []
    for k in range(n):
        t = rng.integers(0, 4, size=int(rng.integers(lo, hi + 1)), dtype=np.uint8)
        if k % 97 == 0:                       # some targets hold most of the query
            at = int(rng.integers(0, max(1, t.size - 60)))
            t[at:at + 60] = q[10:70][: t.size - at]
        seqs.append(t)
    return seqs


def assert_packed(recs, ops):
    """Target k's ops start where target k - 1's end, from offset 0, inside the buffer."""
    n_ops = recs["n_ops"].astype(np.uint64)
    starts = np.concatenate([[0], np.cumsum(n_ops)[:-1]]).astype(np.uint64)
    assert np.array_equal(recs["ops_offset"], starts)
    assert int(n_ops.sum()) <= ops.size


def test_large_per_target_results_through_the_pinned_slots():
    rng = np.random.default_rng(1300)
    q = rng.integers(0, 4, size=100, dtype=np.uint8)
    seqs = big_dna_database(rng, 450_000, 60, 140, q)
    al, mat, go, ge = dna_scheme()
    db = (hy.Database.builder().sequences(seqs).scoring(hy.Scoring(al, mat, go, ge))
          .mode(hy.Mode.Sw).search_type(hy.SearchType.Alignment((1 << 64) - 1))
          .max_query_len(100).backend(hy.BackendChoice.force(hy.Backend.Hip)).build())
    sc = hy.Scratch(db)

    scores, qe, te = db.scan_all_arrays(sc, q)          # 9 MB: through the slots
    assert scores.nbytes + qe.nbytes + te.nbytes >= 8 << 20
    plain = db.scan_scores_array(sc, q)                 # 1.8 MB: a plain copy
    assert np.array_equal(scores, plain)

    for _ in range(2):                                  # the second call reuses the slots
        recs, ops = db.scan_all_aligned_arrays(sc, q)
        assert recs.nbytes >= 8 << 20
        assert_packed(recs, ops)
        assert np.array_equal(recs["score"], scores)
        for k in list(range(0, len(seqs), 9973)) + [len(seqs) - 1]:
            r = recs[k]
            got = (int(r["score"]), int(r["query_start"]), int(r["query_end"]),
                   int(r["target_start"]), int(r["target_end"]),
                   ops[int(r["ops_offset"]):int(r["ops_offset"]) + int(r["n_ops"])].tolist())
            want = orc.traceback(q, seqs[k], al, mat, go, ge, MODE_SW)
            assert got[:5] == tuple(want[:5]) and got[5] == list(want[5]), k
    for k in range(0, len(seqs), 4999):
        s, wq, wt = orc.align_core(q, seqs[k], al, mat, go, ge, MODE_SW)
        assert (scores[k], qe[k], te[k]) == (s, -1 if wq is None else wq,
                                               -1 if wt is None else wt), k

    # Global alignments span the whole target: their ops fill most of each op region.
    del db, sc
    db = (hy.Database.builder().sequences(seqs).scoring(hy.Scoring(al, mat, go, ge))
          .mode(hy.Mode.Nw).search_type(hy.SearchType.Alignment((1 << 64) - 1))
          .max_query_len(100).backend(hy.BackendChoice.force(hy.Backend.Hip)).build())
    sc = hy.Scratch(db)
    for _ in range(2):
        recs, ops = db.scan_all_aligned_arrays(sc, q)
        assert recs.nbytes + int(recs["n_ops"].sum()) >= 64 << 20
        assert_packed(recs, ops)
        for k in list(range(0, len(seqs), 9973)) + [len(seqs) - 1]:
            r = recs[k]
            got = (int(r["score"]), int(r["query_start"]), int(r["query_end"]),
                   int(r["target_start"]), int(r["target_end"]),
                   ops[int(r["ops_offset"]):int(r["ops_offset"]) + int(r["n_ops"])].tolist())
            want = orc.traceback(q, seqs[k], al, mat, go, ge, MODE_NW)
            assert got[:5] == tuple(want[:5]) and got[5] == list(want[5]), k
