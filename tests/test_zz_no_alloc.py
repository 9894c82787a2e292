"""hy_scan_scores_static, the allocation-free form of hy_scan_scores: the same scores as
hy_scan_scores and the oracle, over two calls in a row (the second with a smaller query) and a
smaller database after them, with every kernel family it uses (packed groups, 32-bit groups,
long targets in the systolic kernel, queries past one systolic pass); a problem beyond its static
array is refused and leaves the next call unaffected."""
import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from common import ALL_MODES, identity_matrix

pytestmark = pytest.mark.gpu

# Entries large enough that the longer targets fail the packed width proof at max_query_len 600,
# so one database holds packed groups, 32-bit groups and (beyond 1,024 residues) long targets.
SCHEME = (5, identity_matrix(5, 60, -30), 40, 6)


def build(seqs, mode, mql, scheme=SCHEME):
    al, mat, go, ge = scheme
    return (hy.Database.builder().sequences(seqs).scoring(hy.Scoring(al, mat, go, ge))
            .mode(hy.Mode(mode)).search_type(hy.SearchType.Score).max_query_len(mql)
            .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())


def oracle_scores(seqs, q, mode, scheme=SCHEME):
    al, mat, go, ge = scheme
    flat, off = orc.flatten(seqs)
    return orc.scan_all_flat(flat, off, q, al, mat, go, ge, mode, ends=False, threads=8)


@pytest.mark.parametrize("mode", ALL_MODES)
def test_static_scan_scores_equals_scan_scores(mode):
    rng = np.random.default_rng(900 + mode)
    seqs = [rng.integers(0, 5, size=int(n), dtype=np.uint8) for n in rng.integers(0, 400, 1500)]
    seqs += [rng.integers(0, 5, size=int(n), dtype=np.uint8) for n in (1100, 1300, 1500)]
    db = build(seqs, mode, 600)
    parts = db.partition_sizes()
    assert parts[hy.ScoreWidth.I32] > 0 and parts[hy.ScoreWidth.I16] > 0
    sc = hy.Scratch(db)
    for qlen in (600, 40):  # past one systolic pass, then a smaller query
        q = rng.integers(0, 5, size=qlen, dtype=np.uint8)
        got = db.scan_scores_static(q)
        assert np.array_equal(got, db.scan_scores_array(sc, q)), qlen
        assert np.array_equal(got, oracle_scores(seqs, q, mode)), qlen
    # A smaller database after the larger one: the array keeps the last call's contents.
    small = seqs[:70]
    sdb = build(small, mode, 600)
    q = rng.integers(0, 5, size=33, dtype=np.uint8)
    got = sdb.scan_scores_static(q)
    assert got.shape == (70,) and np.array_equal(got, oracle_scores(small, q, mode))


def test_static_limit_is_refused():
    rng = np.random.default_rng(990)
    lens = rng.integers(150, 250, size=100_000)
    off = np.zeros(lens.size + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    flat = rng.integers(0, 4, size=int(off[-1]), dtype=np.uint8)
    scheme = (4, identity_matrix(4, 2, -1), 2, 1)
    big = (hy.Database.builder().sequences_flat(flat, off)
           .scoring(hy.Scoring(*scheme[:2], *scheme[2:])).mode(hy.Mode.Sw)
           .max_query_len(100).backend(hy.BackendChoice.force(hy.Backend.Hip)).build())
    q = rng.integers(0, 4, size=100, dtype=np.uint8)
    with pytest.raises(hy.StaticScratchLimit):
        big.scan_scores_static(q)
    seqs = [flat[int(off[k]):int(off[k + 1])] for k in range(300)]
    db = build(seqs, 0, 100, scheme)
    assert np.array_equal(db.scan_scores_static(q), oracle_scores(seqs, q, 0, scheme))
