"""GPU: a database saved and loaded back gives the same hits as the original, and the library's
writer produces exactly the bytes of the independent writer (tests/image_format.py)."""
import numpy as np
import pytest

import hyalite_amd as hy
import oracle as orc
from common import BLOSUM62, dna_scheme
from image_format import image_bytes

pytestmark = pytest.mark.gpu


def build(seqs, scheme, mode, st, mql):
    al, mat, go, ge = scheme
    return (hy.Database.builder().sequences(seqs).scoring(hy.Scoring(al, mat, go, ge))
            .mode(hy.Mode(mode)).search_type(st).max_query_len(mql)
            .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())


def test_saved_bytes_equal_the_independent_writer(tmp_path):
    seqs = [[0, 1, 2, 3], [], [3, 3, 1], [2] * 9]
    al, mat, go, ge = dna_scheme()
    db = build(seqs, dna_scheme(), 0, hy.SearchType.ScoreEnd, 12)
    p = tmp_path / "a.img"
    db.save(str(p))
    assert p.read_bytes() == image_bytes(seqs, al, mat, go, ge, 0, 1, 0, 12)
    adb = build(seqs, dna_scheme(), 1, hy.SearchType.Alignment(1 << 40), 12)
    adb.save(str(p))
    assert p.read_bytes() == image_bytes(seqs, al, mat, go, ge, 1, 2, 1 << 40, 12)


@pytest.mark.parametrize("mode", [0, 1, 3])
def test_loaded_database_scans_like_the_original(tmp_path, mode):
    rng = np.random.default_rng(30 + mode)
    lens = np.concatenate([rng.integers(0, 600, size=3000), [5000, 0, 1]])
    seqs = [rng.integers(0, 20, size=int(L), dtype=np.uint8) for L in lens]
    scheme = (20, BLOSUM62, 11, 1)
    db = build(seqs, scheme, mode, hy.SearchType.ScoreEnd, 400)
    p = tmp_path / "p.img"
    db.save(str(p))
    ld = hy.Database.load(str(p), hy.BackendChoice.force(hy.Backend.Hip))
    assert (ld.mode(), ld.search_type(), ld.sequence_count(), ld.max_query_len(),
            ld.max_target_len(), ld.score_width()) == (db.mode(), db.search_type(),
                                                       db.sequence_count(), db.max_query_len(),
                                                       db.max_target_len(), db.score_width())
    assert np.array_equal(ld.scoring().matrix(), db.scoring().matrix())
    assert ld.partition_sizes() == db.partition_sizes()
    s1, s2 = hy.Scratch(db), hy.Scratch(ld)
    flat, off = orc.flatten(seqs)
    for qlen in (1, 90, 400):
        q = rng.integers(0, 20, size=qlen, dtype=np.uint8)
        a = db.scan(s1, q)
        assert ld.scan(s2, q) == a
        want = orc.scan_flat(flat, off, q, 20, BLOSUM62, 11, 1, mode, True, threads=4)
        assert (a.score, a.db_index, a.query_end, a.target_end) == tuple(want)
        for x, y in zip(db.scan_all_arrays(s1, q), ld.scan_all_arrays(s2, q)):
            assert np.array_equal(x, y)
        assert np.array_equal(db.scan_scores_array(s1, q), ld.scan_scores_array(s2, q))
    qs = [rng.integers(0, 20, size=int(L), dtype=np.uint8) for L in (30, 200, 399)]
    assert ld.scan_batch(s2, qs) == db.scan_batch(s1, qs)


def test_loaded_alignment_database_traces_back(tmp_path):
    rng = np.random.default_rng(4)
    seqs = [rng.integers(0, 4, size=int(L), dtype=np.uint8) for L in rng.integers(0, 300, 500)]
    db = build(seqs, dna_scheme(), 0, hy.SearchType.Alignment(1 << 40), 150)
    p = tmp_path / "t.img"
    db.save(str(p))
    ld = hy.Database.load(str(p))
    assert ld.search_type().max_bytes() == 1 << 40
    q = rng.integers(0, 4, size=150, dtype=np.uint8)
    a, b = db.scan_aligned(hy.Scratch(db), q), ld.scan_aligned(hy.Scratch(ld), q)
    assert a.db_index == b.db_index and a.alignment == b.alignment
    t = seqs[a.db_index]
    assert a.alignment.as_tuple() == orc.traceback(q, t, *dna_scheme(), 0)


def test_alphabet_above_256_round_trips(tmp_path):
    """Scoring::new takes alphabets above 256 (symbols past 255 never occur in a byte
    sequence); such a database saves, loads back and scans like the original."""
    al = 300
    mat = [-2] * (al * al)
    for i in range(al):
        mat[i * al + i] = 3 + (i % 4)
    rng = np.random.default_rng(303)
    seqs = [rng.integers(0, 256, size=int(L), dtype=np.uint8) for L in rng.integers(0, 90, 200)]
    db = build(seqs, (al, mat, 5, 2), 0, hy.SearchType.ScoreEnd, 60)
    p = tmp_path / "big.img"
    db.save(str(p))
    ld = hy.Database.load(str(p), hy.BackendChoice.force(hy.Backend.Hip))
    assert ld.scoring().alphabet_len() == al
    assert np.array_equal(ld.scoring().matrix(), db.scoring().matrix())
    flat, off = orc.flatten(seqs)
    s1, s2 = hy.Scratch(db), hy.Scratch(ld)
    for qlen in (0, 7, 60):
        q = rng.integers(0, 256, size=qlen, dtype=np.uint8)
        want = orc.scan_flat(flat, off, q, al, mat, 5, 2, 0, True, threads=4)
        for d, s in ((db, s1), (ld, s2)):
            h = d.scan(s, q)
            assert (h.score, h.db_index, h.query_end, h.target_end) == tuple(want)
