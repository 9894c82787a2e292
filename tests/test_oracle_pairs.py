"""Pins the oracle's pairwise extras before the GPU pair kernels are checked against them:
align_pair_position_max (src/kernel.rs:474-575) and score2 (src/kernel.rs:741-792).

Checks come from the reference's own tests (hand vectors of src/kernel.rs:1028-1059 and
1245-1276), from enumeration (column maxima as the best local alignment ending in each column,
by brute force over start windows), from the relation to the pinned scan oracle (best score and
target end equal align_core in SW mode), and, for score2, from the reference's structurally
different transcription (build the explicit peak list, then filter; src/kernel.rs:1207-1243).
"""
import random

import pytest

import oracle as orc
from common import all_sequences, dna_scheme, identity_matrix

MODE_SW = 0


def pm(q, t, scheme=None):
    al, mat, go, ge = scheme or dna_scheme()
    return orc.position_max(q, t, al, mat, go, ge)


# ---- position_max: src/kernel.rs:1028-1059 ---------------------------------------------------

def test_position_max_hand_computed_cases():
    assert pm([0, 1, 2, 3], [0, 1, 2, 3]) == ([2, 4, 6, 8], 8, 3)
    assert pm([0], [1, 1, 0]) == ([0, 0, 2], 2, 2)
    assert pm([0, 0], [1, 1]) == ([0, 0], 0, None)
    assert pm([0, 1], []) == ([], 0, None)
    assert pm([], [0, 1, 2]) == ([0, 0, 0], 0, None)


SCHEMES = [
    (2, identity_matrix(2, 2, -1), 2, 1),
    (2, identity_matrix(2, 1, -1), 3, 3),
    (3, identity_matrix(3, 3, -2), 4, 2),
    (3, identity_matrix(3, 2, -1), 1, 0),
]


@pytest.mark.parametrize("scheme", SCHEMES)
def test_position_max_equals_enumeration_over_short_pairs(scheme):
    al, mat, go, ge = scheme
    seqs = all_sequences(al, 3)
    for q in seqs:
        for t in seqs:
            col, s, te = orc.position_max(q, t, al, mat, go, ge)
            assert col == orc.brute_colmax(q, t, al, mat, go, ge), (q, t)
            assert all(v >= 0 for v in col) and s == max(col + [0])


def test_position_max_agrees_with_the_scan_oracle_on_random_schemes():
    """src/kernel.rs:1076-1102: score and target end equal align_pair(SW, ScoreEnd); the array
    maximum is the score. Asymmetric random matrices catch a transposed lookup."""
    rng = random.Random(11)
    for _ in range(400):
        al = rng.randint(2, 5)
        mat = [rng.randint(-6, 6) for _ in range(al * al)]
        go = rng.randint(0, 7)
        ge = rng.randint(0, go)
        q = [rng.randrange(al) for _ in range(rng.randrange(0, 14))]
        t = [rng.randrange(al) for _ in range(rng.randrange(0, 14))]
        col, s, te = orc.position_max(q, t, al, mat, go, ge)
        assert len(col) == len(t)
        ws, _, wte = orc.align_core(q, t, al, mat, go, ge, MODE_SW)
        assert (s, te) == (ws, wte), (q, t, al, mat, go, ge)
        if len(q) <= 5 and len(t) <= 5:
            assert col == orc.brute_colmax(q, t, al, mat, go, ge)


# ---- score2: src/kernel.rs:1245-1276 and the peak-list transcription of 1207-1243 --------------

def score2_peak_list(colmax, best_score, te, matrix_max, min_score):
    if matrix_max <= 0:
        return None
    peaks = []
    for i, v in enumerate(colmax):
        if v < min_score:
            continue
        if peaks and peaks[-1][1] + 1 == i:
            if peaks[-1][0] < v:
                peaks[-1] = (v, i)
        else:
            peaks.append((v, i))
    w = max((max(best_score, 0) + matrix_max - 1) // matrix_max, 0)
    lo, hi = te - w, te + w
    best2 = None
    for score, pos in peaks:
        if (pos < lo or pos > hi) and (best2 is None or score > best2[0]):
            best2 = (score, pos)
    return best2


def test_score2_hand_computed_cases():
    assert orc.score2([0, 5, 0, 0, 0, 0, 0, 0, 4, 0], 5, 1, 1, 1) == (4, 8)
    assert orc.score2([0, 3, 5, 4, 0], 5, 2, 5, 1) is None
    assert orc.score2([3, 5, 4, 6], 1, 20, 1, 1) == (6, 3)
    assert orc.score2([7, 0, 7], 1, 10, 1, 1) == (7, 0)
    assert orc.score2([0, 0, 0], 0, 0, 1, 1) is None
    assert orc.score2([5, 5], 5, 0, 0, 1) is None
    colmax = [2, 0, 0, 0, 0, 0, 9]
    assert orc.score2(colmax, 9, 6, 9, 5) is None
    assert orc.score2(colmax, 9, 6, 9, 1) == (2, 0)


def test_score2_matches_the_peak_list_transcription():
    rng = random.Random(5)
    for _ in range(3000):
        colmax = [rng.randint(0, 20) for _ in range(rng.randint(0, 40))]
        args = (rng.randint(0, 40), rng.randint(0, 40), rng.randint(-2, 8), rng.randint(1, 6))
        assert orc.score2(colmax, *args) == score2_peak_list(colmax, *args), (colmax, args)


# ---- the striped CPU baseline of the single-pair path --------------------------------------------

def test_sse_striped_baseline_matches_the_scalar_oracle():
    """The SSE4.1 striped (Farrar) restatement used as the pairs CPU baseline: every score equals
    the scalar oracle's, including the lazy-F carries across stripes (queries of 1-80 rows over 8
    or 16 lanes), empty sides, and pairs at both widths the proof selects (short ones at 8 bits,
    the rest at 16)."""
    import numpy as np
    rng = np.random.default_rng(2)
    for al, m_hi, go, ge in ((4, 80, 2, 1), (4, 80, 5, 1), (20, 60, 11, 1), (3, 40, 1, 1),
                             (4, 40, 3, 0)):
        mat = rng.integers(-4, 6, size=al * al).astype(np.int32)
        lq = rng.integers(0, m_hi, size=300)
        lt = rng.integers(0, 120, size=300)
        qf = rng.integers(0, al, size=int(lq.sum()), dtype=np.uint8)
        tf = rng.integers(0, al, size=int(lt.sum()), dtype=np.uint8)
        qo = np.concatenate([[0], np.cumsum(lq)]).astype(np.uint64)
        to = np.concatenate([[0], np.cumsum(lt)]).astype(np.uint64)
        got, _ = orc.sse_striped_sw_pairs(qf, qo, tf, to, al, mat, go, ge)
        for k in range(300):
            q, t = qf[qo[k]:qo[k + 1]], tf[to[k]:to[k + 1]]
            assert got[k] == orc.align_core(q, t, al, mat, go, ge, MODE_SW)[0], (al, k)


def test_sse_striped_baseline_widths_and_limits():
    """The reference's pairwise bench shapes (DNA +2/-1, gaps 2/1): 30 x 30 proves 8 bits, 150 x
    150 and 150 x 1,400 16 bits; scores equal the scalar oracle at each. A pair needing 32 bits
    is refused rather than computed at a narrower width."""
    import numpy as np
    rng = np.random.default_rng(9)
    mat = np.array([2 if i == j else -1 for i in range(4) for j in range(4)], dtype=np.int32)
    for m, n in ((30, 30), (63, 63), (64, 200), (150, 150), (150, 1400)):
        qs = [rng.integers(0, 4, size=m, dtype=np.uint8) for _ in range(20)]
        ts = [rng.integers(0, 4, size=n, dtype=np.uint8) for _ in range(20)]
        qs[0] = ts[0][:m].copy() if n >= m else qs[0]        # a high-scoring pair
        qo = np.arange(21, dtype=np.uint64) * m
        to = np.arange(21, dtype=np.uint64) * n
        got, _ = orc.sse_striped_sw_pairs(np.concatenate(qs), qo, np.concatenate(ts), to, 4, mat,
                                          2, 1)
        for k in range(20):
            assert got[k] == orc.align_core(qs[k], ts[k], 4, mat, 2, 1, MODE_SW)[0], (m, n, k)
    big = np.array([20000 if i == j else -1 for i in range(4) for j in range(4)], dtype=np.int32)
    with pytest.raises(ValueError):
        orc.sse_striped_sw_pairs(np.zeros(5, np.uint8), np.array([0, 5], np.uint64),
                                 np.zeros(5, np.uint8), np.array([0, 5], np.uint64), 4, big, 2, 1)
