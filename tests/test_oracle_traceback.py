"""Pins the oracle's traceback (src/align.rs:169-420) and its budget arithmetic (align.rs:228-268)
before the GPU traceback is checked against it.

Three independent kinds of check:
  * the reference's own hand vectors (src/align/tests.rs:137-342, tests/traceback.rs:87-104,
    tests/database_alignment.rs:84-156);
  * the reference's consistency properties (align/tests.rs:377-412): the ops re-score to the
    reported score with per-run affine penalties, Match/Mismatch agree with symbol equality, and
    the op counts fill the reported spans;
  * the canonical choice among co-optimal alignments, fixed by enumerating the paths of the affine
    state machine (H, horizontal gap, vertical gap) from the mode's answer region and taking the
    best score, then the smallest (column, row) end, then the first path under the walk's
    preference order: stop (local) < diagonal < target-consuming gap (Del) < query-consuming gap
    (Ins), and close a gap before extending it. The enumeration scores each path by its own moves
    and never touches the DP matrices, so a swapped preference, a wrong border or a wrong close
    test in the oracle fails it.
"""
import functools
import random

import pytest

import oracle as orc
from common import (ALL_MODES, MODE_HW, MODE_NW, MODE_OV, MODE_SHW, MODE_SW, all_sequences,
                    dna_scheme, identity_matrix)

M, X, I, D = orc.OP_MATCH, orc.OP_MISMATCH, orc.OP_INS, orc.OP_DEL

# Flags::for_mode (src/kernel.rs:57-97): top_row_free, left_col_free, answer_last_row,
# answer_last_col, local.
FLAGS = {
    MODE_NW: (False, False, False, False, False),
    MODE_HW: (True, False, True, False, False),
    MODE_OV: (True, True, True, True, False),
    MODE_SW: (True, True, False, False, True),
    MODE_SHW: (False, True, False, True, False),
}


def scorings():
    """src/align/tests.rs:29-36."""
    return [
        dna_scheme(),
        (4, identity_matrix(4, 1, -3), 10, 5),
        (4, identity_matrix(4, 3, -2), 4, 0),
        (4, identity_matrix(4, 5, -1), 1, 1),
    ]


def tb(q, t, mode, scheme=None):
    al, mat, go, ge = scheme or dna_scheme()
    return orc.traceback(q, t, al, mat, go, ge, mode)


def gap(go, ge, n):
    return 0 if n == 0 else go + (n - 1) * ge


def rescore(res, q, t, al, mat, go, ge):
    """Per-maximal-run affine re-scoring (src/align/tests.rs:38-75)."""
    score, qs, _, ts, _, ops = res
    qi, tj, total, k = qs, ts, 0, 0
    while k < len(ops):
        if ops[k] in (M, X):
            total += mat[q[qi] * al + t[tj]]
            qi, tj, k = qi + 1, tj + 1, k + 1
        else:
            kind, run = ops[k], 0
            while k < len(ops) and ops[k] == kind:
                run, k = run + 1, k + 1
                if kind == I:
                    qi += 1
                else:
                    tj += 1
            total -= gap(go, ge, run)
    return total


def assert_consistent(res, q, t, al, mat, go, ge):
    """src/align/tests.rs:77-127 and the property of 380-410."""
    score, qs, qe, ts, te, ops = res
    assert rescore(res, q, t, al, mat, go, ge) == score
    assert qe <= len(q) and te <= len(t) and qs <= qe and ts <= te
    qi, tj = qs, ts
    for op in ops:
        if op in (M, X):
            assert (q[qi] == t[tj]) == (op == M)
            qi, tj = qi + 1, tj + 1
        elif op == I:
            qi += 1
        else:
            tj += 1
    assert (qi, tj) == (qe, te)


def canonical_by_enumeration(q, t, al, mat, go, ge, mode):
    """The canonical alignment from the affine state machine, by exhaustive search over choices.

    best(state, i, j) is the best-scoring way to finish a backward path that stands at cell (i, j)
    in `state`, as (score, choices, ops_rev, start). Choices are compared lexicographically after
    the score, so among co-optimal completions the earliest preferred move wins; ties on the
    first move are impossible (distinct moves), so the lexicographic minimum over whole paths is
    the preferred-first choice at each step among moves that can still reach the optimum.
    """
    top_free, left_free, last_row, last_col, local = FLAGS[mode]
    m, n = len(q), len(t)

    def better(a, b):
        if b is None:
            return True
        return a[0] > b[0] or (a[0] == b[0] and a[1] < b[1])

    @functools.lru_cache(maxsize=None)
    def best(state, i, j):
        if state == "H":
            if local and (i == 0 or j == 0):
                return (0, (0,), (), (i, j))
            if i == 0 and j == 0:
                return (0, (), (), (0, 0))
            if i == 0:
                if top_free:
                    return (0, (), (), (0, j))
                return (-gap(go, ge, j), (), (D,) * j, (0, 0))
            if j == 0:
                if left_free:
                    return (0, (), (), (i, 0))
                return (-gap(go, ge, i), (), (I,) * i, (0, 0))
            cands = []
            if local:
                cands.append((0, (0,), (), (i, j)))
            s, c, o, st = best("H", i - 1, j - 1)
            op = M if q[i - 1] == t[j - 1] else X
            cands.append((s + mat[q[i - 1] * al + t[j - 1]], (1,) + c, (op,) + o, st))
            s, c, o, st = best("E", i, j)
            cands.append((s, (2,) + c, o, st))
            s, c, o, st = best("F", i, j)
            cands.append((s, (3,) + c, o, st))
        elif state == "E":  # emit Del at column j, then close (this Del opened) or extend
            cands = []
            s, c, o, st = best("H", i, j - 1)
            cands.append((s - go, (0,) + c, (D,) + o, st))
            if j >= 2:
                s, c, o, st = best("E", i, j - 1)
                cands.append((s - ge, (1,) + c, (D,) + o, st))
        else:  # F: emit Ins at row i
            cands = []
            s, c, o, st = best("H", i - 1, j)
            cands.append((s - go, (0,) + c, (I,) + o, st))
            if i >= 2:
                s, c, o, st = best("F", i - 1, j)
                cands.append((s - ge, (1,) + c, (I,) + o, st))
        out = None
        for cand in cands:
            if better(cand, out):
                out = cand
        return out

    if local:
        region = [(i, j) for i in range(m + 1) for j in range(n + 1)]
    else:
        region = {(m, n)}
        if last_col:
            region |= {(i, n) for i in range(m + 1)}
        if last_row:
            region |= {(m, j) for j in range(n + 1)}
    end = None
    for (i, j) in region:
        s = best("H", i, j)[0]
        if end is None or s > end[0] or (s == end[0] and (j, i) < (end[2], end[1])):
            end = (s, i, j)
    score, ei, ej = end
    _, _, ops_rev, (si, sj) = best("H", ei, ej)
    return score, si, ei, sj, ej, list(reversed(ops_rev))


# ---- src/align/tests.rs:137-312 hand vectors ---------------------------------------------------

def test_nw_identical_is_all_match():
    res = tb([0, 1, 2, 3], [0, 1, 2, 3], MODE_NW)
    assert res == (8, 0, 4, 0, 4, [M] * 4)
    assert orc.cigar(res[5]) == "4M" and orc.cigar(res[5], True) == "4="


def test_nw_single_substitution():
    res = tb([0, 1, 2, 3], [0, 1, 1, 3], MODE_NW)
    assert res[5] == [M, M, X, M] and res[0] == 5
    assert orc.cigar(res[5]) == "4M" and orc.cigar(res[5], True) == "2=1X1="


def test_nw_needs_one_gap():
    q, t = [0, 1, 2, 3], [0, 1, 3]
    res = tb(q, t, MODE_NW)
    assert res[0] == 4
    assert_consistent(res, q, t, *dna_scheme())
    assert orc.cigar(res[5]).count("I") == 1


def test_sw_recovers_exact_substring():
    res = tb([2, 2, 0, 1, 2, 3, 2, 2], [0, 1, 2, 3], MODE_SW)
    assert res == (8, 2, 6, 0, 4, [M] * 4)
    assert orc.cigar(res[5]) == "4M"


def test_hw_query_is_a_free_window_of_the_target():
    assert tb([1, 2], [0, 1, 2, 3], MODE_HW) == (4, 0, 2, 1, 3, [M, M])


def test_shw_target_is_a_free_window_of_the_query():
    assert tb([0, 1, 2, 3], [1, 2], MODE_SHW) == (4, 1, 3, 0, 2, [M, M])


def test_ov_prefers_smaller_target_end_on_a_tie():
    q, t = [2, 3], [2, 3, 2, 3]
    res = tb(q, t, MODE_OV)
    assert res[0] == 4 and (res[3], res[4]) == (0, 2)
    assert_consistent(res, q, t, *dna_scheme())


def test_in_alphabet_unknown_symbol_traces_back():
    mat = identity_matrix(5, 2, -1)
    mat[4 * 5 + 4] = -1
    for mode in ALL_MODES:
        q = t = [4, 0, 1]
        res = orc.traceback(q, t, 5, mat, 2, 1, mode)
        assert_consistent(res, q, t, 5, mat, 2, 1)
        if res[5]:
            assert res[5][0] == M, mode


def test_empty_inputs():
    for mode in ALL_MODES:
        res = tb([], [], mode)
        assert res == (0, 0, 0, 0, 0, [])
        assert orc.cigar(res[5]) == "" and orc.cigar(res[5], True) == ""


def test_empty_query_against_target():
    t = [0, 1, 2]
    res = tb([], t, MODE_NW)
    assert res[5] == [D] * 3 and res[0] == -gap(2, 1, 3) and orc.cigar(res[5]) == "3D"
    for mode in (MODE_SW, MODE_HW, MODE_OV):
        res = tb([], t, mode)
        assert res[0] == 0 and res[5] == [], mode


def test_single_symbol_match_and_mismatch():
    for mode in ALL_MODES:
        res = tb([2], [2], mode)
        assert res[5] == [M] and res[0] == 2, mode
        assert_consistent(tb([2], [0], mode), [2], [0], *dna_scheme())


def test_all_mismatch_local_is_empty():
    res = tb([0, 0, 0], [1, 1, 1], MODE_SW)
    assert res[0] == 0 and res[5] == []


def test_cigar_run_length_and_operators():
    ops = [M, M, X, I, I, D, M]
    assert orc.cigar(ops) == "3M2I1D1M"
    assert orc.cigar(ops, True) == "2=1X2I1D1="


def test_known_global_alignment_with_a_deletion():
    """tests/traceback.rs:87-104."""
    q, t = [0, 2, 3], [0, 1, 2, 3]
    res = tb(q, t, MODE_NW)
    assert res[0] == 4 and orc.cigar(res[5]).count("D") == 1
    assert (res[1], res[2], res[3], res[4]) == (0, 3, 0, 4)
    assert_consistent(res, q, t, *dna_scheme())


def test_database_alignment_known_answers():
    """tests/database_alignment.rs:84-156: the 8-long NW self match and the SW per-target case."""
    seq = [0, 1, 2, 3, 0, 1, 2, 3]
    res = tb(seq, seq, MODE_NW)
    assert res[0] == 16 and orc.cigar(res[5]) == "8M"
    db = [[0, 0, 0, 0], [0, 1, 2, 3], [3, 3, 3, 3]]
    outs = [tb([0, 1, 2, 3], t, MODE_SW) for t in db]
    assert outs[1][0] == 8 and orc.cigar(outs[1][5]) == "4M"
    assert max(range(3), key=lambda k: (outs[k][0], -k)) == 1


# ---- consistency and score equality over every short pair (align/tests.rs:377-412) -------------

@pytest.mark.parametrize("scheme_ix", range(4))
def test_traceback_is_consistent_and_scores_like_the_scan(scheme_ix):
    al, mat, go, ge = scorings()[scheme_ix]
    seqs = all_sequences(3, 3)
    for q in seqs:
        for t in seqs:
            for mode in ALL_MODES:
                res = orc.traceback(q, t, al, mat, go, ge, mode)
                s, qe, te = orc.align_core(q, t, al, mat, go, ge, mode)
                assert res[0] == s == orc.brute_score(mode, q, t, al, mat, go, ge), (mode, q, t)
                # The traceback's end is the scan's end (ends are inclusive, None for row/col 0).
                assert (res[2], res[4]) == (0 if qe is None else qe + 1,
                                            0 if te is None else te + 1), (mode, q, t)
                assert_consistent(res, q, t, al, mat, go, ge)


# ---- the canonical co-optimal choice ------------------------------------------------------------

@pytest.mark.parametrize("scheme_ix", range(4))
def test_canonical_path_matches_state_machine_enumeration_exhaustively(scheme_ix):
    al, mat, go, ge = scorings()[scheme_ix]
    seqs = all_sequences(2, 4)
    for q in seqs[::3]:
        for t in seqs:
            for mode in ALL_MODES:
                want = canonical_by_enumeration(q, t, al, mat, go, ge, mode)
                assert orc.traceback(q, t, al, mat, go, ge, mode) == want, (mode, q, t)


def test_canonical_path_matches_enumeration_on_random_pairs():
    rng = random.Random(20261203)
    for _ in range(300):
        al, mat, go, ge = rng.choice(scorings())
        q = [rng.randrange(4) for _ in range(rng.randrange(0, 11))]
        t = [rng.randrange(4) for _ in range(rng.randrange(0, 11))]
        for mode in ALL_MODES:
            want = canonical_by_enumeration(q, t, al, mat, go, ge, mode)
            assert orc.traceback(q, t, al, mat, go, ge, mode) == want, (mode, q, t)


def test_gap_preference_is_target_consuming_first():
    """AC vs CA with cheap gaps: 'Ins A, Match C, Del A' and 'Del C, Match A, Ins C' both score
    -1 + 5 - 1 = 3 while the two mismatches score -2. Backward from (2, 2) the walk takes the
    target-consuming gap first, so the forward ops end with the Del; an Ins-first walk would give
    [D, M, I]."""
    al, mat, go, ge = 4, identity_matrix(4, 5, -1), 1, 1
    q, t = [0, 1], [1, 0]
    res = orc.traceback(q, t, al, mat, go, ge, MODE_NW)
    assert res == (3, 0, 2, 0, 2, [I, M, D])
    assert res == canonical_by_enumeration(q, t, al, mat, go, ge, MODE_NW)


# ---- budget arithmetic (align.rs:228-268; align/tests.rs:344-362, 497-566) ----------------------

def test_budget_known_values():
    # (0, 0): one cell of three i32 matrices.
    assert orc.traceback_min_bytes_for_database(0, 0) == 12
    # 8 x 8: full = 81 cells * 12 = 972; checkpoint k = isqrt(8) = 2: (2 * 5 + 3 * 3) * 9 * 4 = 684.
    assert orc.traceback_min_bytes_for_database(8, 8) == 684
    # 400 x 400: k = 20, (2 * 21 + 3 * 21) * 401 * 4 = 168420 < full (1929612); the reference
    # reports a budget of 8 bytes as too small and 1 MiB as enough.
    assert orc.traceback_min_bytes_for_database(400, 400) == 168420
    # Degenerate edges: a long query against tiny targets is bounded by the empty-target column.
    assert orc.traceback_min_bytes_for_database(30, 1) == max(
        min(31 * 2 * 12, (2 * (30 // 5 + 1) + 3 * 6) * 2 * 4), 31 * 12, 2 * 12)
    assert orc.traceback_min_bytes_for_database(1, 30) == 2 * 31 * 12


def test_budget_is_monotonic():
    for n in (0, 1, 2, 3, 5, 8, 21, 64, 100):
        prev = 0
        for m in range(401):
            v = orc.traceback_min_bytes_for_database(m, n)
            assert v >= prev, (m, n)
            prev = v
    for m in (0, 1, 2, 3, 5, 8, 21, 64, 100):
        prev = 0
        for n in range(401):
            v = orc.traceback_min_bytes_for_database(m, n)
            assert v >= prev, (m, n)
            prev = v


def test_budget_covers_every_sub_problem():
    """The budget for a (max_m, max_n) box is at least the full-matrix bytes of any sub-problem
    that must take the full path (m == 0 or n == 0) and at least the cheaper footprint of every
    other one (align/tests.rs:531-566)."""
    def full(m, n):
        return (m + 1) * (n + 1) * 12

    def ckpt(m, n):
        k = max(1, int(m ** 0.5))
        while k * k > m:
            k -= 1
        while (k + 1) * (k + 1) <= m:
            k += 1
        k = max(k, 1)
        return (2 * (m // k + 1) + 3 * (k + 1)) * (n + 1) * 4

    for bm, bn in ((40, 55), (30, 1), (1, 30), (25, 2)):
        budget = orc.traceback_min_bytes_for_database(bm, bn)
        for m in range(bm + 1):
            for n in range(bn + 1):
                need = full(m, n) if (m == 0 or n == 0 or full(m, n) <= budget) else ckpt(m, n)
                assert need <= budget, (m, n, bm, bn)


def test_budget_saturates_instead_of_wrapping():
    big = (1 << 64) - 1
    assert orc.traceback_min_bytes_for_database(1 << 40, 1 << 40) == big
