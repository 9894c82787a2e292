"""Pins the C oracle before it is trusted as the parity checker.

Every check here comes from the reference's own tests (hand vectors, known answers, width-proof
boundaries), from exhaustive brute-force enumeration with a structure independent of the DP
(tests/common/mod.rs:65-171), or from invariants the mathematics fixes (mode ordering, symmetry,
self-alignment sums). A dropped term, a wrong sign, a transposed operand or a wrong tie-break fails
at least one of them.
"""
import itertools

import numpy as np
import pytest

import oracle as orc
from common import (ALL_MODES, BLOSUM62, AMINO_ACIDS, MODE_HW, MODE_NW, MODE_OV, MODE_SHW, MODE_SW,
                    UBIQUITIN, all_sequences, cr4_adapters, cr4_scheme, dna_scheme,
                    encode_protein, identity_matrix, phix)


def score(mode, q, t, scheme=None):
    al, mat, go, ge = scheme or dna_scheme()
    return orc.align_core(q, t, al, mat, go, ge, mode)[0]


# ---- tests/alignment.rs:184-214: scalar == brute force over all short pairs -------------------

SCHEMES = [
    (2, 3, 2, -1, 2, 1),
    (2, 3, 1, -1, 3, 3),
    (2, 3, 3, -2, 4, 2),
    (3, 2, 2, -1, 2, 1),
]


@pytest.mark.parametrize("al,max_len,m,x,go,ge", SCHEMES)
def test_scores_match_brute_force_over_all_short_pairs(al, max_len, m, x, go, ge):
    mat = identity_matrix(al, m, x)
    seqs = all_sequences(al, max_len)
    for q in seqs:
        for t in seqs:
            for mode in ALL_MODES:
                want = orc.brute_score(mode, q, t, al, mat, go, ge)
                got = orc.align_core(q, t, al, mat, go, ge, mode)[0]
                assert got == want, (mode, q, t)


@pytest.mark.parametrize("al,max_len,m,x,go,ge", SCHEMES)
def test_ends_match_brute_force_over_all_short_pairs(al, max_len, m, x, go, ge):
    """Ends by enumerating every end cell's best window and reducing with the tie-break."""
    mat = identity_matrix(al, m, x)
    seqs = all_sequences(al, max_len)
    for q in seqs:
        for t in seqs:
            for mode in ALL_MODES:
                want = orc.brute_end(mode, q, t, al, mat, go, ge)
                got = orc.align_core(q, t, al, mat, go, ge, mode)
                assert got == want, (mode, q, t)


def test_asymmetric_random_matrices_match_brute_force():
    """Non-symmetric matrices catch a transposed matrix index (matrix[q*al+t] vs [t*al+q])."""
    rng = np.random.default_rng(7)
    for _ in range(40):
        al = int(rng.integers(2, 4))
        mat = [int(v) for v in rng.integers(-4, 5, size=al * al)]
        go = int(rng.integers(0, 5))
        ge = int(rng.integers(0, go + 1))
        for _ in range(10):
            q = [int(v) for v in rng.integers(0, al, size=int(rng.integers(0, 5)))]
            t = [int(v) for v in rng.integers(0, al, size=int(rng.integers(0, 5)))]
            for mode in ALL_MODES:
                assert orc.align_core(q, t, al, mat, go, ge, mode) == \
                    orc.brute_end(mode, q, t, al, mat, go, ge), (al, mat, go, ge, q, t, mode)


def test_small_penalty_affine_init_matches_brute():
    """tests/opal_regressions.rs:84-121 (gap regimes (0,0),(1,0),...,(2,2) and deeper NW)."""
    seqs = all_sequences(2, 2)
    for go, ge in [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]:
        mat = identity_matrix(2, 2, -2)
        for q in seqs:
            for t in seqs:
                for mode in ALL_MODES:
                    assert score(mode, q, t, (2, mat, go, ge)) == \
                        orc.brute_score(mode, q, t, 2, mat, go, ge)
    deep = all_sequences(2, 4)
    for go, ge in [(1, 1), (2, 1), (0, 0)]:
        mat = identity_matrix(2, 3, -2)
        for q in deep[::3]:
            for t in deep[::2]:
                assert score(MODE_NW, q, t, (2, mat, go, ge)) == \
                    orc.brute_score(MODE_NW, q, t, 2, mat, go, ge)


# ---- tests/alignment.rs:231-296 hand vectors ---------------------------------------------------

def test_perfect_match_scores_full_match_in_every_mode():
    seq = [0, 1, 2, 3]
    for mode in ALL_MODES:
        assert orc.align_core(seq, seq, *dna_scheme(), mode) == (8, 3, 3)


def test_single_internal_mismatch():
    assert score(MODE_NW, [0, 1, 2, 3], [0, 1, 0, 3]) == 5
    assert score(MODE_SW, [0, 1, 2, 3], [0, 1, 0, 3]) == 5


def test_single_gap_and_two_base_gap():
    assert score(MODE_NW, [0, 1, 2, 3], [0, 2, 3]) == 4
    assert score(MODE_NW, [0, 1, 2, 3], [0, 3]) == 1


def test_hw_fits_query_into_longer_target():
    assert orc.align_core([1, 2], [0, 1, 2, 3], *dna_scheme(), MODE_HW) == (4, 1, 2)
    assert score(MODE_NW, [1, 2], [0, 1, 2, 3]) == 0


def test_ov_suffix_prefix_overlap():
    assert score(MODE_OV, [0, 0, 2, 3], [2, 3, 1, 1]) == 4


def test_tie_break_prefers_smallest_target_then_query_end():
    """tests/alignment.rs:334-362."""
    assert orc.align_core([0], [0, 0, 0], *dna_scheme(), MODE_SW) == (2, 0, 0)
    s, _, te = orc.align_core([0], [0, 0, 0], *dna_scheme(), MODE_HW)
    assert (s, te) == (2, 0)
    for run in range(1, 7):
        assert orc.align_core([0], [0] * run, *dna_scheme(), MODE_SW)[2] == 0


def test_empty_inputs():
    """tests/alignment.rs:369-390."""
    for mode in ALL_MODES:
        assert orc.align_core([], [], *dna_scheme(), mode) == (0, None, None)
    assert orc.align_core([], [0, 1, 2, 3], *dna_scheme(), MODE_NW) == (-5, None, 3)
    assert score(MODE_SW, [0, 1, 2], []) == 0
    assert score(MODE_OV, [], [0, 1, 2]) == 0
    assert score(MODE_HW, [], [0, 1, 2]) == 0


# ---- invariants: mode ordering and symmetry (tests/alignment.rs:303-331, 397-424) --------------

def test_mode_ordering_and_symmetry_over_all_short_pairs():
    seqs = all_sequences(4, 3)
    for q in seqs[::2]:
        for t in seqs[::3]:
            sw, ov, hw, nw = (score(m, q, t) for m in (MODE_SW, MODE_OV, MODE_HW, MODE_NW))
            assert sw >= 0 and sw >= ov >= hw >= nw
            for mode in (MODE_SW, MODE_NW, MODE_OV):
                assert score(mode, q, t) == score(mode, t, q)
            # SHW is the transpose of HW (src/kernel.rs:89-97).
            assert score(MODE_SHW, q, t) == score(MODE_HW, t, q)


def test_hw_is_generally_not_symmetric():
    assert score(MODE_HW, [1, 2], [0, 1, 2, 3]) != score(MODE_HW, [0, 1, 2, 3], [1, 2])


# ---- width proof KATs (src/width.rs:177-414) ----------------------------------------------------

def width(mode, min_e, max_e, go, ge, m, n):
    rc, bound, bits = orc.required_width(mode, min_e, max_e, go, ge, m, n)
    return (bits if rc == orc.OK else None), bound


def test_width_boundaries():
    # narrowest_for boundaries through SW's negative reach == gap_open
    for go, want in [(0, 8), (127, 8), (128, 16), (32767, 16), (32768, 32), (536870911, 32)]:
        assert width(MODE_SW, -1, 1, go, 1, 4, 4)[0] == want
    assert width(MODE_SW, -1, 1, 536870912, 1, 4, 4)[0] is None


def test_width_i32_sentinel_ceiling():
    assert width(MODE_NW, 0, 0, 2_000_000_000, 1, 1, 1) == (None, 2_000_000_001)
    assert width(MODE_SW, -1, 1, 1_100_000_000, 1_100_000_000, 9, 4)[0] is None
    assert width(MODE_NW, -2_000_000_000, 1, 1, 1, 2, 2)[0] is None
    assert width(MODE_NW, -1, 1, 2, 1, 30_000, 30_000)[0] == 32


def test_width_mode_aware_negative_reach():
    assert width(MODE_SW, -1, 2, 3, 1, 10, 10)[0] == 8
    assert width(MODE_SW, -100, 1, 2, 1, 8, 8)[0] == 8
    for mode in (MODE_NW, MODE_HW, MODE_OV):
        assert width(mode, -100, 1, 2, 1, 8, 8)[0] > 8
    assert width(MODE_SW, -1, 1, 200, 1, 8, 8)[0] == 16
    assert width(MODE_OV, -2, 1, 2, 2, 91, 33)[0] == 8
    assert width(MODE_NW, -2, 1, 2, 2, 91, 33)[0] == 16


def test_width_escalation_and_min_length():
    for mode in ALL_MODES:
        assert width(mode, 0, 100, 0, 0, 1, 1)[0] == 8
        assert width(mode, 0, 100, 0, 0, 327, 327)[0] == 16
        assert width(mode, 0, 100, 0, 0, 328, 328)[0] == 32
        assert width(mode, 0, 100, 0, 0, 50_000, 50_000)[0] == 32
        assert width(mode, -5, 5, 4, 2, 0, 0)[0] == 8
    assert width(MODE_SW, 0, 100, 0, 0, 2, 1_000_000)[0] == 16
    assert width(MODE_SW, 0, 100, 0, 0, 1, 1_000_000)[0] == 8


def test_width_pathological_lengths():
    bits, bound = width(MODE_NW, -1, 1, 0, 1, 2_000_000_000, 2_000_000_000)
    assert bits is None and bound > 2**31 - 1
    bits, bound = width(MODE_OV, -1, 1, 1, 1, 2**64 - 1, 2**64 - 1)
    assert bits is None and bound == 2**63 - 1


def test_width_closed_form_matches_width_rs_cases():
    """chosen_width_actually_contains_the_bound (src/width.rs:389-413): minimal width."""
    cases = [(MODE_SW, -1, 2, 3, 1, 50, 50), (MODE_NW, -4, 5, 8, 2, 200, 200),
             (MODE_HW, -2, 3, 6, 1, 500, 30), (MODE_OV, -7, 7, 10, 3, 4000, 4000),
             (MODE_NW, -1, 1, 0, 1, 40_000, 40_000)]
    lim = {8: 127, 16: 32767, 32: 536870911}
    for c in cases:
        bits, bound = width(*c)
        assert bound <= lim[bits]
        for narrower in [b for b in (8, 16, 32) if b < bits]:
            assert bound > lim[narrower]
    # Closed forms: NW (m+n)*|min| + go + (m+n-1)*ge; OV min(m,n)*|min| + go; positive reach.
    assert width(MODE_NW, -4, 5, 8, 2, 200, 200)[1] == 400 * 4 + 8 + 399 * 2
    assert width(MODE_OV, -7, 7, 10, 3, 4000, 4000)[1] == max(4000 * 7, 4000 * 7 + 10)
    assert width(MODE_HW, -2, 3, 6, 1, 500, 30)[1] == 530 * 2 + 6 + 529


def test_wide_dna_scoring_partitions():
    """src/database.rs:1366-1399: lengths 4,6,10,24,2000 under +20/-5, 8/2 SW, mql 2000."""
    got = [width(MODE_SW, -5, 20, 8, 2, 2000, n)[0] for n in (4, 6, 10, 24, 2000)]
    assert got == [8, 8, 16, 16, 32]


# ---- scoring validation error order (src/scoring.rs:65-100) -------------------------------------

def test_scoring_validation_order():
    assert orc.scoring_check(0, [], 1, 1)[0] == orc.E_EMPTY_ALPHABET
    assert orc.scoring_check(2, [1, 2, 3], -1, 5)[0] == orc.E_MATRIX_SHAPE
    assert orc.scoring_check(2, [1, 2, 3, 4], -1, 5)[0] == orc.E_NEGATIVE_GAP
    assert orc.scoring_check(2, [1, 2, 3, 4], 1, 2)[0] == orc.E_GAP_OPEN_LT_EXT
    assert orc.scoring_check(2, [1, -2, 3, 4], 2, 2) == (orc.OK, -2, 4)


# ---- database-level known answers (src/database.rs:1258-1319; tests/real_data.rs) --------------

def test_db_known_answers():
    al, mat, go, ge = dna_scheme()
    assert orc.scan([[2, 2, 2, 2], [0, 1, 2, 3]], [0, 1, 2, 3], al, mat, go, ge, MODE_SW, True) \
        == (8, 1, 3, 3)
    s = orc.scan_scores([[2, 2, 2, 2], [0, 1, 2, 3], [0, 1, 2, 2]], [0, 1, 2, 3], al, mat, go, ge,
                        MODE_SW)
    assert s[1] == 8 and s[1] > s[2] > s[0]
    assert orc.scan([[0, 1, 2, 3], [0, 1, 2, 3], [3]], [0, 1, 2, 3], al, mat, go, ge, MODE_SW,
                    True)[:2] == (8, 0)
    assert orc.scan([[0, 1, 2, 3]], [0, 1, 2, 3], al, mat, go, ge, MODE_SW, False) == \
        (8, 0, None, None)


def test_cr4_adapter_scan_known_answers():
    """tests/real_data.rs:145-201."""
    adapters = cr4_adapters()
    assert adapters[0][0] == "TSO_10x"
    seqs = [s for _, s in adapters]
    genome = phix()
    al, mat, go, ge = cr4_scheme()
    tso_read = seqs[0] + genome[800:840]
    s, idx, _, _ = orc.scan(seqs, tso_read, al, mat, go, ge, MODE_OV, True)
    assert (idx, s) == (0, 30)
    polya_read = genome[1500:1530] + [0] * 25
    s, idx, _, _ = orc.scan(seqs, polya_read, al, mat, go, ge, MODE_OV, True)
    assert idx == 1 and s >= 25


def test_phix_exact_substring_is_recovered():
    """tests/real_data.rs:46-61."""
    genome = phix()
    al, mat, go, ge = dna_scheme()
    window = genome[1234:1284]
    for mode in (MODE_SW, MODE_HW):
        assert orc.align_core(window, genome, al, mat, go, ge, mode) == (100, 49, 1234 + 49)


def test_blosum62_transcription_and_ubiquitin_self_alignment():
    """tests/real_data.rs:249-287."""
    for i in range(20):
        for j in range(20):
            assert BLOSUM62[i * 20 + j] == BLOSUM62[j * 20 + i]
    idx = AMINO_ACIDS.index
    s = lambda a, b: BLOSUM62[idx(a) * 20 + idx(b)]  # noqa: E731
    assert (s("W", "W"), s("C", "C"), s("A", "A"), s("L", "I"), s("W", "A"), s("P", "W")) == \
        (11, 9, 4, 2, -3, -4)
    ub = encode_protein(UBIQUITIN)
    assert len(ub) == 76
    want = sum(BLOSUM62[a * 20 + a] for a in ub)
    for mode in ALL_MODES:
        assert orc.align_core(ub, ub, 20, BLOSUM62, 11, 1, mode)[0] == want


def test_proptest_regression_seeds():
    """proptest-regressions/inter.txt:7 and tests/properties.proptest-regressions:7-8."""
    for mode in ALL_MODES:
        got = orc.scan([[]], [0], 2, [0, 0, 0, 0], 0, 0, mode, True)
        assert got == orc.brute_end(mode, [0], [], 2, [0, 0, 0, 0], 0, 0)[:1] + (0,) + \
            orc.brute_end(mode, [0], [], 2, [0, 0, 0, 0], 0, 0)[1:]
    cases = [([0, 0, 0, -5], 8, 0, [0, 0, 1, 0, 1, 1, 0, 1, 1, 0, 0, 1], [1, 1, 0, 1, 1]),
             ([0, 0, 11, 0], 1, 0, [0] * 12, [1])]
    for mat, go, ge, t, q in cases:
        for mode in ALL_MODES:
            assert orc.align_core(q, t, 2, mat, go, ge, mode)[0] == \
                orc.brute_score(mode, q, t, 2, mat, go, ge)


def test_scan_all_threads_agree():
    rng = np.random.default_rng(3)
    seqs = [rng.integers(0, 4, size=int(n), dtype=np.uint8) for n in rng.integers(0, 40, 300)]
    q = rng.integers(0, 4, size=30, dtype=np.uint8)
    al, mat, go, ge = dna_scheme()
    for mode in ALL_MODES:
        a = orc.scan_all(seqs, q, al, mat, go, ge, mode, threads=1)
        b = orc.scan_all(seqs, q, al, mat, go, ge, mode, threads=5)
        assert a == b
        best = max(range(len(a)), key=lambda k: (a[k][0], -k))
        assert orc.scan(seqs, q, al, mat, go, ge, mode, True, threads=3) == \
            (a[best][0], best, a[best][1], a[best][2])


# ---- the AVX2 CPU baseline must agree with the scalar oracle before it is timed -----------------

@pytest.mark.parametrize("sort", [False, True])
def test_avx2_baseline_matches_scalar_oracle(sort):
    rng = np.random.default_rng(41)
    from common import BLOSUM62 as B62
    seqs = [rng.integers(0, 20, size=int(n), dtype=np.uint8) for n in rng.integers(0, 120, 333)]
    flat, off = orc.flatten(seqs)
    for q in (rng.integers(0, 20, size=90, dtype=np.uint8), np.zeros(0, dtype=np.uint8),
              rng.integers(0, 20, size=1, dtype=np.uint8)):
        for mode in ALL_MODES:
            want = orc.scan_all_flat(flat, off, q, 20, B62, 11, 1, mode, ends=False, threads=3)
            got, secs = orc.avx2_scan_scores_flat(flat, off, q, 20, B62, 11, 1, mode, threads=3,
                                                  sort=sort)
            assert np.array_equal(got, want), (mode, sort, len(q))
            assert secs >= 0.0
    al, mat, go, ge = dna_scheme()
    for mode in ALL_MODES:
        q = rng.integers(0, 4, size=40, dtype=np.uint8)
        want = orc.scan_all_flat(flat % 4, off, q, al, mat, go, ge, mode, ends=False)
        got, _ = orc.avx2_scan_scores_flat(flat % 4, off, q, al, mat, go, ge, mode, threads=2,
                                           sort=sort)
        assert np.array_equal(got, want), mode
