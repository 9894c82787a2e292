"""CPU-side checks of the C ABI library: it loads, exports every declared symbol, and reproduces
the reference builder's validation order and error payloads (src/database.rs:737-877,
src/scoring.rs:65-100, src/backend.rs:143-204) — all of which run before any device is touched."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import hyalite_amd as hy
from common import dna_scheme

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hyalite_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(hy_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = hy.load_library()
    names = declared_symbols()
    assert "hy_scan" in names and "hy_db_build" in names
    for name in names:
        assert hasattr(lib, name), name


def scoring():
    al, mat, go, ge = dna_scheme()
    return hy.Scoring(al, mat, go, ge)


def test_scoring_validation_order():
    with pytest.raises(hy.EmptyAlphabet):
        hy.Scoring(0, [], 1, 1)
    with pytest.raises(hy.MatrixShape):
        hy.Scoring(2, [1, 2, 3], -1, 5)
    with pytest.raises(hy.NegativeGapPenalty):
        hy.Scoring(2, [1, 2, 3, 4], -1, 5)
    with pytest.raises(hy.GapOpenLessThanExtend):
        hy.Scoring(2, [1, 2, 3, 4], 1, 2)


def test_builder_incomplete_fields_in_order():
    b = hy.Database.builder
    with pytest.raises(hy.IncompleteBuilder) as e:
        b().build()
    assert e.value.field == "sequences"
    with pytest.raises(hy.IncompleteBuilder) as e:
        b().sequences([[0]]).build()
    assert e.value.field == "scoring"
    with pytest.raises(hy.IncompleteBuilder) as e:
        b().sequences([[0]]).scoring(scoring()).build()
    assert e.value.field == "mode"
    with pytest.raises(hy.IncompleteBuilder) as e:
        b().sequences([[0]]).scoring(scoring()).mode(hy.Mode.Sw).build()
    assert e.value.field == "max_query_len"


def test_builder_empty_symbol_width_errors():
    b = lambda seqs: (hy.Database.builder().sequences(seqs).scoring(scoring())  # noqa: E731
                      .mode(hy.Mode.Nw).max_query_len(8))
    with pytest.raises(hy.EmptyDatabase):
        b([]).build()
    with pytest.raises(hy.SymbolOutOfRange) as e:
        b([[0, 1], [2, 7, 9]]).build()
    assert (e.value.symbol, e.value.alphabet_len) == (7, 4)
    huge = hy.Scoring(2, [0, 0, 0, 0], 2_000_000_000, 1)
    with pytest.raises(hy.ScoreRangeTooWide) as e:
        hy.Database.builder().sequences([[0]]).scoring(huge).mode(hy.Mode.Nw) \
            .max_query_len(1).build()
    assert e.value.bound == 2_000_000_001


def test_alignment_budget_is_validated_at_build():
    """tests/database_alignment.rs:84-104: a one-byte budget cannot cover an 8 x 8 traceback.
    The budget check runs before the backend is resolved (database.rs:773-794), so it is decided
    here without a device. The needed bytes are the database budget of the (8, 8) box: the
    checkpoint footprint (2 * 5 + 3 * 3) * 9 * 4 = 684."""
    seqs = [[0, 1, 2, 3, 0, 1, 2, 3]]
    with pytest.raises(hy.TracebackBudgetExceeded) as e:
        hy.Database.builder().sequences(seqs).scoring(scoring()).mode(hy.Mode.Nw) \
            .search_type(hy.SearchType.Alignment(1)).max_query_len(8).build()
    assert (e.value.needed_bytes, e.value.budget_bytes) == (684, 1)
    assert "traceback needs 684 bytes but the budget is 1" in str(e.value)
    # One byte short of the exact budget fails; the exact budget passes validation (and then
    # stops at the backend, which this container lacks, or builds on a device).
    with pytest.raises(hy.TracebackBudgetExceeded):
        hy.Database.builder().sequences(seqs).scoring(scoring()).mode(hy.Mode.Nw) \
            .search_type(hy.SearchType.Alignment(683)).max_query_len(8).build()
    try:
        hy.Database.builder().sequences(seqs).scoring(scoring()).mode(hy.Mode.Nw) \
            .search_type(hy.SearchType.Alignment(684)).max_query_len(8).build()
    except hy.BackendUnavailable:
        pass
    # Width and symbol errors come first (database.rs:756-771).
    with pytest.raises(hy.SymbolOutOfRange):
        hy.Database.builder().sequences([[9]]).scoring(scoring()).mode(hy.Mode.Nw) \
            .search_type(hy.SearchType.Alignment(1)).max_query_len(8).build()
    st = hy.SearchType.Alignment(4096)
    assert st.tracks_end() and st.max_bytes() == 4096
    assert hy.SearchType.ScoreEnd.max_bytes() is None


def test_alignment_budget_matches_the_oracle_over_a_box_sweep():
    """The library's budget arithmetic against the pinned oracle, through the build check: for
    each box the oracle's exact budget passes validation and one byte less fails."""
    import oracle as orc
    for m, n in [(0, 0), (0, 5), (5, 0), (1, 30), (30, 1), (8, 8), (40, 55), (400, 400),
                 (1000, 35000)]:
        need = orc.traceback_min_bytes_for_database(m, n)
        seqs = [[0] * n] if n else [[]]
        b = lambda budget: (hy.Database.builder().sequences(seqs).scoring(scoring())  # noqa
                            .mode(hy.Mode.Sw).search_type(hy.SearchType.Alignment(budget))
                            .max_query_len(m))
        with pytest.raises(hy.TracebackBudgetExceeded) as e:
            b(need - 1).build()
        assert e.value.needed_bytes == need, (m, n)
        try:
            b(need).build()
        except hy.BackendUnavailable:
            pass


def test_backend_parse_and_env_var():
    assert hy.BackendChoice.parse("AUTO") == hy.BackendChoice.AUTO
    assert hy.BackendChoice.parse("  Hip ") == hy.BackendChoice.force(hy.Backend.Hip)
    for bad in ["", "sse2", "ssse3", "avx512", "gpu", "x"]:
        with pytest.raises(hy.InvalidBackendName):
            hy.BackendChoice.parse(bad)
    for cpu in ["scalar", "sse4.1", "sse41", "avx2", "neon"]:
        with pytest.raises(hy.BackendUnavailable):
            hy.BackendChoice.parse(cpu)


def _child(env_value):
    code = (
        "import sys; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import hyalite_amd as hy\n"
        "from common import dna_scheme\n"
        "al, mat, go, ge = dna_scheme()\n"
        "try:\n"
        "    hy.Database.builder().sequences([[0,1]]).scoring(hy.Scoring(al, mat, go, ge))"
        ".mode(hy.Mode.Sw).max_query_len(4).build()\n"
        "    print('built')\n"
        "except hy.HyaliteError as e:\n"
        "    print(type(e).__name__)\n" % (ROOT, os.path.join(ROOT, "tests")))
    env = dict(os.environ)
    env["HYALITE_BACKEND"] = env_value
    # No device in the child, as on a CPU-only host.
    env["HIP_VISIBLE_DEVICES"] = "-1"
    env["ROCR_VISIBLE_DEVICES"] = "-1"
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=120)
    return out.stdout.strip().splitlines()[-1]


def test_env_var_is_read_in_its_own_process():
    """tests/backend_env.rs:1-97: an unknown name is InvalidBackendName before anything else; a
    CPU backend name or an absent device is BackendUnavailable."""
    assert _child("definitely-not-a-backend") == "InvalidBackendName"
    assert _child("scalar") == "BackendUnavailable"
    assert _child("hip") == "BackendUnavailable"


def test_strerror_covers_every_code():
    lib = hy.load_library()
    for code in list(range(0, 12)) + list(range(100, 109)):
        assert lib.hy_strerror(code) not in (None, b"unknown status code")


@pytest.mark.parametrize("flat, offsets, msg", [
    (np.zeros(8, np.uint8), np.array([1, 4, 8]), "start at 0"),
    (np.zeros(8, np.uint8), np.array([0, 5, 3, 8]), "not decrease"),
    (np.zeros(8, np.uint8), np.array([0, 4, 9]), "past the buffer"),
    (np.zeros(8, np.uint8), np.array([], dtype=np.uint64), "at least one"),
    (np.zeros(8, np.uint8), np.array([0, -1, 8]), "non-negative"),
    (np.zeros((2, 4), np.uint8), np.array([0, 4, 8]), "1-D"),
    (np.zeros(8, np.uint8), np.array([[0, 4, 8]]), "1-D"),
])
def test_sequences_flat_rejects_bad_offsets(flat, offsets, msg):
    """Offsets that would make the build read outside the caller's buffer (or wrap a length to
    nearly 2^64) are refused before anything reaches the library."""
    with pytest.raises(ValueError, match=msg):
        hy.Database.builder().sequences_flat(flat, offsets)


def test_sequences_flat_accepts_good_offsets():
    b = hy.Database.builder().sequences_flat(np.arange(8, dtype=np.uint8) % 4,
                                             np.array([0, 0, 3, 8], dtype=np.int64))
    assert b._flat[1].dtype == np.uint64 and list(b._flat[1]) == [0, 0, 3, 8]
