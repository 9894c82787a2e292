"""CPU oracle for the database scan — TEST INFRASTRUCTURE ONLY.

ctypes bindings over ``liboracle.so`` (built from ``oracle.c`` by ``oracle/Makefile``), a plain-C
restatement of hyalite's scalar Gotoh kernel (src/kernel.rs:39-277), width proof
(src/width.rs:36-169), scoring validation (src/scoring.rs:65-100) and scalar database
reductions (src/database.rs:115-144, 405-414, 511-513), plus the brute-force enumerators of
tests/common/mod.rs:65-171.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may import
this package, and only as the checker. The product path (``hyalite_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MODE_SW, MODE_NW, MODE_HW, MODE_OV, MODE_SHW = 0, 1, 2, 3, 4
ALL_MODES = (MODE_SW, MODE_NW, MODE_HW, MODE_OV, MODE_SHW)

OK = 0
E_EMPTY_ALPHABET = 1
E_MATRIX_SHAPE = 2
E_NEGATIVE_GAP = 3
E_GAP_OPEN_LT_EXT = 4
E_SCORE_RANGE = 5

_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.hyo_gap_penalty.restype = ctypes.c_int32
    lib.hyo_gap_penalty.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64]
    lib.hyo_scoring_check.restype = ctypes.c_int
    lib.hyo_scoring_check.argtypes = [ctypes.c_uint32, _i32p, ctypes.c_uint64, ctypes.c_int32,
                                      ctypes.c_int32, _i32p, _i32p]
    lib.hyo_required_width.restype = ctypes.c_int
    lib.hyo_required_width.argtypes = [ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                                       ctypes.c_uint64, _i64p, _i32p]
    lib.hyo_align_core.restype = ctypes.c_int
    lib.hyo_align_core.argtypes = [_u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64, ctypes.c_uint32,
                                   _i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _i32p,
                                   _i64p, _i64p]
    lib.hyo_brute_score.restype = ctypes.c_int64
    lib.hyo_brute_score.argtypes = [ctypes.c_uint32, _u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64,
                                    ctypes.c_uint32, _i32p, ctypes.c_int32, ctypes.c_int32]
    lib.hyo_brute_end.restype = None
    lib.hyo_brute_end.argtypes = [ctypes.c_uint32, _u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64,
                                  ctypes.c_uint32, _i32p, ctypes.c_int32, ctypes.c_int32, _i32p,
                                  _i64p, _i64p]
    lib.hyo_scan_all.restype = ctypes.c_int
    lib.hyo_scan_all.argtypes = [_u8p, _u64p, ctypes.c_uint64, _u8p, ctypes.c_uint64,
                                 ctypes.c_uint32, _i32p, ctypes.c_int32, ctypes.c_int32,
                                 ctypes.c_uint32, _i32p, _i64p, _i64p, ctypes.c_int]
    lib.hyo_scan.restype = ctypes.c_int
    lib.hyo_scan.argtypes = [_u8p, _u64p, ctypes.c_uint64, _u8p, ctypes.c_uint64,
                             ctypes.c_uint32, _i32p, ctypes.c_int32, ctypes.c_int32,
                             ctypes.c_uint32, ctypes.c_uint32, _i32p, _u64p, _i64p, _i64p,
                             ctypes.c_int]
    lib.hyo_traceback.restype = ctypes.c_int
    lib.hyo_traceback.argtypes = [_u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64, ctypes.c_uint32,
                                  _i32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _i32p,
                                  _u64p, _u64p, _u64p, _u64p, _u8p, _u64p]
    lib.hyo_position_max.restype = ctypes.c_int
    lib.hyo_position_max.argtypes = [_u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64, ctypes.c_uint32,
                                     _i32p, ctypes.c_int32, ctypes.c_int32, _i32p,
                                     ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]
    lib.hyo_score2.restype = ctypes.c_int
    lib.hyo_score2.argtypes = [_i32p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_uint64,
                               ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                               ctypes.POINTER(ctypes.c_uint64)]
    lib.hyo_sse_striped_sw_pairs.restype = ctypes.c_int
    lib.hyo_sse_striped_sw_pairs.argtypes = [_u8p, _u64p, _u8p, _u64p, ctypes.c_uint64,
                                             ctypes.c_uint32, _i32p, ctypes.c_int32,
                                             ctypes.c_int32, _i32p,
                                             ctypes.POINTER(ctypes.c_double)]
    lib.hyo_brute_colmax.restype = None
    lib.hyo_brute_colmax.argtypes = [_u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64,
                                     ctypes.c_uint32, _i32p, ctypes.c_int32, ctypes.c_int32, _i32p]
    lib.hyo_traceback_min_bytes_for_database.restype = ctypes.c_uint64
    lib.hyo_traceback_min_bytes_for_database.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    if hasattr(lib, "hyo_avx2_scan_scores_timed"):
        lib.hyo_avx2_scan_scores_timed.restype = ctypes.c_int
        lib.hyo_avx2_scan_scores_timed.argtypes = [
            _u8p, _u64p, ctypes.c_uint64, _u8p, ctypes.c_uint64, ctypes.c_uint32, _i32p,
            ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _i32p, ctypes.c_int, ctypes.c_int,
            ctypes.POINTER(ctypes.c_double)]
    _lib = lib
    return lib


def _u8(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint8))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _opt(v: int):
    return None if v < 0 else int(v)


def gap_penalty(go: int, ge: int, length: int) -> int:
    return int(_load().hyo_gap_penalty(go, ge, length))


def scoring_check(al: int, matrix, go: int, ge: int):
    """Returns (status, min_entry, max_entry)."""
    m = _i32(matrix) if len(matrix) else np.zeros(1, dtype=np.int32)
    lo, hi = ctypes.c_int32(0), ctypes.c_int32(0)
    rc = _load().hyo_scoring_check(al, _ptr(m, _i32p), len(matrix), go, ge, ctypes.byref(lo),
                                   ctypes.byref(hi))
    return rc, lo.value, hi.value


def required_width(mode: int, min_e: int, max_e: int, go: int, ge: int, mql: int, mtl: int):
    """Returns (status, bound, width_bits) where width_bits is 8, 16, 32 (0 when too wide)."""
    bound, bits = ctypes.c_int64(0), ctypes.c_int32(0)
    rc = _load().hyo_required_width(mode, min_e, max_e, go, ge, mql, mtl, ctypes.byref(bound),
                                    ctypes.byref(bits))
    return rc, bound.value, bits.value


def align_core(q, t, al: int, matrix, go: int, ge: int, mode: int):
    """(score, query_end or None, target_end or None) for one pair."""
    qa, ta, ma = _u8(q), _u8(t), _i32(matrix)
    s, qe, te = ctypes.c_int32(0), ctypes.c_int64(0), ctypes.c_int64(0)
    rc = _load().hyo_align_core(_ptr(qa, _u8p), len(qa), _ptr(ta, _u8p), len(ta), al,
                                _ptr(ma, _i32p), go, ge, mode, ctypes.byref(s), ctypes.byref(qe),
                                ctypes.byref(te))
    if rc != OK:
        raise MemoryError("oracle allocation failed")
    return s.value, _opt(qe.value), _opt(te.value)


# Alignment operations (src/align.rs:47-57).
OP_MATCH, OP_MISMATCH, OP_INS, OP_DEL = 0, 1, 2, 3


def traceback(q, t, al: int, matrix, go: int, ge: int, mode: int):
    """The canonical alignment of one pair (src/align.rs:169-420): (score, query_start, query_end,
    target_start, target_end, ops) with half-open spans and ops in forward order."""
    qa = _u8(q) if len(q) else np.zeros(1, dtype=np.uint8)
    ta = _u8(t) if len(t) else np.zeros(1, dtype=np.uint8)
    ma = _i32(matrix)
    ops = np.zeros(len(q) + len(t) + 1, dtype=np.uint8)
    s = ctypes.c_int32(0)
    qs, qe, ts, te, k = (ctypes.c_uint64(0) for _ in range(5))
    rc = _load().hyo_traceback(_ptr(qa, _u8p), len(q), _ptr(ta, _u8p), len(t), al,
                               _ptr(ma, _i32p), go, ge, mode, ctypes.byref(s), ctypes.byref(qs),
                               ctypes.byref(qe), ctypes.byref(ts), ctypes.byref(te),
                               _ptr(ops, _u8p), ctypes.byref(k))
    if rc != OK:
        raise MemoryError("oracle allocation failed")
    return s.value, qs.value, qe.value, ts.value, te.value, [int(x) for x in ops[:k.value]]


def traceback_min_bytes_for_database(max_m: int, max_n: int) -> int:
    return int(_load().hyo_traceback_min_bytes_for_database(max_m, max_n))


def position_max(q, t, al: int, matrix, go: int, ge: int):
    """SW per-target-position maxima (src/kernel.rs:474-575): (colmax list, score, target_end
    or None)."""
    qa = _u8(q) if len(q) else np.zeros(1, dtype=np.uint8)
    ta = _u8(t) if len(t) else np.zeros(1, dtype=np.uint8)
    out = np.zeros(max(len(t), 1), dtype=np.int32)
    s, te = ctypes.c_int32(0), ctypes.c_int64(0)
    rc = _load().hyo_position_max(_ptr(qa, _u8p), len(q), _ptr(ta, _u8p), len(t), al,
                                  _ptr(_i32(matrix), _i32p), go, ge, _ptr(out, _i32p),
                                  ctypes.byref(s), ctypes.byref(te))
    if rc != OK:
        raise MemoryError("oracle allocation failed")
    return [int(v) for v in out[:len(t)]], s.value, _opt(te.value)


def score2(colmax, best_score: int, best_target_end: int, matrix_max: int, min_score: int):
    """src/kernel.rs:741-792: (score2, position) or None."""
    a = np.ascontiguousarray(np.asarray(colmax, dtype=np.int32))
    if a.size == 0:
        a = np.zeros(1, dtype=np.int32)
    s, p = ctypes.c_int32(0), ctypes.c_uint64(0)
    found = _load().hyo_score2(_ptr(a, _i32p), len(colmax), best_score, best_target_end,
                               matrix_max, min_score, ctypes.byref(s), ctypes.byref(p))
    return (s.value, p.value) if found else None


def sse_striped_sw_pairs(qflat, qoff, tflat, toff, al: int, matrix, go: int, ge: int):
    """(SW scores of the pairs, seconds): the SSE4.1 striped restatement of the reference's
    single-pair path, one thread, each pair at the width its proof gives (16 x i8 or 8 x i16);
    ValueError when a pair needs 32 bits."""
    qf = _u8(qflat) if len(qflat) else np.zeros(1, dtype=np.uint8)
    tf = _u8(tflat) if len(tflat) else np.zeros(1, dtype=np.uint8)
    qo = np.ascontiguousarray(qoff, dtype=np.uint64)
    to = np.ascontiguousarray(toff, dtype=np.uint64)
    n = len(qo) - 1
    out = np.zeros(max(n, 1), dtype=np.int32)
    secs = ctypes.c_double(0)
    rc = _load().hyo_sse_striped_sw_pairs(_ptr(qf, _u8p), _ptr(qo, _u64p), _ptr(tf, _u8p),
                                          _ptr(to, _u64p), n, al, _ptr(_i32(matrix), _i32p), go,
                                          ge, _ptr(out, _i32p), ctypes.byref(secs))
    if rc == -2:
        raise ValueError("a pair needs 32-bit cells; the striped baseline covers 8 and 16 bits")
    if rc != 0:
        raise MemoryError("oracle allocation failed")
    return out[:n], secs.value


def brute_colmax(q, t, al: int, matrix, go: int, ge: int):
    qa = _u8(q) if len(q) else np.zeros(1, dtype=np.uint8)
    ta = _u8(t) if len(t) else np.zeros(1, dtype=np.uint8)
    out = np.zeros(max(len(t), 1), dtype=np.int32)
    _load().hyo_brute_colmax(_ptr(qa, _u8p), len(q), _ptr(ta, _u8p), len(t), al,
                             _ptr(_i32(matrix), _i32p), go, ge, _ptr(out, _i32p))
    return [int(v) for v in out[:len(t)]]


def cigar(ops, extended: bool = False) -> str:
    """Run-length CIGAR of the ops (src/align.rs:96-160): M/I/D, or =/X/I/D when extended."""
    letters = ("=", "X", "I", "D") if extended else ("M", "M", "I", "D")
    out, prev, run = [], None, 0
    for op in ops:
        c = letters[op]
        if c == prev:
            run += 1
        else:
            if prev is not None:
                out.append(f"{run}{prev}")
            prev, run = c, 1
    if prev is not None:
        out.append(f"{run}{prev}")
    return "".join(out)


def brute_score(mode: int, q, t, al: int, matrix, go: int, ge: int) -> int:
    qa, ta, ma = _u8(q), _u8(t), _i32(matrix)
    return int(_load().hyo_brute_score(mode, _ptr(qa, _u8p), len(qa), _ptr(ta, _u8p), len(ta), al,
                                       _ptr(ma, _i32p), go, ge))


def brute_end(mode: int, q, t, al: int, matrix, go: int, ge: int):
    qa, ta, ma = _u8(q), _u8(t), _i32(matrix)
    s, qe, te = ctypes.c_int32(0), ctypes.c_int64(0), ctypes.c_int64(0)
    _load().hyo_brute_end(mode, _ptr(qa, _u8p), len(qa), _ptr(ta, _u8p), len(ta), al,
                          _ptr(ma, _i32p), go, ge, ctypes.byref(s), ctypes.byref(qe),
                          ctypes.byref(te))
    return s.value, _opt(qe.value), _opt(te.value)


def flatten(seqs):
    """(flat uint8 residues, uint64 offsets of length n+1) for a list of sequences."""
    lens = np.fromiter((len(s) for s in seqs), dtype=np.uint64, count=len(seqs))
    off = np.zeros(len(seqs) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    flat = np.empty(int(off[-1]) if len(seqs) else 0, dtype=np.uint8)
    for k, s in enumerate(seqs):
        if len(s):
            flat[int(off[k]):int(off[k + 1])] = np.asarray(s, dtype=np.uint8)
    return flat, off


def scan_all_flat(flat, off, q, al, matrix, go, ge, mode, ends=True, threads=1):
    """Per-target (scores, query_ends, target_ends) in database order; ends are -1 for None."""
    n = len(off) - 1
    flat = _u8(flat) if len(flat) else np.zeros(1, dtype=np.uint8)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    qa, ma = _u8(q) if len(q) else np.zeros(1, dtype=np.uint8), _i32(matrix)
    scores = np.empty(max(n, 1), dtype=np.int32)
    qe = np.empty(max(n, 1), dtype=np.int64) if ends else None
    te = np.empty(max(n, 1), dtype=np.int64) if ends else None
    rc = _load().hyo_scan_all(_ptr(flat, _u8p), _ptr(off, _u64p), n, _ptr(qa, _u8p), len(q), al,
                              _ptr(ma, _i32p), go, ge, mode, _ptr(scores, _i32p),
                              _ptr(qe, _i64p) if ends else None,
                              _ptr(te, _i64p) if ends else None, threads)
    if rc != OK:
        raise MemoryError("oracle allocation failed")
    if ends:
        return scores[:n], qe[:n], te[:n]
    return scores[:n]


def scan_all(seqs, q, al, matrix, go, ge, mode, threads=1):
    """List of (score, query_end, target_end) per target, None for absent ends."""
    flat, off = flatten(seqs)
    s, qe, te = scan_all_flat(flat, off, q, al, matrix, go, ge, mode, ends=True, threads=threads)
    return [(int(a), _opt(int(b)), _opt(int(c))) for a, b, c in zip(s, qe, te)]


def scan_scores(seqs, q, al, matrix, go, ge, mode, threads=1):
    flat, off = flatten(seqs)
    return scan_all_flat(flat, off, q, al, matrix, go, ge, mode, ends=False, threads=threads)


def scan_flat(flat, off, q, al, matrix, go, ge, mode, tracks_end, threads=1):
    """Single best hit (score, db_index, query_end, target_end)."""
    n = len(off) - 1
    flat = _u8(flat) if len(flat) else np.zeros(1, dtype=np.uint8)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    qa, ma = _u8(q) if len(q) else np.zeros(1, dtype=np.uint8), _i32(matrix)
    s, idx = ctypes.c_int32(0), ctypes.c_uint64(0)
    qe, te = ctypes.c_int64(0), ctypes.c_int64(0)
    rc = _load().hyo_scan(_ptr(flat, _u8p), _ptr(off, _u64p), n, _ptr(qa, _u8p), len(q), al,
                          _ptr(ma, _i32p), go, ge, mode, 1 if tracks_end else 0, ctypes.byref(s),
                          ctypes.byref(idx), ctypes.byref(qe), ctypes.byref(te), threads)
    if rc != OK:
        raise ValueError(f"oracle scan failed with status {rc}")
    return s.value, idx.value, _opt(qe.value), _opt(te.value)


def scan(seqs, q, al, matrix, go, ge, mode, tracks_end, threads=1):
    flat, off = flatten(seqs)
    return scan_flat(flat, off, q, al, matrix, go, ge, mode, tracks_end, threads)


def avx2_available() -> bool:
    return hasattr(_load(), "hyo_avx2_scan_scores_timed")


def avx2_scan_scores_flat(flat, off, q, al, matrix, go, ge, mode, threads=1, sort=False):
    """(per-target scores, scan seconds) from the AVX2 16-lane i16 restatement of the reference's
    inter-sequence kernel (the CPU baseline). Only databases whose width proof gives <= 16 bits
    are accepted; the per-batch tables are built before the timed scan, as the reference builds
    them once per database."""
    n = len(off) - 1
    flat = _u8(flat) if len(flat) else np.zeros(1, dtype=np.uint8)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    qa, ma = _u8(q) if len(q) else np.zeros(1, dtype=np.uint8), _i32(matrix)
    scores = np.empty(max(n, 1), dtype=np.int32)
    secs = ctypes.c_double(0.0)
    rc = _load().hyo_avx2_scan_scores_timed(_ptr(flat, _u8p), _ptr(off, _u64p), n,
                                            _ptr(qa, _u8p), len(q), al, _ptr(ma, _i32p), go, ge,
                                            mode, _ptr(scores, _i32p), threads,
                                            1 if sort else 0, ctypes.byref(secs))
    if rc != OK:
        raise ValueError(f"avx2 baseline failed with status {rc}")
    return scores[:n], secs.value
