"""hyalite_amd — the MI355X (gfx950) database-scan backend of hyalite, driven from Python.

A thin ctypes layer over the C ABI in ``include/hyalite_hip.h`` (``libhyalite_hip.so``, built from
``hyalite_amd/csrc``). It mirrors hyalite's public API for the database-scan path so that parity
tests read like the reference's own:

    Scoring.new(...)                                  src/scoring.rs:65-100
    Database.builder().sequences(..).scoring(..).mode(..).search_type(..)
            .max_query_len(..).backend(..).build()    src/database.rs:657-877
    Scratch(db)                                       src/database.rs:954-999
    db.scan / db.scan_all / db.scan_scores            src/database.rs:83-515
    BestHit, Mode, SearchType, ScoreWidth, Backend, BackendChoice, errors
                                                      src/hit.rs, mode.rs, search.rs, width.rs,
                                                      backend.rs, error.rs

Every result comes from the HIP kernels; there is no CPU fallback. Importing works without a GPU,
but building a database raises ``BackendUnavailable`` unless a gfx950 device is present, and any
use raises ``HipLibraryMissing`` when the shared library has not been built.
"""
from __future__ import annotations

import ctypes
import enum
import os
import subprocess
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

__all__ = [
    "Mode", "SearchType", "ScoreWidth", "Backend", "BackendChoice", "BestHit", "Scoring",
    "Database", "DatabaseBuilder", "Scratch", "HyaliteError", "EmptyAlphabet", "MatrixShape",
    "NegativeGapPenalty", "GapOpenLessThanExtend", "ScoreRangeTooWide", "SymbolOutOfRange",
    "IncompleteBuilder", "EmptyDatabase", "BackendUnavailable", "InvalidBackendName",
    "ScanContractViolation", "DeviceError", "HipLibraryMissing", "BACKEND_ENV_VAR", "build",
    "library_path", "load_library", "Tuning", "Server", "ServedHit", "ServerStats",
    "ServerTimeout", "AlignOp", "Alignment", "AlignedHit", "TracebackBudgetExceeded",
    "PairScratch", "align_pair", "align_pair_with", "align_pairs", "align_pair_position_max",
    "align_pair_position_max_with", "score2", "align_pairs_flat", "DatabaseImageError",
    "StaticScratchLimit",
]

BACKEND_ENV_VAR = "HYALITE_BACKEND"

_PKG = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_PKG)
_LIB_NAME = "libhyalite_hip.so"
# Translation units: the host library, the autotuner, the mode-independent kernels and dispatch,
# and the mode-templated kernels compiled once per alignment mode (SW, NW, HW, OV, SHW).
_UNITS = [("csrc/hy_host.cpp", []), ("csrc/hy_tune.cpp", []), ("csrc/hy_serve.cpp", []),
          ("csrc/hy_align.cpp", []), ("csrc/hy_pair.cpp", []), ("csrc/hy_image.cpp", []),
          ("csrc/hy_xfer.cpp", []), ("csrc/hy_launch.hip", []), ("csrc/hy_static.hip", [])] + [
    (src, [f"-DHY_KMODE={k}"])
    for src in ("csrc/hy_kernels.hip", "csrc/hy_trace.hip", "csrc/hy_pair.hip")
    for k in range(5)]
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def library_path() -> str:
    return os.path.join(_PKG, _LIB_NAME)


def build(verbose: bool = False, jobs: Optional[int] = None) -> str:
    """Compile libhyalite_hip.so for gfx950 in-tree with hipcc (translation units in parallel)."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc):
        hipcc = "hipcc"
    objdir = os.path.join(_PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    cmds, objs = [], []
    for src, extra in _UNITS:
        stem = os.path.splitext(os.path.basename(src))[0] + "".join(
            "_" + e.split("=")[-1] for e in extra)
        obj = os.path.join(objdir, stem + ".o")
        objs.append(obj)
        cmds.append([hipcc, *HIPCC_FLAGS, *extra, "-c", os.path.join(_PKG, src), "-o", obj])
    jobs = jobs or max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0)
                              or os.cpu_count() or 1))

    def run(cmd):
        return cmd, subprocess.run(cmd, capture_output=True, text=True)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for cmd, res in pool.map(run, cmds):
            if res.returncode != 0:
                raise RuntimeError("hipcc failed: " + " ".join(cmd) + "\n" + res.stdout +
                                   res.stderr)
            if verbose:
                print(" ".join(cmd))
    out = library_path()
    tmp = out + ".tmp"
    link = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(" ".join(link))
    os.replace(tmp, out)
    return out


# ---- errors (src/error.rs:16-96) ---------------------------------------------------------------

class HyaliteError(Exception):
    """Base of every construction-time error (the reference's `Error` enum)."""


class EmptyAlphabet(HyaliteError):
    pass


class MatrixShape(HyaliteError):
    pass


class NegativeGapPenalty(HyaliteError):
    pass


class GapOpenLessThanExtend(HyaliteError):
    pass


class ScoreRangeTooWide(HyaliteError):
    def __init__(self, bound: int):
        super().__init__(f"reachable score magnitude bound ({bound}) exceeds the i32 range; "
                         "reduce sequence lengths or penalty magnitudes")
        self.bound = bound


class SymbolOutOfRange(HyaliteError):
    def __init__(self, symbol: int, alphabet_len: int):
        super().__init__(f"encoded symbol {symbol} is out of range for alphabet_len "
                         f"{alphabet_len} (must be in 0..{alphabet_len})")
        self.symbol = symbol
        self.alphabet_len = alphabet_len


class IncompleteBuilder(HyaliteError):
    def __init__(self, field: str):
        super().__init__(f"database builder is missing required field: {field}")
        self.field = field


class EmptyDatabase(HyaliteError):
    pass


class TracebackBudgetExceeded(HyaliteError):
    """src/error.rs:90-95, 147-154."""

    def __init__(self, needed_bytes: int, budget_bytes: int):
        super().__init__(f"traceback needs {needed_bytes} bytes but the budget is {budget_bytes}; "
                         "raise max_bytes (the linear-space path is not yet available)")
        self.needed_bytes = needed_bytes
        self.budget_bytes = budget_bytes


class BackendUnavailable(HyaliteError):
    def __init__(self, backend: str = "hip"):
        super().__init__(f"backend {backend} is not available on this build/device")
        self.backend = backend


class InvalidBackendName(HyaliteError):
    def __init__(self, name: str):
        super().__init__(f"unrecognised backend name {name!r}; expected one of: "
                         "auto, scalar, sse4.1, avx2, neon, hip")
        self.name = name


class ScanContractViolation(Exception):
    """A scan-time contract breach (over-long query, symbol out of range). The reference panics
    here (src/database.rs:79-89)."""


class DatabaseImageError(ValueError):
    """Database.load: the file is not a database image, or a damaged one (size or checksum)."""


class DeviceError(RuntimeError):
    """A HIP runtime failure or a failed device-side self-check."""


class StaticScratchLimit(RuntimeError):
    """Database.scan_scores_static: the problem needs more than the library's static device
    array holds (about 4 bytes per database residue plus 4 per target, 64 MiB in all)."""


class HipLibraryMissing(RuntimeError):
    """libhyalite_hip.so has not been built; run hyalite_amd.build()."""


# ---- plain types --------------------------------------------------------------------------------

class Mode(enum.IntEnum):
    """src/mode.rs:11-27."""
    Sw = 0
    Nw = 1
    Hw = 2
    Ov = 3
    Shw = 4

    def is_local(self) -> bool:
        return self is Mode.Sw

    def code(self) -> str:
        return {0: "SW", 1: "NW", 2: "HW", 3: "OV", 4: "SHW"}[int(self)]

    def __str__(self) -> str:
        return self.code()


class SearchType(enum.IntEnum):
    """src/search.rs:13-44. ``SearchType.Score``, ``SearchType.ScoreEnd``, or
    ``SearchType.Alignment(max_bytes)`` (an :class:`AlignmentSearch`)."""
    Score = 0
    ScoreEnd = 1

    def tracks_end(self) -> bool:
        return self is SearchType.ScoreEnd

    def max_bytes(self) -> Optional[int]:
        return None


@dataclass(frozen=True)
class AlignmentSearch:
    """SearchType::Alignment { max_bytes } (src/search.rs:23-26): ends are tracked and the
    aligned scans are available; the build proves the traceback budget once."""
    budget: int

    def tracks_end(self) -> bool:
        return True

    def max_bytes(self) -> Optional[int]:
        return self.budget

    def __int__(self) -> int:
        return 2


SearchType.Alignment = staticmethod(lambda max_bytes: AlignmentSearch(int(max_bytes)))  # type: ignore


class AlignOp(enum.IntEnum):
    """src/align.rs:47-57. Ins consumes a query symbol, Del a target symbol."""
    Match = 0
    Mismatch = 1
    Ins = 2
    Del = 3


@dataclass(frozen=True, eq=False)
class Alignment:
    """src/align.rs:58-160: score, half-open spans, ops in forward order (a uint8 array of
    AlignOp values; ``ops`` gives them as AlignOp)."""
    score: int
    query_start: int
    query_end: int
    target_start: int
    target_end: int
    op_codes: np.ndarray

    @property
    def ops(self) -> list:
        return [AlignOp(int(v)) for v in self.op_codes]

    def _encode(self, extended: bool) -> str:
        letters = "=XID" if extended else "MMID"
        out, prev, run = [], None, 0
        for v in self.op_codes.tolist():
            c = letters[v]
            if c == prev:
                run += 1
            else:
                if prev is not None:
                    out.append(f"{run}{prev}")
                prev, run = c, 1
        if prev is not None:
            out.append(f"{run}{prev}")
        return "".join(out)

    def cigar(self) -> str:
        return self._encode(False)

    def cigar_extended(self) -> str:
        return self._encode(True)

    def __eq__(self, other) -> bool:
        if not isinstance(other, Alignment):
            return NotImplemented
        return ((self.score, self.query_start, self.query_end, self.target_start,
                 self.target_end) == (other.score, other.query_start, other.query_end,
                                      other.target_start, other.target_end)
                and np.array_equal(self.op_codes, other.op_codes))

    def as_tuple(self):
        """(score, query_start, query_end, target_start, target_end, [op values])."""
        return (self.score, self.query_start, self.query_end, self.target_start,
                self.target_end, [int(v) for v in self.op_codes])


@dataclass(frozen=True)
class AlignedHit:
    """src/align.rs:83-90."""
    db_index: int
    alignment: Alignment


class ScoreWidth(enum.IntEnum):
    """src/width.rs:26-34."""
    I8 = 8
    I16 = 16
    I32 = 32

    def max_abs(self) -> int:
        return {8: 127, 16: 32767, 32: 536_870_911}[int(self)]

    def bytes(self) -> int:
        return int(self) // 8

    def __str__(self) -> str:
        return f"i{int(self)}"


class Backend(enum.Enum):
    """The backend this library provides (a new variant of src/backend.rs:27-36)."""
    Hip = "hip"

    def name_str(self) -> str:
        return self.value

    def is_available(self) -> bool:
        return bool(load_library().hy_backend_available())

    def __str__(self) -> str:
        return self.value


@dataclass(frozen=True)
class BackendChoice:
    """src/backend.rs:126-155. ``BackendChoice.AUTO`` or ``BackendChoice.force(Backend.Hip)``."""
    forced: Optional[Backend] = None

    @staticmethod
    def force(backend: Backend) -> "BackendChoice":
        return BackendChoice(backend)

    @staticmethod
    def parse(name: str) -> "BackendChoice":
        out = ctypes.c_int32(0)
        rc = load_library().hy_backend_parse(name.encode(), ctypes.byref(out))
        if rc != 0:
            raise InvalidBackendName(name)
        if out.value == 9:
            return BackendChoice(Backend.Hip)
        if out.value == 0:
            return BackendChoice(None)
        raise BackendUnavailable(name.strip().lower())


BackendChoice.AUTO = BackendChoice(None)  # type: ignore[attr-defined]


@dataclass(frozen=True)
class BestHit:
    """src/hit.rs:16-25. Ends are 0-based inclusive, None when nothing is aligned."""
    score: int
    db_index: int
    query_end: Optional[int]
    target_end: Optional[int]

    def __str__(self) -> str:
        s = f"score={self.score} db={self.db_index}"
        if self.query_end is not None and self.target_end is not None:
            s += f" query_end={self.query_end} target_end={self.target_end}"
        return s


class Scoring:
    """Substitution matrix plus affine gaps, validated like src/scoring.rs:65-100."""

    def __init__(self, alphabet_len: int, matrix: Sequence[int], gap_open: int, gap_ext: int):
        if alphabet_len == 0:
            raise EmptyAlphabet("alphabet length must be greater than zero")
        if len(matrix) != alphabet_len * alphabet_len:
            raise MatrixShape(f"substitution matrix must have {alphabet_len * alphabet_len} "
                              f"entries (alphabet_len {alphabet_len} squared), but got "
                              f"{len(matrix)}")
        if gap_open < 0 or gap_ext < 0:
            raise NegativeGapPenalty(f"gap penalties must be non-negative magnitudes, but got "
                                     f"gap_open={gap_open}, gap_ext={gap_ext}")
        if gap_open < gap_ext:
            raise GapOpenLessThanExtend(f"gap_open ({gap_open}) must be >= gap_ext ({gap_ext}) "
                                        "(Opal issue #28)")
        self._al = int(alphabet_len)
        self._matrix = np.ascontiguousarray(np.asarray(matrix, dtype=np.int32))
        self._go = int(gap_open)
        self._ge = int(gap_ext)

    new = classmethod(lambda cls, al, matrix, go, ge: cls(al, matrix, go, ge))

    def alphabet_len(self) -> int:
        return self._al

    def gap_open(self) -> int:
        return self._go

    def gap_ext(self) -> int:
        return self._ge

    def entry_bounds(self) -> tuple[int, int]:
        return int(self._matrix.min()), int(self._matrix.max())

    def score(self, q: int, t: int) -> int:
        if not (0 <= q < self._al and 0 <= t < self._al):
            raise IndexError(f"symbol index out of range: q={q}, t={t}, alphabet_len={self._al}")
        return int(self._matrix[q * self._al + t])

    def matrix(self) -> np.ndarray:
        return self._matrix


# ---- the library --------------------------------------------------------------------------------

_lib = None
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_sizep = ctypes.POINTER(ctypes.c_size_t)


class _HyAlignment(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int32), ("reserved", ctypes.c_uint32),
                ("query_start", ctypes.c_uint64), ("query_end", ctypes.c_uint64),
                ("target_start", ctypes.c_uint64), ("target_end", ctypes.c_uint64),
                ("n_ops", ctypes.c_uint64), ("ops_offset", ctypes.c_uint64)]


_ALIGNMENT_DTYPE = np.dtype([("score", "<i4"), ("reserved", "<u4"), ("query_start", "<u8"),
                             ("query_end", "<u8"), ("target_start", "<u8"),
                             ("target_end", "<u8"), ("n_ops", "<u8"), ("ops_offset", "<u8")])


class _HyServerConfig(ctypes.Structure):
    _fields_ = [("scratch_budget_bytes", ctypes.c_uint64), ("max_lanes", ctypes.c_uint32),
                ("max_batch", ctypes.c_uint32), ("prealloc", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32 * 3)]


class _HyServerResult(ctypes.Structure):
    _fields_ = [("score", ctypes.c_int32), ("status", ctypes.c_int32),
                ("db_index", ctypes.c_uint64), ("query_end", ctypes.c_int64),
                ("target_end", ctypes.c_int64), ("submit_ns", ctypes.c_uint64),
                ("admit_ns", ctypes.c_uint64), ("done_ns", ctypes.c_uint64),
                ("batch_size", ctypes.c_uint32), ("lane", ctypes.c_uint32)]


class _HyServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("query_residues", ctypes.c_uint64),
                ("lanes", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("lane_bytes", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64),
                ("peak_queries", ctypes.c_uint64), ("inflight_bytes", ctypes.c_uint64)]


class _HyTuning(ctypes.Structure):
    _fields_ = [("tile_rows", ctypes.c_int32), ("waves", ctypes.c_int32),
                ("batch_tile_rows", ctypes.c_int32), ("batch_wgs_per_cu", ctypes.c_int32),
                ("batch_chunk", ctypes.c_int32), ("batch_pairs", ctypes.c_int32)]


@dataclass(frozen=True)
class Tuning:
    """Launch shapes of the packed kernels held by a Scratch (hy_tuning). Shapes change speed
    only: every shape gives the same hits."""
    tile_rows: int = 32
    waves: int = 1
    batch_tile_rows: int = 32
    batch_wgs_per_cu: int = 0
    batch_chunk: int = 256
    batch_pairs: int = 1

    def _c(self) -> _HyTuning:
        return _HyTuning(self.tile_rows, self.waves, self.batch_tile_rows, self.batch_wgs_per_cu,
                         self.batch_chunk, self.batch_pairs)

    @staticmethod
    def _from(c: _HyTuning) -> "Tuning":
        return Tuning(c.tile_rows, c.waves, c.batch_tile_rows, c.batch_wgs_per_cu, c.batch_chunk,
                      c.batch_pairs)


def load_library():
    """Load libhyalite_hip.so (raises HipLibraryMissing when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise HipLibraryMissing(f"{path} not found: build it with hyalite_amd.build() "
                                "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.hy_backend_available.restype = ctypes.c_int
    lib.hy_backend_available.argtypes = []
    lib.hy_backend_parse.restype = ctypes.c_int
    lib.hy_backend_parse.argtypes = [ctypes.c_char_p, _i32p]
    lib.hy_db_build.restype = ctypes.c_int
    lib.hy_db_build.argtypes = [ctypes.POINTER(_u8p), _sizep, ctypes.c_size_t, ctypes.c_uint32,
                                _i32p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int32,
                                ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                ctypes.c_size_t, ctypes.c_uint32, ctypes.POINTER(vp), _i64p]
    lib.hy_db_retain.restype = ctypes.c_int
    lib.hy_db_retain.argtypes = [vp]
    lib.hy_db_free.restype = None
    lib.hy_db_free.argtypes = [vp]
    lib.hy_db_info.restype = ctypes.c_int
    lib.hy_db_info.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), _u64p, _u64p, _u64p,
                               ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    lib.hy_db_save.restype = ctypes.c_int
    lib.hy_db_save.argtypes = [vp, ctypes.c_char_p]
    lib.hy_db_load.restype = ctypes.c_int
    lib.hy_db_load.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(vp), _i64p]
    lib.hy_db_scoring.restype = ctypes.c_int
    lib.hy_db_scoring.argtypes = [vp, _i32p, ctypes.POINTER(ctypes.c_uint32), _i32p, _i32p, _u64p]
    lib.hy_db_partitions.restype = ctypes.c_int
    lib.hy_db_partitions.argtypes = [vp, _u64p, _u64p, _u64p]
    lib.hy_scratch_new.restype = ctypes.c_int
    lib.hy_scratch_new.argtypes = [vp, ctypes.POINTER(vp)]
    lib.hy_scratch_free.restype = None
    lib.hy_scratch_free.argtypes = [vp]
    lib.hy_scan.restype = ctypes.c_int
    lib.hy_scan.argtypes = [vp, vp, _u8p, ctypes.c_size_t, _i32p, _u64p, _i64p, _i64p]
    lib.hy_scan_all.restype = ctypes.c_int
    lib.hy_scan_all.argtypes = [vp, vp, _u8p, ctypes.c_size_t, _i32p, _i64p, _i64p]
    lib.hy_scan_scores.restype = ctypes.c_int
    lib.hy_scan_scores.argtypes = [vp, vp, _u8p, ctypes.c_size_t, _i32p]
    lib.hy_scan_scores_static.restype = ctypes.c_int
    lib.hy_scan_scores_static.argtypes = [vp, _u8p, ctypes.c_size_t, _i32p]
    lib.hy_scan_aligned.restype = ctypes.c_int
    lib.hy_scan_aligned.argtypes = [vp, vp, _u8p, ctypes.c_size_t, _u64p,
                                    ctypes.POINTER(_HyAlignment), _u8p, ctypes.c_size_t]
    lib.hy_db_aligned_ops_bytes.restype = ctypes.c_int
    lib.hy_db_aligned_ops_bytes.argtypes = [vp, ctypes.c_size_t, _u64p]
    lib.hy_scan_all_aligned.restype = ctypes.c_int
    lib.hy_scan_all_aligned.argtypes = [vp, vp, _u8p, ctypes.c_size_t, ctypes.c_void_p, _u8p,
                                        ctypes.c_size_t]
    lib.hy_pair_scratch_new.restype = ctypes.c_int
    lib.hy_pair_scratch_new.argtypes = [ctypes.POINTER(vp)]
    lib.hy_pair_scratch_free.restype = None
    lib.hy_pair_scratch_free.argtypes = [vp]
    lib.hy_align_pair.restype = ctypes.c_int
    lib.hy_align_pair.argtypes = [vp, _u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t, _i32p,
                                  ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int32,
                                  ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, _i32p, _i64p,
                                  _i64p, _i64p]
    lib.hy_align_pairs.restype = ctypes.c_int
    lib.hy_align_pairs.argtypes = [vp, ctypes.POINTER(_u8p), _sizep, ctypes.POINTER(_u8p), _sizep,
                                   ctypes.c_size_t, _i32p, ctypes.c_size_t, ctypes.c_uint32,
                                   ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                   ctypes.c_uint32, _i32p, _i64p, _i64p, _i64p, _u64p]
    lib.hy_align_pairs_flat.restype = ctypes.c_int
    lib.hy_align_pairs_flat.argtypes = [vp, _u8p, _u64p, _u8p, _u64p, ctypes.c_size_t, _i32p,
                                        ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32, _i32p,
                                        _i64p, _i64p, _i64p, _u64p]
    lib.hy_pair_scratch_profile_read.restype = ctypes.c_int
    lib.hy_pair_scratch_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_double), _u64p,
                                                 ctypes.c_int]
    lib.hy_align_pair_position_max.restype = ctypes.c_int
    lib.hy_align_pair_position_max.argtypes = [vp, _u8p, ctypes.c_size_t, _u8p, ctypes.c_size_t,
                                               _i32p, ctypes.c_size_t, ctypes.c_uint32,
                                               ctypes.c_int32, ctypes.c_int32, _i32p, _i32p,
                                               _i64p, _i64p]
    lib.hy_score2.restype = ctypes.c_int
    lib.hy_score2.argtypes = [_i32p, ctypes.c_size_t, ctypes.c_int32, ctypes.c_uint64,
                              ctypes.c_int32, ctypes.c_int32, _i32p, _i32p, _u64p]
    lib.hy_scan_batch.restype = ctypes.c_int
    lib.hy_scan_batch.argtypes = [vp, vp, ctypes.POINTER(_u8p), _sizep, ctypes.c_size_t, _i32p,
                                  _u64p, _i64p, _i64p]
    lib.hy_scratch_profile.restype = ctypes.c_int
    lib.hy_scratch_profile.argtypes = [vp, ctypes.c_int]
    lib.hy_scratch_profile_read.restype = ctypes.c_int
    lib.hy_scratch_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double), _u64p, ctypes.c_int]
    lib.hy_scratch_reserve_batch.restype = ctypes.c_int
    lib.hy_scratch_reserve_batch.argtypes = [vp, ctypes.c_size_t]
    lib.hy_scratch_batch_profile_read.restype = ctypes.c_int
    lib.hy_scratch_batch_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_double), _u64p,
                                                  ctypes.c_int]
    lib.hy_scratch_last_kernel.restype = ctypes.c_int
    lib.hy_scratch_last_kernel.argtypes = [vp, _i32p]
    lib.hy_scratch_get_tuning.restype = ctypes.c_int
    lib.hy_scratch_get_tuning.argtypes = [vp, ctypes.POINTER(_HyTuning)]
    lib.hy_scratch_set_tuning.restype = ctypes.c_int
    lib.hy_scratch_set_tuning.argtypes = [vp, ctypes.POINTER(_HyTuning)]
    lib.hy_scratch_autotune.restype = ctypes.c_int
    lib.hy_scratch_autotune.argtypes = [vp, vp, ctypes.POINTER(_u8p), _sizep, ctypes.c_size_t,
                                        ctypes.c_uint32, ctypes.POINTER(_HyTuning)]
    lib.hy_server_new.restype = ctypes.c_int
    lib.hy_server_new.argtypes = [vp, ctypes.POINTER(_HyServerConfig), ctypes.POINTER(vp)]
    lib.hy_server_submit.restype = ctypes.c_int
    lib.hy_server_submit.argtypes = [vp, _u8p, ctypes.c_size_t, _u64p]
    lib.hy_server_wait.restype = ctypes.c_int
    lib.hy_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int64,
                                   ctypes.POINTER(_HyServerResult)]
    lib.hy_server_stats_read.restype = ctypes.c_int
    lib.hy_server_stats_read.argtypes = [vp, ctypes.POINTER(_HyServerStats)]
    lib.hy_server_free.restype = None
    lib.hy_server_free.argtypes = [vp]
    lib.hy_strerror.restype = ctypes.c_char_p
    lib.hy_strerror.argtypes = [ctypes.c_int]
    _lib = lib
    return lib


_FIELDS = ["sequences", "scoring", "mode", "max_query_len"]


def _raise_build_error(rc: int, detail: int, alphabet_len: int, env_name: Optional[str],
                       budget: int = 0):
    lib = load_library()
    msg = lib.hy_strerror(rc).decode()
    if rc == 1:
        raise EmptyAlphabet(msg)
    if rc == 2:
        raise MatrixShape(msg)
    if rc == 3:
        raise NegativeGapPenalty(msg)
    if rc == 4:
        raise GapOpenLessThanExtend(msg)
    if rc == 5:
        raise ScoreRangeTooWide(detail)
    if rc == 6:
        raise SymbolOutOfRange(detail, alphabet_len)
    if rc == 7:
        raise IncompleteBuilder(_FIELDS[detail])
    if rc == 8:
        raise EmptyDatabase(msg)
    if rc == 9:
        raise BackendUnavailable()
    if rc == 10:
        raise InvalidBackendName(env_name or "")
    if rc == 11:
        raise TracebackBudgetExceeded(detail & ((1 << 64) - 1), budget)
    raise DeviceError(f"hy_db_build failed: {msg} (status {rc})")


def _check_scan(rc: int):
    if rc == 0:
        return
    msg = load_library().hy_strerror(rc).decode()
    if rc in (100, 101, 108):
        raise ScanContractViolation(msg)
    if rc == 110:
        raise StaticScratchLimit(msg)
    raise DeviceError(f"{msg} (status {rc})")


def _as_u8(q) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(q, dtype=np.uint8))
    return a if a.size else np.zeros(1, dtype=np.uint8)


class DatabaseBuilder:
    """src/database.rs:657-735. Required: sequences, scoring, mode, max_query_len."""

    def __init__(self):
        self._seqs = None
        self._scoring = None
        self._mode = None
        self._st = SearchType.Score
        self._mql = None
        self._backend = None

    def sequences(self, seqs) -> "DatabaseBuilder":
        self._seqs = [np.ascontiguousarray(np.asarray(s, dtype=np.uint8)) for s in seqs]
        self._flat = None
        return self

    def sequences_flat(self, flat: np.ndarray, offsets: np.ndarray) -> "DatabaseBuilder":
        """Sequences given as one residue buffer plus n+1 offsets (no per-sequence objects).
        Sequence k is flat[offsets[k]:offsets[k+1]]: the offsets start at 0, never decrease and
        end within the buffer (ValueError otherwise, before anything is built)."""
        buf = np.asarray(flat)
        off = np.asarray(offsets)
        if buf.ndim != 1 or off.ndim != 1:
            raise ValueError("sequences_flat takes a 1-D residue buffer and 1-D offsets")
        if off.size == 0:
            raise ValueError("sequences_flat needs n + 1 offsets (at least one)")
        if off.dtype.kind not in "ui" or (off.dtype.kind == "i" and (off < 0).any()):
            raise ValueError("offsets must be non-negative integers")
        off = off.astype(np.uint64)
        if int(off[0]) != 0:
            raise ValueError("offsets must start at 0")
        if off.size > 1 and (off[1:] < off[:-1]).any():
            raise ValueError("offsets must not decrease")
        if int(off[-1]) > buf.size:
            raise ValueError(f"offsets end at {int(off[-1])}, past the buffer's {buf.size} bytes")
        self._flat = (np.ascontiguousarray(buf, dtype=np.uint8), np.ascontiguousarray(off))
        self._seqs = []
        return self

    def scoring(self, scoring: Scoring) -> "DatabaseBuilder":
        self._scoring = scoring
        return self

    def mode(self, mode: Mode) -> "DatabaseBuilder":
        self._mode = Mode(mode)
        return self

    def search_type(self, st) -> "DatabaseBuilder":
        self._st = st if isinstance(st, AlignmentSearch) else SearchType(st)
        return self

    def max_query_len(self, n: int) -> "DatabaseBuilder":
        self._mql = int(n)
        return self

    def backend(self, choice: BackendChoice) -> "DatabaseBuilder":
        self._backend = choice
        return self

    def build(self) -> "Database":
        lib = load_library()
        have = 0
        if self._seqs is not None:
            have |= 1
        if self._scoring is not None:
            have |= 2
        if self._mode is not None:
            have |= 4
        if self._mql is not None:
            have |= 8
        flat = getattr(self, "_flat", None)
        keep = []
        if flat is not None:
            buf, off = flat
            if buf.size == 0:
                buf = np.zeros(1, dtype=np.uint8)
            keep.append(buf)
            n = len(off) - 1
            ptr_arr = (np.uint64(buf.ctypes.data) + off[:-1]).astype(np.uint64)
            len_arr = np.diff(off).astype(np.uint64)
            keep += [ptr_arr, len_arr]
            ptrs = ptr_arr.ctypes.data_as(ctypes.POINTER(_u8p))
            lens = len_arr.ctypes.data_as(_sizep)
            seq_lens = len_arr
        else:
            seqs = self._seqs or []
            n = len(seqs)
            ptrs = (_u8p * max(n, 1))()
            lens = (ctypes.c_size_t * max(n, 1))()
            for k, s in enumerate(seqs):
                a = s if s.size else np.zeros(1, dtype=np.uint8)
                keep.append(a)
                ptrs[k] = a.ctypes.data_as(_u8p)
                lens[k] = int(s.size)
            seq_lens = [int(s.size) for s in seqs]
        sc = self._scoring
        mat = sc.matrix() if sc is not None else np.zeros(1, dtype=np.int32)
        backend = 0
        if self._backend is not None and self._backend.forced is not None:
            backend = 9
        out = ctypes.c_void_p(None)
        detail = ctypes.c_int64(0)
        rc = lib.hy_db_build(ptrs, lens, n, have, mat.ctypes.data_as(_i32p),
                             mat.size if sc is not None else 0,
                             sc.alphabet_len() if sc is not None else 0,
                             sc.gap_open() if sc is not None else 0,
                             sc.gap_ext() if sc is not None else 0,
                             int(self._mode) if self._mode is not None else 0, int(self._st),
                             min(self._st.max_bytes() or 0, (1 << 64) - 1),
                             self._mql if self._mql is not None else 0, backend,
                             ctypes.byref(out), ctypes.byref(detail))
        if rc != 0:
            _raise_build_error(rc, detail.value, sc.alphabet_len() if sc else 0,
                               os.environ.get(BACKEND_ENV_VAR), self._st.max_bytes() or 0)
        return Database(out, sc, self._mode, self._st, seq_lens)


class Database:
    """An immutable, device-resident database (src/database.rs:44-63). Share it freely across
    threads; give each thread its own Scratch."""

    def __init__(self, handle: ctypes.c_void_p, scoring: Scoring, mode: Mode, st: SearchType,
                 lens: list[int]):
        self._h = handle
        self._scoring = scoring
        self._mode = mode
        self._st = st
        self._lens = lens
        lib = load_library()
        w, n, mql, mtl = ctypes.c_uint32(0), ctypes.c_uint64(0), ctypes.c_uint64(0), \
            ctypes.c_uint64(0)
        lib.hy_db_info(handle, ctypes.byref(w), ctypes.byref(n), ctypes.byref(mql),
                       ctypes.byref(mtl), None, None)
        self._width = ScoreWidth(w.value)
        self._n = n.value
        self._mql = mql.value
        self._mtl = mtl.value

    @staticmethod
    def builder() -> DatabaseBuilder:
        return DatabaseBuilder()

    def save(self, path: str) -> None:
        """Write the database's image (scheme, builder fields, residues; checksummed)."""
        rc = load_library().hy_db_save(self._h, os.fsencode(path))
        if rc != 0:
            raise OSError(f"cannot write database image {path!r}: "
                          f"{load_library().hy_strerror(rc).decode()}")

    @staticmethod
    def load(path: str, backend: Optional[BackendChoice] = None) -> "Database":
        """Rebuild a database from an image written by save(); the build's validation runs
        again on its contents."""
        lib = load_library()
        out = ctypes.c_void_p(None)
        detail = ctypes.c_int64(0)
        b = 9 if backend is not None and backend.forced is not None else 0
        rc = lib.hy_db_load(os.fsencode(path), b, ctypes.byref(out), ctypes.byref(detail))
        if rc == 109:
            raise DatabaseImageError(f"{path!r}: {lib.hy_strerror(rc).decode()}")
        if rc == 103:
            raise OSError(f"cannot read database image {path!r}")
        if rc != 0:
            with open(path, "rb") as f:  # the header names the alphabet for the error message
                head = f.read(16)
            al0 = int.from_bytes(head[12:16], "little") if len(head) == 16 else 0
            _raise_build_error(rc, detail.value, al0, os.environ.get(BACKEND_ENV_VAR))
        al, go, ge, budget = ctypes.c_uint32(0), ctypes.c_int32(0), ctypes.c_int32(0), \
            ctypes.c_uint64(0)
        lib.hy_db_scoring(out, None, ctypes.byref(al), ctypes.byref(go), ctypes.byref(ge),
                          ctypes.byref(budget))
        mat = np.zeros(al.value * al.value, dtype=np.int32)
        lib.hy_db_scoring(out, mat.ctypes.data_as(_i32p), None, None, None, None)
        mode, st = ctypes.c_uint32(0), ctypes.c_uint32(0)
        lib.hy_db_info(out, None, None, None, None, ctypes.byref(mode), ctypes.byref(st))
        stype = AlignmentSearch(budget.value) if st.value == 2 else SearchType(st.value)
        return Database(out, Scoring(al.value, mat, go.value, ge.value), Mode(mode.value), stype,
                        None)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.hy_db_free(h)
            self._h = None

    def clone(self) -> "Database":
        load_library().hy_db_retain(self._h)
        return Database(self._h, self._scoring, self._mode, self._st, self._lens)

    # accessors (src/database.rs:581-656)
    def sequence_count(self) -> int:
        return self._n

    def mode(self) -> Mode:
        return self._mode

    def search_type(self) -> SearchType:
        return self._st

    def scoring(self) -> Scoring:
        return self._scoring

    def score_width(self) -> ScoreWidth:
        return self._width

    def backend(self) -> Backend:
        return Backend.Hip

    def max_query_len(self) -> int:
        return self._mql

    def max_target_len(self) -> int:
        return self._mtl

    def partition_sizes(self) -> dict:
        a, b, c = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_uint64(0)
        load_library().hy_db_partitions(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {ScoreWidth.I8: a.value, ScoreWidth.I16: b.value, ScoreWidth.I32: c.value}

    # scans
    def scan(self, scratch: "Scratch", query) -> BestHit:
        q = np.asarray(query, dtype=np.uint8)
        qa = _as_u8(q)
        s, idx = ctypes.c_int32(0), ctypes.c_uint64(0)
        qe, te = ctypes.c_int64(0), ctypes.c_int64(0)
        _check_scan(load_library().hy_scan(self._h, scratch._h, qa.ctypes.data_as(_u8p), q.size,
                                           ctypes.byref(s), ctypes.byref(idx), ctypes.byref(qe),
                                           ctypes.byref(te)))
        return BestHit(s.value, idx.value, None if qe.value < 0 else qe.value,
                       None if te.value < 0 else te.value)

    def scan_all_arrays(self, scratch: "Scratch", query):
        """(scores int32[n], query_ends int64[n], target_ends int64[n]); -1 means None."""
        q = np.asarray(query, dtype=np.uint8)
        qa = _as_u8(q)
        n = self._n
        scores = np.empty(n, dtype=np.int32)
        qe = np.empty(n, dtype=np.int64)
        te = np.empty(n, dtype=np.int64)
        _check_scan(load_library().hy_scan_all(self._h, scratch._h, qa.ctypes.data_as(_u8p),
                                               q.size, scores.ctypes.data_as(_i32p),
                                               qe.ctypes.data_as(_i64p), te.ctypes.data_as(_i64p)))
        return scores, qe, te

    def scan_all(self, scratch: "Scratch", query, out: Optional[list] = None) -> list:
        s, qe, te = self.scan_all_arrays(scratch, query)
        hits = [BestHit(int(a), k, None if b < 0 else int(b), None if c < 0 else int(c))
                for k, (a, b, c) in enumerate(zip(s, qe, te))]
        if out is not None:
            out.clear()
            out.extend(hits)
            return out
        return hits

    def scan_scores_array(self, scratch: "Scratch", query) -> np.ndarray:
        q = np.asarray(query, dtype=np.uint8)
        qa = _as_u8(q)
        scores = np.empty(self._n, dtype=np.int32)
        _check_scan(load_library().hy_scan_scores(self._h, scratch._h, qa.ctypes.data_as(_u8p),
                                                  q.size, scores.ctypes.data_as(_i32p)))
        return scores

    def scan_scores_static(self, query) -> np.ndarray:
        """scan_scores without a Scratch and without any device allocation (hy_scan_scores_static:
        a static device array per GPU, the null stream); StaticScratchLimit when the database and
        query need more than it holds. Calls on one GPU must not overlap."""
        q = np.asarray(query, dtype=np.uint8)
        qa = _as_u8(q)
        scores = np.empty(self._n, dtype=np.int32)
        _check_scan(load_library().hy_scan_scores_static(
            self._h, qa.ctypes.data_as(_u8p), q.size, scores.ctypes.data_as(_i32p)))
        return scores

    def scan_scores(self, scratch: "Scratch", query, out: Optional[list] = None) -> list:
        s = [int(v) for v in self.scan_scores_array(scratch, query)]
        if out is not None:
            out.clear()
            out.extend(s)
            return out
        return s

    def scan_aligned(self, scratch: "Scratch", query) -> AlignedHit:
        """src/database.rs:528-540: the winner of the score pass with its canonical alignment
        (the alignment align::traceback gives for the pair), traced back on the device."""
        q = np.asarray(query, dtype=np.uint8)
        qa = _as_u8(q)
        ops = np.empty(max(int(q.size) + int(self._mtl), 1), dtype=np.uint8)
        idx = ctypes.c_uint64(0)
        al = _HyAlignment()
        _check_scan(load_library().hy_scan_aligned(
            self._h, scratch._h, qa.ctypes.data_as(_u8p), q.size, ctypes.byref(idx),
            ctypes.byref(al), ops.ctypes.data_as(_u8p), ops.size))
        return AlignedHit(int(idx.value), Alignment(
            int(al.score), int(al.query_start), int(al.query_end), int(al.target_start),
            int(al.target_end), ops[:al.n_ops].copy()))

    def scan_all_aligned_arrays(self, scratch: "Scratch", query):
        """(records, ops): one hy_alignment record per target in database order (a structured
        array with score, spans, n_ops, ops_offset) and the op bytes they index."""
        q = np.asarray(query, dtype=np.uint8)
        qa = _as_u8(q)
        lib = load_library()
        need = ctypes.c_uint64(0)
        lib.hy_db_aligned_ops_bytes(self._h, q.size, ctypes.byref(need))
        ops = np.empty(max(need.value, 1), dtype=np.uint8)
        recs = np.zeros(self._n, dtype=_ALIGNMENT_DTYPE)
        _check_scan(lib.hy_scan_all_aligned(self._h, scratch._h, qa.ctypes.data_as(_u8p), q.size,
                                            recs.ctypes.data, ops.ctypes.data_as(_u8p),
                                            ops.size))
        return recs, ops

    def scan_all_aligned(self, scratch: "Scratch", query, out: Optional[list] = None) -> list:
        """src/database.rs:551-559: one Alignment per database sequence, in database order."""
        recs, ops = self.scan_all_aligned_arrays(scratch, query)
        res = [Alignment(int(r["score"]), int(r["query_start"]), int(r["query_end"]),
                         int(r["target_start"]), int(r["target_end"]),
                         ops[int(r["ops_offset"]):int(r["ops_offset"]) + int(r["n_ops"])])
               for r in recs]
        if out is not None:
            out.clear()
            out.extend(res)
            return out
        return res

    def scan_batch_arrays(self, scratch: "Scratch", flat, offsets):
        """Batched form over queries given as one residue buffer plus n+1 offsets (query k =
        flat[offsets[k]:offsets[k+1]]), with no per-query Python objects: (scores int32[n],
        db_indices uint64[n], query_ends int64[n], target_ends int64[n]), -1 meaning None."""
        buf = np.ascontiguousarray(np.asarray(flat, dtype=np.uint8).ravel())
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.uint64))
        if off.ndim != 1 or off.size == 0 or int(off[0]) != 0 or \
                (off.size > 1 and (off[1:] < off[:-1]).any()) or int(off[-1]) > buf.size:
            raise ValueError("offsets must start at 0, never decrease and end within the buffer")
        nq = off.size - 1
        keep = buf if buf.size else np.zeros(1, dtype=np.uint8)
        ptrs = (np.uint64(keep.ctypes.data) + off[:-1]).astype(np.uint64)
        lens = np.diff(off).astype(np.uint64)
        s = np.empty(max(nq, 1), dtype=np.int32)
        idx = np.empty(max(nq, 1), dtype=np.uint64)
        qe = np.empty(max(nq, 1), dtype=np.int64)
        te = np.empty(max(nq, 1), dtype=np.int64)
        _check_scan(load_library().hy_scan_batch(
            self._h, scratch._h, ptrs.ctypes.data_as(ctypes.POINTER(_u8p)),
            lens.ctypes.data_as(_sizep), nq, s.ctypes.data_as(_i32p),
            idx.ctypes.data_as(_u64p), qe.ctypes.data_as(_i64p), te.ctypes.data_as(_i64p)))
        return s[:nq], idx[:nq], qe[:nq], te[:nq]

    def scan_batch(self, scratch: "Scratch", queries) -> list:
        """Batched form: one hit per query, each equal to scan() of that query."""
        qs = [np.ascontiguousarray(np.asarray(q, dtype=np.uint8)) for q in queries]
        nq = len(qs)
        ptrs = (_u8p * max(nq, 1))()
        lens = (ctypes.c_size_t * max(nq, 1))()
        keep = []
        for k, q in enumerate(qs):
            a = q if q.size else np.zeros(1, dtype=np.uint8)
            keep.append(a)
            ptrs[k] = a.ctypes.data_as(_u8p)
            lens[k] = q.size
        s = np.empty(max(nq, 1), dtype=np.int32)
        idx = np.empty(max(nq, 1), dtype=np.uint64)
        qe = np.empty(max(nq, 1), dtype=np.int64)
        te = np.empty(max(nq, 1), dtype=np.int64)
        _check_scan(load_library().hy_scan_batch(self._h, scratch._h, ptrs, lens, nq,
                                                 s.ctypes.data_as(_i32p), idx.ctypes.data_as(_u64p),
                                                 qe.ctypes.data_as(_i64p),
                                                 te.ctypes.data_as(_i64p)))
        return [BestHit(int(s[k]), int(idx[k]), None if qe[k] < 0 else int(qe[k]),
                        None if te[k] < 0 else int(te[k])) for k in range(nq)]


class Scratch:
    """Per-thread working memory: a HIP stream plus device buffers (src/database.rs:954-999)."""

    def __init__(self, db: Database):
        out = ctypes.c_void_p(None)
        rc = load_library().hy_scratch_new(db._h, ctypes.byref(out))
        if rc != 0:
            raise DeviceError(load_library().hy_strerror(rc).decode())
        self._h = out
        self._db = db  # keeps the database alive while the scratch exists

    def profile(self, enable: bool = True) -> None:
        """Time the per-target kernels of every scan with HIP events on this scratch's stream."""
        rc = load_library().hy_scratch_profile(self._h, 1 if enable else 0)
        if rc != 0:
            raise DeviceError(load_library().hy_strerror(rc).decode())

    def reserve_batch(self, max_queries: int) -> None:
        """Reserve buffers for batches of up to max_queries queries (batched form)."""
        rc = load_library().hy_scratch_reserve_batch(self._h, max_queries)
        if rc != 0:
            raise DeviceError(load_library().hy_strerror(rc).decode())

    def batch_profile_read(self, reset: bool = False) -> tuple[float, int]:
        """(persistent batch kernel ms, batches timed), accumulated while profiling."""
        a, n = ctypes.c_double(0), ctypes.c_uint64(0)
        load_library().hy_scratch_batch_profile_read(self._h, ctypes.byref(a), ctypes.byref(n),
                                                     1 if reset else 0)
        return a.value, n.value

    def last_kernel(self) -> str:
        """The packed kernel of this scratch's last one-query scan: "none", "packed",
        "pipelined", "chained" or "ends" (hy_scratch_last_kernel)."""
        k = ctypes.c_int32(-1)
        rc = load_library().hy_scratch_last_kernel(self._h, ctypes.byref(k))
        if rc != 0:
            raise DeviceError(load_library().hy_strerror(rc).decode())
        return {-1: "none", 0: "packed", 1: "pipelined", 2: "chained", 3: "ends"}[k.value]

    def tuning(self) -> Tuning:
        c = _HyTuning()
        rc = load_library().hy_scratch_get_tuning(self._h, ctypes.byref(c))
        if rc != 0:
            raise DeviceError(load_library().hy_strerror(rc).decode())
        return Tuning._from(c)

    def set_tuning(self, t: Tuning) -> None:
        """Set launch shapes; invalid shapes raise ValueError."""
        rc = load_library().hy_scratch_set_tuning(self._h, ctypes.byref(t._c()))
        if rc == 103:
            raise ValueError(f"invalid tuning {t}")
        if rc != 0:
            raise DeviceError(load_library().hy_strerror(rc).decode())

    def autotune(self, queries, plain: bool = True, batch: bool = True) -> Tuning:
        """Time candidate shapes on a sample of `queries` against this scratch's database and keep
        the fastest (hy_scratch_autotune; HYALITE_HIP_TUNE_CACHE names an optional cache file)."""
        raw = [np.asarray(q, dtype=np.uint8) for q in queries]
        qs = [_as_u8(q) for q in raw]
        n = len(qs)
        ptrs = (_u8p * max(n, 1))()
        lens = (ctypes.c_size_t * max(n, 1))()
        for k in range(n):
            ptrs[k] = qs[k].ctypes.data_as(_u8p)
            lens[k] = raw[k].size
        c = _HyTuning()
        forms = (1 if plain else 0) | (2 if batch else 0)
        _check_scan(load_library().hy_scratch_autotune(self._db._h, self._h, ptrs, lens, n, forms,
                                                       ctypes.byref(c)))
        return Tuning._from(c)

    def profile_read(self, reset: bool = False) -> tuple[float, float, int]:
        """(inter-kernel ms, second-stream ms, scans timed), accumulated; the second stream runs
        the long targets and, beside the packed kernel, the targets that fail its width proof."""
        a, b, n = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_uint64(0)
        load_library().hy_scratch_profile_read(self._h, ctypes.byref(a), ctypes.byref(b),
                                               ctypes.byref(n), 1 if reset else 0)
        return a.value, b.value, n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.hy_scratch_free(h)
            self._h = None


# ---- serving form -------------------------------------------------------------------------------

@dataclass(frozen=True)
class ServedHit:
    """One served query: its hit (equal to Database.scan of the query alone) and its timeline
    in seconds on the steady clock: queued (submit -> admitted to a batch) and latency
    (submit -> result ready)."""
    hit: BestHit
    queued_s: float
    latency_s: float
    batch_size: int
    lane: int


@dataclass(frozen=True)
class ServerStats:
    submitted: int
    completed: int
    batches: int
    query_residues: int
    lanes: int
    max_batch: int
    lane_bytes: int
    peak_bytes: int = 0
    peak_queries: int = 0
    inflight_bytes: int = 0


class ServerTimeout(TimeoutError):
    pass


# ---- pairwise alignment (src/kernel.rs:285-575, 741-792) ------------------------------------------

class PairScratch:
    """PairScratch (src/kernel.rs:285-311): a HIP stream and device buffers that grow to the
    largest problem seen and are reused. One per thread."""

    def __init__(self):
        out = ctypes.c_void_p(None)
        rc = load_library().hy_pair_scratch_new(ctypes.byref(out))
        if rc == 9:
            raise BackendUnavailable()
        if rc != 0:
            raise DeviceError(load_library().hy_strerror(rc).decode())
        self._h = out

    def profile_read(self, reset: bool = False) -> tuple[float, int]:
        """(device ms, batches) of the align_pairs calls through this scratch: HIP events around
        the kernels, after the uploads and before the copy back."""
        ms, n = ctypes.c_double(0), ctypes.c_uint64(0)
        load_library().hy_pair_scratch_profile_read(self._h, ctypes.byref(ms), ctypes.byref(n),
                                                    1 if reset else 0)
        return ms.value, n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.hy_pair_scratch_free(h)
            self._h = None


def _raise_pair_error(rc: int, detail: int, scoring: Scoring):
    if rc in (1, 2, 3, 4, 5, 6):
        _raise_build_error(rc, detail, scoring.alphabet_len(), None)
    raise DeviceError(f"{load_library().hy_strerror(rc).decode()} (status {rc})")


def _st_code(search_type) -> int:
    return int(search_type)


def align_pair_with(scratch: PairScratch, query, target, scoring: Scoring, mode: Mode,
                    search_type) -> BestHit:
    """src/kernel.rs:354-420: one pair on the device (one workgroup of the band kernel, or one
    wavefront for alphabets above 63 symbols); errors as the reference's (SymbolOutOfRange,
    ScoreRangeTooWide)."""
    q, t = np.asarray(query, dtype=np.uint8), np.asarray(target, dtype=np.uint8)
    qa, ta = _as_u8(q), _as_u8(t)
    mat = scoring.matrix()
    s, qe, te, det = ctypes.c_int32(0), ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_int64(0)
    rc = load_library().hy_align_pair(
        scratch._h, qa.ctypes.data_as(_u8p), q.size, ta.ctypes.data_as(_u8p), t.size,
        mat.ctypes.data_as(_i32p), mat.size, scoring.alphabet_len(), scoring.gap_open(),
        scoring.gap_ext(), int(Mode(mode)), _st_code(search_type), ctypes.byref(s),
        ctypes.byref(qe), ctypes.byref(te), ctypes.byref(det))
    if rc != 0:
        _raise_pair_error(rc, det.value, scoring)
    return BestHit(s.value, 0, None if qe.value < 0 else qe.value,
                   None if te.value < 0 else te.value)


def align_pair(query, target, scoring: Scoring, mode: Mode, search_type) -> BestHit:
    """src/kernel.rs:331-352 (a fresh PairScratch per call; reuse one with align_pair_with)."""
    return align_pair_with(PairScratch(), query, target, scoring, mode, search_type)


def align_pairs_flat(queries, query_offsets, targets, target_offsets, scoring: Scoring,
                     mode: Mode, search_type, scratch: Optional[PairScratch] = None):
    """align_pairs over concatenated sequences (pair k = queries[query_offsets[k]:
    query_offsets[k+1]] against the same slice of targets): (scores int32[n], query_ends
    int64[n], target_ends int64[n]), -1 standing for None."""
    qf = _as_u8(np.asarray(queries, dtype=np.uint8).ravel())
    tf = _as_u8(np.asarray(targets, dtype=np.uint8).ravel())
    qo = np.ascontiguousarray(query_offsets, dtype=np.uint64)
    to = np.ascontiguousarray(target_offsets, dtype=np.uint64)
    if qo.size != to.size or qo.size == 0:
        raise ValueError("query and target offsets need n + 1 entries each")
    n = qo.size - 1
    scores = np.empty(max(n, 1), dtype=np.int32)
    qe = np.empty(max(n, 1), dtype=np.int64)
    te = np.empty(max(n, 1), dtype=np.int64)
    mat = scoring.matrix()
    det, bad = ctypes.c_int64(0), ctypes.c_uint64(0)
    sc = scratch or PairScratch()
    rc = load_library().hy_align_pairs_flat(
        sc._h, qf.ctypes.data_as(_u8p), qo.ctypes.data_as(_u64p), tf.ctypes.data_as(_u8p),
        to.ctypes.data_as(_u64p), n, mat.ctypes.data_as(_i32p), mat.size, scoring.alphabet_len(),
        scoring.gap_open(), scoring.gap_ext(), int(Mode(mode)), _st_code(search_type),
        scores.ctypes.data_as(_i32p), qe.ctypes.data_as(_i64p), te.ctypes.data_as(_i64p),
        ctypes.byref(det), ctypes.byref(bad))
    if rc != 0:
        _raise_pair_error(rc, det.value, scoring)
    return scores[:n], qe[:n], te[:n]


def _flatten(seqs):
    arrs = [np.asarray(x, dtype=np.uint8).ravel() for x in seqs]
    off = np.zeros(len(arrs) + 1, dtype=np.uint64)
    if arrs:
        np.cumsum([a.size for a in arrs], out=off[1:])
    flat = np.concatenate(arrs) if arrs else np.zeros(0, dtype=np.uint8)
    return flat, off


def align_pairs(pairs, scoring: Scoring, mode: Mode, search_type, out: Optional[list] = None,
                scratch: Optional[PairScratch] = None) -> list:
    """src/kernel.rs:429-452: one BestHit per (query, target) pair, db_index = pair index; the
    first failing pair's error is raised."""
    if out is not None:
        out.clear()
    qf, qo = _flatten([q for q, _ in pairs])
    tf, to = _flatten([t for _, t in pairs])
    scores, qe, te = align_pairs_flat(qf, qo, tf, to, scoring, mode, search_type, scratch)
    n = len(pairs)
    hits = [BestHit(int(scores[k]), k, None if qe[k] < 0 else int(qe[k]),
                    None if te[k] < 0 else int(te[k])) for k in range(n)]
    if out is not None:
        out.extend(hits)
        return out
    return hits


def align_pair_position_max_with(scratch: PairScratch, query, target, scoring: Scoring,
                                 out: Optional[list] = None) -> BestHit:
    """src/kernel.rs:490-575: SW per-target-position maxima into `out` (cleared and refilled);
    the hit's score is their maximum, its target_end the first position holding it (None when
    the best is 0); query_end is not tracked."""
    q, t = np.asarray(query, dtype=np.uint8), np.asarray(target, dtype=np.uint8)
    qa, ta = _as_u8(q), _as_u8(t)
    col = np.zeros(max(t.size, 1), dtype=np.int32)
    mat = scoring.matrix()
    s, te, det = ctypes.c_int32(0), ctypes.c_int64(0), ctypes.c_int64(0)
    if out is not None:
        out.clear()
    rc = load_library().hy_align_pair_position_max(
        scratch._h, qa.ctypes.data_as(_u8p), q.size, ta.ctypes.data_as(_u8p), t.size,
        mat.ctypes.data_as(_i32p), mat.size, scoring.alphabet_len(), scoring.gap_open(),
        scoring.gap_ext(), col.ctypes.data_as(_i32p), ctypes.byref(s), ctypes.byref(te),
        ctypes.byref(det))
    if rc != 0:
        _raise_pair_error(rc, det.value, scoring)
    if out is not None:
        out.extend(int(v) for v in col[:t.size])
    return BestHit(s.value, 0, None, None if te.value < 0 else te.value)


def align_pair_position_max(query, target, scoring: Scoring,
                            out: Optional[list] = None) -> BestHit:
    return align_pair_position_max_with(PairScratch(), query, target, scoring, out)


def score2(colmax, best_score: int, best_target_end: int, matrix_max: int,
           min_score: int) -> Optional[tuple]:
    """src/kernel.rs:741-792: (score2, target position) of the best secondary peak, or None."""
    a = np.ascontiguousarray(np.asarray(colmax, dtype=np.int32))
    n = a.size
    if n == 0:
        a = np.zeros(1, dtype=np.int32)
    found, s2, pos = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_uint64(0)
    rc = load_library().hy_score2(a.ctypes.data_as(_i32p), n, best_score, best_target_end,
                                  matrix_max, min_score, ctypes.byref(found), ctypes.byref(s2),
                                  ctypes.byref(pos))
    if rc != 0:
        raise DeviceError(load_library().hy_strerror(rc).decode())
    return (s2.value, pos.value) if found.value else None


class Server:
    """Serving form over a resident database (hy_server_*): queries are queued, admitted by free
    lanes in batches of up to max_batch and scanned while other batches run. Admission is by
    device bytes: scratch_budget_bytes (0 = no limit) covers the lanes and each request in flight
    (its query copy, key and result, and beyond 512 rows its winner boundary row); lanes are
    added on demand up to max_lanes (prealloc=True builds every lane the budget holds at once)."""

    def __init__(self, db: Database, scratch_budget_bytes: int = 0, max_lanes: int = 2,
                 max_batch: int = 64, prealloc: bool = False):
        cfg = _HyServerConfig(scratch_budget_bytes, max_lanes, max_batch, int(bool(prealloc)))
        out = ctypes.c_void_p(None)
        rc = load_library().hy_server_new(db._h, ctypes.byref(cfg), ctypes.byref(out))
        if rc == 103:
            raise ValueError("invalid server configuration")
        if rc != 0:
            raise DeviceError(load_library().hy_strerror(rc).decode())
        self._h = out
        self._db = db

    def submit(self, query) -> int:
        q = np.asarray(query, dtype=np.uint8)
        buf = _as_u8(q)
        t = ctypes.c_uint64(0)
        _check_scan(load_library().hy_server_submit(self._h, buf.ctypes.data_as(_u8p), q.size,
                                                    ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout: Optional[float] = None) -> ServedHit:
        r = _HyServerResult()
        us = -1 if timeout is None else int(timeout * 1e6)
        rc = load_library().hy_server_wait(self._h, ticket, us, ctypes.byref(r))
        if rc == 107:
            raise ServerTimeout(f"ticket {ticket} not completed within {timeout} s")
        if rc == 103:
            raise KeyError(ticket)
        _check_scan(rc)
        _check_scan(r.status)
        hit = BestHit(int(r.score), int(r.db_index),
                      None if r.query_end < 0 else int(r.query_end),
                      None if r.target_end < 0 else int(r.target_end))
        return ServedHit(hit, (r.admit_ns - r.submit_ns) / 1e9, (r.done_ns - r.submit_ns) / 1e9,
                         int(r.batch_size), int(r.lane))

    def stats(self) -> ServerStats:
        st = _HyServerStats()
        load_library().hy_server_stats_read(self._h, ctypes.byref(st))
        return ServerStats(st.submitted, st.completed, st.batches, st.query_residues, st.lanes,
                           st.max_batch, st.lane_bytes, st.peak_bytes, st.peak_queries,
                           st.inflight_bytes)

    def close(self) -> None:
        """Finish every queued query, then stop the lanes."""
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.hy_server_free(h)
        self._h = None

    def __enter__(self) -> "Server":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        self.close()
