"""Shared test helpers: scoring schemes, sequence generators, FASTA fixtures.

Mirrors the helpers of hyalite's tests/common/mod.rs (identity matrices, the DNA test scheme, all
short sequences, FASTA parsing, CR4 scoring, BLOSUM62) so the parity tests read like the
reference's own.
"""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

MODE_SW, MODE_NW, MODE_HW, MODE_OV, MODE_SHW = 0, 1, 2, 3, 4
ALL_MODES = (MODE_SW, MODE_NW, MODE_HW, MODE_OV, MODE_SHW)
MODE_NAMES = {MODE_SW: "SW", MODE_NW: "NW", MODE_HW: "HW", MODE_OV: "OV", MODE_SHW: "SHW"}


def identity_matrix(al: int, m: int, x: int) -> list[int]:
    """tests/common/mod.rs:13-20."""
    v = [x] * (al * al)
    for i in range(al):
        v[i * al + i] = m
    return v


def dna_scheme():
    """(al, matrix, go, ge): match +2, mismatch -1, gap_open 2, gap_ext 1 (tests/common/mod.rs:23-25)."""
    return 4, identity_matrix(4, 2, -1), 2, 1


def all_sequences(alphabet: int, max_len: int) -> list[list[int]]:
    """Every sequence of length 0..=max_len over `alphabet` symbols (tests/common/mod.rs:28-45)."""
    out = [[]]
    frontier = [[]]
    for _ in range(max_len):
        nxt = []
        for s in frontier:
            for sym in range(alphabet):
                nxt.append(s + [sym])
        out.extend(nxt)
        frontier = nxt
    return out


def lcg_next(state: list[int]) -> int:
    """The deterministic PRNG of tests/sweeps.rs:23-28 (state held in a one-element list)."""
    state[0] = (state[0] * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
    return state[0] >> 33


def lcg_seq(state: list[int], length: int, al: int) -> list[int]:
    return [lcg_next(state) % al for _ in range(length)]


def encode_dna(seq: str) -> list[int]:
    """A,C,G,T,N -> 0..4, skipping whitespace (tests/common/mod.rs:174-187)."""
    table = {"A": 0, "C": 1, "G": 2, "T": 3, "N": 4}
    return [table[c.upper()] for c in seq if not c.isspace()]


def parse_fasta(path: str) -> list[tuple[str, list[int]]]:
    """(id, encoded sequence) per record (tests/common/mod.rs:189-210)."""
    records, ident, seq = [], None, []
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if line.startswith(">"):
                if ident is not None:
                    records.append((ident, encode_dna("".join(seq))))
                    seq = []
                ident = line[1:].split()[0] if line[1:].split() else ""
            else:
                seq.append(line.strip())
    if ident is not None:
        records.append((ident, encode_dna("".join(seq))))
    return records


def phix() -> list[int]:
    recs = parse_fasta(os.path.join(GOLDEN, "phix174_NC_001422.1.fasta"))
    assert len(recs) == 1 and len(recs[0][1]) == 5386
    return recs[0][1]


def cr4_adapters():
    return parse_fasta(os.path.join(GOLDEN, "cr4_adapters.fa"))


def cr4_scheme():
    """STAR CellRanger4 scoring over A,C,G,T,N (tests/common/mod.rs:214-227)."""
    mat = [
        1, -2, -2, -2, -2,
        -2, 1, -2, -2, -2,
        -2, -2, 1, -2, -2,
        -2, -2, -2, 1, -2,
        -2, -2, -2, -2, 0,
    ]
    return 5, mat, 2, 2


AMINO_ACIDS = "ARNDCQEGHILKMFPSTWYV"

# The published NCBI BLOSUM62 log-odds matrix (Henikoff & Henikoff 1992), AMINO_ACIDS order.
BLOSUM62 = [
    4, -1, -2, -2, 0, -1, -1, 0, -2, -1, -1, -1, -1, -2, -1, 1, 0, -3, -2, 0,
    -1, 5, 0, -2, -3, 1, 0, -2, 0, -3, -2, 2, -1, -3, -2, -1, -1, -3, -2, -3,
    -2, 0, 6, 1, -3, 0, 0, 0, 1, -3, -3, 0, -2, -3, -2, 1, 0, -4, -2, -3,
    -2, -2, 1, 6, -3, 0, 2, -1, -1, -3, -4, -1, -3, -3, -1, 0, -1, -4, -3, -3,
    0, -3, -3, -3, 9, -3, -4, -3, -3, -1, -1, -3, -1, -2, -3, -1, -1, -2, -2, -1,
    -1, 1, 0, 0, -3, 5, 2, -2, 0, -3, -2, 1, 0, -3, -1, 0, -1, -2, -1, -2,
    -1, 0, 0, 2, -4, 2, 5, -2, 0, -3, -3, 1, -2, -3, -1, 0, -1, -3, -2, -2,
    0, -2, 0, -1, -3, -2, -2, 6, -2, -4, -4, -2, -3, -3, -2, 0, -2, -2, -3, -3,
    -2, 0, 1, -1, -3, 0, 0, -2, 8, -3, -3, -1, -2, -1, -2, -1, -2, -2, 2, -3,
    -1, -3, -3, -3, -1, -3, -3, -4, -3, 4, 2, -3, 1, 0, -3, -2, -1, -3, -1, 3,
    -1, -2, -3, -4, -1, -2, -3, -4, -3, 2, 4, -2, 2, 0, -3, -2, -1, -2, -1, 1,
    -1, 2, 0, -1, -3, 1, 1, -2, -1, -3, -2, 5, -1, -3, -1, 0, -1, -3, -2, -2,
    -1, -1, -2, -3, -1, 0, -2, -3, -2, 1, 2, -1, 5, 0, -2, -1, -1, -1, -1, 1,
    -2, -3, -3, -3, -2, -3, -3, -3, -1, 0, 0, -3, 0, 6, -4, -2, -2, 1, 3, -1,
    -1, -2, -2, -1, -3, -1, -1, -2, -2, -3, -3, -1, -2, -4, 7, -1, -1, -4, -3, -2,
    1, -1, 1, 0, -1, 0, 0, 0, -1, -2, -2, 0, -1, -2, -1, 4, 1, -3, -2, -2,
    0, -1, 0, -1, -1, -1, -1, -2, -2, -1, -1, -1, -1, -2, -1, 1, 5, -2, -2, 0,
    -3, -3, -4, -4, -2, -2, -3, -2, -2, -3, -2, -3, -1, 1, -4, -3, -2, 11, 2, -3,
    -2, -2, -2, -3, -2, -1, -2, -3, 2, -1, -1, -2, -1, 3, -3, -2, -2, 2, 7, -2,
    0, -3, -3, -3, -1, -2, -2, -3, -3, 3, 1, -2, 1, -1, -2, -2, 0, -3, -2, 4,
]


def encode_protein(seq: str) -> list[int]:
    return [AMINO_ACIDS.index(c.upper()) for c in seq if not c.isspace()]


UBIQUITIN = "MQIFVKTLTGKTITLEVEPSDTIENVKAKIQDKEGIPPDQQRLIFAGKQLEDGRTLSDYNIQKESTLHLVLRLRGG"


def random_db(rng: np.random.Generator, n: int, lo: int, hi: int, al: int) -> list[np.ndarray]:
    lens = rng.integers(lo, hi + 1, size=n)
    return [rng.integers(0, al, size=int(L), dtype=np.uint8) for L in lens]
