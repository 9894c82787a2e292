"""Seeded synthetic workloads of the benchmark configurations.

Each generator returns a ``Workload`` with the scoring scheme, mode, search type, the database as
flat residues plus offsets, and the queries. Choices the configurations leave open are fixed here
and printed by the benchmark: RNG = numpy PCG64 seeded per configuration; protein background
frequencies = Robinson & Robinson (1991) for the 20 amino acids plus four rare symbols at 0.1%
each; log-normal target lengths use sigma = 0.75.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# Robinson & Robinson (1991) amino-acid background frequencies, order ARNDCQEGHILKMFPSTWYV.
AA_FREQ = np.array([7.805, 5.129, 4.487, 5.364, 1.925, 4.264, 6.295, 7.377, 2.199, 5.142,
                    9.019, 5.744, 2.243, 3.856, 5.203, 7.120, 5.841, 1.330, 3.216, 6.441])

LOGNORMAL_SIGMA = 0.75


@dataclass
class Workload:
    name: str
    al: int
    matrix: np.ndarray
    go: int
    ge: int
    mode: int
    search_type: int
    mql: int
    flat: np.ndarray
    off: np.ndarray
    queries: list = field(default_factory=list)
    note: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.off) - 1

    @property
    def residues(self) -> int:
        return int(self.off[-1])

    def lens(self) -> np.ndarray:
        return np.diff(self.off).astype(np.int64)

    def subsample(self, k: int, seed: int = 0) -> "Workload":
        """k targets drawn without replacement (for the bounded CPU baseline)."""
        rng = np.random.default_rng(seed)
        idx = np.sort(rng.choice(self.n, size=min(k, self.n), replace=False))
        lens = self.lens()[idx]
        off = np.zeros(len(idx) + 1, dtype=np.uint64)
        np.cumsum(lens, out=off[1:])
        flat = np.concatenate([self.flat[int(self.off[i]):int(self.off[i + 1])] for i in idx]) \
            if len(idx) else np.zeros(0, dtype=np.uint8)
        return Workload(self.name + f"[sample {len(idx)}]", self.al, self.matrix, self.go,
                        self.ge, self.mode, self.search_type, self.mql, flat.astype(np.uint8),
                        off, self.queries, self.note)


def dna_matrix() -> np.ndarray:
    m = np.full((4, 4), -1, dtype=np.int32)
    np.fill_diagonal(m, 2)
    return m.reshape(-1)


def protein_freqs() -> np.ndarray:
    p = np.concatenate([AA_FREQ / AA_FREQ.sum() * 0.996, np.full(4, 0.001)])
    return p / p.sum()


def protein_matrix(seed: int) -> np.ndarray:
    """Seeded symmetric 24x24 matrix: diagonal U{4..11}, off-diagonal U{-4..3}."""
    rng = np.random.default_rng(seed)
    m = rng.integers(-4, 4, size=(24, 24))
    m = np.triu(m, 1)
    m = m + m.T
    np.fill_diagonal(m, rng.integers(4, 12, size=24))
    return m.astype(np.int32).reshape(-1)


def _draw(rng, n, p=None, al=4) -> np.ndarray:
    if p is None:
        return rng.integers(0, al, size=n, dtype=np.uint8)
    cdf = np.cumsum(p)
    cdf[-1] = 1.0
    u = rng.random(n, dtype=np.float32) if n > 0 else np.zeros(0, dtype=np.float32)
    return np.searchsorted(cdf.astype(np.float32), u, side="right").astype(np.uint8).clip(0, len(p) - 1)


def _flat_from_lens(rng, lens, p=None, al=4):
    off = np.zeros(len(lens) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    return _draw(rng, int(off[-1]), p, al), off


def lognormal_lens(rng, n, median, lo, hi, sigma=LOGNORMAL_SIGMA) -> np.ndarray:
    x = rng.lognormal(np.log(median), sigma, size=n)
    return np.clip(np.rint(x), lo, hi).astype(np.int64)


def c1_dna(seed: int = 1) -> Workload:
    """Plain DNA SW ScoreEnd: 100-nt query vs 10,000 targets U{50..300}, +2/-1, 2/1."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(50, 301, size=10_000)
    flat, off = _flat_from_lens(rng, lens)
    q = _draw(rng, 100)
    return Workload("dna_sw_scoreend_100x10k", 4, dna_matrix(), 2, 1, 0, 1, 100, flat, off, [q])


def c2_protein(seed: int = 2, n_targets: int = 1_000_000, mql: int = 360) -> Workload:
    """Protein SW ScoreEnd: 360-aa query vs 1M log-normal(median 300) targets in [30, 5000]."""
    rng = np.random.default_rng(seed)
    p = protein_freqs()
    lens = lognormal_lens(rng, n_targets, 300, 30, 5000)
    flat, off = _flat_from_lens(rng, lens, p)
    q = _draw(rng, 360, p)
    return Workload(f"protein_sw_scoreend_360x{n_targets}", 24, protein_matrix(seed), 11, 1, 0,
                    1, mql, flat, off, [q],
                    note=f"lognormal sigma {LOGNORMAL_SIGMA}, Robinson-Robinson background")


def c3_dna_modes(seed: int = 3, n_targets: int = 2_000_000, mode: int = 1) -> Workload:
    """DNA short reads: 150-nt query vs targets U{100..1000}; 5% are mutated query copies."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(100, 1001, size=n_targets)
    q = _draw(rng, 150)
    n_mut = n_targets // 20
    mut_idx = np.sort(rng.choice(n_targets, size=n_mut, replace=False))
    muts = []
    for _ in range(n_mut):
        s = q.copy()
        k = rng.random(150) < 0.05
        s[k] = (s[k] + rng.integers(1, 4, size=int(k.sum()))) % 4
        nd = int(rng.integers(1, 4))
        pos = int(rng.integers(0, 150))
        if rng.random() < 0.5:
            s = np.concatenate([s[:pos], s[pos + nd:]])
        else:
            s = np.concatenate([s[:pos], _draw(rng, nd), s[pos:]])
        muts.append(s.astype(np.uint8))
    lens = lens.copy()
    for k, i in enumerate(mut_idx):
        lens[i] = len(muts[k])
    flat, off = _flat_from_lens(rng, lens)
    for k, i in enumerate(mut_idx):
        flat[int(off[i]):int(off[i + 1])] = muts[k]
    return Workload(f"dna_mode{mode}_scores_150x{n_targets}", 4, dna_matrix(), 2, 1, mode, 0,
                    150, flat, off, [q])


def c3_scaled(base: Workload, mode: int, scale: int = 25) -> Workload:
    """The DNA short-read database of c3_dna_modes with the scheme scaled by `scale` (+2s / -s,
    gap_open 2s, gap_ext s): at scale 25 and a 150-nt query, the NW/HW/SHW width proof gives i16
    up to 504-nt targets and i32 beyond, so one database splits across both widths."""
    name = base.name.replace("mode1", {1: "NW", 2: "HW", 3: "OV", 4: "SHW"}.get(mode, str(mode)))
    return Workload(name.replace("_scores_", f"_scaled{scale}_scores_"), base.al,
                    (base.matrix * scale).astype(np.int32), base.go * scale, base.ge * scale,
                    mode, 1, base.mql, base.flat, base.off, base.queries)


def c4_batch_queries(seed: int = 4, nq: int = 256) -> list:
    rng = np.random.default_rng(seed)
    p = protein_freqs()
    lens = np.rint(np.exp(rng.uniform(np.log(20), np.log(2000), size=nq))).astype(int)
    return [_draw(rng, int(L), p) for L in lens]


def c5_serving_queries(seed: int = 5, nq: int = 64) -> list:
    rng = np.random.default_rng(seed)
    p = protein_freqs()
    lens = lognormal_lens(rng, nq, 350, 20, 3000)
    return [_draw(rng, int(L), p) for L in lens]


def c4_batch(n_targets: int = 1_000_000, nq: int = 256) -> Workload:
    """Batched form: 256 protein queries, lengths log-uniform in [20, 2000], against the protein
    database of the single-query protein workload rebuilt with max_query_len 2000."""
    w = c2_protein(n_targets=n_targets, mql=2000)
    w.name = f"protein_sw_scoreend_batch{nq}x{n_targets}"
    w.extra["single_query"] = w.queries[0]  # the 360-aa query of the single-query workload
    w.queries = c4_batch_queries(nq=nq)
    w.note += "; queries log-uniform [20, 2000] (seed 4)"
    return w


def c5_serving(n_targets: int = 2_000_000, nq: int = 64) -> Workload:
    """Serving form: protein queries with log-normal lengths (median 350, clipped [20, 3000])
    against 2M targets drawn like the single-query protein database (seed 5), max_query_len 3000,
    so targets of 2979 residues or more need the 32-bit width."""
    w = c2_protein(seed=5, n_targets=n_targets, mql=3000)
    w.name = f"protein_sw_scoreend_serving{nq}x{n_targets}"
    w.queries = c5_serving_queries(nq=nq)
    w.note += "; queries log-normal median 350 [20, 3000] (seed 5)"
    return w


@dataclass
class PairBatch:
    """Independent (query, target) pairs as one buffer per side with n + 1 offsets each."""
    name: str
    al: int
    matrix: np.ndarray
    go: int
    ge: int
    queries: np.ndarray
    qoff: np.ndarray
    targets: np.ndarray
    toff: np.ndarray

    @property
    def n(self) -> int:
        return len(self.qoff) - 1

    @property
    def cells(self) -> int:
        return int((np.diff(self.qoff).astype(np.int64) * np.diff(self.toff).astype(np.int64)).sum())

    def pair(self, k: int):
        return (self.queries[int(self.qoff[k]):int(self.qoff[k + 1])],
                self.targets[int(self.toff[k]):int(self.toff[k + 1])])


def dna_read_pairs(n: int = 200_000, seed: int = 6, read_len: int = 150,
                   window_len: int = 500) -> PairBatch:
    """Mate-rescue-shaped DNA pairs for align_pairs: reads of read_len against windows of
    window_len (uniform i.i.d. ACGT), every other window holding its read at a random offset
    with 5% substitutions; the DNA scheme (+2 / -1, gap_open 2, gap_ext 1)."""
    rng = np.random.default_rng(seed)
    q = rng.integers(0, 4, size=n * read_len, dtype=np.uint8)
    t = rng.integers(0, 4, size=n * window_len, dtype=np.uint8)
    for k in range(0, n, 2):
        at = k * window_len + int(rng.integers(0, window_len - read_len + 1))
        r = q[k * read_len:(k + 1) * read_len].copy()
        sub = rng.random(read_len) < 0.05
        r[sub] = (r[sub] + rng.integers(1, 4, size=int(sub.sum()), dtype=np.uint8)) % 4
        t[at:at + read_len] = r
    qoff = (np.arange(n + 1, dtype=np.uint64) * read_len)
    toff = (np.arange(n + 1, dtype=np.uint64) * window_len)
    return PairBatch(f"dna_read_pairs_{read_len}x{window_len}x{n}", 4, dna_matrix(), 2, 1, q,
                     qoff, t, toff)


def protein_pairs(n: int = 100_000, seed: int = 7) -> PairBatch:
    """Protein pairs: queries of 250-350 and targets of 300-400 residues drawn at background
    frequencies, the protein matrix of the single-query workload (gap_open 11, gap_ext 1)."""
    rng = np.random.default_rng(seed)
    p = protein_freqs()
    lq = rng.integers(250, 351, size=n)
    lt = rng.integers(300, 401, size=n)
    q = _draw(rng, int(lq.sum()), p)
    t = _draw(rng, int(lt.sum()), p)
    qoff = np.zeros(n + 1, dtype=np.uint64)
    toff = np.zeros(n + 1, dtype=np.uint64)
    np.cumsum(lq, out=qoff[1:])
    np.cumsum(lt, out=toff[1:])
    return PairBatch(f"protein_pairs_250-350x300-400x{n}", 24, protein_matrix(2), 11, 1, q, qoff,
                     t, toff)


# ---- The reference's own scan benchmark shapes (tiny databases, many reads) -------------------

CR4_TSO = "AAGCAGTGGTATCAACGCAGAGTACATGGG"  # the 10x template-switch oligo
CR4_READ_LEN = 91


def _lcg(state: list) -> int:
    """The corpus PRNG of benches/cr4_scan.rs:54-60 (state in a one-element list)."""
    state[0] = (state[0] * 6364136223846793005 + 1442695040888963407) & ((1 << 64) - 1)
    return state[0] >> 33


def _encode_acgtn(s: str) -> np.ndarray:
    return np.array(["ACGT".find(c) if c in "ACGT" else 4 for c in s], dtype=np.uint8)


def cr4_adapters(count: int = 64) -> list:
    """benches/cr4_scan.rs:66-74: the TSO, poly-A/T/G 30-mers, then random 28-33 nt fillers."""
    v = [_encode_acgtn(CR4_TSO), np.zeros(30, np.uint8), np.full(30, 3, np.uint8),
         np.full(30, 2, np.uint8)]
    st = [0x00C0FFEE]
    while len(v) < count:
        n = 28 + _lcg(st) % 6
        v.append(np.array([_lcg(st) % 4 for _ in range(n)], dtype=np.uint8))
    return v


def cr4_reads(count: int) -> list:
    """benches/cr4_scan.rs:78-93: 91-nt reads, every third TSO-led, every third ending in 25 A."""
    tso = _encode_acgtn(CR4_TSO)
    st = [0x0000BEEF]
    out = []
    for i in range(count):
        r = np.array([_lcg(st) % 4 for _ in range(CR4_READ_LEN)], dtype=np.uint8)
        if i % 3 == 0:
            r[:len(tso)] = tso
        elif i % 3 == 1:
            r[CR4_READ_LEN - 25:] = 0
        out.append(r)
    return out


def cr4_matrix(scale: int = 1) -> np.ndarray:
    """STAR's CellRanger4 scoring (benches/cr4_scan.rs:26-36): +1 match, -2 mismatch or N, N-N 0."""
    m = np.full((5, 5), -2, dtype=np.int32)
    np.fill_diagonal(m, 1)
    m[4, 4] = 0
    return (m * scale).reshape(-1)
