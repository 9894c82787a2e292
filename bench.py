"""Benchmark of the MI355X database scan (one GPU).

Workload (default): protein SW ScoreEnd scan — one 360-aa query against 1,000,000 targets with
log-normal lengths (median 300, sigma 0.75, clipped to 30..5000), 24-letter alphabet at
Robinson-Robinson background frequencies plus 4 rare symbols, seeded symmetric matrix
(diagonal 4..11, off-diagonal -4..3), gap_open 11, gap_ext 1. A step is one Database::scan of
the query over the whole resident database (best hit with its end positions).

Prints one JSON line: GCUPS = |q| * sum|t| / seconds / 1e9 over real cells only, the VALU roofline
of the dominant kernel (timed with HIP events on its own stream), and the CPU oracle timed on a
bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import hyalite_amd as hy  # noqa: E402
from hyalite_amd import workloads as wl  # noqa: E402

# MI355X: 256 CUs x 64 lanes x 2.4 GHz = 39.3e12 32-bit VALU lane-ops/s; packed 16-bit
# instructions carry two elements, 78.6e12 element-ops/s.
VALU_PEAK_TOPS_I32 = 256 * 64 * 2.4e9 / 1e12
VALU_PEAK_TOPS_PK16 = 2 * VALU_PEAK_TOPS_I32
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dominant kernel, measured with rocprofv3 --pmc (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, the factors a 256 MiB device copy gives); the run and the counters are in
# profiles/k_inter16_protein_sw_360x1M.md.
MEASURED_TRAFFIC_BYTES = {
    ("protein_sw_scoreend_360x1000000", "k_inter16<SW,24,1>"): 50.14e9,
}


def ops_per_cell(mode: int) -> int:
    """Algorithmic integer ops per DP cell: E = 2 sub + max, F = 2 sub + max, diag add,
    H = 2 max, plus the local-mode clamp."""
    return 10 if mode == 0 else 9


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def host_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def cpu_baseline(w: wl.Workload, budget_s: float = 3.0, max_targets: int = 60_000):
    """The AVX2 restatement of the reference's avx2 i16 inter-sequence backend (16 saturating
    lanes, database-order batches, Precomputed table built before the timed scan) on a bounded
    random sample of the same database, all host threads, plus a 1-thread and a length-sorted
    run. Every timed result is checked against the scalar oracle on the same sample first."""
    import oracle
    threads = host_threads()
    q = w.queries[0]
    k = 4000
    while True:
        sub = w.subsample(k, seed=17)
        scores, secs = oracle.avx2_scan_scores_flat(sub.flat, sub.off, q, w.al, w.matrix, w.go,
                                                    w.ge, w.mode, threads=threads)
        if secs >= budget_s / 3 or k >= min(w.n, max_targets):
            break
        k = min(w.n, max_targets, int(k * max(2.0, budget_s / max(secs, 1e-3))))
    want = oracle.scan_all_flat(sub.flat, sub.off, q, w.al, w.matrix, w.go, w.ge, w.mode,
                                ends=False, threads=threads)
    if not np.array_equal(scores, want):
        raise RuntimeError("AVX2 baseline disagrees with the scalar oracle")
    cells = len(q) * sub.residues
    sorted_scores, sorted_secs = oracle.avx2_scan_scores_flat(
        sub.flat, sub.off, q, w.al, w.matrix, w.go, w.ge, w.mode, threads=threads, sort=True)
    if not np.array_equal(sorted_scores, want):
        raise RuntimeError("sorted AVX2 baseline disagrees with the scalar oracle")
    one = sub.subsample(max(200, sub.n // max(threads, 1)), seed=3)
    one_scores, one_secs = oracle.avx2_scan_scores_flat(one.flat, one.off, q, w.al, w.matrix,
                                                        w.go, w.ge, w.mode, threads=1)
    return {
        "value": round(cells / secs / 1e9, 3),
        "unit": "GCUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{sub.n} random targets ({sub.residues} residues) of the same database; AVX2 "
                  f"16 x i16 restatement of the reference avx2 backend, database-order batches, "
                  f"{threads} threads on {cpu_model()}, scan {secs:.2f} s; checked against the "
                  f"scalar oracle",
        "one_thread_gcups": round(len(q) * one.residues / one_secs / 1e9, 3),
        "length_sorted_gcups": round(cells / sorted_secs / 1e9, 3),
    }


def copy_calibration(mib: int) -> int:
    """One device-to-device copy of `mib` MiB (for scaling HBM byte counters: a plain copy
    fetches and writes exactly that many bytes)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": n}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--targets", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (the scan would "
              "shard targets across GPUs and all-reduce one 64-bit key per query)",
              file=sys.stderr)
        return 2

    t0 = time.perf_counter()
    w = wl.c2_protein(n_targets=args.targets)
    gen_s = time.perf_counter() - t0
    q = w.queries[0]
    t0 = time.perf_counter()
    db = (hy.Database.builder().sequences_flat(w.flat, w.off)
          .scoring(hy.Scoring(w.al, w.matrix, w.go, w.ge)).mode(hy.Mode(w.mode))
          .search_type(hy.SearchType(w.search_type)).max_query_len(w.mql)
          .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())
    build_s = time.perf_counter() - t0
    sc = hy.Scratch(db)
    print(f"# workload {w.name}: {w.n} targets, {w.residues} residues, query {len(q)}, "
          f"gen {gen_s:.1f}s, build+upload {build_s:.1f}s, partitions "
          f"{ {str(k): v for k, v in db.partition_sizes().items()} }", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        hit = db.scan(sc, q)
    # Device-wide synchronisation around the timed region (what torch.cuda.synchronize does),
    # called directly so the benchmark does not pay for importing torch.
    import ctypes
    _hip = ctypes.CDLL("libamdhip64.so")

    def sync():
        if _hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    sc.profile(True)
    sc.profile_read(reset=True)
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hit = db.scan(sc, q)
    sync()
    dt = time.perf_counter() - t0
    inter_ms, tail_ms, nscan = sc.profile_read(reset=True)
    print(f"# hit {hit}", file=sys.stderr, flush=True)

    cells = len(q) * w.residues
    ms_per_step = dt * 1e3 / args.steps
    gcups = cells / (dt / args.steps) / 1e9
    # Dominant kernel: the inter-sequence kernel over all but the long-target tail.
    lens = w.lens()
    inter_ms_avg = inter_ms / max(nscan, 1)
    tile = (24, 1)
    if os.environ.get("HYALITE_HIP_TILE"):
        tile = tuple(int(x) for x in os.environ["HYALITE_HIP_TILE"].split(","))
    tail_ms_avg = tail_ms / max(nscan, 1)
    roof = None
    if inter_ms_avg > 0:
        # Cells of the inter kernel = all targets except the long tail it hands to the systolic
        # kernel (same threshold as the library: max(1024, 8 * mean length)).
        thr = max(1024, int(8.0 * lens.mean()))
        inter_cells = len(q) * int(lens[lens <= thr].sum())
        achieved = inter_cells * ops_per_cell(w.mode) / (inter_ms_avg / 1e3) / 1e12
        kernel = f"k_inter16<SW,{tile[0]},{tile[1]}>"
        roof = {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(VALU_PEAK_TOPS_PK16, 2),
            "unit": "TOP/s",
            "frac": round(achieved / VALU_PEAK_TOPS_PK16, 4),
            "traffic": MEASURED_TRAFFIC_BYTES.get((w.name, kernel)),
            "kernel": kernel,
            "peak_note": "packed 16-bit VALU element-ops (2 per lane per instruction)",
            "kernel_ms": round(inter_ms_avg, 3),
            "ops_per_cell": ops_per_cell(w.mode),
            "cells_per_launch": inter_cells,
            "hbm_gbs_algorithmic": round(int(lens[lens <= thr].sum()) / (inter_ms_avg / 1e3) / 1e9,
                                         2),
        }
    out = {
        "metric": "GCUPS on 1 MI355X (plain scan kernel, protein SW ScoreEnd)",
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int16 (packed pairs, saturating)",
        "data": "synthetic (seeded; see hyalite_amd/workloads.py)",
        "config": {"workload": w.name, "query_len": len(q), "targets": w.n,
                   "residues": w.residues, "mode": "SW", "search_type": "ScoreEnd",
                   "lengths": "lognormal median 300 sigma 0.75 clip 30..5000",
                   "form": "plain (one query per launch)"},
        "roofline": roof,
        "tail_kernel_ms": round(tail_ms_avg, 3),
        "hit": {"score": hit.score, "db_index": hit.db_index, "query_end": hit.query_end,
                "target_end": hit.target_end},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
        out["gpu_vs_cpu"] = round(gcups / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
