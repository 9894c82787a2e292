"""Benchmark of the MI355X database scan (one GPU): batched form vs the plain kernel, and the
serving form's latency at a fixed arrival rate.

Headline workload (the `value`): the batched form — 256 protein queries with lengths log-uniform
in [20, 2000] against 1,000,000 protein targets (log-normal lengths, median 300, sigma 0.75,
clipped to 30..5000; 24-letter alphabet at Robinson-Robinson background frequencies plus 4 rare
symbols; seeded symmetric matrix, diagonal 4..11, off-diagonal -4..3; gap_open 11, gap_ext 1;
SW ScoreEnd; max_query_len 2000). A step is one Database::scan_batch of all 256 queries over the
resident database (best hit with end positions per query).

Beside it, on the same GPU in the same run (each part can be skipped with a flag):
  * plain: the same 256 queries scanned one Database::scan at a time (one query per launch);
  * single: the 360-aa query of the single-query protein workload on the same targets;
  * serving: Poisson arrivals at a fixed QPS of 64 queries (log-normal lengths, median 350,
    clipped [20, 3000]) against 2,000,000 targets (max_query_len 3000, so the longest targets
    need 32 bits), p50/p99 latency and sustained GCUPS, every served hit checked against the
    plain scan of that query alone;
  * other_configs: the DNA configurations — 100-nt SW ScoreEnd scan of 10,000 targets and the
    150-nt NW/HW/SHW/OV scan_scores sweep over 2,000,000 targets (5% mutated query copies);
  * traceback: the same 1,000,000 targets built with SearchType::Alignment — scan_aligned of
    the 360-aa query (winner plus its canonical alignment) against scan alone, and
    scan_all_aligned (one alignment per target, ops copied to the host);
  * db_image: the database saved to an on-disk image and loaded back (sizes and times);
  * pairs: align_pairs over 200,000 mate-rescue-shaped DNA pairs (150-nt reads x 500-nt
    windows) and 100,000 protein pairs (250-350 x 300-400), Score and ScoreEnd, device and
    host-inclusive rates; align_pair and align_pair_position_max latency on single pairs;
  * cpu_baseline: the AVX2 16 x i16 restatement of the reference avx2 backend on a bounded
    sample of the batched workload, on this host's cores; beside it the scalar traceback
    restatement on a sample of the traceback pairs, and for the pairs the SSE4.1 striped
    restatement of the reference's single-pair Score path plus the scalar pair DP for
    ScoreEnd, each checked against the device's result.

GCUPS = sum over queries of |q| * sum|t| / seconds / 1e9, real cells only. The roofline is the
VALU throughput of the persistent batch kernel, timed with HIP events on its stream.
"""
from __future__ import annotations

import argparse
import ctypes
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

METRIC = "GCUPS on 1 MI355X (batched vs plain kernel) + p50/p99 query latency at QPS"

# MI355X: 256 CUs x 64 lanes x 2.4 GHz = 39.3e12 32-bit VALU lane-ops/s; packed 16-bit
# instructions carry two elements, 78.6e12 element-ops/s.
VALU_PEAK_TOPS_I32 = 256 * 64 * 2.4e9 / 1e12
VALU_PEAK_TOPS_PK16 = 2 * VALU_PEAK_TOPS_I32

# HBM bytes per launch of the dominant kernel come from rocprofv3 --pmc runs (FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0, the factors a 256 MiB device copy gives on this GPU). Each measurement is
# recorded in its profile summary as a line "traffic record: {json}" naming the workload, the
# kernel, the bytes and the SHA-256 of the kernel sources it was measured on; a record whose
# sources have changed since is reported as stale instead of used.
TRAFFIC_SOURCES = ("hyalite_amd/csrc/hy_kernels.hip", "hyalite_amd/csrc/hy_device.h",
                   "hyalite_amd/csrc/hy_internal.h")


def kernel_sources_sha() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def traffic_record(workload: str, kernel: str) -> tuple:
    """(bytes or None, note) from the traffic records in profiles/*.md."""
    import glob
    sha = kernel_sources_sha()
    found = None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic record: "):
                    rec = json.loads(line[len("traffic record: "):].strip().strip("`"))
                    if rec.get("workload") == workload and rec.get("kernel") == kernel:
                        found = (rec, os.path.basename(path))
    if found is None:
        return None, "no traffic record for this kernel and workload"
    rec, where = found
    if rec.get("sources_sha256") != sha:
        return None, (f"stale: {where} measured kernel sources {rec.get('sources_sha256')}, "
                      f"these are {sha}")
    return float(rec["bytes"]), f"{where}, kernel sources {sha}"

SERVING_QPS = 10.0


def ops_per_cell(mode: int) -> int:
    """Integer ops per DP cell of the Gotoh recurrence with H - go computed once and shared by E
    and F: diag add, HG = H - go, E = max(HG, E - ge) (2), F = max(HG, F - ge) (2), H = 2 max,
    plus the local-mode clamp. (Counting H - go separately for E and F, as a scalar restatement
    does, gives 10 and 9.)"""
    return 9 if mode == 0 else 8


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


def physical_cores() -> tuple:
    """(physical cores, logical CPUs) of this process's CPU set, from the sysfs topology (a
    physical core = a distinct (package, core id) pair); (logical, logical) when it is absent."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f1, open(base + "core_id") as f2:
                cores.add((f1.read().strip(), f2.read().strip()))
        except OSError:
            return len(cpus), len(cpus)
    return max(1, len(cores)), len(cpus)


def spread(queries, k):
    """k queries spread evenly over the length-sorted list."""
    order = sorted(range(len(queries)), key=lambda i: len(queries[i]))
    if len(order) <= k:
        return [queries[i] for i in order]
    return [queries[order[int(round(j * (len(order) - 1) / (k - 1)))]] for j in range(k)]


def build_db(w: wl.Workload) -> hy.Database:
    return (hy.Database.builder().sequences_flat(w.flat, w.off)
            .scoring(hy.Scoring(w.al, w.matrix, w.go, w.ge)).mode(hy.Mode(w.mode))
            .search_type(hy.SearchType(w.search_type)).max_query_len(w.mql)
            .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())


def cpu_baseline(w: wl.Workload, queries, budget_s: float = 12.0, max_targets: int = 40_000):
    """The AVX2 restatement of the reference's avx2 i16 inter-sequence backend (16 saturating
    lanes, database-order batches, Precomputed table built before the timed scan) on a bounded
    random sample of the same database, for queries spread over the batch's length range, on all
    host threads; plus a one-thread and a length-sorted run. Every timed result is checked
    against the scalar oracle on the same sample first."""
    import oracle
    threads, logical = physical_cores()
    qs = spread(queries, 6)
    qres = sum(len(q) for q in qs)

    def timed(sub, thr, sort=False):
        total = 0.0
        out = []
        for q in qs:
            s, secs = oracle.avx2_scan_scores_flat(sub.flat, sub.off, q, w.al, w.matrix, w.go,
                                                   w.ge, w.mode, threads=thr, sort=sort)
            out.append(s)
            total += secs
        return out, total

    # Correctness first, on a small sample: the AVX2 scores equal the scalar oracle's.
    check = w.subsample(600, seed=5)
    for q, s in zip(qs, timed(check, threads)[0]):
        want = oracle.scan_all_flat(check.flat, check.off, q, w.al, w.matrix, w.go, w.ge, w.mode,
                                    ends=False, threads=threads)
        if not np.array_equal(s, want):
            raise RuntimeError("AVX2 baseline disagrees with the scalar oracle")
    # Then a sample sized from a short probe so the timed scan takes about budget_s / 3.
    probe = w.subsample(1500, seed=17)
    _, psecs = timed(probe, threads)
    k = int(min(w.n, max_targets, max(1500, 1500 * budget_s / 3 / max(psecs, 1e-3))))
    sub = w.subsample(k, seed=17)
    scores, secs = timed(sub, threads)
    cells = qres * sub.residues
    sorted_scores, sorted_secs = timed(sub, threads, sort=True)
    if any(not np.array_equal(a, b) for a, b in zip(scores, sorted_scores)):
        raise RuntimeError("sorted AVX2 baseline disagrees with the unsorted one")
    one = sub.subsample(max(200, sub.n // max(threads, 1)), seed=3)
    _, one_secs = timed(one, 1)
    return {
        "value": round(cells / secs / 1e9, 3),
        "unit": "GCUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(qs)} batch queries spread over lengths {len(qs[0])}..{len(qs[-1])} "
                  f"against {sub.n} random targets ({sub.residues} residues) of the same "
                  f"database; AVX2 16 x i16 restatement of the reference avx2 backend, "
                  f"database-order batches, one thread per physical core ({threads} cores "
                  f"of {logical} logical CPUs in this process's CPU set) on {cpu_model()}, "
                  f"scan {secs:.2f} s; checked against the scalar oracle",
        "logical_cpus": logical,
        "one_thread_gcups": round(qres * one.residues / one_secs / 1e9, 3),
        "length_sorted_gcups": round(cells / sorted_secs / 1e9, 3),
    }


def copy_calibration(mib: int) -> int:
    """One device-to-device copy of `mib` MiB (for scaling HBM byte counters: a plain copy
    fetches and writes exactly that many bytes)."""
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


def log(msg: str) -> None:
    print("# " + msg, file=sys.stderr, flush=True)


def serving(args, sync) -> dict:
    from hyalite_amd import loadgen
    t0 = time.perf_counter()
    w = wl.c5_serving(n_targets=args.serving_targets, nq=args.serving_queries)
    db = build_db(w)
    parts = {str(k): v for k, v in db.partition_sizes().items()}
    log(f"serving workload {w.name}: {w.n} targets, {w.residues} residues, partitions {parts}, "
        f"gen+build {time.perf_counter() - t0:.1f}s")
    qs = w.queries
    sc = hy.Scratch(db)
    db.scan(sc, qs[0])
    sync()
    t0 = time.perf_counter()
    alone = [db.scan(sc, q) for q in qs]
    sync()
    plain_s = time.perf_counter() - t0
    del sc
    cells = sum(len(q) for q in qs) * w.residues

    def run(prealloc: bool) -> dict:
        """The Poisson loads on one server: lanes built on demand within the byte budget, or
        (prealloc) every lane built at construction."""
        loads = []
        with hy.Server(db, max_lanes=args.serving_lanes, max_batch=args.serving_batch,
                       prealloc=prealloc) as srv:
            srv.wait(srv.submit(qs[0]))
            for qps in [args.qps] + list(args.qps_sweep):
                before = srv.stats()
                r = loadgen.run_poisson(srv, qs, qps=qps, seed=7)
                st = srv.stats()
                if [x.hit for x in r.served] != alone:
                    raise RuntimeError(f"a served hit at {qps} QPS differs from the plain scan")
                loads.append({
                    "qps": qps, "p50_ms": round(r.percentile_ms(50), 2),
                    "p99_ms": round(r.percentile_ms(99), 2),
                    "mean_ms": round(float(r.latencies().mean() * 1e3), 2),
                    "mean_batch": round(float(np.mean([x.batch_size for x in r.served])), 2),
                    "sustained_gcups": round(cells / r.wall_s / 1e9, 1),
                    "wall_s": round(r.wall_s, 3),
                    "lanes_built": st.lanes, "lanes_before": before.lanes,
                    "peak_queries_in_flight": st.peak_queries,
                    "peak_gb": round(st.peak_bytes / 1e9, 2),
                    "hits_equal_alone": True})
                log(f"serving ({'prealloc' if prealloc else 'on demand'}) at {qps} QPS: p50 "
                    f"{loads[-1]['p50_ms']} ms, p99 {loads[-1]['p99_ms']} ms, mean batch "
                    f"{loads[-1]['mean_batch']}, lanes {st.lanes}, {loads[-1]['sustained_gcups']}"
                    f" GCUPS")
            return {"lanes_max": args.serving_lanes, "lane_gb": round(st.lane_bytes / 1e9, 2),
                    "max_batch": st.max_batch, "loads": loads}

    on_demand = run(False)
    fixed = run(True)
    first = on_demand["loads"][0]
    return {
        "workload": w.name, "targets": w.n, "residues": w.residues, "queries": len(qs),
        "query_lengths": "lognormal median 350 sigma 0.75 clip 20..3000",
        "partitions": parts, "qps": args.qps, "arrivals": "Poisson (seed 7)",
        "p50_ms": first["p50_ms"], "p99_ms": first["p99_ms"], "mean_ms": first["mean_ms"],
        "sustained_gcups": first["sustained_gcups"], "wall_s": first["wall_s"],
        "lanes": first["lanes_built"], "max_batch": on_demand["max_batch"],
        "lane_gb": on_demand["lane_gb"], "mean_batch": first["mean_batch"],
        "plain_loop_gcups": round(cells / plain_s / 1e9, 1),
        "plain_loop_ms_per_query": round(plain_s * 1e3 / len(qs), 2),
        "hits_equal_alone": True,
        "admission": "by device bytes, lanes on demand (up to lanes_max)",
        "on_demand": on_demand,
        "prealloc_lanes": fixed,
    }


def c1_config(sync) -> dict:
    """The 100-nt DNA SW ScoreEnd scan of 10,000 targets (single best with ends), and scan_all
    with ends on the same query. Each is timed as the median of three runs of 100 calls after
    20 untimed ones: the first few dozen calls of a 0.2 ms operation run slower."""
    w = wl.c1_dna()
    db = build_db(w)
    sc = hy.Scratch(db)
    q = w.queries[0]

    def timed(f) -> float:
        for _ in range(20):
            f()
        sync()
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            for _ in range(100):
                f()
            sync()
            runs.append((time.perf_counter() - t0) / 100)
        return sorted(runs)[1]

    dt = timed(lambda: db.scan(sc, q))
    hit = db.scan(sc, q)
    kernel = sc.last_kernel()
    # Every target's score and ends (scan_all), arrays copied to the host.
    dt_all = timed(lambda: db.scan_all_arrays(sc, q))
    return {w.name: {"op": "scan (ScoreEnd)", "ms": round(dt * 1e3, 3),
                     "gcups": round(len(q) * w.residues / dt / 1e9, 1),
                     "kernel": kernel,
                     "scan_all_ends_ms": round(dt_all * 1e3, 3),
                     "scan_all_kernel": sc.last_kernel(),
                     "partitions": {str(k): v for k, v in db.partition_sizes().items()},
                     "hit": {"score": hit.score, "db_index": hit.db_index,
                             "query_end": hit.query_end, "target_end": hit.target_end}}}


def other_configs(sync) -> dict:
    """The DNA configurations on the same GPU: the 100-nt SW ScoreEnd scan of 10,000 targets
    (single best with ends) and the 150-nt mode sweep over 2,000,000 targets (scan_scores per
    mode, the per-target score vector copied to the host)."""
    out = c1_config(sync)
    base = wl.c3_dna_modes()
    for mode, name in ((1, "NW"), (2, "HW"), (4, "SHW"), (3, "OV")):
        # Built for ScoreEnd: scan_scores ignores the search type, scan_all then tracks ends.
        w = wl.Workload(base.name.replace("mode1", name), base.al, base.matrix, base.go, base.ge,
                        mode, 1, base.mql, base.flat, base.off, base.queries)
        db = build_db(w)
        sc = hy.Scratch(db)
        q = w.queries[0]
        for _ in range(2):
            scores = db.scan_scores_array(sc, q)
        sync()
        t0 = time.perf_counter()
        for _ in range(5):
            scores = db.scan_scores_array(sc, q)
        sync()
        dt = (time.perf_counter() - t0) / 5
        # Per-target hits with their ends (scan_all, ScoreEnd): 20 bytes per target come back.
        db.scan_all_arrays(sc, q)
        sync()
        t0 = time.perf_counter()
        for _ in range(5):
            s_all, _, _ = db.scan_all_arrays(sc, q)
        sync()
        dta = (time.perf_counter() - t0) / 5
        if not np.array_equal(s_all, scores):
            raise RuntimeError(f"{w.name}: scan_all scores differ from scan_scores")
        out[w.name] = {"op": "scan_scores", "ms": round(dt * 1e3, 3),
                       "gcups": round(len(q) * w.residues / dt / 1e9, 1),
                       "scan_all_ends_ms": round(dta * 1e3, 3),
                       "scan_all_ends_gcups": round(len(q) * w.residues / dta / 1e9, 1),
                       "partitions": {str(k): v for k, v in db.partition_sizes().items()},
                       "best_score": int(scores.max()),
                       "targets_at_best": int((scores == scores.max()).sum())}
        log(f"{w.name}: {out[w.name]['gcups']} GCUPS scan_scores, "
            f"{out[w.name]['scan_all_ends_gcups']} GCUPS scan_all with ends")
        del sc, db
    # The same database with the scheme scaled by 25: the longer half of the targets prove i32
    # and run in the 32-bit kernel beside the packed one.
    for mode in (1, 2, 4):
        w = wl.c3_scaled(base, mode)
        db = build_db(w)
        sc = hy.Scratch(db)
        sc.profile(True)
        q = w.queries[0]
        db.scan_scores_array(sc, q)
        sc.profile_read(reset=True)
        sync()
        t0 = time.perf_counter()
        for _ in range(5):
            scores = db.scan_scores_array(sc, q)
        sync()
        dt = (time.perf_counter() - t0) / 5
        inter_ms, tail_ms, n = sc.profile_read(reset=True)
        sc.profile(False)
        parts = db.partition_sizes()
        lens = w.lens()
        i32_cells = int(lens[lens > 504].sum()) * len(q)
        out[w.name] = {"op": "scan_scores", "ms": round(dt * 1e3, 3),
                       "gcups": round(len(q) * w.residues / dt / 1e9, 1),
                       "scheme": f"+{int(w.matrix[0])}/{int(w.matrix[1])}, gaps {w.go}/{w.ge}",
                       "partitions": {str(k): v for k, v in parts.items()},
                       "i32_cell_share": round(i32_cells / (len(q) * w.residues), 3),
                       "packed_kernel_ms": round(inter_ms / max(n, 1), 3),
                       "side_kernel_ms": round(tail_ms / max(n, 1), 3),
                       "best_score": int(scores.max())}
        log(f"{w.name}: {out[w.name]['gcups']} GCUPS, partitions {out[w.name]['partitions']}")
        del sc, db
    return out


def reference_shapes(sync) -> dict:
    """The reference's own scan benchmark shapes: a CR4-style database of 64 adapters (~30 nt)
    against 91-nt reads (benches/cr4_scan.rs), as the CHANGELOG quotes them (CHANGELOG.md:112-121:
    a 64-target x 2000-read i16 SW Score scan and an i16 HW ScoreEnd scan_all, their AVX2 and
    SSE4.1 speed-ups over scalar). The i16 width comes from the CR4 matrix and gaps scaled by 4 (the
    CHANGELOG leaves the scheme open; unscaled, this database proves i8). On the device: one scan
    (or scan_all) call per read, the batched form over all reads, and for scan_all the cross
    product of reads and adapters as one align_pairs batch (the same per-target hits). On one host
    thread beside it: the scalar oracle and the AVX2 restatement (scores only). Every device
    result is checked against the oracle."""
    import oracle
    adapters = wl.cr4_adapters(64)
    flat, off = oracle.flatten(adapters)
    residues = int(off[-1])
    out = {}

    def timed(f, reps=1):
        f()
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = f()
        sync()
        return r, (time.perf_counter() - t0) / reps

    shapes = (("cr4_ov_scoreend_scan_64x256", 3, 1, 1, 256, "scan"),
              ("sw_score_i16_scan_64x2000", 0, 0, 4, 2000, "scan"),
              ("hw_scoreend_i16_scan_all_64x2000", 2, 1, 4, 2000, "scan_all"))
    for name, mode, st, scale, nreads, op in shapes:
        reads = wl.cr4_reads(nreads)
        mat = wl.cr4_matrix(scale)
        go = ge = 2 * scale
        cells = sum(len(r) for r in reads) * residues
        db = (hy.Database.builder().sequences(adapters).scoring(hy.Scoring(5, mat, go, ge))
              .mode(hy.Mode(mode)).search_type(hy.SearchType(st)).max_query_len(91)
              .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())
        sc = hy.Scratch(db)
        width = str(db.score_width())
        row = {"mode": hy.Mode(mode).code(), "search_type": "ScoreEnd" if st else "Score",
               "targets": 64, "reads": nreads, "width": width, "cells": cells}
        t_s = time.perf_counter()
        if op == "scan":
            want = [oracle.scan_flat(flat, off, r, 5, mat, go, ge, mode, bool(st)) for r in reads]
            scalar_s = time.perf_counter() - t_s
            loop, loop_s = timed(lambda: [db.scan(sc, r) for r in reads])
            rflat, roff = oracle.flatten(reads)
            batch, batch_s = timed(lambda: db.scan_batch_arrays(sc, rflat, roff), 3)
            got = [(h.score, h.db_index, h.query_end, h.target_end) for h in loop]
            bat = [(int(a), int(b), None if c < 0 else int(c), None if d < 0 else int(d))
                   for a, b, c, d in zip(*batch)]
            if got != want or bat != want:
                raise RuntimeError(f"{name}: device hits differ from the oracle")
            row["batch_chunk"] = sc.tuning().batch_chunk
            row.update({"device_scan_loop_ms": round(loop_s * 1e3, 3),
                        "device_scan_batch_ms": round(batch_s * 1e3, 3)})
            best_s = batch_s
        else:
            want = [oracle.scan_all_flat(flat, off, r, 5, mat, go, ge, mode) for r in reads]
            scalar_s = time.perf_counter() - t_s
            loop, loop_s = timed(lambda: [db.scan_all_arrays(sc, r) for r in reads])
            qf = np.concatenate([np.repeat(r[None, :], 64, axis=0).ravel() for r in reads])
            qo = np.arange(0, 64 * nreads + 1, dtype=np.uint64) * 91
            tf = np.tile(flat, nreads)
            to = np.concatenate([off[:-1] + k * residues for k in range(nreads)] +
                                [np.array([nreads * residues], dtype=np.uint64)])
            psc = hy.PairScratch()
            pairs, pairs_s = timed(lambda: hy.align_pairs_flat(
                qf, qo, tf, to, hy.Scoring(5, mat, go, ge), hy.Mode(mode), hy.SearchType(st),
                psc), 3)
            for k, w3 in enumerate(want):
                for got in (loop[k], tuple(a[64 * k:64 * (k + 1)] for a in pairs)):
                    if not all(np.array_equal(a, b) for a, b in zip(got, w3)):
                        raise RuntimeError(f"{name}: device hits differ from the oracle")
            row.update({"device_scan_all_loop_ms": round(loop_s * 1e3, 3),
                        "device_align_pairs_cross_product_ms": round(pairs_s * 1e3, 3)})
            best_s = pairs_s
        avx2_s = 0.0
        for r in reads:
            avx2_s += oracle.avx2_scan_scores_flat(flat, off, r, 5, mat, go, ge, mode,
                                                  threads=1)[1]
        row.update({"cpu_scalar_ms": round(scalar_s * 1e3, 2),
                    "cpu_avx2_scores_ms": round(avx2_s * 1e3, 2),
                    "avx2_over_scalar": round(scalar_s / avx2_s, 1),
                    "device_gcups": round(cells / best_s / 1e9, 1),
                    "device_over_avx2": round(avx2_s / best_s, 1),
                    "note": "one host thread each; the AVX2 restatement computes scores only "
                            "(no argmax or ends), the scalar oracle the full result"})
        out[name] = row
        log(f"{name}: device {row['device_gcups']} GCUPS, scalar {row['cpu_scalar_ms']} ms, "
            f"avx2 {row['cpu_avx2_scores_ms']} ms")
        del sc, db
    return out


def traceback(w: wl.Workload, sync) -> tuple[dict, dict]:
    """Aligned scans on the batched workload's targets: an Alignment-search database of the same
    sequences, the single-query workload's 360-aa query. Returns the report and a sample of the
    device's per-target alignments for the CPU leg to check and time against."""
    q = w.extra["single_query"]
    t0 = time.perf_counter()
    db = (hy.Database.builder().sequences_flat(w.flat, w.off)
          .scoring(hy.Scoring(w.al, w.matrix, w.go, w.ge)).mode(hy.Mode(w.mode))
          .search_type(hy.SearchType.Alignment(1 << 62)).max_query_len(len(q))
          .backend(hy.BackendChoice.force(hy.Backend.Hip)).build())
    build_s = time.perf_counter() - t0
    sc = hy.Scratch(db)

    def timed(f, k):
        f()
        sync()
        t = time.perf_counter()
        for _ in range(k):
            r = f()
        sync()
        return (time.perf_counter() - t) / k, r

    scan_s, hit = timed(lambda: db.scan(sc, q), 5)
    al_s, ah = timed(lambda: db.scan_aligned(sc, q), 5)
    if (ah.db_index, ah.alignment.score) != (hit.db_index, hit.score):
        raise RuntimeError("scan_aligned disagrees with scan")
    all_s, (recs, ops) = timed(lambda: db.scan_all_aligned_arrays(sc, q), 2)
    if recs["score"][hit.db_index] != hit.score:
        raise RuntimeError("scan_all_aligned disagrees with scan")
    cells = len(q) * w.residues
    rng = np.random.default_rng(23)
    pick = np.concatenate([[hit.db_index], rng.choice(w.n, size=299, replace=False)])
    sample = {int(k): (int(recs["score"][k]), int(recs["query_start"][k]),
                       int(recs["query_end"][k]), int(recs["target_start"][k]),
                       int(recs["target_end"][k]),
                       ops[int(recs["ops_offset"][k]):int(recs["ops_offset"][k]) +
                           int(recs["n_ops"][k])].tolist()) for k in pick}
    a = ah.alignment
    out = {
        "database": f"{w.n} targets of {w.name}, SearchType::Alignment, max_query_len {len(q)}",
        "build_s": round(build_s, 2),
        "query_len": len(q),
        "scan_ms": round(scan_s * 1e3, 3),
        "scan_aligned_ms": round(al_s * 1e3, 3),
        "traceback_overhead_ms": round((al_s - scan_s) * 1e3, 3),
        "winner": {"db_index": ah.db_index, "score": a.score,
                   "query_span": [a.query_start, a.query_end],
                   "target_span": [a.target_start, a.target_end], "ops": len(a.op_codes),
                   "cigar_prefix": a.cigar()[:60]},
        "scan_all_aligned_s": round(all_s, 4),
        "alignments_per_s": round(w.n / all_s, 1),
        "gcups_equivalent": round(cells / all_s / 1e9, 1),
        # The ops come back packed in database order: their total, not the caller's buffer.
        "ops_bytes_to_host": int(recs["n_ops"].sum()),
        "record_bytes_to_host": int(recs.nbytes),
    }
    log(f"traceback: scan_aligned {out['scan_aligned_ms']} ms (scan {out['scan_ms']} ms), "
        f"scan_all_aligned {all_s:.3f} s = {out['alignments_per_s']:.0f} alignments/s "
        f"({out['ops_bytes_to_host'] / 1e6:.0f} MB of ops, "
        f"{out['record_bytes_to_host'] / 1e6:.0f} MB of records)")
    del sc, db
    return out, {"query": q, "sample": sample}


def cpu_traceback(w: wl.Workload, tb: dict) -> dict:
    """The scalar restatement of the reference's traceback (full H/E/F matrices and the canonical
    walk) on the sampled pairs, one thread, each result checked against the device's."""
    import oracle
    q = tb["query"]
    secs, cells = 0.0, 0
    for k, got in tb["sample"].items():
        t = w.flat[w.off[k]:w.off[k + 1]]
        t0 = time.perf_counter()
        want = oracle.traceback(q, t, w.al, w.matrix, w.go, w.ge, w.mode)
        secs += time.perf_counter() - t0
        cells += len(q) * len(t)
        if tuple(want) != got:
            raise RuntimeError(f"device traceback of target {k} disagrees with the oracle")
    n = len(tb["sample"])
    return {"value": round(cells / secs / 1e9, 4), "unit": "GCUPS", "cores": 1, "kind": "port",
            "alignments_per_s": round(n / secs, 1),
            "sample": f"{n} (query, target) pairs of the traceback database (the winner and "
                      f"{n - 1} random targets), scalar full-matrix traceback restatement, "
                      f"{secs:.2f} s; every alignment equal to the device's"}


def pairs(sync) -> tuple[dict, dict]:
    """The pairwise functions: align_pairs on the DNA read-pair and protein pair batches (flat
    buffers; device time from HIP events around the kernels, wall time including the upload of
    the sequences and the copy back), and single-pair latencies."""
    out = {}
    keep = {}
    sc = hy.PairScratch()
    for b in (wl.dna_read_pairs(), wl.protein_pairs()):
        s = hy.Scoring(b.al, b.matrix, b.go, b.ge)
        cells = b.cells
        entry = {"pairs": b.n, "cells": cells, "mode": "SW"}
        for st in (hy.SearchType.Score, hy.SearchType.ScoreEnd):
            f = lambda: hy.align_pairs_flat(b.queries, b.qoff, b.targets, b.toff, s, hy.Mode.Sw,
                                            st, scratch=sc)
            f()
            sync()
            sc.profile_read(reset=True)
            t0 = time.perf_counter()
            for _ in range(3):
                res = f()
            sync()
            wall = (time.perf_counter() - t0) / 3
            dev_ms, calls = sc.profile_read(reset=True)
            dev = dev_ms / max(calls, 1) / 1e3
            key = "score" if st == hy.SearchType.Score else "score_end"
            entry[key] = {"device_ms": round(dev * 1e3, 3), "wall_ms": round(wall * 1e3, 3),
                          "gcups_device": round(cells / dev / 1e9, 1),
                          "gcups_wall": round(cells / wall / 1e9, 1),
                          "pairs_per_s": round(b.n / wall, 1)}
            if st == hy.SearchType.ScoreEnd:
                keep[b.name] = (b, res)
        out[b.name] = entry
        log(f"pairs {b.name}: {entry['score']['gcups_device']} / "
            f"{entry['score_end']['gcups_device']} GCUPS on the device (Score / ScoreEnd)")
    single, small = single_pairs(sc, sync)
    out["single_pairs"] = single
    keep["_single"] = small
    return out, keep


# The reference's pairwise bench (benches/pairwise.rs): DNA, match +2 / mismatch -1, gap_open 2,
# gap_ext 1, random sequences; align_pair at 30 x 30, 150 x 150 and 150 x 1,400 (Score and
# ScoreEnd), align_pair_position_max at 150 x 1,400, align_pairs over 256 pairs of 30 x 30. One
# long pair (1,000 x 10,000) is added.
REF_PAIR_SHAPES = ((30, 30), (150, 150), (150, 1400), (1000, 10000))


def dna_pair_scoring():
    return hy.Scoring(4, [2 if i == j else -1 for i in range(4) for j in range(4)], 2, 1)


def single_pairs(sc, sync):
    """Median latency (host call to result, us) of align_pair_with (Score, ScoreEnd) and
    align_pair_position_max_with per shape, after a second of warm-up, and of align_pairs over
    the 256 x 30 x 30 batch. Returns the rows and the sequences for the CPU leg."""
    rng = np.random.default_rng(8)
    s = dna_pair_scoring()
    q0 = rng.integers(0, 4, 1000, dtype=np.uint8)
    t0 = rng.integers(0, 4, 20000, dtype=np.uint8)
    end = time.time() + 1.0
    while time.time() < end:
        hy.align_pair_with(sc, q0, t0, s, hy.Mode.Sw, hy.SearchType.Score)

    def med(f, k):
        f()
        sync()
        ts = []
        for _ in range(k):
            t1 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t1)
        return round(float(np.median(ts)) * 1e6, 1)

    rows, small = [], {}
    for m, n in REF_PAIR_SHAPES:
        q = rng.integers(0, 4, m, dtype=np.uint8)
        t = rng.integers(0, 4, n, dtype=np.uint8)
        k = 50 if m * n < 1_000_000 else 10
        row = {"pair": [m, n],
               "score_us": med(lambda: hy.align_pair_with(sc, q, t, s, hy.Mode.Sw,
                                                          hy.SearchType.Score), k),
               "score_end_us": med(lambda: hy.align_pair_with(sc, q, t, s, hy.Mode.Sw,
                                                              hy.SearchType.ScoreEnd), k),
               "position_max_us": med(lambda: hy.align_pair_position_max_with(sc, q, t, s, []),
                                      k)}
        row["score_end_gcups"] = round(m * n / row["score_end_us"] / 1e3, 3)
        rows.append(row)
        small[(m, n)] = (q, t)
    qs = rng.integers(0, 4, 256 * 30, dtype=np.uint8)
    ts = rng.integers(0, 4, 256 * 30, dtype=np.uint8)
    off = np.arange(257, dtype=np.uint64) * 30
    rows.append({"batch": "align_pairs 256 x 30 x 30", "score_us": med(
        lambda: hy.align_pairs_flat(qs, off, ts, off, s, hy.Mode.Sw, hy.SearchType.Score,
                                    scratch=sc), 50)})
    small["batch"] = (qs, ts, off)
    log("single pairs: " + "; ".join(
        f"{r['pair'][0]}x{r['pair'][1]} {r['score_us']}/{r['score_end_us']}/"
        f"{r['position_max_us']} us" for r in rows if "pair" in r))
    return rows, small


def cpu_single_pairs(small: dict) -> dict:
    """The CPU side of single_pairs on the same sequences, one thread: Score by the SSE4.1
    striped restatement at the width the proof gives (repeated to 50 ms), ScoreEnd and the column
    maxima by the scalar restatements (the reference's CPU path for them at these widths is
    scalar for ScoreEnd; its position maxima run striped, so that figure is a scalar bound)."""
    import oracle
    s = dna_pair_scoring()
    al, mat = 4, np.asarray(s.matrix(), dtype=np.int32)
    out = []
    for (m, n), (q, t) in ((k, v) for k, v in small.items() if k != "batch"):
        qo = np.array([0, m], dtype=np.uint64)
        to = np.array([0, n], dtype=np.uint64)
        reps, secs = 0, 0.0
        while secs < 0.05:
            _, dt = oracle.sse_striped_sw_pairs(q, qo, t, to, al, mat, 2, 1)
            secs += dt
            reps += 1
        t1 = time.perf_counter()
        oracle.align_core(q, t, al, mat, 2, 1, 0)
        se = time.perf_counter() - t1
        t1 = time.perf_counter()
        oracle.position_max(q, t, al, mat, 2, 1)
        pm = time.perf_counter() - t1
        out.append({"pair": [m, n], "score_striped_us": round(secs / reps * 1e6, 1),
                    "score_end_scalar_us": round(se * 1e6, 1),
                    "position_max_scalar_us": round(pm * 1e6, 1)})
    qs, ts, off = small["batch"]
    reps, secs = 0, 0.0
    while secs < 0.05:
        _, dt = oracle.sse_striped_sw_pairs(qs, off, ts, off, al, mat, 2, 1)
        secs += dt
        reps += 1
    out.append({"batch": "align_pairs 256 x 30 x 30",
                "score_striped_us": round(secs / reps * 1e6, 1)})
    return {"rows": out, "cores": 1, "kind": "port",
            "sample": "the single pairs and 256-pair batch of the device rows, same sequences"}


def cpu_pairs(keep: dict) -> dict:
    """Per batch, two one-thread CPU legs, every result checked against the device's ScoreEnd
    run: Score by the SSE4.1 striped (Farrar, 8 x i16) restatement of the reference's x86
    single-pair path on a contiguous slice of pairs, and ScoreEnd by the scalar pair DP (the
    oracle's align_core) on 1,000 pairs spread over the batch."""
    import oracle
    res = {}
    if "_single" in keep:
        res["single_pairs"] = cpu_single_pairs(keep["_single"])
    for name, (b, (scores, qe, te)) in ((k, v) for k, v in keep.items() if k != "_single"):
        n = min(b.n, 40_000 if b.al <= 4 else 20_000)
        qo, to = b.qoff[:n + 1], b.toff[:n + 1]
        got, ssecs = oracle.sse_striped_sw_pairs(b.queries[:int(qo[-1])], qo,
                                                 b.targets[:int(to[-1])], to, b.al, b.matrix,
                                                 b.go, b.ge)
        if not np.array_equal(got, np.asarray(scores[:n])):
            k = int(np.flatnonzero(got != np.asarray(scores[:n]))[0])
            raise RuntimeError(f"device align_pairs disagrees with the striped oracle at pair {k}")
        scells = int((np.diff(qo).astype(np.int64) * np.diff(to).astype(np.int64)).sum())
        secs, cells = 0.0, 0
        for k in np.linspace(0, b.n - 1, 1000).astype(int):
            q, t = b.pair(int(k))
            t0 = time.perf_counter()
            s, wq, wt = oracle.align_core(q, t, b.al, b.matrix, b.go, b.ge, 0)
            secs += time.perf_counter() - t0
            cells += len(q) * len(t)
            if (s, -1 if wq is None else wq, -1 if wt is None else wt) != (scores[k], qe[k],
                                                                           te[k]):
                raise RuntimeError(f"device align_pairs disagrees with the oracle at pair {k}")
        res[name] = {
            "score": {"value": round(scells / ssecs / 1e9, 3), "unit": "GCUPS", "cores": 1,
                      "kind": "port", "sample": f"pairs 0-{n - 1}, SSE4.1 striped 8 x i16 local "
                      f"score, {ssecs:.2f} s; every score equal to the device's"},
            "score_end": {"value": round(cells / secs / 1e9, 4), "unit": "GCUPS", "cores": 1,
                          "kind": "port", "sample": f"1000 pairs spread over the batch, scalar "
                          f"full DP with ends, {secs:.2f} s; every result equal to the device's"}}
        log(f"cpu pairs {name}: striped Score {res[name]['score']['value']} GCUPS, scalar "
            f"ScoreEnd {res[name]['score_end']['value']} GCUPS")
    return res


def db_image(db: hy.Database, w: wl.Workload, build_s: float) -> dict:
    """The on-disk image of the batched workload's database: save, load back (checksum, the
    build's validation, residue upload, layouts gathered on the device), one scan compared."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.img")
        t0 = time.perf_counter()
        db.save(path)
        save_s = time.perf_counter() - t0
        size = os.path.getsize(path)
        t0 = time.perf_counter()
        ld = hy.Database.load(path)
        load_s = time.perf_counter() - t0
    q = w.extra["single_query"]
    if ld.scan(hy.Scratch(ld), q) != db.scan(hy.Scratch(db), q):
        raise RuntimeError("loaded database disagrees with the original")
    log(f"image: {size / 1e6:.0f} MB, save {save_s:.2f} s, load {load_s:.2f} s "
        f"(build from memory {build_s:.2f} s)")
    return {"image_bytes": size, "save_s": round(save_s, 3), "load_s": round(load_s, 3),
            "build_from_memory_s": round(build_s, 3), "targets": w.n, "residues": w.residues}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--targets", type=int, default=1_000_000)
    ap.add_argument("--queries", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-query loop")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-traceback", action="store_true", help="skip the aligned scans")
    ap.add_argument("--no-pairs", action="store_true", help="skip the pairwise functions")
    ap.add_argument("--no-image", action="store_true", help="skip the database image round trip")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the DNA configurations (10k-target scan, 2M-target mode sweep)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of running the autotuner "
                         "(it picks them before timing, from the batch's own queries)")
    ap.add_argument("--qps", type=float, default=SERVING_QPS)
    ap.add_argument("--qps-sweep", type=float, nargs="*", default=[20.0, 40.0],
                    help="further serving loads on the same server (continuous batching)")
    ap.add_argument("--serving-targets", type=int, default=2_000_000)
    ap.add_argument("--serving-queries", type=int, default=64)
    ap.add_argument("--serving-lanes", type=int, default=3)
    ap.add_argument("--serving-batch", type=int, default=16)
    ap.add_argument("--only", choices=("configs", "c1", "shapes", "serving"),
                    help="only run the DNA configurations (c1: the 10,000-target scan alone), "
                         "the reference's scan shapes or the serving form and print their JSON "
                         "(for profiling)")
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

    # Device-wide synchronisation around timed regions (what torch.cuda.synchronize does), called
    # directly so the benchmark does not pay for importing torch.
    _hip = ctypes.CDLL("libamdhip64.so")

    def sync():
        if _hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")

    if args.only == "configs":
        print(json.dumps({"other_configs": other_configs(sync)}), flush=True)
        return 0
    if args.only == "c1":
        print(json.dumps({"other_configs": c1_config(sync)}), flush=True)
        return 0
    if args.only == "shapes":
        print(json.dumps({"reference_shapes": reference_shapes(sync)}), flush=True)
        return 0
    if args.only == "serving":
        print(json.dumps({"serving": serving(args, sync)}), flush=True)
        return 0

    t0 = time.perf_counter()
    w = wl.c4_batch(n_targets=args.targets, nq=args.queries)
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    db = build_db(w)
    build_s = time.perf_counter() - t0
    parts = {str(k): v for k, v in db.partition_sizes().items()}
    qs = w.queries
    image = None if args.no_image else db_image(db, w, build_s)
    sc = hy.Scratch(db)
    log(f"workload {w.name}: {w.n} targets, {w.residues} residues, {len(qs)} queries "
        f"({sum(len(q) for q in qs)} residues), gen {gen_s:.1f}s, build+upload {build_s:.1f}s, "
        f"partitions {parts}")
    tune_s = None
    default_shapes = sc.tuning()
    if not args.no_autotune:
        t0 = time.perf_counter()
        sc.autotune(qs)
        tune_s = time.perf_counter() - t0
    tuning = sc.tuning()
    log(f"launch shapes {tuning} (defaults {default_shapes}), autotune {tune_s} s")

    # Warm-up submissions are graph-captured and replayed; the timed steps record HIP events
    # around the persistent kernel (profiling submits the same sequence without the graph).
    for _ in range(args.warmup):
        hits = db.scan_batch(sc, qs)
    sc.profile(True)
    sc.batch_profile_read(reset=True)
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hits = db.scan_batch(sc, qs)
    sync()
    dt = time.perf_counter() - t0
    kernel_ms, launches = sc.batch_profile_read(reset=True)
    sc.profile(False)
    cells = sum(len(q) for q in qs) * w.residues
    ms_per_step = dt * 1e3 / args.steps
    gcups = cells / (dt / args.steps) / 1e9
    log(f"batched: {ms_per_step:.1f} ms per step, {gcups:.1f} GCUPS")

    plain = None
    if not args.no_plain:
        sync()
        t0 = time.perf_counter()
        plain_hits = [db.scan(sc, q) for q in qs]
        sync()
        plain_s = time.perf_counter() - t0
        if plain_hits != hits:
            raise RuntimeError("batched hits differ from the plain per-query scans")
        plain = {"gcups": round(cells / plain_s / 1e9, 1), "seconds": round(plain_s, 3),
                 "batched_over_plain": round(plain_s / (dt / args.steps), 3),
                 "hits_identical": True}
        log(f"plain loop: {plain_s:.2f} s, {plain['gcups']} GCUPS")

    # The single-query protein workload's 360-aa query on the same targets.
    q360 = w.extra["single_query"]
    sc.profile(True)
    sc.profile_read(reset=True)
    for _ in range(2):
        db.scan(sc, q360)
    sc.profile_read(reset=True)
    sync()
    t0 = time.perf_counter()
    for _ in range(10):
        h360 = db.scan(sc, q360)
    sync()
    s360 = (time.perf_counter() - t0) / 10
    inter_ms, tail_ms, n360 = sc.profile_read(reset=True)
    sc.profile(False)
    # Per-target hits with their ends (scan_all, ScoreEnd) for the same query.
    db.scan_all_arrays(sc, q360)
    sync()
    t0 = time.perf_counter()
    for _ in range(5):
        all_s, all_qe, all_te = db.scan_all_arrays(sc, q360)
    sync()
    sall = (time.perf_counter() - t0) / 5
    if (all_s[h360.db_index], all_te[h360.db_index]) != (h360.score, h360.target_end):
        raise RuntimeError("scan_all disagrees with scan on the winner")
    single = {"workload": f"protein_sw_scoreend_360x{w.n}", "gcups": round(
        len(q360) * w.residues / s360 / 1e9, 1), "ms_per_scan": round(s360 * 1e3, 3),
        "inter_kernel_ms": round(inter_ms / max(n360, 1), 3),
        "tail_kernel_ms": round(tail_ms / max(n360, 1), 3),
        "scan_all_ends_ms": round(sall * 1e3, 3),
        "scan_all_ends_gcups": round(len(q360) * w.residues / sall / 1e9, 1),
        "hit": {"score": h360.score, "db_index": h360.db_index, "query_end": h360.query_end,
                "target_end": h360.target_end}}

    kernel = (f"k_inter16q_batch<{tuning.batch_tile_rows}>" if w.mode == 0 and tuning.batch_pairs
              else f"k_inter16_batch<SW,{tuning.batch_tile_rows}>")
    roof = None
    if launches:
        k_ms = kernel_ms / launches
        # One launch per step; every target is in the packed partition here.
        k_cells = cells * args.steps / launches
        achieved = k_cells * ops_per_cell(w.mode) / (k_ms / 1e3) / 1e12
        traffic, traffic_note = traffic_record(w.name, kernel)
        roof = {
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": round(VALU_PEAK_TOPS_PK16, 2),
            "unit": "TOP/s",
            "frac": round(achieved / VALU_PEAK_TOPS_PK16, 4),
            "traffic": traffic,
            "traffic_source": traffic_note,
            "kernel": kernel,
            "peak_note": "packed 16-bit VALU element-ops (2 per lane per instruction)",
            "kernel_ms": round(k_ms, 3),
            "ops_per_cell": ops_per_cell(w.mode),
            "cells_per_launch": int(k_cells),
            "hbm_gbs_algorithmic": round(w.residues / (k_ms / 1e3) / 1e9, 2),
        }
    del sc, db

    tb, tb_sample = (None, None) if args.no_traceback else traceback(w, sync)
    pr, pr_keep = (None, None) if args.no_pairs else pairs(sync)
    serve = None if args.no_serving else serving(args, sync)
    others = None if args.no_configs else other_configs(sync)
    ref_shapes = None if args.no_configs else reference_shapes(sync)

    out = {
        "metric": METRIC,
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
        "data": "synthetic (seeded; hyalite_amd/workloads.py)",
        "config": {"workload": w.name, "form": "batched (persistent queue, one launch per step)",
                   "queries": len(qs), "query_lengths": "log-uniform 20..2000",
                   "targets": w.n, "residues": w.residues, "mode": "SW",
                   "search_type": "ScoreEnd", "max_query_len": w.mql,
                   "target_lengths": "lognormal median 300 sigma 0.75 clip 30..5000",
                   "partitions": parts, "launch_shapes": tuning.__dict__,
                   "default_shapes": default_shapes.__dict__,
                   "autotune_s": None if tune_s is None else round(tune_s, 1)},
        "plain": plain,
        "single_query": single,
        "serving": serve,
        "other_configs": others,
        "reference_shapes": ref_shapes,
        "traceback": tb,
        "pairs": pr,
        "db_image": image,
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, qs)
        out["gpu_vs_cpu"] = round(gcups / out["cpu_baseline"]["value"], 1)
        if tb_sample is not None:
            out["cpu_baseline"]["traceback"] = cpu_traceback(w, tb_sample)
        if pr_keep is not None:
            out["cpu_baseline"]["pairs"] = cpu_pairs(pr_keep)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
