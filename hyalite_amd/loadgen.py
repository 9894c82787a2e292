"""Open-loop load generator for the serving form: Poisson arrivals at a fixed rate.

Queries are submitted at exponentially distributed gaps (rate `qps`) from the calling thread
while a collector thread waits for the results in submission order, so a slow query never delays
the arrivals behind it. Latencies come from the server's own steady-clock timestamps
(submit -> result ready).
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import ServedHit, Server


@dataclass(frozen=True)
class LoadResult:
    served: List[ServedHit]
    wall_s: float          # first submission -> last result
    offered_qps: float

    def latencies(self) -> np.ndarray:
        return np.array([s.latency_s for s in self.served])

    def percentile_ms(self, p: float) -> float:
        return float(np.percentile(self.latencies(), p) * 1e3)


def run_poisson(server: Server, queries: Sequence, qps: float, seed: int = 0) -> LoadResult:
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=len(queries))
    handoff: "queue.Queue" = queue.Queue()
    served: List = [None] * len(queries)
    errors: List[BaseException] = []

    def collect():
        try:
            for _ in range(len(queries)):
                k, ticket = handoff.get()
                served[k] = server.wait(ticket)
        except BaseException as e:  # surfaced to the caller below
            errors.append(e)

    th = threading.Thread(target=collect, daemon=True)
    th.start()
    start = time.perf_counter()
    due = start
    for k, q in enumerate(queries):
        due += gaps[k]
        delay = due - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        handoff.put((k, server.submit(q)))
    th.join()
    wall = time.perf_counter() - start
    if errors:
        raise errors[0]
    return LoadResult(served, wall, qps)
