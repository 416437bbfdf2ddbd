"""Schedule autotuner for the batched pipeline.

Sweeps the inflate kernel variants (output ring per member x waves per
workgroup, which set how many members each CU keeps in flight), the work-queue
order and the boundary-scan window on a sample of the resident file, timing
each stage with HIP events, and returns the fastest configuration. Every
schedule produces the same bytes and columns; only time differs.
"""
from __future__ import annotations

import time

import numpy as np

from . import hip


def sample_prefix(compressed: np.ndarray, max_bytes: int) -> np.ndarray:
    """The longest whole-member prefix of a BGZF file within max_bytes, plus an EOF member."""
    if compressed.size <= max_bytes:
        return compressed
    spans, _ = hip.discover(compressed[:max_bytes])
    if len(spans) == 0:
        return compressed
    end = int(spans["compressed_offset"][-1])  # drop the last (possibly partial-record) member
    eof = np.frombuffer(bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000"), np.uint8)
    return np.concatenate([compressed[:end], eof])


def autotune(ctx: hip.HipContext, compressed: np.ndarray, *, sample_bytes=512 << 20, repeats=2,
             schedules=None, orders=(0, 1), windows=(2048,), log=None) -> dict:
    sample = sample_prefix(compressed, sample_bytes)
    table = hip.inflate_schedules()
    schedules = range(len(table)) if schedules is None else schedules
    results = []
    t0 = time.time()
    for sched in schedules:
        for order in orders:
            for window in windows:
                p = hip.BatchedPipeline(ctx, sample, inflate_schedule=sched, order=order, bam_window=window)
                best = None
                for _ in range(repeats):
                    p.run()
                    ms = p.stage_ms()
                    best = ms if best is None or ms["total"] < best["total"] else best
                p.close()
                rec = {"inflate_schedule": sched, "ring": table[sched][0], "waves_per_group": table[sched][1],
                       "waves_per_cu": table[sched][2], "tokens": table[sched][3], "order": order, "bam_window": window,
                       "inflate_ms": best["inflate"], "total_ms": best["total"]}
                results.append(rec)
                if log:
                    log(f"autotune {rec}")
    best = min(results, key=lambda r: r["total_ms"])
    return {"best": best, "sample_bytes": int(sample.size), "seconds": time.time() - t0, "results": results}
