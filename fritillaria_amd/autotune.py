"""Schedule autotuner for the batched pipeline.

Sweeps the inflate kernel variants (output ring per member x waves per
workgroup, which set how many members each CU keeps in flight), the work-queue
order and the boundary-scan window on a sample of the resident file, timing
each stage with HIP events. The one-wave-per-member variants reach their
speed on any sample that fills the GPU, but the two-phase variants run one
member per lane and need many more members in flight, so a sample
understates them: when the file is larger than the sample, the best
configuration of each variant family is timed again on the whole file (the
two-phase family with every work-queue order: on a sample, largest-first
waves finish together, while on a whole file, whose waves share SIMDs, the
order that wins depends on which waves end up sharing one) and the fastest
there wins. Every schedule produces
the same bytes and columns; only time differs.
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
             schedules=None, orders=(0, 1), windows=(2048,), whole_file_orders=(0, 1, 2, 3), log=None) -> dict:
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
    if sample.size < compressed.size:
        finals = []
        for fam in sorted({r["tokens"] for r in results}):
            r = min((x for x in results if x["tokens"] == fam), key=lambda x: x["total_ms"])
            for order in (whole_file_orders if fam == 3 else (r["order"],)):
                p = hip.BatchedPipeline(ctx, compressed, inflate_schedule=r["inflate_schedule"], order=order,
                                        bam_window=r["bam_window"])
                full = None
                for _ in range(repeats):
                    p.run()
                    ms = p.stage_ms()
                    full = ms if full is None or ms["total"] < full["total"] else full
                p.close()
                rec = dict(r, order=order, inflate_ms=full["inflate"], total_ms=full["total"], whole_file=True)
                finals.append(rec)
                if log:
                    log(f"autotune whole file {rec}")
        best = min(finals, key=lambda r: r["total_ms"])
        results += finals
    return {"best": best, "sample_bytes": int(sample.size), "seconds": time.time() - t0, "results": results}
