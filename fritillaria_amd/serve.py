"""Serving loop over a resident BAM: Poisson arrivals of virtual-offset range
queries, continuous batching with admission by free HBM, latency percentiles.

Requests follow the config-5 workload: each is a [vstart, vend) range whose
compressed size is log-uniform over 64 KiB - 256 MiB, snapped to record-start
virtual offsets (noodles csi Query semantics,
crates/fritillaria-csi/src/io/query.rs:66-82). The loop runs in wall-clock time
and batches continuously: whenever one of the server's slots is free and
requests have arrived, those requests (in arrival order, while they fit the
HBM budget left by the batches still on the device) go to the GPU as one
batch, which runs alongside the batches already there; a request's latency is
its batch's completion time minus its arrival time.
"""
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import hip


# A batch joins others on the device only while those hold fewer compressed
# bytes than this: a batch at least this large keeps the device busy by itself,
# and a second one beside it only slows both (measured on the batched large
# BAM's range queries: p99 at 30% / 60% load 73.6 / 75.2 ms with this rule,
# 78.6 / 77.5 ms one batch at a time, 109 / 79 ms with every free slot used).
OVERLAP_BELOW = 256 << 20


@dataclass
class Request:
    arrival: float          # seconds from the start of the run
    vstart: int
    vend: int
    compressed_bytes: int


@dataclass
class ServeStats:
    latencies_ms: np.ndarray
    makespan_s: float
    inflated_bytes: int
    records: int
    batches: int
    batch_sizes: List[int] = field(default_factory=list)
    failures: int = 0
    # device interval (start, end) of every batch, ms, and how many batches
    # were launched while another was still on the device
    intervals: List[tuple] = field(default_factory=list)
    overlapped_launches: int = 0

    def overlap_ms(self) -> float:
        """Device time during which two or more batches were running."""
        ev = sorted([(a, 1) for a, _ in self.intervals] + [(b, -1) for _, b in self.intervals])
        depth, last, total = 0, 0.0, 0.0
        for t, d in ev:
            if depth >= 2:
                total += t - last
            depth += d
            last = t
        return total

    def summary(self) -> dict:
        lat = self.latencies_ms
        return {
            "requests": int(lat.size),
            "p50_ms": round(float(np.percentile(lat, 50)), 3) if lat.size else None,
            "p99_ms": round(float(np.percentile(lat, 99)), 3) if lat.size else None,
            "throughput_gbps": round(self.inflated_bytes / self.makespan_s / 1e9, 3) if self.makespan_s else None,
            "records": self.records,
            "batches": self.batches,
            "mean_batch": round(float(np.mean(self.batch_sizes)), 2) if self.batch_sizes else 0.0,
            "failures": self.failures,
            "overlapped_launches": self.overlapped_launches,
            "device_overlap_ms": round(self.overlap_ms(), 3),
        }


def make_requests(voffs: np.ndarray, coffsets: np.ndarray, n: int, rate_hz: float, seed: int,
                  lo_bytes: int = 64 << 10, hi_bytes: int = 256 << 20) -> List[Request]:
    """n requests, Poisson arrivals at rate_hz, each [vstart, vend) with a
    log-uniform compressed extent, both ends record-start virtual offsets."""
    rng = np.random.default_rng(seed)
    rec_coff = voffs >> np.uint64(16)
    total_c = int(coffsets[-1]) if len(coffsets) else 0
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate_hz)
        size = int(np.exp(rng.uniform(np.log(lo_bytes), np.log(hi_bytes))))
        size = min(size, max(total_c, 1))
        i = int(rng.integers(0, len(voffs)))
        c_end = int(rec_coff[i]) + size
        j = int(np.searchsorted(rec_coff, np.uint64(c_end), side="left"))
        if j <= i:
            j = i + 1
        vend = int(voffs[j]) if j < len(voffs) else hip.TO_END
        c_bytes = (int(rec_coff[j]) if j < len(voffs) else total_c) - int(rec_coff[i])
        out.append(Request(t, int(voffs[i]), vend, max(c_bytes, 0)))
    return out


def serve(server: "hip.QueryServer", requests: List[Request], budget_bytes: int,
          max_batch: Optional[int] = None, free_bytes=None, slots: Optional[int] = None,
          overlap_below: Optional[int] = OVERLAP_BELOW) -> ServeStats:
    """Continuous batching over the server's slots. Whenever a slot is free and
    requests have arrived, they join a new batch in arrival order while its
    device footprint fits what is left of the budget: budget_bytes less the
    footprints of the batches still on the device, and no more than the free
    HBM read (free_bytes(), hipMemGetInfo by default) at that moment. A request
    larger than the budget runs alone once nothing else is in flight. The batch
    is launched at once and overlaps the ones already running; finished slots
    are collected as soon as the device reports them. `slots` caps the slots
    used (default: all); with overlap_below, a batch joins others on the device
    only while those hold fewer than that many compressed bytes (a device not
    yet full). Blocks until every request is served."""
    n = len(requests)
    k_slots = min(server.slots(), slots or server.slots())
    if free_bytes is None:
        free_bytes = lambda: hip.device_memory(server.ctx)[0]  # noqa: E731
    # device footprint of each request alone; a batch is charged their sum
    fp = [server.footprint([(r.vstart, r.vend)]) for r in requests]
    lat = np.zeros(n)
    pending: List[int] = []
    in_flight = {}  # slot -> (request indices, charged bytes)
    nxt = 0
    done = 0
    inflated = records = batches = failures = overlapped = 0
    sizes, intervals = [], []
    t0 = time.perf_counter()
    while done < n:
        now = time.perf_counter() - t0
        while nxt < n and requests[nxt].arrival <= now:
            pending.append(nxt)
            nxt += 1
        progressed = False
        for slot in list(in_flight):
            if not server.query(slot):
                continue
            batch, _ = in_flight.pop(slot)
            res = server.finish(slot)
            end = time.perf_counter() - t0
            intervals.append(server.batch_interval(slot))
            for i, r in zip(batch, res):
                lat[i] = (end - requests[i].arrival) * 1e3
                if r["status"] != 0:
                    failures += 1
                inflated += int(r["inflated_bytes"])
                records += int(r["records"])
            done += len(batch)
            progressed = True
        free_slots = [k for k in range(k_slots) if k not in in_flight]
        if in_flight and overlap_below is not None and \
                sum(requests[i].compressed_bytes for b, _ in in_flight.values() for i in b) >= overlap_below:
            free_slots = []
        if pending and free_slots:
            charged_now = sum(c for _, c in in_flight.values())
            room = min(budget_bytes - charged_now, int(free_bytes()))
            batch, charged = [], 0
            for k in pending:
                if (max_batch and len(batch) >= max_batch) or charged + fp[k] > room:
                    break
                batch.append(k)
                charged += fp[k]
            if not batch and not in_flight:
                batch, charged = [pending[0]], fp[pending[0]]  # too large for the budget: alone
            if batch:
                pending = pending[len(batch):]
                slot = free_slots[0]
                if in_flight:
                    overlapped += 1
                server.launch(slot, [(requests[i].vstart, requests[i].vend) for i in batch])
                in_flight[slot] = (batch, charged)
                batches += 1
                sizes.append(len(batch))
                progressed = True
        if not progressed:
            if not in_flight and not pending and nxt < n:
                time.sleep(max(0.0, requests[nxt].arrival - (time.perf_counter() - t0)))
            else:
                time.sleep(0.0001)
    return ServeStats(lat, time.perf_counter() - t0, inflated, records, batches, sizes, failures, intervals,
                      overlapped)


def serve_plain(ctx: "hip.HipContext", compressed: np.ndarray, requests: List[Request], coffsets: np.ndarray,
                blocks_per_batch: int = 256) -> ServeStats:
    """The same arrivals through the per-call drop-in, one request at a time:
    DeviceBgzfReader over the request's compressed members (host-fed) and
    BamDecoder per batch, the loop of crates/fritillaria-cuda/tests/bam_reader.rs:76-115."""
    codec = hip.HipCodec(ctx)
    decoder = hip.BamDecoder(ctx)
    n = len(requests)
    lat = np.zeros(n)
    inflated = records = 0
    total_c = len(compressed)
    t0 = time.perf_counter()
    for k, r in enumerate(requests):
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        c0 = r.vstart >> 16
        m_end = len(coffsets) if r.vend == hip.TO_END else int(np.searchsorted(coffsets, r.vend >> 16))
        if r.vend != hip.TO_END and (r.vend & 0xFFFF):
            m_end += 1
        c1 = int(coffsets[m_end]) if m_end < len(coffsets) else total_c
        reader = hip.DeviceBgzfReader(compressed[c0:c1], codec, blocks_per_batch=blocks_per_batch)
        first = True
        while True:
            try:
                b = reader.next_batch()
            except hip.FrError:
                # the range's last member ends inside a record that starts at or
                # after vend: everything below vend has been served
                break
            if b is None:
                break
            start = (r.vstart & 0xFFFF) if first else b.start
            first = False
            recs = decoder.decode(b.data, start)
            records += len(recs)
            inflated += b.data.byte_len()
            tail = recs.tail()
            del recs
            reader.carry_from(tail)
        reader.close()
        lat[k] = (time.perf_counter() - t0 - r.arrival) * 1e3
    return ServeStats(lat, time.perf_counter() - t0, inflated, records, n, [1] * n, 0)
