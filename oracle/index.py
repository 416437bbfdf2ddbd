"""BAI/CSI binning indexes and region queries -- TEST INFRASTRUCTURE ONLY.

A CPU restatement, in plain Python and numpy, of how the reference reads a BAI
or CSI index, maps a region to BGZF chunks, builds an index from a coordinate-
sorted BAM, and decides which records a region query returns. Only ``tests/``
(and the checker legs that may use the oracle) import it; the product reads
indexes in ``fritillaria_amd/csrc/index_host.cpp`` and filters records on the
device.

Reference functions restated here (crates/...):
  read_bai           fritillaria-bam/src/bai/io/reader/index.rs:9-37,
                     .../index/reference_sequences.rs:13-33, .../bins.rs:8-58,
                     .../intervals.rs:7-22
  read_csi           fritillaria-csi/src/io/reader/index.rs:72-120,
                     .../reference_sequences.rs:50-86, .../bins.rs:69-108,
                     .../metadata.rs (n_chunk must be 2), .../header.rs:80-93
  reg2bin, reg2bins  fritillaria-csi/src/binning_index/index/reference_sequence.rs:352-397
  optimize_chunks    fritillaria-csi/src/binning_index.rs:102-160
  min_offset         .../reference_sequence/index/linear_index.rs:11-15 (BAI),
                     .../reference_sequence/index/binned_index.rs:12-28 (CSI)
  resolve_interval   fritillaria-csi/src/binning_index/index.rs:131-167
  query              fritillaria-csi/src/binning_index/index.rs:93-121,
                     .../index/reference_sequence.rs:89-109
  Indexer            fritillaria-csi/src/binning_index/indexer.rs:40-101,
                     .../index/reference_sequence.rs:152-216 (update, optimize),
                     .../bin.rs:67-76 (add_chunk), linear_index.rs:21-29,
                     metadata.rs:128-137; fritillaria-bam/src/fs/index.rs:22-80.
                     A CSI index's per-bin loffsets are built as htslib builds
                     them (the linear offset of the bin's first window), not as
                     noodles' Indexer does; see build_index. Only test indexes
                     are built here: the product reads indexes, it does not write them.
  write_bai          fritillaria-bam/src/bai/io/writer/index.rs (same layout as read)
  intersects         fritillaria-bam/src/io/reader/query.rs:104-132 with
                     fritillaria-sam/src/alignment/record.rs:108-135 (alignment
                     span from the raw CIGAR) and fritillaria-core/src/region/
                     interval.rs:96-108
  query_records      fritillaria-csi/src/io/query.rs:66-82 (seek to each chunk
                     start, read while the virtual position is below its end)

What pins each function (tests/test_index.py):
  - the reference's own unit vectors: read_bins' bytes (bins.rs tests),
    optimize_chunks / merge_chunks doc examples, the Indexer build test
    (indexer.rs test_build: bins 4681 and 73, linear index, metadata), the
    linear and binned min_offset tests;
  - the samtools-written testdata/htslib_multiblock.bam.bai: the Indexer run
    over htslib_multiblock.bam must rebuild the same bins, chunks, linear
    index, metadata and unplaced count;
  - brute force: for any region, the records an index query returns equal the
    records of a full scan that intersect the region (an index is only a way
    to skip records), over random regions on the fixture and on synthetic
    multi-reference files.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

BAI_MIN_SHIFT = 14
BAI_DEPTH = 5


class IndexError_(ValueError):
    """An io::Error of kind InvalidData / InvalidInput in the reference."""


def bin_limit(depth: int) -> int:
    # bin.rs:80-83
    return (1 << ((depth + 1) * 3)) // 7


def metadata_id(depth: int) -> int:
    return bin_limit(depth) + 1


def max_position(min_shift: int, depth: int) -> int:
    return (1 << (min_shift + 3 * depth)) - 1


@dataclass
class RefIndex:
    bins: dict = field(default_factory=dict)        # bin id -> [(start, end), ...]
    linear: list = field(default_factory=list)      # BAI: window -> virtual offset
    loffsets: dict = field(default_factory=dict)    # CSI: bin id -> virtual offset
    metadata: tuple | None = None                   # (ref_beg, ref_end, n_mapped, n_unmapped)


@dataclass
class Index:
    kind: str                   # "bai" or "csi"
    min_shift: int
    depth: int
    refs: list
    unplaced_unmapped: int | None
    aux: bytes = b""


class _Reader:
    def __init__(self, data: bytes):
        self.d = bytes(data)
        self.p = 0

    def take(self, n):
        if self.p + n > len(self.d):
            raise IndexError_("unexpected end of index")
        b = self.d[self.p:self.p + n]
        self.p += n
        return b

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def at_end(self):
        return self.p == len(self.d)


def _read_metadata(r: _Reader):
    n_chunk = r.u32()
    if n_chunk != 2:
        raise IndexError_(f"invalid metadata chunk count {n_chunk}")
    return (r.u64(), r.u64(), r.u64(), r.u64())


def _read_chunks(r: _Reader, signed: bool = True):
    # n_chunk is an i32 in both formats (csi .../bins/chunks.rs, which the BAI reader reuses)
    n = r.i32() if signed else r.u32()
    if n < 0:
        raise IndexError_("invalid chunk count")
    return [(r.u64(), r.u64()) for _ in range(n)]


def _read_unplaced(r: _Reader):
    if r.at_end():
        return None
    if len(r.d) - r.p < 8:
        raise IndexError_("truncated unplaced unmapped record count")
    return r.u64()


def read_bai(data: bytes) -> Index:
    r = _Reader(data)
    if r.take(4) != b"BAI\x01":
        raise IndexError_("invalid BAI magic number")
    n_ref = r.u32()
    mid = metadata_id(BAI_DEPTH)
    refs = []
    for _ in range(n_ref):
        ref = RefIndex()
        for _ in range(r.u32()):
            bid = r.u32()
            if bid == mid:
                if ref.metadata is not None:
                    raise IndexError_(f"duplicate bin ID: {bid}")
                ref.metadata = _read_metadata(r)
            else:
                if bid in ref.bins:
                    raise IndexError_(f"duplicate bin ID: {bid}")
                ref.bins[bid] = _read_chunks(r, signed=True)
        ref.linear = [r.u64() for _ in range(r.u32())]
        refs.append(ref)
    return Index("bai", BAI_MIN_SHIFT, BAI_DEPTH, refs, _read_unplaced(r))


def read_csi(data: bytes) -> Index:
    r = _Reader(data)
    if r.take(4) != b"CSI\x01":
        raise IndexError_("invalid CSI magic number")
    min_shift, depth = r.i32(), r.i32()
    if not 0 <= min_shift <= 255 or not 0 <= depth <= 255:
        raise IndexError_("invalid min shift or depth")
    l_aux = r.i32()
    if l_aux < 0:
        raise IndexError_("invalid aux length")
    aux = r.take(l_aux)
    n_ref = r.i32()
    if n_ref < 0:
        raise IndexError_("invalid reference sequence count")
    mid = metadata_id(depth)
    refs = []
    for _ in range(n_ref):
        ref = RefIndex()
        n_bin = r.i32()
        if n_bin < 0:
            raise IndexError_("invalid bin count")
        for _ in range(n_bin):
            bid = r.u32()
            loff = r.u64()
            if bid == mid:
                if ref.metadata is not None:
                    raise IndexError_(f"duplicate bin: {bid}")
                ref.metadata = _read_metadata(r)
            else:
                if bid in ref.bins:
                    raise IndexError_(f"duplicate bin: {bid}")
                ref.loffsets[bid] = loff
                ref.bins[bid] = _read_chunks(r, signed=True)
        refs.append(ref)
    return Index("csi", min_shift, depth, refs, _read_unplaced(r), aux)


def read_index(data: bytes) -> Index:
    magic = bytes(data[:4])
    if magic == b"BAI\x01":
        return read_bai(data)
    if magic == b"CSI\x01":
        return read_csi(data)
    raise IndexError_("neither a BAI nor a CSI index")


def reg2bin(start: int, end: int, min_shift: int, depth: int) -> int:
    """Bin of the 1-based closed interval [start, end] (CSIv1 reg2bin)."""
    beg, end = start - 1, end - 1
    s, t = min_shift, ((1 << (depth * 3)) - 1) // 7
    for l in range(depth, 0, -1):
        if beg >> s == end >> s:
            return t + (beg >> s)
        s += 3
        t -= 1 << ((l - 1) * 3)
    return 0


def reg2bins(start: int, end: int, min_shift: int, depth: int) -> set:
    """Every bin that may hold a record overlapping [start, end] (CSIv1 reg2bins)."""
    beg, end = start - 1, end - 1
    out = set()
    t, s = 0, min_shift + depth * 3
    for l in range(depth + 1):
        out.update(range(t + (beg >> s), t + (end >> s) + 1))
        s -= 3
        t += 1 << (l * 3)
    return out


def parent_id(bid: int):
    return (bid - 1) // 8 if bid > 0 else None


def optimize_chunks(chunks, min_offset: int = 0):
    """Drop chunks ending at or before min_offset, sort by start, merge
    overlapping and touching chunks."""
    cs = sorted((c for c in chunks if c[1] > min_offset), key=lambda c: c[0])
    if not cs:
        return []
    out = []
    cur = cs[0]
    for nxt in cs[1:]:
        if nxt[0] > cur[1]:
            out.append(cur)
            cur = nxt
        elif cur[1] < nxt[1]:
            cur = (cur[0], nxt[1])
    out.append(cur)
    return out


def min_offset(ix: Index, ref: RefIndex, start: int) -> int:
    if ix.kind == "bai":
        i = (start - 1) // (1 << ix.min_shift)
        return ref.linear[i] if i < len(ref.linear) else 0
    bid = reg2bin(start, start, ix.min_shift, ix.depth)
    while bid is not None:
        if bid in ref.loffsets:
            return ref.loffsets[bid]
        bid = parent_id(bid)
    return 0


def resolve_interval(ix: Index, start, end):
    mp = max_position(ix.min_shift, ix.depth)
    s = 1 if start is None else start
    if s > mp:
        raise IndexError_("invalid start bound")
    e = mp if end is None else end
    if e > mp:
        raise IndexError_("invalid end bound")
    return s, e


def query(ix: Index, ref_id: int, start=None, end=None):
    """Merged chunks [(vstart, vend), ...] a region query reads, in order."""
    if not 0 <= ref_id < len(ix.refs):
        raise IndexError_(f"invalid reference sequence ID: {ref_id}")
    ref = ix.refs[ref_id]
    s, e = resolve_interval(ix, start, end)
    want = reg2bins(s, e, ix.min_shift, ix.depth)
    chunks = [c for bid, cs in ref.bins.items() if bid in want for c in cs]
    return optimize_chunks(chunks, min_offset(ix, ref, s))


# ---------------------------------------------------------------------------
# building an index (noodles' Indexer over a coordinate-sorted BAM)


def _merge_chunks_by_block(chunks):
    # reference_sequence.rs:303-330: merge when the current chunk's end block
    # reaches the next chunk's start block
    cs = sorted(chunks, key=lambda c: c[0])
    if not cs:
        return cs
    out = []
    cur = cs[0]
    for nxt in cs[1:]:
        if (cur[1] >> 16) >= (nxt[0] >> 16):
            if cur[1] < nxt[1]:
                cur = (cur[0], nxt[1])
        else:
            out.append(cur)
            cur = nxt
    out.append(cur)
    return out


def build_index(contexts, n_ref: int, min_shift=BAI_MIN_SHIFT, depth=BAI_DEPTH, kind="bai") -> Index:
    """contexts: per record in file order, (ref_id, start, end, is_mapped,
    chunk_start, chunk_end) with ref_id None when the record has no
    reference, start or end (then it counts as unplaced unmapped)."""
    refs: list = []
    unplaced = 0
    linear_kind = kind == "bai"
    lins: dict = {}
    for ref_id, start, end, mapped, c0, c1 in contexts:
        if ref_id is None:
            unplaced += 1
            continue
        if not refs:
            refs.append(RefIndex())
        if ref_id < len(refs) - 1:
            raise IndexError_("invalid reference sequence ID")
        while len(refs) <= ref_id:
            refs.append(RefIndex())
        ref = refs[ref_id]
        bid = reg2bin(start, end, min_shift, depth)
        chunks = ref.bins.setdefault(bid, [])
        if chunks and c0 <= chunks[-1][1]:
            chunks[-1] = (chunks[-1][0], c1)
        else:
            chunks.append((c0, c1))
        if linear_kind:
            w = (end - 1) // (1 << min_shift)
            if w + 1 > len(ref.linear):
                ref.linear.extend([c0] * (w + 1 - len(ref.linear)))
        else:
            # windows the record overlaps take its offset if unset
            lin = lins.setdefault(ref_id, {})
            for w in range((start - 1) >> min_shift, ((end - 1) >> min_shift) + 1):
                lin.setdefault(w, c0)
        m = ref.metadata or ((1 << 64) - 1, 0, 0, 0)
        ref.metadata = (min(m[0], c0), max(m[1], c1), m[2] + (1 if mapped else 0), m[3] + (0 if mapped else 1))
    while len(refs) < n_ref:
        refs.append(RefIndex())
    for ref in refs:
        # optimize: fold bins spanning under 64 KiB of compressed data into
        # their parent, deepest first, then merge each bin's chunks
        work = {k: list(v) for k, v in ref.bins.items()}
        kept = {}
        for bid in sorted(work, reverse=True):
            cs = work.pop(bid)
            pid = parent_id(bid)
            if pid is None:
                kept[bid] = cs
                continue
            if cs:
                lo = min(cs, key=lambda c: c[0])
                hi = max(cs, key=lambda c: c[0])
                span = (hi[1] >> 16) - (lo[0] >> 16)
            else:
                span = 0
            if span < (1 << 16) and pid in work:
                work[pid].extend(cs)
            else:
                kept[bid] = cs
        ref.bins = {k: _merge_chunks_by_block(v) for k, v in kept.items()}
    if not linear_kind:
        # A CSI bin's loffset is the linear offset of its first window, as
        # htslib writes it (hts.c update_loff: windows no record overlaps take
        # the previous window's offset, leading ones the reference's first
        # record). noodles' Indexer instead keeps the least chunk start of the
        # bin's own subtree (binned_index.rs:32-50), which can sit past a long
        # record of a neighbouring bin that overlaps the window; the query
        # side (min_offset, binned_index.rs:12-28) is the same for both.
        for rid, ref in enumerate(refs):
            lin = lins.get(rid, {})
            n = max(lin) + 1 if lin else 0
            arr, prev = [], ref.metadata[0] if ref.metadata else 0
            for w in range(n):
                prev = lin.get(w, prev)
                arr.append(prev)
            ref.loffsets = {}
            for bid in ref.bins:
                bot = bin_first_window(bid, depth)
                ref.loffsets[bid] = arr[bot] if bot < n else 0
    return Index(kind, min_shift, depth, refs, unplaced)


def bin_first_window(bid: int, depth: int) -> int:
    """First bottom-level window a bin covers."""
    level, first = 0, 0
    while first + (1 << (3 * level)) <= bid:
        first += 1 << (3 * level)
        level += 1
    return (bid - first) << (3 * (depth - level))


def write_bai(ix: Index) -> bytes:
    assert ix.kind == "bai"
    out = [b"BAI\x01", struct.pack("<I", len(ix.refs))]
    mid = metadata_id(ix.depth)
    for ref in ix.refs:
        n_bin = len(ref.bins) + (1 if ref.metadata is not None else 0)
        out.append(struct.pack("<I", n_bin))
        for bid in sorted(ref.bins):
            cs = ref.bins[bid]
            out.append(struct.pack("<II", bid, len(cs)))
            out.extend(struct.pack("<QQ", a, b) for a, b in cs)
        if ref.metadata is not None:
            out.append(struct.pack("<II", mid, 2) + struct.pack("<QQQQ", *ref.metadata))
        out.append(struct.pack("<I", len(ref.linear)))
        out.extend(struct.pack("<Q", v) for v in ref.linear)
    if ix.unplaced_unmapped is not None:
        out.append(struct.pack("<Q", ix.unplaced_unmapped))
    return b"".join(out)


def write_csi(ix: Index) -> bytes:
    assert ix.kind == "csi"
    out = [b"CSI\x01", struct.pack("<iii", ix.min_shift, ix.depth, len(ix.aux)), ix.aux,
           struct.pack("<i", len(ix.refs))]
    mid = metadata_id(ix.depth)
    for ref in ix.refs:
        n_bin = len(ref.bins) + (1 if ref.metadata is not None else 0)
        out.append(struct.pack("<i", n_bin))
        for bid in sorted(ref.bins):
            cs = ref.bins[bid]
            out.append(struct.pack("<IQi", bid, ref.loffsets.get(bid, 0), len(cs)))
            out.extend(struct.pack("<QQ", a, b) for a, b in cs)
        if ref.metadata is not None:
            out.append(struct.pack("<IQI", mid, 0, 2) + struct.pack("<QQQQ", *ref.metadata))
    if ix.unplaced_unmapped is not None:
        out.append(struct.pack("<Q", ix.unplaced_unmapped))
    return b"".join(out)


# ---------------------------------------------------------------------------
# records


REF_OPS = np.array([1, 0, 1, 1, 0, 0, 0, 1, 1] + [0] * 7, dtype=np.int64)  # M D N = X consume the reference


def record_fields(u: np.ndarray, rec: np.ndarray) -> dict:
    """Per record: ref_id, pos, flag, the raw CIGAR's reference span, and
    whether the CIGAR holds an op kind above 8 (an error in the reference's
    span)."""
    u = np.asarray(u, dtype=np.uint8)
    rec = np.asarray(rec, dtype=np.int64)
    def le(off, width, dt):
        idx = off[:, None] + np.arange(width)[None, :]
        return u[idx].copy().view(dt).reshape(-1) if len(off) else np.zeros(0, dt)
    ref_id = le(rec + 4, 4, "<i4").astype(np.int64)
    pos = le(rec + 8, 4, "<i4").astype(np.int64)
    l_name = u[rec + 12].astype(np.int64) if len(rec) else np.zeros(0, np.int64)
    n_cig = le(rec + 16, 2, "<u2").astype(np.int64)
    flag = le(rec + 18, 2, "<u2").astype(np.int64)
    c0 = rec + 36 + l_name
    tot = int(n_cig.sum())
    span = np.zeros(len(rec), dtype=np.int64)
    bad_op = np.zeros(len(rec), dtype=bool)
    if tot:
        owner = np.repeat(np.arange(len(rec)), n_cig)
        first = np.cumsum(n_cig) - n_cig
        k = np.arange(tot) - np.repeat(first, n_cig)
        ops = le(np.repeat(c0, n_cig) + 4 * k, 4, "<u4").astype(np.int64)
        kind = ops & 0xF
        np.add.at(span, owner, np.where(kind <= 8, REF_OPS[kind] * (ops >> 4), 0))
        np.logical_or.at(bad_op, owner, kind > 8)
    return {"ref_id": ref_id, "pos": pos, "flag": flag, "span": span, "bad_op": bad_op}


def alignment_contexts(fields: dict, vo_start: np.ndarray, vo_end: np.ndarray):
    """The indexer's per-record context (fritillaria-bam/src/fs/index.rs:45-80):
    (ref_id, start, end, is_mapped, chunk_start, chunk_end), ref_id None
    without a reference, start or end."""
    out = []
    for i in range(len(vo_start)):
        rid, pos, span = int(fields["ref_id"][i]), int(fields["pos"][i]), int(fields["span"][i])
        if rid < 0 or pos < 0:
            out.append((None, 0, 0, False, int(vo_start[i]), int(vo_end[i])))
            continue
        start = pos + 1
        end = start + span - 1 if span > 0 else start
        out.append((rid, start, end, not (int(fields["flag"][i]) & 0x4), int(vo_start[i]), int(vo_end[i])))
    return out


def intersects(fields: dict, ref_id: int, start=None, end=None):
    """(keep, bad) per record for a region query: keep when the record is on
    the region's reference and, for a bounded region, its alignment [start,
    end] overlaps the region's; bad where the reference's check raises (a
    negative ref_id or pos below -1, or a CIGAR op kind above 8, on a record
    of the region's reference)."""
    rid = fields["ref_id"]
    pos = fields["pos"]
    bad = rid < -1
    same = rid == ref_id
    if start is None and end is None:
        return same & ~bad, bad
    bad = bad | (same & (pos < -1)) | (same & (pos >= 0) & fields["bad_op"])
    a_start = pos + 1
    span = fields["span"]
    a_end = np.where(span > 0, a_start + span - 1, a_start)
    rs = 1 if start is None else start
    re = (1 << 63) - 1 if end is None else end
    keep = same & (pos >= 0) & (a_start <= re) & (rs <= a_end) & ~bad
    return keep, bad


def query_records(chunks, vo: np.ndarray) -> np.ndarray:
    """Indices of the records a csi Query over `chunks` reads: from each
    chunk's start (a record start), every record whose virtual offset is below
    the chunk's end, chunk after chunk."""
    vo = np.asarray(vo, dtype=np.uint64)
    out = []
    for a, b in chunks:
        i = int(np.searchsorted(vo, np.uint64(a), side="left"))
        if i >= len(vo) or int(vo[i]) != a:
            raise IndexError_(f"chunk start {a:#x} is not a record start")
        j = int(np.searchsorted(vo, np.uint64(b), side="left"))
        out.append(np.arange(i, max(i, j)))
    return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)


def region_records(ix: Index, fields: dict, vo: np.ndarray, ref_id: int, start=None, end=None):
    """(record indices, bad) of a region query: the chunk reads filtered by
    intersects, and whether a read record made the reference's check fail."""
    rows = query_records(query(ix, ref_id, start, end), vo)
    keep, bad = intersects(fields, ref_id, start, end)
    return rows[keep[rows]], bool(bad[rows].any())
