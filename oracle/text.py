"""Text framing (SAM, VCF, BED, GFF, GTF) and FASTA compaction -- TEST
INFRASTRUCTURE ONLY.

A CPU restatement, in numpy, of the reference's host decoders for the two
device paths; only ``tests/`` imports it. The product runs
``fritillaria_amd/csrc/text.hip`` and ``fasta.hip``.

Reference functions restated here (crates/...):
  scan_lines, decode      fritillaria-text/src/columnar/batch.rs:56-86 (lines from
                          `start`, blank lines skipped, a line without a newline
                          ends the scan and is the tail; a line is a header when
                          its first byte is the dialect's comment or secondary
                          byte), :236-258 (a record's end excludes its newline;
                          tabs of record lines only, in order; field starts)
  fasta bounds_at         fritillaria-fasta/src/columnar/record.rs:58-97
  fasta scan_records      fritillaria-fasta/src/columnar/record.rs:99-126
  fasta compaction        fritillaria-fasta/src/columnar/record.rs:177-205 and
                          batch.rs:120-156, as FastaCompactor::compact uses them
                          (fritillaria-cuda/src/fasta.rs: contig starts are '>' at
                          a line start from offset 0, the layout is decoded from the
                          first of them with at_eof = true)

What pins them (tests/test_oracle_text.py):
  - the reference's unit vectors (batch.rs tests: SAMPLE, ragged tables,
    unterminated final line, blank lines, quotes, dialects; record.rs tests:
    wrapped records, non-uniform wrapping, blank line between records, final
    line without a newline, empty record);
  - the fixtures' published counts (testdata/README.md: header lines, records
    and fields of reads.sam, calls.vcf, lambda.gff3, lambda.gtf, cytoband.bed)
    and controls.fa.fai (samtools faidx: name, length, sequence offset, line
    bases, line width of both contigs);
  - other formats: reads.sam is `samtools view -h` of htslib_multiblock.bam and
    calls.vcf is `bcftools view` of giab_hg002.bcf, so each SAM record's QNAME
    and POS, and each VCF record's POS, equal the BAM / BCF oracles' fields;
  - brute force: a plain split-on-newline restatement over random buffers.

Compaction where the reference is undefined: the reference compacts a contig
it calls uniform by the formula i - i / line_width. Its uniformity test skips
empty lines and never looks at the last line, so a contig with a blank line
inside its sequence, or a last line longer than the others, passes it and the
formula then writes past the contig or over its own bases (the host path
panics on the out-of-range index). The device path here removes the newlines
of every contig, which is the formula's result wherever the formula is in
range and consistent; ``fasta_compact`` below does the same, and
``formula_compaction`` keeps the reference's formula so tests can show the two
agree on every input where it is well defined.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SAM = (ord("@"), None)
VCF = (ord("#"), None)
GFF = (ord("#"), None)
BED = (ord("#"), None)


@dataclass
class TextBatch:
    record_offsets: np.ndarray
    record_ends: np.ndarray
    header_offsets: np.ndarray
    tabs: np.ndarray
    field_starts: np.ndarray   # records + 1 entries (just [0] without records)
    tail: int


def _u8(buf) -> np.ndarray:
    if isinstance(buf, np.ndarray):
        return buf.astype(np.uint8, copy=False)
    return np.frombuffer(bytes(buf), dtype=np.uint8)


def text_decode(buf, start: int = 0, dialect=SAM) -> TextBatch:
    """RecordBatch::decode over buf[start:]."""
    u = _u8(buf)
    n = u.size
    comment, secondary = dialect
    if start >= n:
        e = np.zeros(0, np.uint64)
        return TextBatch(e, e, e, e, np.zeros(1, np.uint64), start if start < n else n)
    nl = np.flatnonzero(u == 10)
    nl = nl[nl >= start]
    # line starts: start, and every position after a newline, from start on
    cand = np.concatenate([[start], nl + 1])
    cand = cand[cand < n]
    cand = cand[u[cand] != 10]                    # blank lines are skipped
    # a line needs a newline at or after its start
    last_nl = nl[-1] if nl.size else -1
    tail = n
    if cand.size and cand[-1] > last_nl:
        tail = int(cand[-1])
        cand = cand[:-1]
    first = u[cand]
    hdr = first == comment
    if secondary is not None:
        hdr |= first == secondary
    rec = cand[~hdr]
    # each line ends at the first newline at or after its start
    ends = nl[np.searchsorted(nl, rec, side="left")] if rec.size else np.zeros(0, np.int64)
    tabs_all = np.flatnonzero(u == 9)
    tabs_all = tabs_all[tabs_all >= start]
    # a tab belongs to the record whose [offset, end) holds it
    k = np.searchsorted(rec, tabs_all, side="right") - 1
    ok = k >= 0
    ok[ok] &= tabs_all[ok] < ends[k[ok]]
    tabs = tabs_all[ok]
    per = np.bincount(k[ok], minlength=rec.size) if rec.size else np.zeros(0, np.int64)
    starts = np.concatenate([[0], np.cumsum(per)]).astype(np.uint64)
    return TextBatch(rec.astype(np.uint64), ends.astype(np.uint64), cand[hdr].astype(np.uint64),
                     tabs.astype(np.uint64), starts, tail)


def text_decode_naive(buf: bytes, start: int = 0, dialect=SAM) -> TextBatch:
    """The same decode written as a loop over split lines (brute force)."""
    comment, secondary = dialect
    data = bytes(buf)
    rec, ends, hdrs, tabs, starts = [], [], [], [], [0]
    pos = start
    tail = len(data) if start < len(data) else len(data)
    if start >= len(data):
        tail = min(start, len(data))
    while pos < len(data):
        if data[pos] == 10:
            pos += 1
            continue
        e = data.find(b"\n", pos)
        if e < 0:
            tail = pos
            break
        if data[pos] == comment or (secondary is not None and data[pos] == secondary):
            hdrs.append(pos)
        else:
            rec.append(pos)
            ends.append(e)
            tabs.extend(pos + i for i, b in enumerate(data[pos:e]) if b == 9)
            starts.append(len(tabs))
        pos = e + 1
    else:
        tail = max(pos, tail) if start < len(data) else tail
    a = lambda v: np.asarray(v, dtype=np.uint64)
    return TextBatch(a(rec), a(ends), a(hdrs), a(tabs), a(starts), tail)


# ---------------------------------------------------------------------------
# FASTA


@dataclass
class FastaBounds:
    sequence_start: int   # from the record start
    sequence_len: int
    line_bases: int
    line_width: int
    record_end: int       # from the record start
    uniform: bool


def fasta_bounds(u: np.ndarray, pos: int):
    """bounds_at: None when pos is not '>' or its definition line has no newline."""
    n = u.size
    if pos >= n or u[pos] != ord(">"):
        return None
    rest = u[pos:]
    nl = np.flatnonzero(rest == 10)
    if nl.size == 0:
        return None
    seq0 = int(nl[0]) + 1
    # the record ends at the first line (after the definition) that starts with '>'
    line_starts = np.concatenate([[seq0], nl[nl >= seq0] + 1])
    line_starts = line_starts[line_starts < rest.size]
    gt = line_starts[rest[line_starts] == ord(">")]
    end = int(gt[0]) if gt.size else rest.size
    span = rest[seq0:end]
    snl = np.flatnonzero(span == 10)
    # line lengths: [prev + 1, nl) for each newline, and the final piece if unterminated
    bounds = np.concatenate([[-1], snl])
    lengths = (snl - bounds[:-1] - 1).tolist()
    if span.size and (snl.size == 0 or snl[-1] != span.size - 1):
        lengths.append(int(span.size - (snl[-1] + 1 if snl.size else 0)))
    seq_len = int(span.size - snl.size)
    nonempty = [l for l in lengths if l > 0]
    lb = nonempty[0] if nonempty else 0
    uniform = len(nonempty) < 2 or all(l == lb for l in nonempty[:-1])
    return FastaBounds(seq0, seq_len, lb, lb + 1, end, uniform)


def fasta_scan_records(u: np.ndarray, start: int, at_eof: bool = True):
    """scan_records: (record offsets, tail); raises ValueError at a position
    that is neither a newline nor '>'."""
    offsets = []
    pos = start
    while pos < u.size:
        if u[pos] == 10:
            pos += 1
            continue
        if u[pos] != ord(">"):
            raise ValueError(f"expected a record to start with '>' at {pos}")
        b = fasta_bounds(u, pos)
        if b is None:
            break
        offsets.append(pos)
        pos += b.record_end
    if not at_eof and offsets:
        last = offsets.pop()
        return offsets, last
    return offsets, pos


@dataclass
class FastaReference:
    record_offsets: np.ndarray
    sequence_offsets: np.ndarray
    sequence_lengths: np.ndarray
    sequence: np.ndarray
    bounds: list


def fasta_compact(buf) -> FastaReference:
    """FastaCompactor::compact: contigs from the first '>' at a line start,
    each contig's sequence with its newlines removed, back to back."""
    u = _u8(buf)
    gt = np.flatnonzero(u == ord(">"))
    gt = gt[(gt == 0) | (u[np.maximum(gt - 1, 0)] == 10)]
    if gt.size == 0:
        e = np.zeros(0, np.uint64)
        return FastaReference(e, e, e, np.zeros(0, np.uint8), [])
    offs, _ = fasta_scan_records(u, int(gt[0]), True)
    bounds = [fasta_bounds(u, o) for o in offs]
    lens = np.array([b.sequence_len for b in bounds], dtype=np.uint64)
    soff = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint64) if len(offs) else np.zeros(0, np.uint64)
    seq = []
    for o, b in zip(offs, bounds):
        span = u[o + b.sequence_start:o + b.record_end]
        seq.append(span[span != 10])
    s = np.concatenate(seq) if seq else np.zeros(0, np.uint8)
    return FastaReference(np.asarray(offs, np.uint64), soff, lens, s, bounds)


def formula_compaction(u: np.ndarray, offset: int, b: FastaBounds):
    """The reference's uniform path for one contig: dst[i - i / line_width] for
    every non-newline byte, or None where that index leaves the contig."""
    span = u[offset + b.sequence_start:offset + b.record_end]
    out = np.zeros(b.sequence_len, np.uint8)
    idx = np.flatnonzero(span != 10)
    if b.line_width == 0:
        return None
    dst = idx - idx // b.line_width
    if dst.size and int(dst.max()) >= b.sequence_len:
        return None
    out[dst] = span[idx]
    return out
