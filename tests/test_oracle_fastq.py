"""The FASTQ oracle (oracle/oracle.c or_fastq_*) against what pins it.

- record counts and read-length ranges of the FASTQ fixtures
  (testdata/README.md:319-322);
- the fixtures are `samtools fastq` output of the BAM fixtures
  (testdata/README.md:307-309), so every FASTQ record's name, bases and
  qualities must equal its BAM record's, decoded by the BAM oracle;
- the reference's hand-built cases: the '@' decoy inside a quality line, the
  length-matched decoy only the '+' check rejects, the seam newline, a short
  quality line at the buffer end, an empty read, a resume at a non-record
  offset (crates/fritillaria-fastq/src/columnar/record.rs:291-468).
"""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN

BASES = np.frombuffer(b"=ACMGRSVTWYHKDBN", dtype=np.uint8)
COMPLEMENT = bytes.maketrans(b"ACGTN", b"TGCAN")


def rec(name: bytes, seq: bytes, plus: bytes = b"") -> bytes:
    return b"@" + name + b"\n" + seq + b"\n+" + plus + b"\n" + b"I" * len(seq) + b"\n"


def fixture(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return gzip.decompress(f.read())


@pytest.mark.parametrize("name,count,lo,hi", [
    ("illumina.fastq.gz", 4000, 100, 100),
    ("ont_ultralong.fastq.gz", 3, 5556, 162932),
    ("pacbio_hifi.fastq.gz", 8, 13661, 20499),
    ("pacbio_hifi.plain.fastq.gz", 8, 13661, 20499),
])
def test_fixture_counts_and_lengths(oracle_mod, name, count, lo, hi):
    d = fixture(name)
    rc, offs, bounds, tail, _ = oracle_mod.fastq_scan_records(d, 0)
    assert rc == 0 and tail == len(d)
    assert len(offs) == count
    assert bounds[:, 3].min() == lo and bounds[:, 3].max() == hi
    # every line start that begins a record is on the chain (zero false positives)
    assert oracle_mod.fastq_survivors(d, 0) == count
    # the records tile the file
    ends = offs + bounds[:, 4]
    assert offs[0] == 0 and np.array_equal(ends[:-1], offs[1:]) and ends[-1] == len(d)


def bam_reads(oracle_mod, bam_name):
    d = gzip.decompress(open(os.path.join(GOLDEN, bam_name), "rb").read())
    rc, start, _, _ = oracle_mod.bam_parse_header(d)
    assert rc == 0
    rc, offs, tail = oracle_mod.bam_scan_records(d, start)
    assert rc == 0
    cols = oracle_mod.bam_decode_columns(d, offs)
    a = np.frombuffer(d, dtype=np.uint8)
    out = []
    for i, o in enumerate(offs):
        o = int(o)
        flag = int(cols["flags"][i])
        if flag & 0x900:
            continue
        name = bytes(a[o + 36: o + int(cols["cigar_start"][i]) - 1])
        ls = int(cols["sequence_len"][i])
        s0 = o + int(cols["sequence_start"][i])
        packed = a[s0: s0 + (ls + 1) // 2]
        nib = np.empty(2 * packed.size, dtype=np.uint8)
        nib[0::2] = packed >> 4
        nib[1::2] = packed & 15
        seq = bytes(BASES[nib[:ls]])
        q0 = o + int(cols["quality_start"][i])
        qual = bytes(a[q0: q0 + ls] + 33)
        if flag & 0x10:
            seq = seq.translate(COMPLEMENT)[::-1]
            qual = qual[::-1]
        out.append((name, seq, qual))
    return out


@pytest.mark.parametrize("fq,bam", [
    ("illumina.fastq.gz", "htslib_multiblock.bam"),
    ("ont_ultralong.fastq.gz", "ont_ultralong.bam"),
    ("pacbio_hifi.fastq.gz", "pacbio_hifi.bam"),
])
def test_fastq_records_equal_their_bam_records(oracle_mod, fq, bam):
    d = fixture(fq)
    rc, offs, bounds, tail, _ = oracle_mod.fastq_scan_records(d, 0)
    assert rc == 0
    reads = bam_reads(oracle_mod, bam)
    assert len(reads) == len(offs)
    for (name, seq, qual), o, b in zip(reads, offs, bounds):
        o = int(o)
        definition = d[o + 1: o + int(b[0]) - 1]
        assert definition.split(b" ")[0].split(b"\t")[0] == name
        assert d[o + int(b[0]): o + int(b[0]) + int(b[3])] == seq
        assert d[o + int(b[2]): o + int(b[2]) + int(b[3])] == qual


def test_at_inside_a_quality_line_is_not_a_record(oracle_mod):
    buf = bytearray(rec(b"r0", b"ACGT"))
    q = len(buf) - 5
    buf[q] = ord("@")
    buf += rec(b"r1", b"TTTT")
    rc, offs, _, tail, _ = oracle_mod.fastq_scan_records(bytes(buf), 0)
    assert rc == 0 and len(offs) == 2 and tail == len(buf)
    assert oracle_mod.fastq_bounds(bytes(buf), q) is None


def test_only_the_plus_check_rejects_a_length_matched_decoy(oracle_mod):
    buf = b"@A\nACG\n+\n@BC\n" + b"@XY\nACGT\n+ab\nIIII\n"
    assert buf[9:13] == b"@BC\n"
    assert oracle_mod.fastq_bounds(buf, 9) is None
    rc, offs, _, tail, _ = oracle_mod.fastq_scan_records(buf, 0)
    assert rc == 0 and list(offs) == [0, 13] and tail == len(buf)
    assert oracle_mod.fastq_survivors(buf, 0) == 2


def test_bounds_of_a_minimal_record(oracle_mod):
    buf = rec(b"r0 length=4", b"ACGT", b"r0")
    # '@r0 length=4\n' = 13, 'ACGT\n' = 5, '+r0\n' = 4, 'IIII\n' = 5
    assert oracle_mod.fastq_bounds(buf, 0) == (13, 18, 22, 4, 27)


def test_record_ending_at_a_batch_edge_resumes_cleanly(oracle_mod):
    whole = rec(b"r0", b"ACGT")
    first = whole[:-1]
    rc, offs, b, tail, _ = oracle_mod.fastq_scan_records(first, 0)
    assert rc == 0 and list(offs) == [0] and tail == len(first) and b[0, 4] == len(first)
    second = b"\n" + rec(b"r1", b"TTTT")
    rc, offs, _, tail, _ = oracle_mod.fastq_scan_records(second, 0)
    assert rc == 0 and list(offs) == [1] and tail == len(second)


def test_partial_record_is_the_tail(oracle_mod):
    buf = rec(b"whole", b"ACGT")
    boundary = len(buf)
    buf += rec(b"partial", b"ACGTACGT")
    buf = buf[:-4]
    rc, offs, _, tail, _ = oracle_mod.fastq_scan_records(buf, 0)
    assert rc == 0 and len(offs) == 1 and tail == boundary


def test_short_quality_line_at_the_end_is_rejected(oracle_mod):
    buf = rec(b"r0", b"ACGT")[:-2]
    assert oracle_mod.fastq_bounds(buf, 0) is None
    rc, offs, _, tail, _ = oracle_mod.fastq_scan_records(buf, 0)
    assert rc == 0 and len(offs) == 0 and tail == 0


def test_empty_read_is_legal(oracle_mod):
    buf = rec(b"r0", b"")
    assert oracle_mod.fastq_bounds(buf, 0) == (4, 5, 7, 0, 8)
    rc, offs, _, _, _ = oracle_mod.fastq_scan_records(buf, 0)
    assert rc == 0 and len(offs) == 1


def test_resume_at_a_non_record_offset_is_an_error(oracle_mod):
    buf = rec(b"r0", b"ACGT")
    rc, offs, _, _, err = oracle_mod.fastq_scan_records(buf, 3)
    # offset 3 is the newline ending the name line; it is skipped, then "A" is no record start
    assert rc == oracle_mod.E_MALFORMED and err == 4


def test_blank_lines_between_records_are_skipped(oracle_mod):
    buf = rec(b"a", b"AC") + b"\n\n" + rec(b"b", b"GT") + b"\n"
    rc, offs, _, tail, _ = oracle_mod.fastq_scan_records(buf, 0)
    assert rc == 0 and list(offs) == [0, len(rec(b"a", b"AC")) + 2] and tail == len(buf)
