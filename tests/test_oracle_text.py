"""The text-framing and FASTA oracle pinned: the reference's unit vectors, the
fixtures' published counts and samtools' .fai, the BAM and BCF oracles on the
same data, and a split-on-newline restatement over random buffers."""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_bytes

from oracle import text as ot

SAMPLE = b"@HD\tVN:1.6\n@SQ\tSN:c0\tLN:100\nr0\t0\tc0\t1\t60\t4M\nr1\t16\tc0\t9\t60\t4M\n"


def fields(b: ot.TextBatch, buf: bytes, i: int):
    a, e = int(b.record_offsets[i]), int(b.record_ends[i])
    t0, t1 = int(b.field_starts[i]), int(b.field_starts[i + 1])
    cuts = [a - 1] + [int(x) for x in b.tabs[t0:t1]] + [e]
    return [buf[cuts[k] + 1:cuts[k + 1]] for k in range(len(cuts) - 1)]


def test_text_unit_vectors():
    # fritillaria-text/src/columnar/batch.rs tests
    b = ot.text_decode(SAMPLE, 0, ot.SAM)
    assert len(b.header_offsets) == 2 and len(b.record_offsets) == 2 and b.tail == len(SAMPLE)
    assert fields(b, SAMPLE, 0) == [b"r0", b"0", b"c0", b"1", b"60", b"4M"]
    assert fields(b, SAMPLE, 1)[0] == b"r1" and fields(b, SAMPLE, 1)[5] == b"4M"
    assert SAMPLE[int(b.record_offsets[0]):int(b.record_ends[0])] == b"r0\t0\tc0\t1\t60\t4M"
    b = ot.text_decode(b"a\tb\nc\td\te\nf\n", 0, ot.SAM)
    assert b.field_starts.tolist() == [0, 1, 3, 3]
    b = ot.text_decode(b"a\tb\nc\td", 0, ot.SAM)
    assert len(b.record_offsets) == 1 and b.tail == 4
    assert len(ot.text_decode(b"a\tb\n\n\nc\td\n", 0, ot.SAM).record_offsets) == 2
    q = b"r0\t\"\"\"\t\"a\tb\"\n"
    assert fields(ot.text_decode(q, 0, ot.SAM), q, 0) == [b"r0", b"\"\"\"", b"\"a", b"b\""]
    buf = b"#comment\n@notacomment\n"
    s, v = ot.text_decode(buf, 0, ot.SAM), ot.text_decode(buf, 0, ot.VCF)
    assert (len(s.header_offsets), len(s.record_offsets)) == (1, 1)
    assert (len(v.header_offsets), len(v.record_offsets)) == (1, 1)
    e = ot.text_decode(b"", 0, ot.SAM)
    assert len(e.record_offsets) == 0 and e.tail == 0


@pytest.mark.parametrize("name,dialect,headers,records,nfields", [
    ("reads.sam", ot.SAM, 4, 4000, 12),
    ("calls.vcf", ot.VCF, 235, 275, 10),
    # testdata/README.md lists 309 records; the file holds 308 record lines and
    # a blank line (fritillaria-text/tests/real_files.rs compares with noodles'
    # GFF reader instead of that number)
    ("lambda.gff3", ot.GFF, 5, 308, 9),
    ("lambda.gtf", ot.GFF, 3, 357, 9),
    ("cytoband.bed", ot.BED, 0, 1549, 5),
])
def test_text_fixture_counts(name, dialect, headers, records, nfields):
    # testdata/README.md and fritillaria-text/tests/real_files.rs:240-300:
    # header lines, records and fields per record
    buf = golden_bytes(name)
    b = ot.text_decode(buf, 0, dialect)
    assert len(b.record_offsets) == records and b.tail == len(buf)
    assert len(b.header_offsets) == headers
    assert (np.diff(b.field_starts) == nfields - 1).all()
    n = ot.text_decode_naive(buf, 0, dialect)
    for k in ("record_offsets", "record_ends", "header_offsets", "tabs", "field_starts"):
        assert np.array_equal(getattr(b, k), getattr(n, k)), k


def test_sam_records_are_the_bam_fixture(oracle_mod):
    """reads.sam is samtools view of htslib_multiblock.bam: QNAME and POS of
    every record equal the BAM oracle's read name and position + 1; the quote
    bytes that trap CSV-style scanners are there."""
    buf = golden_bytes("reads.sam")
    b = ot.text_decode(buf, 0, ot.SAM)
    raw = golden_bytes("htslib_multiblock.bam")
    d = oracle_mod.discover(raw)
    r = oracle_mod.inflate_batch(raw, d.spans)
    u = np.frombuffer(r.data, np.uint8)
    rc, start, _, _ = oracle_mod.bam_parse_header(u)
    rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, r.offsets, start)
    cols = oracle_mod.bam_decode_columns(u, rec)
    assert len(rec) == len(b.record_offsets) == 4000
    for i in range(0, 4000, 7):
        f = fields(b, buf, i)
        o = int(rec[i])
        l_name = int(u[o + 12])
        assert f[0] == bytes(u[o + 36:o + 36 + l_name - 1])
        assert int(f[3]) == int(cols["position"][i]) + 1
    assert buf.count(b'"') > 1000


def test_vcf_records_are_the_bcf_fixture(oracle_mod):
    buf = golden_bytes("calls.vcf")
    b = ot.text_decode(buf, 0, ot.VCF)
    d = gzip.decompress(open(os.path.join(GOLDEN, "giab_hg002.bcf"), "rb").read())
    rs, _, _ = oracle_mod.bcf_parse_header(d)
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(d, rs)
    rc, cols, _ = oracle_mod.bcf_decode_columns(d, offs, tail)
    assert len(offs) == len(b.record_offsets) == 275
    pos = [int(fields(b, buf, i)[1]) for i in range(275)]
    assert pos == (cols["position"].astype(np.int64) + 1).tolist()


def test_text_matches_brute_force_on_random_buffers():
    rng = np.random.default_rng(3)
    alphabet = np.frombuffer(b"\n\n\t\t\t@#ACGT\"", np.uint8)
    for trial in range(400):
        n = int(rng.integers(0, 300))
        buf = alphabet[rng.integers(0, alphabet.size, n)].tobytes()
        start = int(rng.integers(0, n + 2))
        for dialect in (ot.SAM, ot.VCF, (ord("#"), ord("@"))):
            a = ot.text_decode(buf, start, dialect)
            e = ot.text_decode_naive(buf, start, dialect)
            for k in ("record_offsets", "record_ends", "header_offsets", "tabs", "field_starts"):
                assert np.array_equal(getattr(a, k), getattr(e, k)), (trial, k)
            assert a.tail == e.tail, (trial, buf, start)


# ---------------------------------------------------------------------------
# FASTA


def build_record(name: bytes, seq: bytes, width: int) -> bytes:
    out = b">" + name + b"\n"
    for i in range(0, len(seq), width):
        out += seq[i:i + width] + b"\n"
    return out


def u8(b):
    return np.frombuffer(b, np.uint8)


def test_fasta_unit_vectors():
    # fritillaria-fasta/src/columnar/record.rs tests
    buf = build_record(b"c0", b"ACGTACGTAC", 4)
    b = ot.fasta_bounds(u8(buf), 0)
    assert (b.sequence_len, b.line_bases, b.line_width, b.uniform) == (10, 4, 5, True)
    assert ot.fasta_compact(buf).sequence.tobytes() == b"ACGTACGTAC"
    seq = bytes(b"ACGT"[i % 4] for i in range(250))
    for width in [1, 2, 7, 60, 70, 249, 250, 300]:
        buf = build_record(b"c0", seq, width)
        b = ot.fasta_bounds(u8(buf), 0)
        assert b.uniform
        assert ot.fasta_compact(buf).sequence.tobytes() == seq
        assert ot.formula_compaction(u8(buf), 0, b).tobytes() == seq
    buf = b">c0\nACGTACGT\nACG\nACGTACGT\n"
    b = ot.fasta_bounds(u8(buf), 0)
    assert not b.uniform and b.sequence_len == 19
    assert ot.fasta_compact(buf).sequence.tobytes() == b"ACGTACGTACGACGTACGT"
    buf = b">a\nACGT\nACGT\nAC\n\n>b\nTTTT\n"
    offs, _ = ot.fasta_scan_records(u8(buf), 0)
    assert len(offs) == 2
    b = ot.fasta_bounds(u8(buf), 0)
    assert b.sequence_len == 10 and b.uniform
    assert ot.fasta_compact(buf).sequence.tobytes() == b"ACGTACGTACTTTT"
    buf = build_record(b"c0", b"ACGTACGT", 4) + build_record(b"c1", b"TTTT", 4) + build_record(b"c2", b"GG", 4)
    offs, tail = ot.fasta_scan_records(u8(buf), 0)
    assert len(offs) == 3 and tail == len(buf)
    first = build_record(b"c0", b"ACGTACGT", 4)
    assert ot.fasta_bounds(u8(first + build_record(b"c1", b"TTTT", 4)), 0).record_end == len(first)
    b = ot.fasta_bounds(u8(b">c0\nACGT\nAC"), 0)
    assert b.sequence_len == 6
    assert ot.fasta_compact(b">c0\nACGT\nAC").sequence.tobytes() == b"ACGTAC"
    assert ot.fasta_bounds(u8(b">c0\n"), 0).sequence_len == 0
    with pytest.raises(ValueError):
        ot.fasta_scan_records(u8(build_record(b"c0", b"ACGT", 4)), 2)
    # a definition line without a newline is not a record
    assert ot.fasta_bounds(u8(b">c0"), 0) is None
    r = ot.fasta_compact(b">a\nAC\n>b")
    assert r.record_offsets.tolist() == [0] and r.sequence.tobytes() == b"AC"


def test_fasta_fixture_matches_the_fai():
    buf = golden_bytes("controls.fa")
    fai = [line.split("\t") for line in golden_bytes("controls.fa.fai").decode().splitlines()]
    r = ot.fasta_compact(buf)
    assert len(r.record_offsets) == len(fai) == 2
    for (name, length, offset, lb, lw), o, b in zip(fai, r.record_offsets, r.bounds):
        o = int(o)
        assert buf[o + 1:o + 1 + len(name)] == name.encode()
        assert b.sequence_len == int(length)
        assert o + b.sequence_start == int(offset)
        assert (b.line_bases, b.line_width, b.uniform) == (int(lb), int(lw), True)
    assert r.sequence_lengths.tolist() == [5386, 48502]
    assert set(np.unique(r.sequence).tobytes()) <= set(b"ACGT")
    assert r.sequence.size == 5386 + 48502


def test_fasta_compaction_agrees_with_the_formula_where_it_is_defined():
    rng = np.random.default_rng(9)
    agreed = undefined = 0
    for trial in range(300):
        parts = []
        for c in range(int(rng.integers(1, 4))):
            L = int(rng.integers(0, 200))
            seq = bytes(rng.choice(list(b"ACGTN"), L).tolist())
            width = int(rng.integers(1, 80))
            rec = build_record(b"c%d" % c, seq, width)
            if rng.random() < 0.3:  # a blank line, a short or long line somewhere
                lines = rec.split(b"\n")
                k = int(rng.integers(1, len(lines)))
                lines.insert(k, b"" if rng.random() < 0.5 else b"ACGTACGTAC"[:int(rng.integers(1, 10))])
                rec = b"\n".join(lines)
            parts.append(rec + (b"\n" if rng.random() < 0.2 else b""))
        buf = b"".join(parts)
        u = u8(buf)
        r = ot.fasta_compact(buf)
        for i, (o, b) in enumerate(zip(r.record_offsets, r.bounds)):
            got = r.sequence[int(r.sequence_offsets[i]):int(r.sequence_offsets[i]) + b.sequence_len]
            span = u[int(o) + b.sequence_start:int(o) + b.record_end]
            assert got.tobytes() == span[span != 10].tobytes()
            if b.uniform:
                f = ot.formula_compaction(u, int(o), b)
                if f is None:
                    undefined += 1
                elif np.array_equal(f, got):
                    agreed += 1
    assert agreed > 100 and undefined > 0
