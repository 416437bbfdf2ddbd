"""Text framing and FASTA compaction on the device against the oracle
(crates/fritillaria-cuda/tests/text_scan.rs and fasta_compact.rs compare the
device columns with the host decoders the same way): the reference's
fixtures, its unit vectors, random buffers whose lines and contigs cross the
4 KiB tiles, and large synthetic inputs."""
import numpy as np
import pytest

from conftest import golden_bytes
from helpers import bgzf

from oracle import text as ot

pytestmark = pytest.mark.gpu

FIXTURES = [("reads.sam", ot.SAM), ("calls.vcf", ot.VCF), ("lambda.gff3", ot.GFF), ("lambda.gtf", ot.GFF),
            ("cytoband.bed", ot.BED)]


def device_bytes(ctx, data: bytes):
    buf = ctx.alloc(len(data))
    if data:
        buf.copy_from_host(np.frombuffer(data, np.uint8))
    return buf


def dialect_of(d):
    import fritillaria_amd as fa
    return fa.Dialect(bytes([d[0]]), bytes([d[1]]) if d[1] is not None else None)


def check_text(scanner, ctx, data: bytes, start: int, d, label=""):
    got = scanner.scan(device_bytes(ctx, data), start, dialect_of(d))
    want = ot.text_decode(data, start, d)
    assert len(got) == len(want.record_offsets), label
    assert got.header_count() == len(want.header_offsets), label
    assert got.tail() == want.tail, label
    h = got.to_host()
    for k in ("record_offsets", "record_ends", "header_offsets", "tabs"):
        assert np.array_equal(h[k], getattr(want, k)), (label, k)
    if len(got):
        assert np.array_equal(h["field_starts"], want.field_starts), label
    else:
        assert h["field_starts"].size == 0
    return got


@pytest.fixture(scope="module")
def scanner(hip_ctx):
    import fritillaria_amd as fa
    return fa.TextScanner(hip_ctx)


@pytest.fixture(scope="module")
def compactor(hip_ctx):
    import fritillaria_amd as fa
    return fa.FastaCompactor(hip_ctx)


@pytest.mark.parametrize("name,d", FIXTURES)
def test_text_fixtures(hip_ctx, scanner, name, d):
    print(f"[gpu] text: {name}", flush=True)
    check_text(scanner, hip_ctx, golden_bytes(name), 0, d, name)


def test_text_through_a_bgzf_batch(hip_ctx, scanner, codec):
    """Staged as the reference's test stages it: BGZF, inflated on the device."""
    print("[gpu] text: reads.sam through a device inflate batch", flush=True)
    import fritillaria_amd as fa
    data = golden_bytes("reads.sam")
    raw = bgzf(data)
    batch = codec.inflate_batch_device(raw, fa.discover(raw)[0])
    got = scanner.scan(batch, 0, fa.Dialect.SAM)
    want = ot.text_decode(data, 0, ot.SAM)
    assert np.array_equal(got.column_to_host("tabs"), want.tabs)
    assert np.array_equal(got.column_to_host("record_ends"), want.record_ends)


def test_text_unit_vectors(hip_ctx, scanner):
    print("[gpu] text: unit vectors", flush=True)
    for buf in [b"@HD\tVN:1.6\n@SQ\tSN:c0\tLN:100\nr0\t0\tc0\t1\t60\t4M\nr1\t16\tc0\t9\t60\t4M\n",
                b"a\tb\nc\td\te\nf\n", b"a\tb\nc\td", b"a\tb\n\n\nc\td\n", b"r0\t\"\"\"\t\"a\tb\"\n",
                b"#comment\n@notacomment\n", b"", b"\n\n\n", b"@HD\tVN:1.6\n@SQ\tSN:c0\tLN:100\n",
                b"x", b"@h\tpartial"]:
        for d in (ot.SAM, ot.VCF):
            for start in (0, 1, 3, len(buf), len(buf) + 5):
                check_text(scanner, hip_ctx, buf, start, d, (buf, start))


def test_text_random_buffers_across_tiles(hip_ctx, scanner):
    print("[gpu] text: random buffers across tiles", flush=True)
    rng = np.random.default_rng(7)
    for trial in range(120):
        n = int(rng.choice([100, 4095, 4096, 4097, 9000, 20000, 70000]))
        # lines of mixed lengths, some longer than a tile
        parts, size = [], 0
        while size < n:
            L = int(rng.choice([0, 1, 5, 30, 200, 5000]))
            first = rng.choice([b"@", b"#", b"r", b"\t"])
            body = rng.choice(np.frombuffer(b"ACGT\t\t\"@#", np.uint8), L).tobytes()
            parts.append(first + body + b"\n" * int(rng.integers(1, 3)))
            size += L + 2
        data = b"".join(parts)[:n]
        start = int(rng.integers(0, n)) if rng.random() < 0.3 else 0
        d = [ot.SAM, ot.VCF, (ord("#"), ord("@"))][trial % 3]
        check_text(scanner, hip_ctx, data, start, d, (trial, n, start))


def test_text_large_vcf_like(hip_ctx, scanner):
    """Lines of 2,504 sample columns (20 KB each), the shape that makes one
    thread per record slow, and a SAM-like file of 600,000 short records."""
    print("[gpu] text: large inputs", flush=True)
    rng = np.random.default_rng(8)
    gts = np.array([b"0|0", b"0|1", b"1|0", b"1|1"])
    lines = [b"##fileformat=VCFv4.2", b"#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t" +
             b"\t".join(b"S%d" % i for i in range(2504))]
    for i in range(400):
        lines.append(b"22\t%d\t.\tA\tG\t50\tPASS\tAC=1\tGT\t" % (16050000 + i) +
                     b"\t".join(gts[rng.integers(0, 4, 2504)].tolist()))
    data = b"\n".join(lines) + b"\n"
    b = check_text(scanner, hip_ctx, data, 0, ot.VCF, "vcf")
    assert len(b) == 400 and b.header_count() == 2
    recs = [b"r%07d\t0\tchr1\t%d\t60\t150M\t*\t0\t0\tACGT\tIIII" % (i, i) for i in range(600_000)]
    data = b"@HD\tVN:1.6\n" + b"\n".join(recs) + b"\nr_partial\t0"
    b = check_text(scanner, hip_ctx, data, 0, ot.SAM, "sam")
    assert len(b) == 600_000 and b.tail() == len(data) - len(b"r_partial\t0")


# ---------------------------------------------------------------------------
# FASTA


def check_fasta(compactor, ctx, data: bytes, label=""):
    got = compactor.compact(device_bytes(ctx, data))
    want = ot.fasta_compact(data)
    assert len(got) == len(want.record_offsets), label
    assert got.total_bases() == want.sequence.size, label
    for k in ("record_offsets", "sequence_offsets", "sequence_lengths"):
        assert np.array_equal(got.column_to_host(k), getattr(want, k)), (label, k)
    assert np.array_equal(got.column_to_host("sequence"), want.sequence), label
    return got


def build_record(name: bytes, seq: bytes, width: int) -> bytes:
    out = b">" + name + b"\n"
    for i in range(0, len(seq), width):
        out += seq[i:i + width] + b"\n"
    return out


def test_fasta_fixture(hip_ctx, compactor):
    print("[gpu] fasta: controls.fa", flush=True)
    data = golden_bytes("controls.fa")
    got = check_fasta(compactor, hip_ctx, data, "controls.fa")
    fai = [line.split("\t") for line in golden_bytes("controls.fa.fai").decode().splitlines()]
    assert got.column_to_host("sequence_lengths").tolist() == [int(f[1]) for f in fai]


def test_fasta_unit_vectors(hip_ctx, compactor):
    print("[gpu] fasta: unit vectors", flush=True)
    seq = bytes(b"ACGT"[i % 4] for i in range(250))
    cases = [build_record(b"c0", b"ACGTACGTAC", 4), b">c0\nACGTACGT\nACG\nACGTACGT\n",
             b">a\nACGT\nACGT\nAC\n\n>b\nTTTT\n", b">c0\nACGT\nAC", b">c0\n", b">a\nAC\n>b", b"", b"no contig here\n",
             b"junk\n>c0\nAC\nGT\n", b">only-a-definition"]
    cases += [build_record(b"c0", seq, w) for w in [1, 2, 7, 60, 70, 249, 250, 300]]
    for c in cases:
        check_fasta(compactor, hip_ctx, c, c[:40])


def test_fasta_random_across_tiles(hip_ctx, compactor):
    print("[gpu] fasta: random contigs across tiles", flush=True)
    rng = np.random.default_rng(11)
    for trial in range(80):
        parts = []
        for c in range(int(rng.integers(1, 12))):
            L = int(rng.choice([0, 3, 100, 4000, 9000, 30000]))
            seq = rng.choice(np.frombuffer(b"ACGTN", np.uint8), L).tobytes()
            rec = build_record(b"c%d desc" % c, seq, int(rng.integers(1, 120)))
            if rng.random() < 0.3:
                lines = rec.split(b"\n")
                k = int(rng.integers(1, len(lines)))
                lines.insert(k, b"" if rng.random() < 0.5 else b"ACGTTT")
                rec = b"\n".join(lines)
            parts.append(rec + (b"\n" if rng.random() < 0.3 else b""))
        data = b"".join(parts)
        if rng.random() < 0.2:
            data = data[:-1]  # no final newline
        if rng.random() < 0.1:
            data += b">dangling"
        check_fasta(compactor, hip_ctx, data, trial)


def test_fasta_large_genome(hip_ctx, compactor):
    """Three contigs of 40, 25 and 0.5 Mbp wrapped at 60 columns: about 66 MB."""
    print("[gpu] fasta: 66 MB genome", flush=True)
    rng = np.random.default_rng(12)
    parts = []
    for name, L in [(b"chr1", 40_000_000), (b"chr2", 25_000_000), (b"chrM", 500_000)]:
        seq = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, L)]
        rows = L // 60
        body = np.full((rows, 61), 10, np.uint8)
        body[:, :60] = seq[:rows * 60].reshape(rows, 60)
        tail = seq[rows * 60:].tobytes()
        parts.append(b">" + name + b" synthetic\n" + body.tobytes() + (tail + b"\n" if tail else b""))
    data = b"".join(parts)
    got = check_fasta(compactor, hip_ctx, data, "genome")
    assert got.total_bases() == 65_500_000
