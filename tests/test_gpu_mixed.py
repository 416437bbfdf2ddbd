"""A BCF stream and a bgzipped FASTQ stream decoded together (one inflate
launch over both streams' members, then both boundary scans and column
decodes), checked against the oracle and against each scanner's per-call
decode of its stream alone."""
import gzip

import numpy as np
import pytest

from conftest import golden_bytes

pytestmark = pytest.mark.gpu

FQ_BOUNDS = ["sequence_start", "plus_start", "quality_start", "sequence_len", "record_end"]


def check_bcf(oracle_mod, recs, u: bytes, start: int):
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(u, start)
    assert rc == 0
    assert len(recs) == len(offs) and recs.tail() == tail
    rc, cols, _ = oracle_mod.bcf_decode_columns(u, offs, tail)
    assert rc == 0
    got = recs.to_host()
    for name in cols:
        assert np.array_equal(got[name], cols[name]), name
    return got


def check_fastq(oracle_mod, recs, u: bytes):
    rc, offs, bounds, tail, _ = oracle_mod.fastq_scan_records(u, 0)
    assert rc == 0
    assert len(recs) == len(offs) and recs.tail() == tail
    got = recs.to_host()
    assert np.array_equal(got["record_offsets"], offs)
    for i, name in enumerate(FQ_BOUNDS):
        assert np.array_equal(got[name], bounds[:, i]), name
    return got


def run_pair(hip_ctx, oracle_mod, bcf_raw: bytes, fq_raw: bytes):
    import fritillaria_amd as fa
    bs, _ = fa.discover(bcf_raw)
    fs, _ = fa.discover(fq_raw)
    r = fa.decode_bcf_fastq(hip_ctx, bcf_raw, bs, fq_raw, fs)
    ub = gzip.decompress(bcf_raw)
    uf = gzip.decompress(fq_raw)
    arena = r.arena.data().to_vec().tobytes()
    assert r.fastq_base % 4096 == 0 and r.fastq_base >= len(ub)
    assert arena[: len(ub)] == ub and arena[r.fastq_base:] == uf
    offs = r.arena.offsets()
    assert offs[len(bs)] == len(ub) and offs[-1] == r.fastq_base + len(uf)
    start, samples, contigs = fa.bcf_parse_header(ub)
    assert (r.bcf_start, r.samples, r.contigs) == (start, samples, contigs)
    gb = check_bcf(oracle_mod, r.bcf, ub, start)
    gf = check_fastq(oracle_mod, r.fastq, uf)
    return r, gb, gf


@pytest.mark.parametrize("two_phase", [False, True])
def test_synthetic_bcf_and_fastq_together(hip_ctx, codec, oracle_mod, two_phase):
    """The mixed BCF + FASTQ shapes (2,504 diploid samples, GT only; 150 bp reads), scaled
    down, on the one-wave inflate and on the two-phase inflate; the columns are
    also those of each stream decoded alone by its own scanner."""
    print(f"[gpu] bcf+fastq together, two_phase={two_phase}", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    b = synth.config_bcf(2000)
    f = synth.config_fastq(60_000)
    run_and_compare(hip_ctx, codec, oracle_mod, b, f, two_phase)


def run_and_compare(hip_ctx, codec, oracle_mod, b, f, two_phase):
    import fritillaria_amd as fa
    default_min = hip_ctx.two_phase_min_members()
    if two_phase:
        hip_ctx.set_inflate_schedule(16)
        hip_ctx.set_two_phase_min_members(0)
    try:
        r, gb, gf = run_pair(hip_ctx, oracle_mod, b.compressed.tobytes(), f.compressed.tobytes())
    finally:
        hip_ctx.set_inflate_schedule(16)
        hip_ctx.set_two_phase_min_members(default_min)
    assert len(gb["record_offsets"]) == b.n_records and len(gf["record_offsets"]) == f.n_records
    bcf_alone = fa.BcfScanner(hip_ctx).decode(
        codec.inflate_batch_device(b.compressed, fa.discover(b.compressed)[0]), b.records_start, b.samples,
        b.contigs).to_host()
    fq_alone = fa.FastqScanner(hip_ctx).decode(
        codec.inflate_batch_device(f.compressed, fa.discover(f.compressed)[0]), 0).to_host()
    for name in bcf_alone:
        assert np.array_equal(gb[name], bcf_alone[name]), name
    for name in fq_alone:
        assert np.array_equal(gf[name], fq_alone[name]), name


def test_mixed_at_twenty_thousand_sites(hip_ctx, codec, oracle_mod):
    """The mixed configuration at a tenth of its full size: 20,000 sites x 2,504
    samples (about 100 MB inflated, 2,000 members) and 500,000 reads, on the
    two-phase inflate (enough members to run it), against the oracle scans."""
    print("[gpu] bcf+fastq together at 20,000 sites / 500,000 reads", flush=True)
    from fritillaria_amd import synth
    b = synth.config_bcf(20_000)
    f = synth.config_fastq(500_000)
    run_and_compare(hip_ctx, codec, oracle_mod, b, f, True)


@pytest.mark.parametrize("bcf_name,fq_name", [("kg_phase3.bcf", "illumina.fastq.gz"),
                                              ("giab_hg002.bcf", "ont_ultralong.fastq.gz")])
def test_fixtures_together(hip_ctx, oracle_mod, bcf_name, fq_name):
    print(f"[gpu] bcf+fastq fixtures {bcf_name} {fq_name}", flush=True)
    _, gb, gf = run_pair(hip_ctx, oracle_mod, golden_bytes(bcf_name), golden_bytes(fq_name))
    assert len(gb["record_offsets"]) > 0 and len(gf["record_offsets"]) > 0


def test_bad_fastq_trailer_is_reported_at_its_member(hip_ctx):
    """A FASTQ member with a wrong trailer CRC fails the whole call with the
    member's offset in the combined batch, as the per-call codec reports it."""
    print("[gpu] bcf+fastq bad trailer", flush=True)
    import fritillaria_amd as fa
    bcf_raw = golden_bytes("kg_phase3.bcf")
    fq = bytearray(golden_bytes("illumina.fastq.gz"))
    fs, _ = fa.discover(bytes(fq))
    k = len(fs) // 2
    end = int(fs["payload_start"][k] + fs["payload_len"][k])
    fq[end] ^= 0xFF  # first CRC byte of member k's trailer
    fs2, _ = fa.discover(bytes(fq))
    with pytest.raises(fa.ChecksumMismatch) as ei:
        fa.decode_bcf_fastq(hip_ctx, bcf_raw, fa.discover(bcf_raw)[0], bytes(fq), fs2)
    assert ei.value.offset == len(bcf_raw) + int(fs["compressed_offset"][k])
