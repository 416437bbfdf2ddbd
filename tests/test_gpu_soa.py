"""GPU parity of the BAM SoA extras (frhip_bam_soa_build: resolved CIGAR ops,
packed sequences, aux tag index) against the oracle's or_bam_soa."""
import numpy as np
import pytest

from conftest import golden_bytes
from helpers import bam_header, bgzf, build_record

pytestmark = pytest.mark.gpu

FIXTURES = ["htslib.bam", "htslib_multiblock.bam", "ont_ultralong.bam", "pacbio_hifi.bam"]


def check(oracle_mod, decoder, batch, start):
    recs = decoder.decode(batch, start)
    u = batch.data().to_vec()
    offs = recs.column_to_host("record_offsets")
    rc, want, _ = oracle_mod.bam_soa(u, offs)
    assert rc == 0
    got = decoder.soa(batch, recs).to_host()
    assert len(got["cigar_offsets"]) == len(offs) + 1
    for k in want:
        if k == "aux_entries":
            for f in ("tag", "type", "subtype", "value_offset", "count", "size"):
                assert np.array_equal(got[k][f], want[k][f]), (k, f)
        else:
            assert np.array_equal(got[k], want[k]), k
    return got


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_soa_matches_oracle(codec, decoder, oracle_mod, name):
    print(f"[gpu] soa fixture {name}", flush=True)
    import fritillaria_amd as fa
    buf = golden_bytes(name)
    spans, _ = fa.discover(buf)
    batch = codec.inflate_batch_device(buf, spans)
    start, _ = fa.bam_parse_header(batch.data().to_vec()[: 1 << 20])
    check(oracle_mod, decoder, batch, start)


@pytest.mark.parametrize("config,n", [(1, 5000), (2, 20000), (3, 40)])
def test_synthetic_soa_matches_oracle(codec, decoder, oracle_mod, config, n):
    print(f"[gpu] soa synthetic config {config}", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(config, n_records=n)
    spans, _ = fa.discover(s.compressed)
    batch = codec.inflate_batch_device(s.compressed, spans)
    got = check(oracle_mod, decoder, batch, s.records_start)
    assert len(got["cigar_offsets"]) == n + 1


def _batch(codec, records):
    import fritillaria_amd as fa
    hdr = bam_header()
    raw = bgzf(hdr + b"".join(records))
    spans, _ = fa.discover(raw)
    return codec.inflate_batch_device(raw, spans), len(hdr)


def test_long_cigar_placeholder_and_errors(codec, decoder, oracle_mod):
    import struct
    import fritillaria_amd as fa
    real = [(10 << 4) | 0, (2 << 4) | 1, (8 << 4) | 0]
    ph = [(20 << 4) | 4, (18 << 4) | 3]
    cg = b"CGBI" + struct.pack("<I", 3) + struct.pack("<3I", *real)
    good = [build_record(b"r%d" % i, b"ACGT" * 5, cigar=ph if i % 3 == 0 else None, aux=b"NMC\x01" + cg)
            for i in range(50)]
    batch, start = _batch(codec, good)
    got = check(oracle_mod, decoder, batch, start)
    assert got["cigar_ops"][:3].tolist() == real
    bad = good[:7] + [build_record(b"bad", b"ACGT" * 5, cigar=ph, aux=b"NMC\x01")] + good[:3]
    batch, start = _batch(codec, bad)
    recs = decoder.decode(batch, start)
    with pytest.raises(fa.Malformed) as ei:
        decoder.soa(batch, recs)
    assert ei.value.position == int(recs.column_to_host("record_offsets")[7])


def test_sequences_of_every_short_length(codec, decoder, oracle_mod):
    """Packed sequences copied in 16-byte chunks of the column: records with no
    sequence (a whole 256-record tile of them), sequences shorter than a chunk,
    chunks across one and several record boundaries, and long sequences, in
    random order, against the oracle."""
    print("[gpu] soa short and empty sequences", flush=True)
    from helpers import random_records
    rng = np.random.default_rng(20261204)
    recs = random_records(rng, 300, min_len=0, max_len=0)
    recs += random_records(rng, 2000, min_len=0, max_len=40, name_prefix=b"s")
    recs += random_records(rng, 200, min_len=200, max_len=3000, name_prefix=b"l")
    order = rng.permutation(len(recs) - 300) + 300
    batch, start = _batch(codec, recs[:300] + [recs[i] for i in order])
    got = check(oracle_mod, decoder, batch, start)
    assert len(got["cigar_offsets"]) == len(recs) + 1


def test_empty_batch(codec, decoder):
    import fritillaria_amd as fa
    recs = decoder.decode(fa.DeviceInflateBatch(), 0)
    soa = decoder.soa(fa.DeviceInflateBatch(), recs)
    assert len(soa) == 0
    assert soa.column_to_host("cigar_offsets").tolist() == [0]
