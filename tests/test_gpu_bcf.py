"""GPU parity of BcfScanner::decode against the oracle's scan_records and
RecordBatch::decode. Mirrors crates/fritillaria-cuda/tests (bcf) and adds
hand-built error cases and a synthetic file."""
import gzip
import struct

import numpy as np
import pytest

from conftest import golden_bytes
from helpers import bgzf

pytestmark = pytest.mark.gpu

FIXTURES = ["kg_phase3.bcf", "giab_hg002.bcf"]


@pytest.fixture(scope="module")
def scanner(hip_ctx):
    import fritillaria_amd as fa
    return fa.BcfScanner(hip_ctx)


def device_batch(codec, raw: bytes):
    import fritillaria_amd as fa
    spans, _ = fa.discover(raw)
    return codec.inflate_batch_device(raw, spans)


def check(oracle_mod, recs, u: bytes, start: int):
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(u, start)
    assert rc == 0
    assert len(recs) == len(offs) and recs.tail() == tail
    rc, cols, _ = oracle_mod.bcf_decode_columns(u, offs, tail)
    assert rc == 0
    got = recs.to_host()
    for name in cols:
        assert np.array_equal(got[name], cols[name]), name
    return got


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_columns_match_oracle(codec, scanner, oracle_mod, name):
    print(f"[gpu] bcf fixture {name}", flush=True)
    import fritillaria_amd as fa
    raw = golden_bytes(name)
    batch = device_batch(codec, raw)
    u = gzip.decompress(raw)
    rs, samples, contigs = fa.bcf_parse_header(u[: 1 << 20])
    recs = scanner.decode(batch, rs, samples, contigs)
    got = check(oracle_mod, recs, u, rs)
    assert recs.tail() == len(u)
    assert len(got["record_offsets"]) == {"kg_phase3.bcf": 715, "giab_hg002.bcf": 275}[name]
    # members never start on a record here; whole-member windows still find every entry
    assert recs.fallbacks() <= 2


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("blocks_per_batch", [1, 4])
def test_columns_match_across_batch_seams(codec, scanner, oracle_mod, name, blocks_per_batch):
    print(f"[gpu] bcf seams {name} {blocks_per_batch}", flush=True)
    import fritillaria_amd as fa
    raw = golden_bytes(name)
    u = gzip.decompress(raw)
    rs, samples, contigs = fa.bcf_parse_header(u[: 1 << 20])
    _, whole, _, _ = oracle_mod.bcf_scan_records(u, rs)
    reader = fa.DeviceBgzfReader(raw, codec, blocks_per_batch=blocks_per_batch)
    first, n, base = True, 0, 0
    while True:
        b = reader.next_batch()
        if b is None:
            break
        start = rs if first else b.start
        first = False
        recs = scanner.decode(b.data, start, samples, contigs)
        host = b.data.data().to_vec().tobytes() if b.data.data() is not None else b""
        check(oracle_mod, recs, host, start)
        n += len(recs)
        tail = recs.tail()
        del recs, b
        reader.carry_from(tail)
    assert n == len(whole)


def test_empty_batch_is_empty(scanner):
    import fritillaria_amd as fa
    recs = scanner.decode(fa.DeviceInflateBatch(), 5, 1, 1)
    assert len(recs) == 0 and recs.tail() == 5


def site(chrom, pos, samples=1, id_=b"\x07", alleles=(b"\x17A", b"\x17C"), ls_override=None):
    shared = struct.pack("<iiiI", chrom, pos, 1, 0x7F800001) + struct.pack("<HH", 0, len(alleles))
    shared += struct.pack("<I", samples)[:3] + b"\x00" + id_ + b"".join(alleles) + b"\x00"
    ls = len(shared) if ls_override is None else ls_override
    return struct.pack("<II", ls, 0) + shared


def test_bad_l_shared_is_malformed_at_the_record(codec, scanner, oracle_mod):
    import fritillaria_amd as fa
    u = site(0, 1) * 40 + site(0, 2, ls_override=20) + site(0, 3) * 5
    rc, _, _, err = oracle_mod.bcf_scan_records(u, 0)
    assert rc == oracle_mod.E_MALFORMED
    batch = device_batch(codec, bgzf(u, cap=300))
    with pytest.raises(fa.Malformed) as ei:
        scanner.decode(batch, 0, 1, 1)
    assert ei.value.position == err


def test_bad_typed_value_is_malformed(codec, scanner, oracle_mod):
    import fritillaria_amd as fa
    u = site(0, 1) * 10 + site(0, 2, id_=bytes([0xF7, 0x11, 3]) + b"abc") + site(0, 3) * 3
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(u, 0)
    assert rc == 0
    rc, _, bad = oracle_mod.bcf_decode_columns(u, offs, tail)
    assert rc == oracle_mod.E_MALFORMED and bad == 10
    batch = device_batch(codec, bgzf(u))
    with pytest.raises(fa.Malformed) as ei:
        scanner.decode(batch, 0, 1, 1)
    assert ei.value.position == int(offs[10])


def test_partial_tail(codec, scanner, oracle_mod):
    u = site(0, 1) * 30
    cut = u[:-7]
    batch = device_batch(codec, bgzf(cut, cap=200))
    check(oracle_mod, scanner.decode(batch, 0, 1, 1), cut, 0)


def test_synthetic_config_bcf(codec, scanner, oracle_mod):
    print("[gpu] bcf synthetic 3000 sites", flush=True)
    from fritillaria_amd import synth
    s = synth.config_bcf(3000)
    batch = device_batch(codec, s.compressed.tobytes())
    got = check(oracle_mod, scanner.decode(batch, s.records_start, s.samples, s.contigs), s.inflated.tobytes(),
                s.records_start)
    assert len(got["record_offsets"]) == 3000
