"""GPU parity of FastqScanner::{scan, decode} against the oracle's serial walk.
Mirrors crates/fritillaria-cuda/tests/fastq_decode.rs and adds hand-built cases
(decoys, blank lines, bytes that are no record start) and a synthetic file."""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_bytes
from helpers import bgzf

pytestmark = pytest.mark.gpu

FIXTURES = ["ont_ultralong.fastq.gz", "pacbio_hifi.fastq.gz", "illumina.fastq.gz"]
BOUNDS = ["sequence_start", "plus_start", "quality_start", "sequence_len", "record_end"]


@pytest.fixture(scope="module")
def scanner(hip_ctx):
    import fritillaria_amd as fa
    return fa.FastqScanner(hip_ctx)


def device_batch(codec, raw: bytes):
    import fritillaria_amd as fa
    spans, _ = fa.discover(raw)
    return codec.inflate_batch_device(raw, spans)


def check(oracle_mod, recs, u: bytes, start: int):
    rc, offs, bounds, tail, _ = oracle_mod.fastq_scan_records(u, start)
    assert rc == 0
    assert len(recs) == len(offs) and recs.tail() == tail
    got = recs.to_host()
    assert np.array_equal(got["record_offsets"], offs)
    for i, name in enumerate(BOUNDS):
        assert np.array_equal(got[name], bounds[:, i]), name
    assert recs.survivors() == oracle_mod.fastq_survivors(u, start)
    return got


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_columns_match_oracle(codec, scanner, oracle_mod, name):
    print(f"[gpu] fastq fixture {name}", flush=True)
    raw = golden_bytes(name)
    batch = device_batch(codec, raw)
    u = gzip.decompress(raw)
    recs = scanner.decode(batch, 0)
    check(oracle_mod, recs, u, 0)
    assert recs.tail() == len(u)
    scan = scanner.scan(batch, 0)
    assert scan.proof == "tiled" and scan.survivors == len(recs)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("blocks_per_batch", [1, 4])
def test_columns_match_across_batch_seams(codec, scanner, oracle_mod, name, blocks_per_batch):
    print(f"[gpu] fastq seams {name} {blocks_per_batch}", flush=True)
    import fritillaria_amd as fa
    raw = golden_bytes(name)
    u = gzip.decompress(raw)
    _, _, whole, _, _ = oracle_mod.fastq_scan_records(u, 0)
    reader = fa.DeviceBgzfReader(raw, codec, blocks_per_batch=blocks_per_batch)
    lengths, batches = [], 0
    while True:
        b = reader.next_batch()
        if b is None:
            break
        recs = scanner.decode(b.data, b.start)
        host = b.data.data().to_vec().tobytes() if b.data.data() is not None else b""
        check(oracle_mod, recs, host, b.start)
        lengths.extend(recs.column_to_host("sequence_len").tolist())
        batches += 1
        tail = recs.tail()
        del recs, b
        reader.carry_from(tail)
    assert lengths == whole[:, 3].tolist()


def test_decoding_past_the_end_is_empty(codec, scanner):
    raw = golden_bytes("pacbio_hifi.fastq.gz")
    batch = device_batch(codec, raw)
    n = batch.byte_len()
    recs = scanner.decode(batch, n)
    assert len(recs) == 0 and recs.tail() == n
    recs = scanner.decode(batch, n + 7)
    assert len(recs) == 0 and recs.tail() == n


def rec(name: bytes, seq: bytes, plus: bytes = b"", qual: bytes = None) -> bytes:
    return b"@" + name + b"\n" + seq + b"\n+" + plus + b"\n" + (qual if qual is not None else b"I" * len(seq)) + b"\n"


CASES = {
    "decoy_in_quality": rec(b"r0", b"ACGT", qual=b"@III") + rec(b"r1", b"TTTT"),
    "length_matched_decoy": b"@A\nACG\n+\n@BC\n" + b"@XY\nACGT\n+ab\nIIII\n",
    "blank_lines": rec(b"a", b"AC") + b"\n\n" + rec(b"b", b"GT") + b"\n\n\n",
    "leading_newline": b"\n" + rec(b"r1", b"TTTT"),
    "no_final_newline": rec(b"a", b"ACGT") + rec(b"b", b"GGCC")[:-1],
    "partial_tail": rec(b"whole", b"ACGT") + rec(b"partial", b"ACGTACGT")[:-4],
    "short_quality_at_end": rec(b"r0", b"ACGT")[:-2],
    "empty_read": rec(b"r0", b"") + rec(b"r1", b"A"),
    "repeated_name_after_plus": rec(b"r0", b"ACGTACGT", b"r0") * 3,
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_hand_built_cases(codec, scanner, oracle_mod, case):
    print(f"[gpu] fastq case {case}", flush=True)
    u = CASES[case]
    batch = device_batch(codec, bgzf(u))
    for start in sorted({0, 1, len(u) // 2}):
        rc, _, _, _, err = oracle_mod.fastq_scan_records(u, start)
        if rc == 0:
            check(oracle_mod, scanner.decode(batch, start), u, start)
        else:
            import fritillaria_amd as fa
            with pytest.raises(fa.Malformed) as ei:
                scanner.decode(batch, start)
            assert ei.value.position == err


def test_garbage_after_records_is_malformed(codec, scanner, oracle_mod):
    import fritillaria_amd as fa
    u = rec(b"r0", b"ACGT") + rec(b"r1", b"GG") + b"garbage\n"
    batch = device_batch(codec, bgzf(u))
    rc, _, _, _, err = oracle_mod.fastq_scan_records(u, 0)
    assert rc == oracle_mod.E_MALFORMED
    with pytest.raises(fa.Malformed) as ei:
        scanner.decode(batch, 0)
    assert ei.value.position == err


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_mixtures(codec, scanner, oracle_mod, seed):
    """Records with '@'/'+' qualities, blank lines and cut records, many segments long."""
    print(f"[gpu] fastq random seed {seed}", flush=True)
    import fritillaria_amd as fa
    rng = np.random.default_rng(seed)
    parts = []
    for j in range(30000):
        r = rng.random()
        n = int(rng.integers(0, 40))
        seq = bytes(np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, n)])
        qual = bytes(np.frombuffer(b"@I+#", np.uint8)[rng.integers(0, 4, n)])
        x = rec(b"r%d" % j, seq, b"" if r < 0.5 else b"r", qual)
        if r < 0.002:
            x = b"\n"
        parts.append(x)
    u = b"".join(parts)
    batch = device_batch(codec, bgzf(u))
    for start in (0, int(rng.integers(0, len(u)))):
        rc, _, _, _, err = oracle_mod.fastq_scan_records(u, start)
        if rc == 0:
            check(oracle_mod, scanner.decode(batch, start), u, start)
        else:
            with pytest.raises(fa.Malformed) as ei:
                scanner.decode(batch, start)
            assert ei.value.position == err


def test_synthetic_config_fastq(codec, scanner, oracle_mod):
    print("[gpu] fastq synthetic 200k", flush=True)
    from fritillaria_amd import synth
    s = synth.config_fastq(200_000)
    batch = device_batch(codec, s.compressed.tobytes())
    u = s.inflated.tobytes()
    got = check(oracle_mod, scanner.decode(batch, 0), u, 0)
    assert len(got["record_offsets"]) == 200_000
    assert np.all(got["sequence_len"] == 150)
