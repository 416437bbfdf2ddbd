"""GPU parity of BamDecoder::decode and the DeviceBgzfReader carry loop against
the oracle. Mirrors crates/fritillaria-cuda/tests/{bam_device,bam_reader}.rs."""
import numpy as np
import pytest

from conftest import golden_bytes
from helpers import bam_header, bgzf, build_record, random_records

pytestmark = pytest.mark.gpu

FIXTURES = ["htslib.bam", "htslib_multiblock.bam", "ont_ultralong.bam", "pacbio_hifi.bam"]
REF_COLS = ["reference_sequence_id", "position", "flags", "mapping_quality", "sequence_len",
            "mate_reference_sequence_id", "mate_position", "template_length", "record_offsets",
            "cigar_start", "sequence_start", "quality_start", "aux_start", "record_end"]


def oracle_decode(oracle_mod, u, offsets, start):
    rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, offsets, start)
    return rc, rec, tail, oracle_mod.bam_decode_columns(u, rec)


def check_batch(oracle_mod, records, u, offsets, coffsets, start):
    rc, rec, tail, cols = oracle_decode(oracle_mod, u, offsets, start)
    assert rc == 0
    assert len(records) == len(rec) and records.tail() == tail
    got = records.to_host()
    for name in REF_COLS:
        assert np.array_equal(got[name], cols[name]), name
    if coffsets is not None:
        want_v = oracle_mod.virtual_offsets(offsets, coffsets, rec)
        assert np.array_equal(got["virtual_offsets"], want_v)
    return got


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_columns_match_oracle(codec, decoder, oracle_mod, name):
    print(f"[gpu] bam fixture {name}", flush=True)
    import fritillaria_amd as fa
    buf = golden_bytes(name)
    spans, _ = fa.discover(buf)
    batch = codec.inflate_batch_device(buf, spans)
    u = np.frombuffer(oracle_mod.inflate_batch(buf, spans).data, np.uint8)
    start, _ = fa.bam_parse_header(u)
    records = decoder.decode(batch, start)
    check_batch(oracle_mod, records, u, batch.offsets(), batch.compressed_offsets(), start)


@pytest.mark.parametrize("config,n", [(1, 10000), (2, 30000), (3, 60)])
def test_synthetic_noodles_framed(codec, decoder, oracle_mod, config, n):
    print(f"[gpu] bam synthetic config {config} n={n}", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(config, n_records=n)
    spans, _ = fa.discover(s.compressed)
    batch = codec.inflate_batch_device(s.compressed, spans)
    u = batch.data().to_vec()
    records = decoder.decode(batch, s.records_start)
    got = check_batch(oracle_mod, records, u, batch.offsets(), batch.compressed_offsets(), s.records_start)
    assert len(got["position"]) == n
    if config != 3:
        # noodles framing puts member starts mid-record; the window speculation
        # must still resolve nearly every member without a fallback walk
        assert records.fallbacks() <= 2 + len(spans) // 50


@pytest.mark.parametrize("window", [64, 300, 2048, 65536])
def test_speculation_window_does_not_change_results(hip_ctx, codec, oracle_mod, window):
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(2, n_records=8000)
    spans, _ = fa.discover(s.compressed)
    batch = codec.inflate_batch_device(s.compressed, spans)
    u = batch.data().to_vec()
    hip_ctx.set_bam_window(window)
    try:
        records = fa.BamDecoder(hip_ctx).decode(batch, s.records_start)
        check_batch(oracle_mod, records, u, batch.offsets(), None, s.records_start)
    finally:
        hip_ctx.set_bam_window(2048)


def test_partial_trailing_record_reports_the_tail(codec, decoder, oracle_mod):
    import fritillaria_amd as fa
    rng = np.random.default_rng(21)
    data = bam_header() + b"".join(random_records(rng, 600))
    for cut in (len(data) - 1, len(data) - 37, len(data) // 2 + 3):
        buf = bgzf(data[:cut], cap=5000)
        spans, _ = fa.discover(buf)
        batch = codec.inflate_batch_device(buf, spans)
        u = batch.data().to_vec()
        start, _ = fa.bam_parse_header(u)
        records = decoder.decode(batch, start)
        check_batch(oracle_mod, records, u, batch.offsets(), batch.compressed_offsets(), start)
        assert records.tail() < len(u)


def test_record_spanning_members_and_every_small_blocking(codec, decoder, oracle_mod):
    import fritillaria_amd as fa
    rng = np.random.default_rng(22)
    recs = random_records(rng, 40, max_len=2000)
    data = b"".join(recs)
    for cap in (37, 100, 333, 1000, 4096):
        buf = bgzf(data, cap=cap, eof=False)
        spans, _ = fa.discover(buf)
        batch = codec.inflate_batch_device(buf, spans)
        u = np.frombuffer(data, np.uint8)
        records = decoder.decode(batch, 0)
        check_batch(oracle_mod, records, u, batch.offsets(), batch.compressed_offsets(), 0)


def test_malformed_record_position_matches_oracle(codec, decoder, oracle_mod):
    import fritillaria_amd as fa
    rng = np.random.default_rng(23)
    recs = random_records(rng, 300)
    bad = bytearray(recs[170])
    bad[20:24] = (1 << 22).to_bytes(4, "little")
    data = b"".join(recs[:170]) + bytes(bad) + b"".join(recs[171:])
    buf = bgzf(data, cap=7000)
    spans, _ = fa.discover(buf)
    batch = codec.inflate_batch_device(buf, spans)
    rc, rec, tail, _ = oracle_mod.bam_scan_blocked(np.frombuffer(data, np.uint8), batch.offsets(), 0)
    assert rc == 1
    with pytest.raises(fa.Malformed) as ei:
        decoder.decode(batch, 0)
    assert ei.value.position == tail and ei.value.format == "bam"
    tiny = bytearray(recs[5])
    tiny[0:4] = (20).to_bytes(4, "little")  # block_size below the 32-byte core
    data = b"".join(recs[:5]) + bytes(tiny) + b"".join(recs[6:])
    buf = bgzf(data, cap=7000)
    spans, _ = fa.discover(buf)
    batch = codec.inflate_batch_device(buf, spans)
    rc, rec, tail, _ = oracle_mod.bam_scan_blocked(np.frombuffer(data, np.uint8), batch.offsets(), 0)
    with pytest.raises(fa.Malformed) as ei:
        decoder.decode(batch, 0)
    assert rc == 1 and ei.value.position == tail


def test_corrupted_block_sizes_walk_like_the_oracle(codec, decoder, oracle_mod):
    """A record's block_size overwritten with random values (below the core,
    inside the stream so the chain walks on through other records' bytes, past
    the end): the device finds the same records, or fails at the same
    position, as the oracle's serial walk."""
    import fritillaria_amd as fa
    rng = np.random.default_rng(24)
    recs = random_records(rng, 300)
    failed = 0
    for trial in range(48):
        k = int(rng.integers(0, len(recs)))
        bad = bytearray(recs[k])
        v = int(rng.choice([int(rng.integers(0, 32)), int(rng.integers(32, 2000)), int(rng.integers(2000, 60000)),
                            int(rng.integers(60000, 1 << 31))]))
        bad[0:4] = v.to_bytes(4, "little")
        data = b"".join(recs[:k]) + bytes(bad) + b"".join(recs[k + 1:])
        buf = bgzf(data, cap=int(rng.choice([900, 7000, 65495])))
        spans, _ = fa.discover(buf)
        batch = codec.inflate_batch_device(buf, spans)
        u = np.frombuffer(data, np.uint8)
        rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, batch.offsets(), 0)
        print(f"[gpu] block_size trial {trial}: record {k} -> {v}, oracle rc {rc}", flush=True)
        if rc != 0:
            failed += 1
            with pytest.raises(fa.Malformed) as ei:
                decoder.decode(batch, 0)
            assert ei.value.position == tail and ei.value.format == "bam", trial
        else:
            records = decoder.decode(batch, 0)
            check_batch(oracle_mod, records, u, batch.offsets(), batch.compressed_offsets(), 0)
    assert 0 < failed < 48


def test_empty_batch_decodes_to_no_records(codec, decoder):
    import fritillaria_amd as fa
    b = codec.inflate_batch_device(b"", np.zeros(0, dtype=fa.SPAN_DTYPE))
    r = decoder.decode(b, 0)
    assert len(r) == 0 and r.tail() == 0
    r = decoder.decode(fa.DeviceInflateBatch(), 17)
    assert len(r) == 0 and r.tail() == 17


def _reader_columns(codec, decoder, data, bpb, read_ahead=0, read_ahead_bytes=0):
    import fritillaria_amd as fa
    reader = fa.DeviceBgzfReader(data, codec, blocks_per_batch=bpb, read_ahead_members=read_ahead,
                                 read_ahead_bytes=read_ahead_bytes)
    first = True
    vo, pos, lseq, n_batches = [], [], [], 0
    while True:
        batch = reader.next_batch()
        if batch is None:
            break
        n_batches += 1
        start = batch.start
        if first:
            start, _ = fa.bam_parse_header(batch.data.data().to_vec()[:1 << 20])
            first = False
        records = decoder.decode(batch.data, start)
        cols = records.to_host()
        vo.append(cols["virtual_offsets"])
        pos.append(cols["position"])
        lseq.append(cols["sequence_len"])
        tail = records.tail()
        del records
        reader.carry_from(tail)
    return np.concatenate(vo), np.concatenate(pos), np.concatenate(lseq), n_batches


@pytest.mark.parametrize("bpb", [1, 2, 3, 4, 5, 8, 256])
def test_reader_loop_matches_whole_file(codec, decoder, oracle_mod, bpb):
    print(f"[gpu] reader loop blocks_per_batch={bpb}", flush=True)
    from fritillaria_amd import synth
    for config, n in ((1, 4000), (3, 12)):
        s = synth.config_bam(config, n_records=n)
        d = oracle_mod.discover(s.compressed)
        r = oracle_mod.inflate_batch(s.compressed, d.spans)
        u = np.frombuffer(r.data, np.uint8)
        rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, r.offsets, s.records_start)
        cols = oracle_mod.bam_decode_columns(u, rec)
        want_v = oracle_mod.virtual_offsets(r.offsets, d.spans["compressed_offset"], rec)
        vo, pos, lseq, nb = _reader_columns(codec, decoder, s.compressed, bpb)
        assert np.array_equal(vo, want_v)
        assert np.array_equal(pos, cols["position"]) and np.array_equal(lseq, cols["sequence_len"])


@pytest.mark.parametrize("bpb", [1, 3, 8])
@pytest.mark.parametrize("read_ahead", [6, 24, 64])
def test_reader_loop_with_small_read_ahead_windows(codec, decoder, oracle_mod, bpb, read_ahead):
    """Windows of a few members, so that many are inflated ahead on the
    reader's stream, taken, and overlapped by the batches that straddle them
    (long records whose carry exceeds the overlap take a refill instead): the
    same records as the whole-file oracle."""
    print(f"[gpu] reader loop blocks_per_batch={bpb} read-ahead {read_ahead} members", flush=True)
    from fritillaria_amd import synth
    for config, n in ((2, 20000), (3, 12)):
        s = synth.config_bam(config, n_records=n)
        d = oracle_mod.discover(s.compressed)
        r = oracle_mod.inflate_batch(s.compressed, d.spans)
        u = np.frombuffer(r.data, np.uint8)
        rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, r.offsets, s.records_start)
        cols = oracle_mod.bam_decode_columns(u, rec)
        want_v = oracle_mod.virtual_offsets(r.offsets, d.spans["compressed_offset"], rec)
        vo, pos, lseq, nb = _reader_columns(codec, decoder, s.compressed, bpb, read_ahead)
        assert np.array_equal(vo, want_v), (config, bpb, read_ahead)
        assert np.array_equal(pos, cols["position"]) and np.array_equal(lseq, cols["sequence_len"])


@pytest.mark.parametrize("bpb", [1, 4])
def test_reader_from_pinned_memory_with_read_ahead(codec, decoder, oracle_mod, bpb):
    """The stream in pinned host memory, so the windows read ahead are
    uploaded asynchronously by the reader's helper thread while batches are
    decoded: the same records as the whole-file oracle."""
    print(f"[gpu] reader loop from pinned memory, blocks_per_batch={bpb}", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(2, n_records=20000)
    d = oracle_mod.discover(s.compressed)
    r = oracle_mod.inflate_batch(s.compressed, d.spans)
    u = np.frombuffer(r.data, np.uint8)
    rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, r.offsets, s.records_start)
    cols = oracle_mod.bam_decode_columns(u, rec)
    want_v = oracle_mod.virtual_offsets(r.offsets, d.spans["compressed_offset"], rec)
    pin = fa.PinnedBuffer(s.compressed.size)
    pin.array[:] = s.compressed
    try:
        vo, pos, lseq, nb = _reader_columns(codec, decoder, pin.array, bpb, read_ahead=16)
    finally:
        pin.close()
    assert np.array_equal(vo, want_v) and np.array_equal(pos, cols["position"])
    assert np.array_equal(lseq, cols["sequence_len"])


def test_reader_random_batch_and_window_bounds(codec, decoder, oracle_mod):
    """Random batch budgets and read-ahead windows (member and byte bounds)
    over the three BAM shapes: every combination reads the same records as
    the whole-file oracle."""
    from fritillaria_amd import synth
    rng = np.random.default_rng(88)
    files = []
    for config, n in ((1, 3000), (2, 8000), (3, 10)):
        s = synth.config_bam(config, n_records=n)
        d = oracle_mod.discover(s.compressed)
        r = oracle_mod.inflate_batch(s.compressed, d.spans)
        u = np.frombuffer(r.data, np.uint8)
        rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, r.offsets, s.records_start)
        cols = oracle_mod.bam_decode_columns(u, rec)
        files.append((config, s.compressed, oracle_mod.virtual_offsets(r.offsets, d.spans["compressed_offset"], rec),
                      cols["position"]))
    for trial in range(18):
        config, data, want_v, want_pos = files[trial % 3]
        bpb = int(rng.integers(1, 20))
        ra = int(rng.integers(2, 80))
        rab = int(rng.choice([0, int(rng.integers(64 << 10, 4 << 20))]))
        print(f"[gpu] reader trial {trial}: config {config}, bpb {bpb}, window {ra} members / {rab} bytes",
              flush=True)
        vo, pos, lseq, nb = _reader_columns(codec, decoder, data, bpb, ra, rab)
        assert np.array_equal(vo, want_v) and np.array_equal(pos, want_pos), (trial, config, bpb, ra, rab)


def test_reader_stream_ending_inside_a_record(codec, decoder):
    import fritillaria_amd as fa
    rng = np.random.default_rng(30)
    data = bam_header() + b"".join(random_records(rng, 300))
    buf = bgzf(data[:-10], cap=9000)
    reader = fa.DeviceBgzfReader(buf, codec, blocks_per_batch=2)
    first = True
    with pytest.raises(fa.Malformed) as ei:
        while True:
            b = reader.next_batch()
            if b is None:
                break
            start = b.start
            if first:
                start, _ = fa.bam_parse_header(b.data.data().to_vec())
                first = False
            rec = decoder.decode(b.data, start)
            reader.carry_from(rec.tail())
    assert ei.value.format == "bgzf"


def test_reader_truncated_bgzf(codec):
    import fritillaria_amd as fa
    buf = bgzf(b"z" * 100000, cap=20000)
    reader = fa.DeviceBgzfReader(buf[:-5], codec, blocks_per_batch=256)
    with pytest.raises(fa.Malformed) as ei:
        reader.next_batch()
    assert "truncated BGZF" in ei.value.reason


def test_carry_past_end_is_rejected(codec):
    import fritillaria_amd as fa
    buf = bgzf(b"z" * 1000)
    reader = fa.DeviceBgzfReader(buf, codec)
    b = reader.next_batch()
    with pytest.raises(fa.InvalidDeviceBatch):
        reader.carry_from(b.data.byte_len() + 1)


def test_columns_report_their_device(codec, decoder):
    import fritillaria_amd as fa
    buf = golden_bytes("htslib.bam")
    batch = codec.inflate_batch_device(buf, fa.discover(buf)[0])
    u = batch.data().to_vec()
    start, _ = fa.bam_parse_header(u)
    r = decoder.decode(batch, start)
    assert r.device_ordinal() == batch.device_ordinal() == 0


def test_speculation_holds_past_65536_members(codec, decoder, oracle_mod):
    """One batch of ~100k small members: the speculation must resolve members
    past the 65,536th like the first ones (no fallback walks)."""
    print("[gpu] bam many small members", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    u, rs = synth.bam_stream(300_000, synth.CONFIG_SEED + 2, read_range=(100, 250), indel_p=0.01,
                             ref_len=250_000_000)
    c = synth.bgzf(u, level=1, payload_cap=1024)
    spans, _ = fa.discover(c)
    assert len(spans) > 70_000
    batch = codec.inflate_batch_device(c, spans)
    records = decoder.decode(batch, rs)
    check_batch(oracle_mod, records, batch.data().to_vec(), batch.offsets(), batch.compressed_offsets(), rs)
    assert records.fallbacks() <= 2 + len(spans) // 50, records.fallbacks()


def test_speculation_holds_past_2_gib(codec, decoder, oracle_mod):
    """One batch of ~2.4 GB inflated: record positions with bit 31 set must
    speculate like the rest (a sign-extended 64-bit broadcast once sent every
    member in [2 GiB, 4 GiB) to a fallback walk)."""
    print("[gpu] bam batch past 2 GiB", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(2, n_records=8_000_000)
    assert s.inflated_len > (1 << 31) + (1 << 28)
    spans, _ = fa.discover(s.compressed)
    batch = codec.inflate_batch_device(s.compressed, spans)
    records = decoder.decode(batch, s.records_start)
    u = batch.data().to_vec()
    rc, want, tail = oracle_mod.bam_scan_records(u, s.records_start)
    assert rc == 0 and len(records) == len(want) == 8_000_000 and records.tail() == tail
    assert np.array_equal(records.column_to_host("record_offsets"), want)
    assert records.fallbacks() <= 2 + len(spans) // 50, records.fallbacks()


@pytest.mark.parametrize("bpb", [1, 3])
@pytest.mark.parametrize("read_ahead", [0, 8])
def test_reader_read_ahead_reports_a_bad_member_where_it_lies(codec, decoder, bpb, read_ahead):
    """The reader inflates ahead of the batches it hands out; a member with a
    bad trailer CRC inside that window still fails exactly the batch that holds
    it (ChecksumMismatch at its offset), after every earlier batch was served
    with the records it holds, as when each batch is inflated on its own."""
    print(f"[gpu] reader read-ahead with a bad member, blocks_per_batch={bpb}, windows {read_ahead}", flush=True)
    import fritillaria_amd as fa
    from helpers import bgzf_member
    rng = np.random.default_rng(77)
    data = bam_header() + b"".join(random_records(rng, 9000))
    cap = 60000
    members = [bgzf_member(data[i:i + cap]) for i in range(0, len(data), cap)]
    bad = 17
    payload = data[bad * cap:(bad + 1) * cap]
    import zlib
    members[bad] = bgzf_member(payload, crc=zlib.crc32(payload) ^ 1)
    blob = b"".join(members) + bgzf(b"", eof=True)
    bad_off = sum(len(m) for m in members[:bad])
    reader = fa.DeviceBgzfReader(blob, codec, blocks_per_batch=bpb, read_ahead_members=read_ahead)
    first, served = True, 0
    with pytest.raises(fa.ChecksumMismatch) as ei:
        while True:
            b = reader.next_batch()
            if b is None:
                break
            start = b.start
            if first:
                start, _ = fa.bam_parse_header(b.data.data().to_vec())
                first = False
            rec = decoder.decode(b.data, start)
            served += len(rec)
            reader.carry_from(rec.tail())
    assert ei.value.offset == bad_off
    # the batches before the failing one hold members before the bad one only;
    # with one 64 KiB budget per batch (a member or two) they held every record
    # that ends before the member carried into the failing batch
    ends = len(bam_header()) + np.cumsum([4 + int.from_bytes(r[:4], "little") for r in _records_of(data)])
    hi = int(np.searchsorted(ends, bad * cap, side="right"))
    assert served <= hi, (served, hi)
    if bpb == 1:
        assert served >= int(np.searchsorted(ends, (bad - 2) * cap, side="right")) > 0


def _records_of(data):
    hdr = len(bam_header())
    out, p = [], hdr
    while p + 4 <= len(data):
        n = 4 + int.from_bytes(data[p:p + 4], "little")
        out.append(data[p:p + n])
        p += n
    return out
