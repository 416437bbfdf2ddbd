"""GPU checks of the drop-in boundary's remaining contract: the ready event a
consumer on its own stream orders against (CudaCodec::ready_event,
crates/fritillaria-cuda/tests/device_resident.rs:417-468), a context on a
caller-owned stream (with_context), copy_to_host's length check
(crates/fritillaria-cuda/src/backend.rs:333-358), select_codec's names
(crates/fritillaria/src/lib.rs:242-260), and the timed calls
(inflate_batch_device_timed, decode_timed) giving the untimed results."""
import numpy as np
import pytest

from helpers import bam_header, bgzf, random_records

pytestmark = pytest.mark.gpu


def payload_bgzf():
    payload = bytes(i % 251 for i in range(80_000))
    return payload, bgzf(payload, cap=4096)


def test_a_consumer_can_order_on_the_ready_event(hip_ctx, codec):
    print("[gpu] boundary: ready event", flush=True)
    import fritillaria_amd as fa
    payload, data = payload_bgzf()
    spans, _ = fa.discover(data)
    batch = codec.inflate_batch_device(data, spans)
    ev = fa.HipCodec.ready_event(batch)
    assert ev is not None
    # a consumer on its own stream: wait on the event, then read through that stream
    theirs = fa.HipContext(0)
    assert theirs.stream != hip_ctx.stream
    fa.event_wait(ev, theirs.stream)
    dst = theirs.alloc(batch.data().byte_len())
    theirs.copy_device(dst, batch.data())
    assert dst.to_vec().tobytes()[: len(payload)] == payload
    fa.event_synchronize(ev)


def test_an_empty_batch_has_no_ready_event():
    import fritillaria_amd as fa
    assert fa.HipCodec.ready_event(fa.DeviceInflateBatch()) is None


def test_decoded_columns_carry_a_ready_event(hip_ctx, codec, decoder):
    import fritillaria_amd as fa
    rng = np.random.default_rng(11)
    data = bgzf(bam_header() + b"".join(random_records(rng, 300)), cap=7000)
    spans, _ = fa.discover(data)
    batch = codec.inflate_batch_device(data, spans)
    recs = decoder.decode(batch, len(bam_header()))
    ev = recs.column("position").ready_event()
    assert ev is not None and ev == recs.column("record_end").ready_event()
    fa.event_synchronize(ev)
    soa = decoder.soa(batch, recs)
    assert soa.column("sequence_packed").ready_event() is not None


def test_a_context_on_a_caller_owned_stream(hip_ctx, oracle_mod):
    print("[gpu] boundary: caller-owned stream", flush=True)
    import fritillaria_amd as fa
    mine = fa.HipContext(0, stream=hip_ctx.stream)
    assert mine.stream == hip_ctx.stream and mine.ordinal == 0
    payload, data = payload_bgzf()
    spans, _ = fa.discover(data)
    got = fa.HipCodec(mine).inflate_batch(data, spans)
    assert bytes(got.data) == payload
    want = oracle_mod.inflate_batch(data, oracle_mod.discover(data).spans)
    assert got.offsets.tolist() == want.offsets.tolist()
    mine.close()
    # the caller's stream outlives the context that borrowed it
    assert bytes(fa.HipCodec(hip_ctx).inflate_batch(data, spans).data) == payload


def test_copy_to_host_rejects_a_wrong_length(codec):
    import fritillaria_amd as fa
    _, data = payload_bgzf()
    spans, _ = fa.discover(data)
    buf = codec.inflate_batch_device(data, spans).data()
    for n in (buf.byte_len() - 1, buf.byte_len() + 1):
        with pytest.raises(fa.InvalidDeviceBatch):
            buf.copy_to_host(np.empty(n, np.uint8))


def test_select_codec_names():
    import fritillaria_amd as fa
    for b in (fa.Backend.AUTO, fa.Backend.HIP):
        codec, resolved = fa.select_codec(b)
        assert resolved is fa.Backend.HIP
        assert codec.name() == "hip" and codec.device_name() == "hip-device"
        assert codec.device_ordinal() == 0


def test_timed_calls_match_untimed_and_accumulate(hip_ctx, codec, decoder, oracle_mod):
    print("[gpu] boundary: timed inflate and decode", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(1, n_records=3000)
    spans, _ = fa.discover(s.compressed)
    it = fa.InflateTimings()
    b1 = codec.inflate_batch_device_timed(s.compressed, spans, it)
    b0 = codec.inflate_batch_device(s.compressed, spans)
    assert b1.data().to_vec().tobytes() == b0.data().to_vec().tobytes()
    host = codec.inflate_batch_timed(s.compressed, spans, it)
    assert bytes(host.data) == b0.data().to_vec().tobytes()
    assert it.blocks == 2 * len(spans)
    assert it.inflated_bytes == 2 * b0.byte_len()
    assert it.compressed_bytes == 2 * int(spans["payload_len"].sum())
    assert it.kernel_ms > 0 and it.upload_ms > 0 and it.download_ms > 0 and it.total() > 0
    dt = fa.DecodeTimings()
    r1 = decoder.decode_timed(b1, s.records_start, dt)
    r0 = decoder.decode(b0, s.records_start)
    for name, _ in fa.BAM_COLUMNS:
        assert np.array_equal(r1.column_to_host(name), r0.column_to_host(name)), name
    decoder.decode_timed(fa.DeviceInflateBatch(), 0, dt)
    assert dt.batches == 2 and dt.records == 3000 and dt.blocks == len(spans)
    assert dt.bytes == b0.byte_len() and dt.scan_ms > 0 and dt.decode_ms > 0


def test_device_queries_and_the_two_phase_default(hip_ctx):
    """The context's compute units, and the two-phase threshold's default: the
    one-wave form's resident waves, 16 per compute unit (4,096 on 256 CUs)."""
    import fritillaria_amd as fa
    ctx = fa.HipContext(0)
    try:
        cu = ctx.compute_units()
        assert cu > 0
        assert ctx.two_phase_min_members() == 16 * cu
        ctx.set_two_phase_min_members(123)
        assert ctx.two_phase_min_members() == 123
    finally:
        ctx.close()
    assert hip_ctx.compute_units() == cu
