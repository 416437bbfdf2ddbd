"""GPU parity of the inflate entry point that allocates nothing
(frhip_inflate_batch_device_noalloc) against frhip_inflate_batch_device, the
entry point whose kernels it shares: same bytes and member boundaries on two
calls in a row (the second smaller, so stale scratch from the first would
show), the same error on a corrupt member, and its own limits."""
import numpy as np
import pytest

from helpers import bam_header, bgzf, bgzf_member, random_records

pytestmark = pytest.mark.gpu

PAD = 4096


def device_batch(ctx, data: bytes):
    """The batch in device memory, followed by PAD zero bytes."""
    buf = ctx.alloc(len(data) + PAD)
    buf.copy_from_host(np.frombuffer(data, np.uint8))
    buf.copy_from_host(np.zeros(PAD, np.uint8), offset=len(data))
    return buf


def run_both(ctx, codec, data: bytes):
    import fritillaria_amd as fa
    spans, _ = fa.discover(data)
    want = codec.inflate_batch_device(data, spans)
    d_in = device_batch(ctx, data)
    total = int(spans["isize"].astype(np.uint64).sum())
    d_out = ctx.alloc(total + 100)
    offs = codec.inflate_batch_device_noalloc(d_in, len(data), spans, d_out)
    assert offs.tolist() == want.offsets().tolist()
    got = d_out.to_vec()[:total]
    assert got.tobytes() == (want.data().to_vec().tobytes() if total else b"")
    return spans


def test_noalloc_matches_allocating_entry_point_twice(hip_ctx, codec):
    print("[gpu] no-alloc inflate: two calls, the second smaller", flush=True)
    from fritillaria_amd import synth
    big = synth.config_bam(2, n_records=30000).compressed
    spans = run_both(hip_ctx, codec, big.tobytes())
    assert len(spans) > 100
    rng = np.random.default_rng(3)
    small = bgzf(bam_header() + b"".join(random_records(rng, 200)), cap=5000)
    run_both(hip_ctx, codec, small)


def test_noalloc_every_schedule_small(hip_ctx, codec):
    print("[gpu] no-alloc inflate: schedules", flush=True)
    import fritillaria_amd as fa
    rng = np.random.default_rng(4)
    data = bgzf(bam_header() + b"".join(random_records(rng, 400)), cap=9000)
    n = len(fa.hip.inflate_schedules())
    try:
        for sched in range(n):
            hip_ctx.set_inflate_schedule(sched)
            run_both(hip_ctx, codec, data)
    finally:
        hip_ctx.set_inflate_schedule(16)


def test_noalloc_reports_the_same_error(hip_ctx, codec):
    import fritillaria_amd as fa
    rng = np.random.default_rng(5)
    good = bgzf(bam_header() + b"".join(random_records(rng, 100)), cap=4000, eof=False)
    blob = good + bgzf_member(b"y" * 300, crc=7) + bgzf_member(b"z" * 50, isize=3)
    spans, _ = fa.discover(blob)
    with pytest.raises(fa.ChecksumMismatch) as want:
        codec.inflate_batch_device(blob, spans)
    d_in = device_batch(hip_ctx, blob)
    d_out = hip_ctx.alloc(int(spans["isize"].astype(np.uint64).sum()) + 16)
    with pytest.raises(fa.ChecksumMismatch) as got:
        codec.inflate_batch_device_noalloc(d_in, len(blob), spans, d_out)
    assert (got.value.offset, got.value.expected, got.value.actual) == \
        (want.value.offset, want.value.expected, want.value.actual)


def test_noalloc_limits(hip_ctx, codec):
    import fritillaria_amd as fa
    data = bgzf(b"a" * 1000, cap=10)  # 100 members + EOF
    spans, _ = fa.discover(data)
    d_in = device_batch(hip_ctx, data)
    # output too small
    with pytest.raises(fa.InvalidDeviceBatch):
        codec.inflate_batch_device_noalloc(d_in, len(data), spans, hip_ctx.alloc(999))
    # more members than the fixed scratch holds
    many = np.concatenate([spans] * (4096 // len(spans) + 1))
    with pytest.raises(fa.InvalidDeviceBatch):
        codec.inflate_batch_device_noalloc(d_in, len(data), many, hip_ctx.alloc(1 << 20))
    # padding after the batch not zero
    bad = hip_ctx.alloc(len(data) + PAD)
    bad.copy_from_host(np.frombuffer(data, np.uint8))
    bad.copy_from_host(np.full(PAD, 7, np.uint8), offset=len(data))
    with pytest.raises(fa.InvalidDeviceBatch):
        codec.inflate_batch_device_noalloc(bad, len(data), spans, hip_ctx.alloc(2000))
    # and a good call after the failures
    d_out = hip_ctx.alloc(1000)
    offs = codec.inflate_batch_device_noalloc(d_in, len(data), spans, d_out)
    assert d_out.to_vec().tobytes() == b"a" * 1000 and int(offs[-1]) == 1000
