"""GPU parity of the batched, device-resident pipeline: the whole file decoded
in one enqueue (and as one hipGraph replay) equals the oracle's whole-file
decode, bit for bit."""
import numpy as np
import pytest

from conftest import golden_bytes
from helpers import bam_header, bgzf, bgzf_member, build_record, random_records

pytestmark = pytest.mark.gpu

REF_COLS = ["reference_sequence_id", "position", "flags", "mapping_quality", "sequence_len",
            "mate_reference_sequence_id", "mate_position", "template_length", "record_offsets",
            "cigar_start", "sequence_start", "quality_start", "aux_start", "record_end"]


def oracle_whole_file(oracle_mod, data):
    d = oracle_mod.discover(data)
    r = oracle_mod.inflate_batch(data, d.spans)
    assert r.kind == 0
    u = np.frombuffer(r.data, np.uint8)
    rc, start, _, _ = oracle_mod.bam_parse_header(u)
    rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, r.offsets, start)
    assert rc == 0
    cols = oracle_mod.bam_decode_columns(u, rec)
    cols["virtual_offsets"] = oracle_mod.virtual_offsets(r.offsets, d.spans["compressed_offset"], rec)
    return r, u, start, tail, cols


def check_soa(got, want):
    for k in want:
        if k == "aux_entries":
            assert len(got[k]) == len(want[k])
            for f in ("tag", "type", "subtype", "value_offset", "count", "size"):
                assert np.array_equal(got[k][f], want[k][f]), (k, f)
        else:
            assert np.array_equal(got[k], want[k]), k


def check_pipeline(pipe, oracle_mod, data, check_bytes=True, soa=False):
    r, u, start, tail, cols = oracle_whole_file(oracle_mod, data)
    assert pipe.records() == len(cols["position"]) and pipe.tail() == tail
    assert pipe.records_start() == start
    offs, co = pipe.member_offsets()
    assert offs.tolist() == r.offsets.tolist()
    if check_bytes:
        assert pipe.inflated().to_vec().tobytes() == r.data
    got = pipe.columns_to_host()
    for name in REF_COLS + ["virtual_offsets"]:
        assert np.array_equal(got[name], cols[name]), name
    if soa:
        rc, want, _ = oracle_mod.bam_soa(u, cols["record_offsets"])
        assert rc == 0
        check_soa(pipe.soa_to_host(), want)
    else:
        assert pipe.soa_column("cigar_ops") is None


@pytest.mark.parametrize("config,n", [(1, 10000), (2, 60000), (3, 40)])
@pytest.mark.parametrize("graph", [False, True])
def test_pipeline_matches_oracle(hip_ctx, oracle_mod, config, n, graph):
    print(f"[gpu] pipeline config {config} n={n} graph={graph}", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(config, n_records=n)
    pipe = fa.BatchedPipeline(hip_ctx, s.compressed, use_graph=graph, soa=True)
    assert pipe.device_discovery()
    for rep in range(3 if graph else 1):
        pipe.run()
        check_pipeline(pipe, oracle_mod, s.compressed, check_bytes=(rep == 0), soa=True)
        # eager work between replays
        hip_ctx_tmp = fa.HipCodec(hip_ctx).inflate_batch_device(s.compressed[:65536 * 2], fa.discover(
            s.compressed[:200000])[0][:1])
        del hip_ctx_tmp
    if not graph:
        ms = pipe.stage_ms()
        assert ms["total"] > 0 and ms["inflate"] > 0


def small_files():
    recs = [build_record(b"r%d" % i, b"ACGTN" * (i + 1), pos=10 * i, cigar=[((5 * (i + 1)) << 4) | 0]) for i in range(3)]
    head = bam_header()
    return [("header only", bgzf(head)),
            ("header only, no EOF member", bgzf(head, eof=False)),
            ("three records in one member, no EOF member", bgzf(head + b"".join(recs), eof=False)),
            ("one record, header and record in separate members", bgzf_member(head) + bgzf(recs[0]))]


@pytest.mark.parametrize("graph", [False, True])
def test_pipeline_small_files(hip_ctx, oracle_mod, graph):
    """Files with no records, a single member, or no EOF member, against the oracle."""
    import fritillaria_amd as fa
    for name, data in small_files():
        print(f"[gpu] pipeline small file: {name}, graph={graph}", flush=True)
        pipe = fa.BatchedPipeline(hip_ctx, np.frombuffer(data, np.uint8), use_graph=graph, soa=True)
        for rep in range(2):
            pipe.run()
            check_pipeline(pipe, oracle_mod, data, check_bytes=True, soa=True)
        pipe.close()


@pytest.mark.parametrize("name", ["htslib.bam", "htslib_multiblock.bam", "ont_ultralong.bam", "pacbio_hifi.bam"])
def test_pipeline_fixtures(hip_ctx, oracle_mod, name):
    import fritillaria_amd as fa
    data = golden_bytes(name)
    pipe = fa.BatchedPipeline(hip_ctx, data, soa=True)
    pipe.run()
    check_pipeline(pipe, oracle_mod, data, soa=True)


@pytest.mark.parametrize("opts", [dict(order=1), dict(waves_per_cu=1), dict(waves_per_cu=8), dict(bam_window=256),
                                  dict(host_discovery=True)])
def test_schedules_do_not_change_results(hip_ctx, oracle_mod, opts):
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(2, n_records=40000)
    pipe = fa.BatchedPipeline(hip_ctx, s.compressed, **opts)
    pipe.run()
    check_pipeline(pipe, oracle_mod, s.compressed, check_bytes=False)


def bgzf_member_with_extra(payload: bytes, before: bytes = b"", after: bytes = b"") -> bytes:
    """A BGZF member whose FEXTRA holds other subfields before and/or after BC
    (RFC 1952 allows any; the BC one gives the block size)."""
    import struct
    import zlib
    from helpers import deflate_raw
    cdata = deflate_raw(payload, 6)
    extra = before + b"BC" + struct.pack("<H", 2) + b"\0\0" + after
    bsize = 12 + len(extra) + len(cdata) + 8
    at = 12 + len(before) + 4
    hdr = bytes([0x1F, 0x8B, 8, 4, 0, 0, 0, 0, 0, 0xFF]) + struct.pack("<H", len(extra)) + extra
    hdr = hdr[:at] + struct.pack("<H", bsize - 1) + hdr[at + 2:]
    return hdr + cdata + struct.pack("<II", zlib.crc32(payload) & 0xFFFFFFFF, len(payload))


def test_headers_with_other_subfields_are_found_on_the_device(hip_ctx, oracle_mod):
    """Device discovery checks the usual 18-byte header from one load; members
    whose FEXTRA carries other subfields take its general walk, with the same
    result as the oracle."""
    print("[gpu] pipeline: headers with other extra subfields", flush=True)
    import fritillaria_amd as fa
    rng = np.random.default_rng(43)
    data = bam_header() + b"".join(random_records(rng, 600))
    layouts = [(b"", b""), (b"XY\x02\x00ab", b""), (b"", b"ZZ\x01\x00q"), (b"AB\x00\x00", b"CD\x03\x00xyz")]
    members, pos, i = [], 0, 0
    while pos < len(data):
        before, after = layouts[i % len(layouts)]
        members.append(bgzf_member_with_extra(data[pos:pos + 7000], before, after))
        pos += 7000
        i += 1
    blob = b"".join(members) + bgzf(b"", eof=True)
    pipe = fa.BatchedPipeline(hip_ctx, blob)
    assert pipe.device_discovery()
    pipe.run()
    assert pipe.members() == len(oracle_mod.discover(blob).spans)
    check_pipeline(pipe, oracle_mod, blob)


def test_false_header_in_payload_falls_back_to_host_walk(hip_ctx, oracle_mod):
    import fritillaria_amd as fa
    rng = np.random.default_rng(41)
    # a stored member whose payload contains a whole BGZF member: the sieve sees
    # a header that is not on the member chain
    inner = bgzf_member(b"hello" * 100)
    data = bam_header() + b"".join(random_records(rng, 200))
    blob = bgzf(data, cap=9000, eof=False) + bgzf_member(inner, level=0) + bgzf(b"", eof=True)
    pipe = fa.BatchedPipeline(hip_ctx, blob)
    assert not pipe.device_discovery()
    pipe.run()
    assert pipe.members() == len(oracle_mod.discover(blob).spans)


def test_pipeline_reports_a_bad_crc_like_the_per_call_path(hip_ctx):
    import fritillaria_amd as fa
    rng = np.random.default_rng(42)
    data = bam_header() + b"".join(random_records(rng, 300))
    good = bgzf(data, cap=20000, eof=False)
    bad = bgzf_member(b"x" * 100, crc=1)
    blob = good + bad + bgzf(b"", eof=True)
    pipe = fa.BatchedPipeline(hip_ctx, blob, soa=True)
    with pytest.raises(fa.ChecksumMismatch) as ei:
        pipe.run()
    assert ei.value.offset == len(good) and ei.value.expected == 1
    # a failed run leaves no columns behind
    assert pipe.records() == 0
    assert pipe.column("position").byte_len() == 0
    assert pipe.soa_column("sequence_packed").byte_len() == 0


def test_pipeline_soa_malformed_aux_is_reported_and_clears_views(hip_ctx, oracle_mod):
    """A record whose aux field does not parse fails the SoA stage of the
    batched run at that record's offset, as frhip_bam_soa_build does; the run
    before it (good file) and the views after it behave as the contract says."""
    print("[gpu] pipeline: malformed aux in the SoA stage", flush=True)
    import fritillaria_amd as fa
    from helpers import build_record
    recs = [build_record(b"r%d" % i, b"ACGT" * 6, aux=b"NMC\x01") for i in range(40)]
    recs[23] = build_record(b"bad", b"ACGT" * 6, aux=b"NMq\x01")  # unknown type byte 'q'
    hdr = bam_header()
    blob = bgzf(hdr + b"".join(recs), cap=1500)
    pipe = fa.BatchedPipeline(hip_ctx, blob, soa=True, use_graph=True)
    with pytest.raises(fa.Malformed) as ei:
        pipe.run()
    u = np.frombuffer(oracle_mod.inflate_batch(blob, oracle_mod.discover(blob).spans).data, np.uint8)
    rc, rec, _ = oracle_mod.bam_scan_records(u, len(hdr))
    assert ei.value.position == int(rec[23])
    assert pipe.records() == 0 and pipe.soa_column("aux_entries").byte_len() == 0


@pytest.mark.parametrize("schedule", [16, 18])
@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_two_phase_pipeline_matches_oracle(hip_ctx, oracle_mod, schedule, order):
    """The two-phase inflate (one member per lane, then one wave per member) in
    the graph-replayed pipeline, each replay against the oracle."""
    print(f"[gpu] two-phase pipeline schedule {schedule} order {order}", flush=True)
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    default_min = hip_ctx.two_phase_min_members()
    hip_ctx.set_two_phase_min_members(0)  # captured into the graph at creation
    try:
        s = synth.config_bam(2, n_records=60000)
        pipe = fa.BatchedPipeline(hip_ctx, s.compressed, inflate_schedule=schedule, order=order, use_graph=True,
                                  soa=True)
    finally:
        hip_ctx.set_two_phase_min_members(default_min)
    for rep in range(3):
        pipe.run()
        check_pipeline(pipe, oracle_mod, s.compressed, check_bytes=(rep == 0), soa=True)


def test_host_fed_pipelines_alternating(hip_ctx, oracle_mod):
    """Two pipelines fed from pinned host memory in turn, each upload
    overlapping the other's run: every run gives the oracle's columns. A
    re-upload whose member table sums to another inflated size fails
    verification (and a good upload afterwards runs again)."""
    print("[gpu] pipeline: host-fed uploads alternating", flush=True)
    import struct
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(2, n_records=40000)
    pinned = fa.PinnedBuffer(s.compressed.size)
    pinned.array[:] = s.compressed
    pipes = [fa.BatchedPipeline(hip_ctx, pinned.array, use_graph=True, soa=True) for _ in range(2)]
    pipes[0].upload(pinned.array)
    for k in range(4):
        cur, nxt = pipes[k % 2], pipes[(k + 1) % 2]
        cur.launch()
        nxt.upload(pinned.array)
        cur.finish()
        check_pipeline(cur, oracle_mod, s.compressed, check_bytes=(k == 0), soa=True)
    # a member's ISIZE changed: the discovered total no longer matches
    bad = s.compressed.copy()
    d = oracle_mod.discover(s.compressed)
    k = len(d.spans) // 2
    end = int(d.spans["payload_start"][k] + d.spans["payload_len"][k])
    isz = int(d.spans["isize"][k])
    bad[end + 4:end + 8] = np.frombuffer(struct.pack("<I", isz - 100), np.uint8)
    p = pipes[0]
    p.upload(bad)
    with pytest.raises(fa.FrError):
        p.run()
    assert p.records() == 0
    p.upload(pinned.array)
    p.run()
    check_pipeline(p, oracle_mod, s.compressed, check_bytes=True, soa=True)
    for p in pipes:
        p.close()
    pinned.close()
