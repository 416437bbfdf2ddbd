"""The unaligned long-read configuration at scale: 2,000 records of log-normal
length (median 20 kb, sigma 1) with every 500th forced to the 1 Mb cap, so
records span dozens of members and whole reader batches. Read through the
DeviceBgzfReader loop (the carry-the-partial-record path of
crates/fritillaria-bgzf/src/device_read.rs:155-182) at 1, 2 and 8 members per
batch, and through the batched pipeline with the SoA extras, against the
oracle's whole-file decode."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 2000
CAP_EVERY = 500


@pytest.fixture(scope="module")
def long_reads(oracle_mod):
    from fritillaria_amd import synth
    s = synth.config_bam(3, n_records=N, cap_every=CAP_EVERY)
    d = oracle_mod.discover(s.compressed)
    r = oracle_mod.inflate_batch(s.compressed, d.spans)
    assert r.kind == 0
    u = np.frombuffer(r.data, np.uint8)
    rc, rec, tail, _ = oracle_mod.bam_scan_blocked(u, r.offsets, s.records_start)
    assert rc == 0 and len(rec) == N and tail == u.size
    cols = oracle_mod.bam_decode_columns(u, rec)
    cols["virtual_offsets"] = oracle_mod.virtual_offsets(r.offsets, d.spans["compressed_offset"], rec)
    capped = np.flatnonzero(cols["sequence_len"] == 1_000_000)
    assert set(range(CAP_EVERY // 2, N, CAP_EVERY)) <= set(capped.tolist())  # one more reaches the cap by chance
    # a capped record holds its bases, qualities and u16 kinetics: dozens of members
    spans = np.diff(np.searchsorted(r.offsets, np.stack([rec[capped], rec[capped] + cols["record_end"][capped]]),
                                    side="right"), axis=0)
    assert (spans >= 40).all()
    return s, d, r, u, rec, cols


@pytest.mark.parametrize("bpb", [1, 2, 8])
def test_reader_loop_carries_long_records(codec, decoder, long_reads, bpb):
    print(f"[gpu] long reads: reader loop blocks_per_batch={bpb}", flush=True)
    import fritillaria_amd as fa
    s, d, r, u, rec, cols = long_reads
    reader = fa.DeviceBgzfReader(s.compressed, codec, blocks_per_batch=bpb)
    first, batches, empty = True, 0, 0
    got = {k: [] for k in ("virtual_offsets", "position", "sequence_len", "record_end", "aux_start")}
    while True:
        batch = reader.next_batch()
        if batch is None:
            break
        batches += 1
        start = batch.start
        if first:
            start, _ = fa.bam_parse_header(batch.data.data().to_vec()[: 1 << 20])
            first = False
        records = decoder.decode(batch.data, start)
        if len(records) == 0:
            empty += 1  # the whole batch is carried into the next, grown window
        for k in got:
            got[k].append(records.column_to_host(k))
        reader.carry_from(records.tail())
        del records
    for k in got:
        assert np.array_equal(np.concatenate(got[k]), cols[k]), k
    # a 1 Mb record spans more members than any of these batches holds
    assert empty >= 4
    assert batches >= len(d.spans) // (2 * bpb)


def test_pipeline_long_reads_with_soa(hip_ctx, oracle_mod, long_reads):
    print("[gpu] long reads: batched pipeline with SoA", flush=True)
    import fritillaria_amd as fa
    from test_gpu_pipeline import REF_COLS, check_soa
    s, d, r, u, rec, cols = long_reads
    pipe = fa.BatchedPipeline(hip_ctx, s.compressed, soa=True, use_graph=True)
    for rep in range(2):
        pipe.run()
        assert pipe.records() == N and pipe.tail() == u.size
        got = pipe.columns_to_host()
        for name in REF_COLS + ["virtual_offsets"]:
            assert np.array_equal(got[name], cols[name]), name
    rc, want, _ = oracle_mod.bam_soa(u, rec)
    assert rc == 0
    soa = pipe.soa_to_host()
    check_soa(soa, want)
    # MM:Z, ML:B:C and the fi:B:S kinetics: three aux fields per record, the
    # arrays addressed in place (zero-copy B arrays)
    assert (np.diff(soa["aux_offsets"]) == 3).all()
    ent = soa["aux_entries"].reshape(N, 3)
    assert (ent["tag"][:, 2] == b"fi").all() and (ent["subtype"][:, 2] == ord("S")).all()
    assert np.array_equal(ent["count"][:, 2], cols["sequence_len"])
    # members inside a long record, or whose first record start lies past the
    # speculation window, are resolved by the exact serial chain: a cost, never
    # a different answer (the columns above are the oracle's)
    print(f"[gpu] long reads: {pipe.members()} members, fallbacks {pipe.fallbacks()}", flush=True)
    assert pipe.fallbacks() == 0
