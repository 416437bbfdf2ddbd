"""Batched virtual-offset range queries (include/frhip_serve.h) against the
oracle's whole-file records: every query's records and columns equal the
oracle's records with virtual offsets in [vstart, vend), and equal what the
same query returns when it runs alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

COLS = ["reference_sequence_id", "position", "flags", "mapping_quality", "sequence_len",
        "mate_reference_sequence_id", "mate_position", "template_length", "cigar_start",
        "sequence_start", "quality_start", "aux_start", "record_end", "virtual_offsets"]


@pytest.fixture(scope="module")
def served(hip_ctx, oracle_mod):
    import fritillaria_amd as fa
    from fritillaria_amd import synth
    s = synth.config_bam(2, n_records=120_000)
    d = oracle_mod.discover(s.compressed)
    r = oracle_mod.inflate_batch(s.compressed, d.spans)
    u = np.frombuffer(r.data, np.uint8)
    rc, offs, tail = oracle_mod.bam_scan_records(u, s.records_start)
    assert rc == 0
    cols = oracle_mod.bam_decode_columns(u, offs)
    coffs = d.spans["compressed_offset"].astype(np.uint64)
    cols["virtual_offsets"] = oracle_mod.virtual_offsets(r.offsets, coffs, offs)
    server = fa.QueryServer(hip_ctx, s.compressed)
    return server, cols, offs, r.offsets


def check(server, res, queries, cols, offs):
    import fritillaria_amd as fa
    got = {c: server.column_to_host(c) for c in COLS}
    v = cols["virtual_offsets"]
    for k, (vs, ve) in enumerate(queries):
        i = int(np.searchsorted(v, vs))
        j = len(v) if ve == fa.TO_END else int(np.searchsorted(v, ve))
        j = max(i, j)
        assert res[k]["status"] == 0, (k, res[k])
        n = int(res[k]["records"])
        assert n == j - i, (k, n, i, j)
        f = int(res[k]["first"])
        for c in COLS:
            assert np.array_equal(got[c][f:f + n], cols[c][i:j]), (k, c)


def sample_queries(v, rng, n):
    import fritillaria_amd as fa
    q = []
    for _ in range(n):
        i, j = sorted(rng.integers(0, len(v), 2).tolist())
        q.append((int(v[i]), int(v[j]) if rng.random() < 0.9 else fa.TO_END))
    q += [(int(v[0]), fa.TO_END), (int(v[5]), int(v[6])), (int(v[7]), int(v[7])), (int(v[100]), int(v[50]))]
    return q


def test_batch_matches_oracle(served):
    print("[gpu] serve batch vs oracle", flush=True)
    server, cols, offs, _ = served
    rng = np.random.default_rng(11)
    queries = sample_queries(cols["virtual_offsets"], rng, 40)
    res = server.run(queries)
    check(server, res, queries, cols, offs)


def test_query_alone_equals_query_in_mix(served):
    print("[gpu] serve alone vs mix", flush=True)
    server, cols, offs, _ = served
    rng = np.random.default_rng(12)
    queries = sample_queries(cols["virtual_offsets"], rng, 12)
    res = server.run(queries)
    mix = {c: server.column_to_host(c) for c in COLS}
    for k, q in enumerate(queries):
        alone = server.run([q])
        assert alone[0]["records"] == res[k]["records"]
        f, n = int(res[k]["first"]), int(res[k]["records"])
        for c in COLS:
            assert np.array_equal(server.column_to_host(c)[:n], mix[c][f:f + n]), (k, c)


def test_invalid_start_is_reported_per_query(served):
    import fritillaria_amd as fa
    server, cols, offs, _ = served
    v = cols["virtual_offsets"]
    bad = ((int(v[10]) >> 16) + 1) << 16  # a compressed offset no member starts at
    res = server.run([(bad, fa.TO_END), (int(v[3]), int(v[9]))])
    assert res[0]["status"] == 1  # FRHIP_E_INVALID_BLOCK
    assert res[1]["status"] == 0 and res[1]["records"] == 6


def check_slot(server, slot, res, queries, cols):
    """Each query's records in the slot's columns equal the oracle's records
    with virtual offsets in [vstart, vend)."""
    import fritillaria_amd as fa
    got = {c: server.slot_column_to_host(slot, c) for c in COLS}
    v = cols["virtual_offsets"]
    for k, (vs, ve) in enumerate(queries):
        i = int(np.searchsorted(v, vs))
        j = max(i, len(v) if ve == fa.TO_END else int(np.searchsorted(v, ve)))
        assert res[k]["status"] == 0 and int(res[k]["records"]) == j - i, (slot, k, res[k])
        f = int(res[k]["first"])
        for c in COLS:
            assert np.array_equal(got[c][f:f + j - i], cols[c][i:j]), (slot, k, c)


def test_two_slots_run_concurrently(served):
    """A batch launched in slot 1 while slot 0's is on the device: both give the
    oracle's records, the same as each query alone, and their device intervals
    overlap."""
    print("[gpu] serve: two slots in flight", flush=True)
    server, cols, offs, _ = served
    assert server.slots() == 2
    rng = np.random.default_rng(21)
    qa = sample_queries(cols["virtual_offsets"], rng, 30)
    qb = sample_queries(cols["virtual_offsets"], rng, 30)
    server.launch(0, qa)
    server.launch(1, qb)
    with pytest.raises(Exception):
        server.launch(1, qb[:1])  # busy
    rb = server.finish(1)
    ra = server.finish(0)
    check_slot(server, 0, ra, qa, cols)
    check_slot(server, 1, rb, qb, cols)
    (a0, a1), (b0, b1) = server.batch_interval(0), server.batch_interval(1)
    assert 0 <= a0 < a1 and 0 <= b0 < b1
    assert b0 < a1 and a0 < b1, ((a0, a1), (b0, b1))
    for k in range(0, 30, 7):
        alone = server.run([qb[k]])
        assert alone[0]["records"] == rb[k]["records"]
        n = int(alone[0]["records"])
        f = int(rb[k]["first"])
        for c in ("position", "virtual_offsets", "record_end"):
            assert np.array_equal(server.column_to_host(c)[:n], server.slot_column_to_host(1, c)[f:f + n])


def test_serve_loop_overlaps_batches(served):
    """The continuous-batching loop with staggered arrivals: every request is
    served without failure, the records served are the oracle's, and batches
    launched while others were on the device overlapped them."""
    print("[gpu] serve loop with staggered arrivals", flush=True)
    from fritillaria_amd import serve as sv
    server, cols, offs, _ = served
    v = cols["virtual_offsets"]
    co, _ = server.member_table()
    reqs = sv.make_requests(v, co, 60, rate_hz=4000.0, seed=7, lo_bytes=64 << 10, hi_bytes=4 << 20)
    st = sv.serve(server, reqs, budget_bytes=8 << 30)
    summ = st.summary()
    assert summ["failures"] == 0 and summ["requests"] == 60
    want = 0
    for r in reqs:
        i = int(np.searchsorted(v, r.vstart))
        j = len(v) if r.vend == (1 << 64) - 1 else int(np.searchsorted(v, r.vend))
        want += max(0, j - i)
    assert summ["records"] == want
    assert st.batches == len(st.intervals)
    print(f"[gpu] serve loop: {summ}", flush=True)
    assert summ["overlapped_launches"] >= 1 and summ["device_overlap_ms"] > 0
