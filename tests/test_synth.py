"""The synthetic writer frames like noodles (65,495-byte payloads, no
record-boundary flush, EOF member) and its records decode with the oracle."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def small():
    from fritillaria_amd import synth
    return synth.config_bam(1, n_records=3000, threads=4)


def test_framing_is_noodles_shaped(small, oracle_mod):
    c = small.compressed
    d = oracle_mod.discover(c)
    assert d.err == 0 and d.position == c.size
    isz = d.spans["isize"]
    assert (isz[:-2] == 65495).all() and isz[-1] == 0 and 0 < isz[-2] <= 65495
    assert c[-28:].tobytes().hex() == "1f8b08040000000000ff0600424302001b0003000000000000000000"
    r = oracle_mod.inflate_batch(c, d.spans)
    assert r.kind == 0 and len(r.data) == small.inflated_len


def test_records_decode_and_span_members(small, oracle_mod):
    d = oracle_mod.discover(small.compressed)
    r = oracle_mod.inflate_batch(small.compressed, d.spans)
    u = np.frombuffer(r.data, np.uint8)
    rc, rs, _, _ = oracle_mod.bam_parse_header(u)
    assert rc == 0 and rs == small.records_start
    rc, rec, tail, fb = oracle_mod.bam_scan_blocked(u, r.offsets, rs)
    assert rc == 0 and len(rec) == 3000 and tail == u.size
    cols = oracle_mod.bam_decode_columns(u, rec)
    assert (np.diff(cols["position"]) >= 0).all()
    assert (cols["sequence_len"] == 150).all() and (cols["mapping_quality"] == 60).all()
    assert set(np.unique(cols["flags"]).tolist()) <= {0, 16}
    # noodles framing: interior member starts are mid-record, so the reference's
    # block-start speculation falls back on nearly every member
    assert fb >= len(d.spans) - 3


def test_generator_is_thread_count_independent():
    from fritillaria_amd import synth
    a = synth.config_bam(2, n_records=4000, threads=1)
    b = synth.config_bam(2, n_records=4000, threads=7)
    assert np.array_equal(a.compressed, b.compressed)


def test_long_read_config_shape(oracle_mod):
    from fritillaria_amd import synth
    s = synth.config_bam(3, n_records=40, threads=4)
    d = oracle_mod.discover(s.compressed)
    r = oracle_mod.inflate_batch(s.compressed, d.spans)
    u = np.frombuffer(r.data, np.uint8)
    rc, rec, tail, fb = oracle_mod.bam_scan_blocked(u, r.offsets, s.records_start)
    assert rc == 0 and len(rec) == 40
    cols = oracle_mod.bam_decode_columns(u, rec)
    assert (cols["reference_sequence_id"] == -1).all() and (cols["flags"] == 4).all()
    assert (cols["record_end"] > 65536).any()
