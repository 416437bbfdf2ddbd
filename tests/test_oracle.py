"""Pins the CPU oracle against things it does not compute itself: zlib, the
reference's known answers and golden member tables, brute force, and the
closed forms of the formats. CPU only."""
import json
import os
import zlib

import numpy as np
import pytest

from conftest import GOLDEN, golden_bytes
from helpers import (EOF_BLOCK, bam_header, bgzf, bgzf_member, build_record, corrupt_streams,
                     deflate_raw, random_records)

KNOWN = json.load(open(os.path.join(GOLDEN, "known_answers.json")))
MEMBERS = json.load(open(os.path.join(GOLDEN, "members.json")))


def test_crc32_known_answers(oracle_mod):
    assert oracle_mod.crc32(b"noodles") == int(KNOWN["crc32"]["noodles"], 16)
    assert oracle_mod.crc32(b"") == 0


def test_crc32_matches_zlib(oracle_mod):
    rng = np.random.default_rng(1)
    for n in [1, 2, 3, 7, 64, 1000, 65536]:
        d = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert oracle_mod.crc32(d) == zlib.crc32(d)


def _payloads(rng):
    yield b""
    yield b"a"
    yield b"ACGT" * 5000
    yield bytes(rng.integers(0, 256, 65536, dtype=np.uint8))            # incompressible
    yield bytes(rng.choice(list(b"ACGTN"), 60000).tolist())             # low entropy
    yield b"\0" * 65536                                                 # long runs: overlapping copies
    yield b"".join(random_records(rng, 200))                           # BAM-shaped


@pytest.mark.parametrize("level", list(range(10)))
def test_inflate_matches_zlib_every_level(oracle_mod, level):
    rng = np.random.default_rng(level)
    for data in _payloads(rng):
        data = data[:65536]
        c = deflate_raw(data, level)
        st, out = oracle_mod.inflate_member(c, len(data))
        assert st == 0 and out == data


@pytest.mark.parametrize("strategy", [zlib.Z_FILTERED, zlib.Z_HUFFMAN_ONLY, zlib.Z_RLE, zlib.Z_FIXED])
def test_inflate_matches_zlib_every_strategy(oracle_mod, strategy):
    rng = np.random.default_rng(7)
    for data in _payloads(rng):
        data = data[:65536]
        c = deflate_raw(data, 6, strategy)
        st, out = oracle_mod.inflate_member(c, len(data))
        assert st == 0 and out == data


def test_inflate_multi_block_streams(oracle_mod):
    rng = np.random.default_rng(3)
    data = b"".join(random_records(rng, 150))[:60000]
    c = deflate_raw(data, 6, flushes=(1, 100, 30000, 59999))
    st, out = oracle_mod.inflate_member(c, len(data))
    assert st == 0 and out == data


def test_inflate_status_codes(oracle_mod):
    for name, payload, cap, want in corrupt_streams():
        st, _ = oracle_mod.inflate_member(payload, cap)
        assert st == want, name


def test_member_tables_match_golden(oracle_mod):
    for name, t in MEMBERS.items():
        buf = golden_bytes(name)
        d = oracle_mod.discover(buf)
        assert d.err == 0 and d.position == len(buf)
        got = [[int(s["compressed_offset"]), int(s["payload_len"]), int(s["isize"]), int(s["crc32"])]
               for s in d.spans]
        assert got == t["members"], name
        r = oracle_mod.inflate_batch(buf, d.spans)
        assert r.kind == 0
        assert len(r.data) == t["inflated_len"] and zlib.crc32(r.data) == t["inflated_crc32"]


def test_eof_block_and_bc_plus_one(oracle_mod):
    assert EOF_BLOCK.hex() == KNOWN["eof_block_hex"]
    d = oracle_mod.discover(EOF_BLOCK)
    assert len(d.spans) == 1
    s = d.spans[0]
    # BSIZE field 0x001b = 27 -> block of 28 bytes: the +1 rule
    assert int(s["isize"]) == 0 and int(s["crc32"]) == 0 and int(s["payload_len"]) == 28 - 18 - 8
    assert d.position == 28


def test_plain_gzip_is_refused(oracle_mod):
    d = oracle_mod.discover(golden_bytes("pacbio_hifi.plain.fastq.gz"))
    assert d.err == 4 and len(d.spans) == 0  # missing FEXTRA


def test_discovery_stops_cleanly_on_partial_members(oracle_mod):
    data = bgzf(b"x" * 200000, cap=30000)
    full = oracle_mod.discover(data)
    ends = [int(s["compressed_offset"]) for s in full.spans] + [len(data)]
    for cut in range(0, len(data), 997):
        d = oracle_mod.discover(data[:cut])
        assert d.err == 0
        assert d.position == max(e for e in ends if e <= cut) if cut >= ends[0] else d.position == 0


def test_batch_verification_order(oracle_mod):
    a = bgzf_member(b"A" * 1000)
    b = bgzf_member(b"B" * 1000, crc=0x12345678)
    c = bgzf_member(b"C" * 1000, isize=999)
    buf = a + b + c
    d = oracle_mod.discover(buf)
    r = oracle_mod.inflate_batch(buf, d.spans)
    assert r.kind == 2 and r.offset == len(a)            # checksum mismatch at member 1
    r = oracle_mod.inflate_batch(a + c, oracle_mod.discover(a + c).spans)
    assert r.kind == 4 and r.offset == len(a)            # ISIZE 999 < produced: output overflow
    cc = bgzf_member(b"C" * 1000, isize=1001)
    r = oracle_mod.inflate_batch(a + cc, oracle_mod.discover(a + cc).spans)
    assert r.kind == 3 and r.expected == 1001 and r.actual == 1000


def _inflate_fixture(oracle_mod, name):
    buf = golden_bytes(name)
    d = oracle_mod.discover(buf)
    r = oracle_mod.inflate_batch(buf, d.spans)
    return np.frombuffer(r.data, np.uint8), r.offsets, d.spans


def test_htslib_known_answers(oracle_mod):
    k = KNOWN["htslib.bam"]
    u, offs, _ = _inflate_fixture(oracle_mod, "htslib.bam")
    rc, start, _, nref = oracle_mod.bam_parse_header(u)
    assert rc == 0 and nref == 2
    rc, rec, tail = oracle_mod.bam_scan_records(u, start)
    assert rc == 0 and len(rec) == k["records"] and tail == len(u)
    cols = oracle_mod.bam_decode_columns(u, rec)
    for f in ("position", "flags", "mapping_quality", "sequence_len", "reference_sequence_id"):
        assert cols[f].tolist() == k[f], f
    # variable-length fields through the column bounds
    for i, name in k["names"].items():
        i = int(i)
        o = int(rec[i])
        assert bytes(u[o + 36:o + int(cols["cigar_start"][i]) - 1]).decode() == name
    for i, seq in k["sequences"].items():
        i = int(i)
        o = int(rec[i])
        packed = u[o + int(cols["sequence_start"][i]):o + int(cols["quality_start"][i])]
        bases = "=ACMGRSVTWYHKDBN"
        got = "".join(bases[(packed[j // 2] >> 4) if j % 2 == 0 else (packed[j // 2] & 15)]
                      for j in range(int(cols["sequence_len"][i])))
        assert got == seq
    o = int(rec[0])
    ops = np.frombuffer(bytes(u[o + int(cols["cigar_start"][0]):o + int(cols["sequence_start"][0])]), "<u4")
    assert [("MIDNSHP=X"[v & 15], int(v >> 4)) for v in ops] == [tuple(x) for x in k["cigar_r001"]]
    q = k["qual_absent_record"]
    o = int(rec[q])
    assert all(b == 0xFF for b in u[o + int(cols["quality_start"][q]):o + int(cols["aux_start"][q])])


def test_multiblock_known_answers(oracle_mod):
    k = KNOWN["htslib_multiblock.bam"]
    u, offs, _ = _inflate_fixture(oracle_mod, "htslib_multiblock.bam")
    rc, start, _, _ = oracle_mod.bam_parse_header(u)
    rc, rec, tail = oracle_mod.bam_scan_records(u, start)
    assert len(rec) == k["records"] and tail == len(u)
    cols = oracle_mod.bam_decode_columns(u, rec)
    assert cols["position"][:2].tolist() == k["position_first"]
    assert int(cols["position"][-1]) == k["position_last"]
    assert (cols["sequence_len"] == k["sequence_len_all"]).all()
    assert (cols["mapping_quality"] == k["mapping_quality_all"]).all()
    o = int(rec[-1])
    assert bytes(u[o + 36:o + int(cols["cigar_start"][-1]) - 1]).decode() == k["last_name"]


def test_ont_record_table(oracle_mod):
    k = KNOWN["ont_ultralong.bam"]
    u, offs, _ = _inflate_fixture(oracle_mod, "ont_ultralong.bam")
    rc, start, _, _ = oracle_mod.bam_parse_header(u)
    rc, rec, tail = oracle_mod.bam_scan_records(u, start)
    cols = oracle_mod.bam_decode_columns(u, rec)
    assert cols["record_end"].tolist() == k["record_bytes"]
    assert cols["sequence_len"].tolist() == k["sequence_len"]
    n_cigar = ((cols["sequence_start"] - cols["cigar_start"]) // 4).tolist()
    assert n_cigar == k["cigar_ops"]
    # the fixture exercises the fallback of the blocked scan
    # (crates/fritillaria-bam/tests/blocked.rs:66-112)
    rc, rec2, tail2, fb = oracle_mod.bam_scan_blocked(u, offs, start)
    assert rc == 0 and fb >= 1 and rec2.tolist() == rec.tolist()


def _blockings(n, step):
    return [np.unique(np.concatenate([np.arange(0, n, b), [n]])) for b in range(1, n + 1, step)]


def test_blocked_scan_equals_plain_scan_every_blocking(oracle_mod):
    rng = np.random.default_rng(11)
    recs = random_records(rng, 25, max_len=40)
    buf = np.frombuffer(b"".join(recs), np.uint8)
    rc, plain, tail = oracle_mod.bam_scan_records(buf, 0)
    for bs in _blockings(buf.size, 1):
        rc2, off, t2, _ = oracle_mod.bam_scan_blocked(buf, bs, 0)
        assert rc2 == 0 and t2 == tail and off.tolist() == plain.tolist()
    # truncated buffers: tail is the first record the buffer cuts short
    for cut in range(0, buf.size, 7):
        b = buf[:cut]
        rc, plain, tail = oracle_mod.bam_scan_records(b, 0)
        rc2, off, t2, _ = oracle_mod.bam_scan_blocked(b, np.array([0, cut], np.uint64), 0)
        assert t2 == tail and off.tolist() == plain.tolist()


def test_blocked_scan_reports_malformed_at_chain_entry(oracle_mod):
    recs = [build_record(b"r%d" % i, b"ACGT" * 5) for i in range(6)]
    bad = bytearray(recs[3])
    bad[20:24] = (1 << 20).to_bytes(4, "little")  # l_seq far beyond block_size: fields_fit fails
    buf = np.frombuffer(b"".join(recs[:3]) + bytes(bad) + b"".join(recs[4:]), np.uint8)
    rc, off, tail, _ = oracle_mod.bam_scan_blocked(buf, np.array([0, buf.size], np.uint64), 0)
    assert rc == 1 and tail == 0   # the chain entry of the member holding the bad record
    mid = len(recs[0]) + len(recs[1]) + 3
    rc, off, tail, _ = oracle_mod.bam_scan_blocked(buf, np.array([0, mid, buf.size], np.uint64), 0)
    assert rc == 1 and tail == len(recs[0]) + len(recs[1]) + len(recs[2])


def test_virtual_offset_vectors(oracle_mod):
    # members at compressed offsets 0, 13, 21 holding 10, 5, 7 bytes
    bo = np.array([0, 10, 15, 22], np.uint64)
    co = np.array([0, 13, 21], np.uint64)
    v = oracle_mod.virtual_offsets(bo, co, [10, 12, 15])
    assert [int(x) for x in v] == [p for p, _, _ in KNOWN["virtual_positions"]]
    for (p, c, u), x in zip(KNOWN["virtual_positions"], v):
        assert int(x) >> 16 == c and int(x) & 0xFFFF == u


def test_mt_baseline_matches_single_thread(oracle_mod):
    rng = np.random.default_rng(5)
    data = bam_header() + b"".join(random_records(rng, 3000))
    f = bgzf(data)
    for t in (1, 3):
        rc, r = oracle_mod.mt_bam_columns(f, t)
        assert rc == 0 and r["records"] == 3000 and r["inflated"] == len(data)


def test_mt_baseline_soa_checksums_match_serial_soa(oracle_mod):
    """The CPU baseline's SoA totals and checksums equal those of the serial
    or_bam_soa over the same records (itself pinned by the fixtures' known
    CIGARs and sequences), whatever the worker count."""
    rng = np.random.default_rng(6)
    data = bam_header() + b"".join(random_records(rng, 2500))
    f = bgzf(data, cap=20000)
    u = np.frombuffer(data, np.uint8)
    rc, off, _ = oracle_mod.bam_scan_records(u, len(bam_header()))
    assert rc == 0
    rc, soa, _ = oracle_mod.bam_soa(u, off)
    assert rc == 0
    want = {"ops": int(soa["cigar_offsets"][-1]), "sum_ops": int(soa["cigar_ops"].astype(np.uint64).sum()),
            "seq_bytes": int(soa["sequence_offsets"][-1]),
            "sum_seq": int(soa["sequence_packed"].astype(np.uint64).sum()),
            "aux": int(soa["aux_offsets"][-1]),
            "sum_aux_value_offset": int(soa["aux_entries"]["value_offset"].astype(np.uint64).sum())}
    assert want["ops"] > 0 and want["seq_bytes"] > 0 and want["aux"] == 2 * 2500
    for t in (1, 4):
        rc, r = oracle_mod.mt_bam_columns(f, t, soa=True)
        assert rc == 0 and r["records"] == 2500
        for k, v in want.items():
            assert r[k] == v, k


def test_mt_inflate_equals_zlib_per_member(oracle_mod):
    """The multithreaded whole-stream inflate (the BCF / FASTQ CPU baseline's)
    gives Python zlib's bytes, and fails on a bad trailer CRC."""
    import zlib
    rng = np.random.default_rng(8)
    data = bytes(rng.integers(0, 4, 300_000, dtype=np.uint8) + 65)
    f = bgzf(data, cap=7000)
    for t in (1, 5):
        rc, u = oracle_mod.mt_inflate(f, t)
        assert rc == 0 and u.tobytes() == data
    bad = bytearray(f)
    d = oracle_mod.discover(f)
    at = int(d.spans["payload_start"][3] + d.spans["payload_len"][3])
    bad[at] ^= 1
    rc, _ = oracle_mod.mt_inflate(bytes(bad), 3)
    assert rc != 0
    assert zlib.decompress(f[18:18 + int(d.spans["payload_len"][0])], -15) == data[:7000]
