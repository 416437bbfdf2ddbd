"""The C-ABI library loads and exports everything include/*.h declares; its
host-only entry points agree with the oracle and zlib. No device needed."""
import ctypes
import json
import os
import subprocess
import zlib

import numpy as np

from conftest import GOLDEN, golden_bytes

MEMBERS = json.load(open(os.path.join(GOLDEN, "members.json")))


def test_library_exports_every_declared_symbol():
    import fritillaria_amd.hip as h
    lib = h.lib()
    names = h.exported_symbols()
    assert len(names) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", h._build.frhip_lib_path()], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert getattr(lib, n) is not None


def test_crc32_combine_matches_zlib():
    from fritillaria_amd import crc32_combine
    rng = np.random.default_rng(2)
    for n1, n2 in [(0, 0), (1, 0), (0, 5), (100, 1), (65536, 4097), (3, 70000)]:
        a = rng.integers(0, 256, n1, dtype=np.uint8).tobytes()
        b = rng.integers(0, 256, n2, dtype=np.uint8).tobytes()
        assert crc32_combine(zlib.crc32(a), zlib.crc32(b), n2) == zlib.crc32(a + b)


def test_discover_matches_golden_tables():
    import fritillaria_amd as fa
    for name, t in MEMBERS.items():
        buf = golden_bytes(name)
        spans, pos = fa.discover(buf)
        assert pos == len(buf)
        got = [[int(s["compressed_offset"]), int(s["payload_len"]), int(s["isize"]), int(s["crc32"])]
               for s in spans]
        assert got == t["members"], name


def test_discover_refuses_plain_gzip():
    import fritillaria_amd as fa
    try:
        fa.discover(golden_bytes("pacbio_hifi.plain.fastq.gz"))
    except fa.InvalidBlock as e:
        assert e.offset == 0 and "FEXTRA" in e.reason
    else:
        raise AssertionError("plain gzip must be refused")


def test_bam_header_matches_oracle(oracle_mod):
    import fritillaria_amd as fa
    for name in ("htslib.bam", "htslib_multiblock.bam", "ont_ultralong.bam", "pacbio_hifi.bam"):
        buf = golden_bytes(name)
        d = oracle_mod.discover(buf)
        u = oracle_mod.inflate_batch(buf, d.spans).data
        rc, rs, _, nref = oracle_mod.bam_parse_header(u)
        assert fa.bam_parse_header(u) == (rs, nref)
    try:
        fa.bam_parse_header(b"BAM\2" + bytes(20))
    except fa.Malformed as e:
        assert e.format == "bam header"
    else:
        raise AssertionError


def test_no_device_means_unavailable_not_fallback():
    import fritillaria_amd as fa
    if fa.device_is_available():
        return
    try:
        fa.HipContext(0)
    except fa.HipUnavailable:
        pass
    else:
        raise AssertionError("a context without a device must raise")


def damaged(rng, good: bytes, n: int):
    """n damaged copies: flipped bits, overwritten 16/32-bit fields, truncations, insertions."""
    out = []
    for _ in range(n):
        b = bytearray(good)
        kind = int(rng.integers(0, 4))
        k = int(rng.integers(0, len(b)))
        if kind == 0:
            for _ in range(int(rng.integers(1, 4))):
                j = int(rng.integers(0, len(b)))
                b[j] ^= 1 << int(rng.integers(0, 8))
        elif kind == 1:
            w = int(rng.choice([2, 4]))
            v = int(rng.integers(0, 1 << (8 * w)))
            b[k:k + w] = v.to_bytes(w, "little")[:len(b[k:k + w])]
        elif kind == 2:
            b = b[:k]
        else:
            b[k:k] = rng.integers(0, 256, int(rng.integers(1, 30)), dtype=np.uint8).tobytes()
        out.append(bytes(b))
    return out


def test_discover_agrees_with_the_oracle_on_damaged_streams(oracle_mod):
    """Damaged BGZF streams (headers, BSIZE, XLEN, subfields, trailers, cuts):
    the host walk finds the same members and stops at the same position as the
    oracle, or refuses the stream at the same offset."""
    import fritillaria_amd as fa
    good = golden_bytes("htslib_multiblock.bam")[:200_000]
    rng = np.random.default_rng(43)
    refused = 0
    for i, b in enumerate(damaged(rng, good, 1500)):
        want = oracle_mod.discover(b)
        try:
            spans, pos = fa.discover(b)
        except fa.InvalidBlock as e:
            assert want.err != 0 and e.offset == want.err_offset, (i, want.err, want.err_offset, e.offset)
            refused += 1
            continue
        assert want.err == 0, (i, want.err, want.err_offset)
        assert pos == want.position and np.array_equal(spans, want.spans), i
    assert refused > 100


def test_bam_header_agrees_with_the_oracle_on_damaged_headers(oracle_mod):
    """Damaged BAM headers (magic, l_text, n_ref, name lengths, cuts): the same
    records start, or a refusal where the oracle refuses."""
    import fritillaria_amd as fa
    buf = golden_bytes("htslib_multiblock.bam")
    d = oracle_mod.discover(buf)
    u = oracle_mod.inflate_batch(buf, d.spans).data
    _, rs, _, _ = oracle_mod.bam_parse_header(u)
    good = bytes(u[:rs + 64])
    rng = np.random.default_rng(47)
    for i, b in enumerate(damaged(rng, good[:rs], 1500)):
        b = b + good[rs:]
        rc, start, _, nref = oracle_mod.bam_parse_header(np.frombuffer(b, np.uint8))
        try:
            got = fa.bam_parse_header(b)
        except fa.Malformed:
            assert rc != 0, i
            continue
        assert rc == 0 and got == (start, nref), (i, rc, got, start, nref)
