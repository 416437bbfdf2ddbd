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
