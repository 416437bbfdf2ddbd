"""The BAM SoA oracle (or_bam_soa: resolved CIGAR ops, packed sequences, aux
index) against what pins it: the source SAM of htslib.bam
(testdata/README.md:125-139: CIGARs, sequences, aux tags and values), the ONT
record table (testdata/README.md:147-152: CIGAR op counts), the reference's
long-CIGAR rule (crates/fritillaria-bam/src/columnar/record.rs:316-361) and its
aux decoding rules (aux.rs) on hand-built records."""
import gzip
import json
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import build_record

BASES = b"=ACMGRSVTWYHKDBN"
OPS = "MIDNSHP=X"


def records(oracle_mod, name):
    d = gzip.decompress(open(os.path.join(GOLDEN, name), "rb").read())
    rc, rs, _, _ = oracle_mod.bam_parse_header(d)
    rc, offs, tail = oracle_mod.bam_scan_records(d, rs)
    return d, offs


def unpack(packed: bytes, n: int) -> bytes:
    out = bytearray()
    for b in packed:
        out += bytes([BASES[b >> 4], BASES[b & 15]])
    return bytes(out[:n])


def test_htslib_against_its_source_sam(oracle_mod):
    d, offs = records(oracle_mod, "htslib.bam")
    rc, soa, _ = oracle_mod.bam_soa(d, offs)
    assert rc == 0
    ka = json.load(open(os.path.join(GOLDEN, "known_answers.json")))["htslib.bam"]
    ops = soa["cigar_ops"]
    co = soa["cigar_offsets"]
    r001 = [(OPS[o & 15], int(o >> 4)) for o in ops[co[0]:co[1]]]
    assert [list(x) for x in r001] == ka["cigar_r001"]
    assert (OPS[ops[co[2]] & 15], int(ops[co[2]] >> 4)) == tuple(ka["cigar_r003_first"])
    assert co[7] - co[6] == 0  # unmapped r007: no CIGAR
    so = soa["sequence_offsets"]
    for i, seq in ka["sequences"].items():
        i = int(i)
        assert unpack(bytes(soa["sequence_packed"][so[i]:so[i + 1]]), len(seq)) == seq.encode()
    # aux tags of the source SAM, with values read at the indexed offsets
    want = {0: [b"RG", b"NM"], 1: [b"NM"], 2: [b"XA", b"XI"], 3: [b"XF"], 4: [], 5: [b"NM"], 6: [], 7: [b"XB"]}
    ao, ae = soa["aux_offsets"], soa["aux_entries"]
    for i, tags in want.items():
        assert [bytes(e["tag"]) for e in ae[ao[i]:ao[i + 1]]] == tags
    base = [int(o) for o in offs]
    rg = ae[ao[0]]
    assert d[base[0] + rg["value_offset"]: base[0] + rg["value_offset"] + rg["count"]] == b"rg1"
    xa = ae[ao[2]]
    assert chr(xa["type"]) == "A" and d[base[2] + xa["value_offset"]] == ord("c")
    xi = ae[ao[2] + 1]
    assert struct.unpack_from("<b", d, base[2] + xi["value_offset"])[0] == -5
    xf = ae[ao[3]]
    assert chr(xf["type"]) == "f" and struct.unpack_from("<f", d, base[3] + xf["value_offset"])[0] == 1.5
    xb = ae[ao[7]]
    assert chr(xb["type"]) == "B" and chr(xb["subtype"]) == "i" and xb["count"] == 3
    assert struct.unpack_from("<3i", d, base[7] + xb["value_offset"]) == (1, 2, 3)


def test_ont_cigar_op_counts(oracle_mod):
    d, offs = records(oracle_mod, "ont_ultralong.bam")
    rc, soa, _ = oracle_mod.bam_soa(d, offs)
    assert rc == 0
    assert np.diff(soa["cigar_offsets"]).tolist() == [2251, 967, 1037, 267]
    assert np.diff(soa["sequence_offsets"]).tolist() == [(n + 1) // 2 for n in (162932, 73028, 0, 5556)]


def aux_b(tag, sub, fmt, values):
    return tag + b"B" + sub + struct.pack("<I", len(values)) + struct.pack("<" + fmt * len(values), *values)


def test_long_cigar_placeholder_is_resolved_from_cg(oracle_mod):
    real = [(10 << 4) | 0, (2 << 4) | 1, (8 << 4) | 0]
    ph = [(20 << 4) | 4, (18 << 4) | 3]  # 20S 18N: the placeholder for a 20-base read
    rec = build_record(b"long", b"A" * 20, cigar=ph, aux=b"NMC\x01" + aux_b(b"CG", b"I", "I", real))
    rc, soa, _ = oracle_mod.bam_soa(rec, np.array([0], np.uint64))
    assert rc == 0 and soa["cigar_ops"].tolist() == real
    # the placeholder without CG is corrupt
    rec = build_record(b"long", b"A" * 20, cigar=ph, aux=b"NMC\x01")
    rc, _, bad = oracle_mod.bam_soa(rec, np.array([0], np.uint64))
    assert rc == oracle_mod.E_MALFORMED and bad == 0
    # CG of another type is corrupt too
    rec = build_record(b"long", b"A" * 20, cigar=ph, aux=aux_b(b"CG", b"i", "i", [1, 2]))
    rc, _, _ = oracle_mod.bam_soa(rec, np.array([0], np.uint64))
    assert rc == oracle_mod.E_MALFORMED
    # an ordinary 2-op CIGAR whose soft clip is not the read length passes through
    ordinary = [(19 << 4) | 4, (18 << 4) | 3]
    rec = build_record(b"x", b"A" * 20, cigar=ordinary, aux=b"")
    rc, soa, _ = oracle_mod.bam_soa(rec, np.array([0], np.uint64))
    assert rc == 0 and soa["cigar_ops"].tolist() == ordinary


@pytest.mark.parametrize("aux,ok", [
    (b"XZZabc\x00", True),
    (b"XHH0aF9\x00", True),
    (b"XHH0aF\x00", False),            # odd-length hex
    (b"XHHzz\x00", False),             # not hex
    (b"XZZabc", False),                # unterminated string
    (b"XBBq\x01\x00\x00\x00\x05", False),  # unknown B subtype
    (b"XBBS\x03\x00\x00\x00\x01\x00", False),  # array shorter than its count
    (b"XQq\x01", False),               # unknown type
    (b"XI", False),                    # truncated field head
])
def test_aux_rules(oracle_mod, aux, ok):
    rec = build_record(b"r", b"ACGT", aux=aux)
    rc, soa, _ = oracle_mod.bam_soa(rec, np.array([0], np.uint64))
    assert (rc == 0) == ok
    if ok:
        e = soa["aux_entries"][0]
        assert bytes(e["tag"]) == aux[:2] and e["size"] == len(aux) - 3
