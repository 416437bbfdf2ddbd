"""The BCF oracle (oracle/oracle.c or_bcf_*) against what pins it.

- fixture known answers of crates/fritillaria-bcf/tests/bcftools.rs: 715 and 275
  records ending exactly at the end of the file, 2,504 samples (first HG00096)
  and 86 contigs, 1 sample (HG002) and 195 contigs; 56 and 3 interior member
  starts, none of them a record start;
- every record's typed values walk exactly to its declared l_shared / l_indiv
  (Record::validate), which a misread descriptor cannot survive;
- the typed-value examples of the BCF spec (crates/fritillaria-bcf/src/columnar/typed.rs
  tests) inside hand-built records, with bounds fixed by construction;
- the scan's error and tail cases (record.rs scan_records).
"""
import gzip
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN


def inflate(name):
    with open(os.path.join(GOLDEN, name), "rb") as f:
        return gzip.decompress(f.read())


@pytest.mark.parametrize("name,records,samples,first,contigs,interior", [
    ("kg_phase3.bcf", 715, 2504, b"HG00096", 86, 56),
    ("giab_hg002.bcf", 275, 1, b"HG002", 195, 3),
])
def test_fixture_known_answers(oracle_mod, name, records, samples, first, contigs, interior):
    d = inflate(name)
    rs, names, n_contigs = oracle_mod.bcf_parse_header(d)
    assert len(names) == samples and names[0] == first and n_contigs == contigs
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(d, rs)
    assert rc == 0 and len(offs) == records and tail == len(d)
    rc, cols, _ = oracle_mod.bcf_decode_columns(d, offs, tail)
    assert rc == 0
    assert np.all(cols["sample_count"] == samples)
    assert all(oracle_mod.bcf_validate(d, int(o)) for o in offs)
    ends = offs + cols["record_end"]
    assert np.array_equal(ends[:-1], offs[1:]) and ends[-1] == len(d)
    # member structure: interior member starts are never record starts
    raw = open(os.path.join(GOLDEN, name), "rb").read()
    spans = oracle_mod.discover(raw).spans
    starts = np.concatenate([[0], np.cumsum(spans["isize"].astype(np.int64))[:-1]])
    inner = [s for s in starts if rs < s < len(d)]
    assert len(inner) == interior
    assert not set(inner) & set(offs.tolist())
    # every true record passes the sieve
    assert all(oracle_mod.bcf_looks_like(d, int(o), samples, contigs) for o in offs)


def typed_str(b: bytes) -> bytes:
    if len(b) < 15:
        return bytes([len(b) << 4 | 7]) + b
    return bytes([0xF7, 0x11, len(b)]) + b


def record(chrom, pos, alleles, samples, id_=b"", filters=b"\x00", info=b"", n_info=0, indiv=b"", n_fmt=0):
    shared = struct.pack("<iiiI", chrom, pos, 1, 0x7F800001)
    shared += struct.pack("<HH", n_info, len(alleles)) + struct.pack("<I", samples)[:3] + bytes([n_fmt])
    shared += typed_str(id_) if id_ else b"\x07"
    for a in alleles:
        shared += typed_str(a)
    shared += filters + info
    return struct.pack("<II", len(shared), len(indiv)) + shared + indiv, shared


def test_bounds_of_hand_built_records(oracle_mod):
    # ID "VariantCallFormatSampleText" escaped (0xF7 0x11 0x1B), alleles ACAC / A,
    # FILTER PASS as an int8 vector, one INFO field, a GT FORMAT field for 2 samples
    info = b"\x11\x01" + b"\x13" + struct.pack("<i", 7)
    indiv = b"\x11\x02" + b"\x21" + bytes([2, 4, 2, 3])
    rec, shared = record(1, 999, [b"ACAC", b"A"], 2, id_=b"VariantCallFormatSampleText",
                         filters=b"\x11\x00", info=info, n_info=1, indiv=indiv, n_fmt=1)
    buf = b"\xAA" * 9 + rec + rec
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(buf, 9)
    assert rc == 0 and list(offs) == [9, 9 + len(rec)] and tail == len(buf)
    rc, cols, _ = oracle_mod.bcf_decode_columns(buf, offs, tail)
    assert rc == 0
    a0 = 32 + 3 + 27            # site core, then the escaped 27-byte ID
    f0 = a0 + 5 + 2             # 0x47 ACAC, 0x17 A
    i0 = f0 + 2                 # 0x11 0x00
    assert cols["alleles_start"][0] == a0 and cols["filters_start"][0] == f0 and cols["info_start"][0] == i0
    assert cols["genotypes_start"][0] == 8 + len(shared) and cols["record_end"][0] == len(rec)
    assert cols["chromosome_id"][0] == 1 and cols["position"][0] == 999 and cols["quality"][0] == 0x7F800001
    assert cols["allele_count"][0] == 2 and cols["sample_count"][0] == 2 and cols["format_count"][0] == 1
    assert oracle_mod.bcf_validate(buf, 9) and oracle_mod.bcf_looks_like(buf, 9, 2, 4)
    assert not oracle_mod.bcf_looks_like(buf, 9, 3, 4) and not oracle_mod.bcf_looks_like(buf, 9, 2, 1)


def test_escaped_vector_length_must_be_at_least_15(oracle_mod):
    # an escaped count below 15 is malformed (typed.rs read_descriptor)
    bad_id = bytes([0xF7, 0x11, 3]) + b"abc"
    shared = struct.pack("<iiiI", 0, 5, 1, 0) + struct.pack("<HH", 0, 1) + b"\x01\x00\x00\x00"
    shared += bad_id + b"\x17A" + b"\x00"
    rec = struct.pack("<II", len(shared), 0) + shared
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(rec, 0)
    assert rc == 0 and len(offs) == 1
    rc, _, bad = oracle_mod.bcf_decode_columns(rec, offs, tail)
    assert rc == oracle_mod.E_MALFORMED and bad == 0
    assert not oracle_mod.bcf_validate(rec, 0)


def test_scan_tail_and_errors(oracle_mod):
    rec, _ = record(0, 10, [b"A", b"C"], 1)
    buf = rec * 3
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(buf[:-5], 0)
    assert rc == 0 and len(offs) == 2 and tail == 2 * len(rec)      # partial record: the tail
    rc, offs, tail, _ = oracle_mod.bcf_scan_records(buf + b"\x01\x02\x03", 0)
    assert rc == 0 and len(offs) == 3 and tail == len(buf)          # fewer than 8 bytes left
    bad = buf[: len(rec)] + struct.pack("<II", 23, 0) + b"\x00" * 40
    rc, offs, tail, err = oracle_mod.bcf_scan_records(bad, 0)
    assert rc == oracle_mod.E_MALFORMED and err == len(rec)        # l_shared below the site core
