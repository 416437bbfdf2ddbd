"""CPU oracle for fritillaria's device path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed CPU
baseline), never as the thing measured or shipped. The product library
(``fritillaria_amd``) does not import it and has no CPU fallback.

The C sources (``oracle.c``, ``oracle_mt.c``) restate the reference algorithm;
each function there cites the reference file:line it follows and names what
pins it. This module is a thin ctypes/numpy wrapper.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

# status codes (crates/fritillaria-cuda/kernels/inflate.cu:29-39)
OK, TRUNCATED, BLOCK_TYPE, STORED_LEN, OUTPUT_OVERFLOW, BAD_CODE, BAD_DISTANCE, BAD_TABLE, \
    BAD_LENGTH_REPEAT, BAD_COUNTS = range(10)
# error kinds (crates/fritillaria-core/src/error.rs)
E_OK, E_INVALID_BLOCK, E_CHECKSUM_MISMATCH, E_SIZE_MISMATCH, E_INFLATE, E_MALFORMED, \
    E_INVALID_DEVICE_BATCH = range(7)


class Span(C.Structure):
    _fields_ = [("compressed_offset", C.c_uint64), ("payload_start", C.c_uint64),
                ("payload_len", C.c_uint64), ("isize", C.c_uint32), ("crc32", C.c_uint32)]


class Err(C.Structure):
    _fields_ = [("kind", C.c_int32), ("code", C.c_uint32), ("offset", C.c_uint64),
                ("expected", C.c_uint64), ("actual", C.c_uint64)]


SPAN_DTYPE = np.dtype([("compressed_offset", "<u8"), ("payload_start", "<u8"),
                       ("payload_len", "<u8"), ("isize", "<u4"), ("crc32", "<u4")])


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        u8p = C.POINTER(C.c_uint8)
        _lib.or_crc32.restype = C.c_uint32
        _lib.or_crc32.argtypes = [C.c_void_p, C.c_uint64]
        _lib.or_inflate_member.restype = C.c_int
        _lib.or_inflate_member.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                           C.POINTER(C.c_uint32)]
        _lib.or_discover.restype = C.c_uint64
        _lib.or_discover.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64,
                                     C.POINTER(C.c_uint64), C.POINTER(C.c_int),
                                     C.POINTER(C.c_uint64)]
        _lib.or_inflate_batch.restype = C.c_int
        _lib.or_inflate_batch.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                          C.c_void_p, C.c_void_p, C.POINTER(Err)]
        _lib.or_bam_parse_header.restype = C.c_int
        _lib.or_bam_parse_header.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
        _lib.or_bam_scan_records.restype = C.c_int
        _lib.or_bam_scan_records.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                             C.c_uint64, C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint64)]
        _lib.or_bam_scan_blocked.restype = C.c_int
        _lib.or_bam_scan_blocked.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                             C.c_uint64, C.c_void_p, C.c_uint64,
                                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                             C.POINTER(C.c_uint64)]
        _lib.or_bam_decode_columns.restype = None
        _lib.or_bam_decode_columns.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64] + [C.c_void_p] * 14
        _lib.or_virtual_offset.restype = C.c_uint64
        _lib.or_virtual_offset.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64]
        _lib.or_fastq_bounds.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]
        _lib.or_fastq_scan_records.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_void_p,
                                               C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_uint64)]
        _lib.or_fastq_survivors.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        _lib.or_fastq_survivors.restype = C.c_uint64
        _lib.or_bcf_looks_like.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        _lib.or_bcf_scan_records.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint64,
                                             C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        _lib.or_bcf_decode_columns.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64] + [C.c_void_p] * 14 + \
            [C.POINTER(C.c_uint64)]
        _lib.or_bcf_validate.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        _lib.or_bam_soa.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64] + [C.c_void_p] * 9 + [C.POINTER(C.c_uint64)]
        _lib.or_mt_bam_columns.restype = C.c_int
        _lib.or_mt_bam_columns.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_int, C.c_void_p]
        _lib.or_mt_inflate.restype = C.c_int
        _lib.or_mt_inflate.argtypes = [C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.POINTER(C.c_uint64)]
        del u8p
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _as_u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def crc32(data) -> int:
    a = _as_u8(data)
    return int(lib().or_crc32(_ptr(a), a.size))


def inflate_member(payload, cap: int):
    """Returns (status, bytes produced) for one raw DEFLATE stream."""
    a = _as_u8(payload)
    out = np.zeros(max(cap, 1), dtype=np.uint8)
    produced = C.c_uint32(0)
    st = lib().or_inflate_member(_ptr(a), a.size, _ptr(out), cap, C.byref(produced))
    return int(st), out[:produced.value].tobytes()


@dataclass
class Discovery:
    spans: np.ndarray  # SPAN_DTYPE
    position: int
    err: int
    err_offset: int


def discover(data, base: int = 0) -> Discovery:
    a = _as_u8(data)
    cap = a.size // 26 + 2
    spans = np.zeros(cap, dtype=SPAN_DTYPE)
    pos = C.c_uint64(0)
    err = C.c_int(0)
    eoff = C.c_uint64(0)
    n = lib().or_discover(_ptr(a), a.size, base, _ptr(spans), cap, C.byref(pos), C.byref(err),
                          C.byref(eoff))
    return Discovery(spans[:n].copy(), pos.value, err.value, eoff.value)


@dataclass
class InflateResult:
    kind: int
    code: int
    offset: int
    expected: int
    actual: int
    data: bytes
    offsets: np.ndarray


def inflate_batch(batch, spans: np.ndarray) -> InflateResult:
    a = _as_u8(batch)
    spans = np.ascontiguousarray(spans, dtype=SPAN_DTYPE)
    total = int(spans["isize"].astype(np.uint64).sum()) if spans.size else 0
    out = np.zeros(max(total, 1), dtype=np.uint8)
    offs = np.zeros(spans.size + 1, dtype=np.uint64)
    e = Err()
    k = lib().or_inflate_batch(_ptr(a), a.size, _ptr(spans), spans.size, _ptr(out), _ptr(offs),
                               C.byref(e))
    if k:
        return InflateResult(k, e.code, e.offset, e.expected, e.actual, b"", np.zeros(1, np.uint64))
    return InflateResult(0, 0, 0, 0, 0, out[:total].tobytes(), offs)


def bam_parse_header(buf):
    a = _as_u8(buf)
    rs = C.c_uint64(0)
    ep = C.c_uint64(0)
    nref = C.c_uint32(0)
    rc = lib().or_bam_parse_header(_ptr(a), a.size, C.byref(rs), C.byref(ep), C.byref(nref))
    return rc, rs.value, ep.value, nref.value


def bam_scan_records(buf, start: int):
    a = _as_u8(buf)
    cap = a.size // 36 + 1
    offs = np.zeros(cap, dtype=np.uint64)
    n = C.c_uint64(0)
    tail = C.c_uint64(0)
    rc = lib().or_bam_scan_records(_ptr(a), a.size, start, _ptr(offs), cap, C.byref(n),
                                   C.byref(tail))
    return rc, offs[:n.value].copy(), tail.value


FASTQ_BOUNDS = ("sequence_start", "plus_start", "quality_start", "sequence_len", "record_end")


def fastq_bounds(buf, pos: int):
    """Line bounds of a record at pos (crates/fritillaria-fastq/src/columnar/record.rs:85-131),
    as a 5-tuple in FASTQ_BOUNDS order, or None."""
    a = _as_u8(buf)
    b = np.zeros(5, dtype=np.uint32)
    if not lib().or_fastq_bounds(_ptr(a), a.size, pos, _ptr(b)):
        return None
    return tuple(int(x) for x in b)


def fastq_scan_records(buf, start: int):
    """Serial walk of record.rs:145-176. Returns (rc, offsets u64[n], bounds u32[n, 5], tail, err_pos);
    rc is 0 or E_MALFORMED (a non-'@', non-newline byte where a record must start)."""
    a = _as_u8(buf)
    cap = max(1, a.size // 6 + 1)
    offs = np.zeros(cap, dtype=np.uint64)
    bounds = np.zeros((cap, 5), dtype=np.uint32)
    n, tail, err = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
    rc = lib().or_fastq_scan_records(_ptr(a), a.size, start, _ptr(offs), _ptr(bounds), cap, C.byref(n),
                                     C.byref(tail), C.byref(err))
    k = n.value
    return rc, offs[:k].copy(), bounds[:k].copy(), tail.value, err.value


def fastq_survivors(buf, start: int) -> int:
    a = _as_u8(buf)
    return int(lib().or_fastq_survivors(_ptr(a), a.size, start))


def bam_scan_blocked(buf, block_starts, start: int):
    """Returns (rc, offsets, tail, fallbacks); rc 1 = malformed at `tail`."""
    a = _as_u8(buf)
    bs = np.ascontiguousarray(block_starts, dtype=np.uint64)
    cap = a.size // 36 + 1
    offs = np.zeros(cap, dtype=np.uint64)
    n = C.c_uint64(0)
    tail = C.c_uint64(0)
    fb = C.c_uint64(0)
    rc = lib().or_bam_scan_blocked(_ptr(a), a.size, _ptr(bs), bs.size - 1, start, _ptr(offs), cap,
                                   C.byref(n), C.byref(tail), C.byref(fb))
    return rc, offs[:n.value].copy(), tail.value, fb.value


# Column order and widths of crates/fritillaria-bam/src/columnar/device.rs:95-110.
BAM_COLUMNS = [
    ("reference_sequence_id", np.int32), ("position", np.int32), ("flags", np.uint16),
    ("mapping_quality", np.uint8), ("sequence_len", np.uint32),
    ("mate_reference_sequence_id", np.int32), ("mate_position", np.int32),
    ("template_length", np.int32), ("record_offsets", np.uint64), ("cigar_start", np.uint32),
    ("sequence_start", np.uint32), ("quality_start", np.uint32), ("aux_start", np.uint32),
    ("record_end", np.uint32),
]


def bam_decode_columns(buf, offsets) -> dict:
    a = _as_u8(buf)
    offs = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = offs.size
    cols = {name: np.zeros(n, dtype=dt) for name, dt in BAM_COLUMNS}
    lib().or_bam_decode_columns(_ptr(a), _ptr(offs), n, *[_ptr(cols[name]) for name, _ in BAM_COLUMNS])
    return cols


def virtual_offsets(block_offsets, coffsets, positions) -> np.ndarray:
    bo = np.ascontiguousarray(block_offsets, dtype=np.uint64)
    co = np.ascontiguousarray(coffsets, dtype=np.uint64)
    out = np.zeros(len(positions), dtype=np.uint64)
    f = lib().or_virtual_offset
    for i, p in enumerate(positions):
        out[i] = f(_ptr(bo), _ptr(co), co.size, int(p))
    return out


def mt_bam_columns(file_bytes, nthreads: int, soa: bool = False):
    """The noodles-structured multithreaded reader over a whole BGZF BAM: record
    count and column checksums (sums mod 2^64); with soa, also the totals and
    checksums of the resolved CIGAR ops, packed sequences and aux entries."""
    a = _as_u8(file_bytes)
    res = np.zeros(13, dtype=np.uint64)
    rc = lib().or_mt_bam_columns(_ptr(a), a.size, nthreads, 1 if soa else 0, _ptr(res))
    out = {"records": int(res[0]), "inflated": int(res[1]), "sum_pos": int(res[2]),
           "sum_flag": int(res[3]), "sum_lseq": int(res[4]), "sum_off": int(res[5]),
           "members": int(res[6])}
    if soa:
        out.update({"ops": int(res[7]), "sum_ops": int(res[8]), "seq_bytes": int(res[9]),
                    "sum_seq": int(res[10]), "aux": int(res[11]), "sum_aux_value_offset": int(res[12])})
    return rc, out


def mt_inflate(file_bytes, nthreads: int):
    """Whole-stream BGZF inflate with zlib on `nthreads` workers (or_mt_inflate):
    (rc, inflated bytes)."""
    a = _as_u8(file_bytes)
    n = C.c_uint64(0)
    rc = lib().or_mt_inflate(_ptr(a), a.size, nthreads, None, C.byref(n))
    if rc != 0:
        return rc, None
    out = np.empty(max(n.value, 1), np.uint8)
    rc = lib().or_mt_inflate(_ptr(a), a.size, nthreads, _ptr(out), C.byref(n))
    return rc, out[: n.value]


# ---------------------------------------------------------------------------
# BCF
# ---------------------------------------------------------------------------

# (name, dtype) in the order of crates/fritillaria-bcf/src/columnar/device.rs:59-98
BCF_COLUMNS = [
    ("chromosome_id", np.int32), ("position", np.int32), ("reference_span", np.int32), ("quality", np.uint32),
    ("info_count", np.uint16), ("allele_count", np.uint16), ("sample_count", np.uint32),
    ("format_count", np.uint8), ("record_offsets", np.uint64), ("alleles_start", np.uint32),
    ("filters_start", np.uint32), ("info_start", np.uint32), ("genotypes_start", np.uint32),
    ("record_end", np.uint32),
]


def bcf_parse_header(buf):
    """records_start, sample names and contig count of a BCF header
    (crates/fritillaria-bcf/src/columnar/header.rs:43-140): magic BCF, major
    version 2, l_text, NUL-stripped text; contigs are ##contig lines with an
    ID, samples the #CHROM columns after FORMAT."""
    b = bytes(_as_u8(buf)[: 1 << 26])
    if len(b) < 9 or b[:3] != b"BCF" or b[3] != 2:
        raise ValueError("not a BCF v2 header")
    l_text = int.from_bytes(b[5:9], "little")
    text = b[9:9 + l_text].rstrip(b"\0")
    contigs, samples = 0, []
    for line in text.split(b"\n"):
        line = line.rstrip(b"\r")
        if line.startswith(b"#CHROM"):
            cols = line.split(b"\t")
            samples = cols[9:] if len(cols) > 9 else []
        elif line.startswith(b"##contig=<") and (b"<ID=" in line or b",ID=" in line):
            contigs += 1
    return 9 + l_text, samples, contigs


def bcf_scan_records(buf, start: int):
    a = _as_u8(buf)
    cap = a.size // 32 + 1
    offs = np.zeros(cap, dtype=np.uint64)
    n, tail, err = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
    rc = lib().or_bcf_scan_records(_ptr(a), a.size, start, _ptr(offs), cap, C.byref(n), C.byref(tail),
                                   C.byref(err))
    return rc, offs[:n.value].copy(), tail.value, err.value


def bcf_decode_columns(buf, offsets, tail: int):
    """Returns (rc, columns dict, index of the first failing record)."""
    a = _as_u8(buf)
    offs = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = offs.size
    cols = {name: np.zeros(max(n, 1), dtype=dt) for name, dt in BCF_COLUMNS}
    bad = C.c_uint64(0)
    rc = lib().or_bcf_decode_columns(_ptr(a), _ptr(offs), n, tail, *[_ptr(cols[k]) for k, _ in BCF_COLUMNS],
                                     C.byref(bad))
    return rc, {k: v[:n] for k, v in cols.items()}, bad.value


def bcf_looks_like(buf, pos: int, samples: int, contigs: int) -> bool:
    a = _as_u8(buf)
    return bool(lib().or_bcf_looks_like(_ptr(a), a.size, pos, samples, contigs))


def bcf_validate(buf, pos: int) -> bool:
    a = _as_u8(buf)
    return bool(lib().or_bcf_validate(_ptr(a), a.size, pos))


AUX_ENTRY_DTYPE = np.dtype([("tag", "S2"), ("type", "u1"), ("subtype", "u1"), ("value_offset", "<u4"),
                            ("count", "<u4"), ("size", "<u4")])


def bam_soa(buf, offsets):
    """Resolved CIGAR ops, packed sequences and the aux tag index of BAM records
    (or_bam_soa). Returns (rc, dict of arrays, first failing record)."""
    a = _as_u8(buf)
    offs = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = offs.size
    n_ops = np.zeros(max(n, 1), np.uint32)
    seq_bytes = np.zeros(max(n, 1), np.uint32)
    n_aux = np.zeros(max(n, 1), np.uint32)
    bad = C.c_uint64(0)
    f = lib().or_bam_soa
    rc = f(_ptr(a), _ptr(offs), n, _ptr(n_ops), _ptr(seq_bytes), _ptr(n_aux), None, None, None, None, None, None,
           C.byref(bad))
    if rc != 0:
        return rc, None, bad.value
    op_off = np.concatenate([[0], np.cumsum(n_ops[:n], dtype=np.uint64)]).astype(np.uint64)
    seq_off = np.concatenate([[0], np.cumsum(seq_bytes[:n], dtype=np.uint64)]).astype(np.uint64)
    aux_off = np.concatenate([[0], np.cumsum(n_aux[:n], dtype=np.uint64)]).astype(np.uint64)
    ops = np.zeros(max(int(op_off[-1]), 1), np.uint32)
    seq = np.zeros(max(int(seq_off[-1]), 1), np.uint8)
    aux = np.zeros(max(int(aux_off[-1]), 1), AUX_ENTRY_DTYPE)
    rc = f(_ptr(a), _ptr(offs), n, _ptr(n_ops), _ptr(seq_bytes), _ptr(n_aux), _ptr(op_off), _ptr(ops),
           _ptr(seq_off), _ptr(seq), _ptr(aux_off), _ptr(aux), C.byref(bad))
    return rc, {"cigar_offsets": op_off, "cigar_ops": ops[: int(op_off[-1])],
                "sequence_offsets": seq_off, "sequence_packed": seq[: int(seq_off[-1])],
                "aux_offsets": aux_off, "aux_entries": aux[: int(aux_off[-1])]}, bad.value
