"""BGZF framing and round trips of compressed output -- TEST INFRASTRUCTURE ONLY.

Compression has no byte-identity oracle: two valid DEFLATE streams of the same
input legitimately differ (crates/fritillaria-core/src/compress.rs:1-11). What
the checks below fix instead, each restating the reference:
  choose_framing, framed_size   crates/fritillaria-core/src/compress.rs:115-142
                                (an empty payload is the canonical empty block;
                                a stream above 65510 bytes, or above the payload
                                plus 5, is stored; a tie keeps the stream)
  frame                         crates/fritillaria-cuda/kernels/bgzf_frame.cu:34-144
                                (gzip header with FEXTRA, XLEN 6, 'BC' SLEN 2,
                                BC = block size - 1, MTIME 0, OS 255; CRC32 and
                                ISIZE of the payload)
  EOF_BLOCK                     crates/fritillaria-bgzf/src/block.rs (the 28-byte
                                marker htslib checks)
  members                       every member of a stream decoded by zlib
                                (gzip wrapper: zlib verifies CRC32 and ISIZE),
                                its BC checked against its length

Pinned by the reference's own vectors (compress.rs tests: 65505 frames to
65536; 148,256-byte streams and 1006 > 1000 + 5 are stored; 1005 and 400 are
kept; an empty payload is 28 bytes whatever is claimed) and by zlib, which
decodes every block independently of this code.
"""
from __future__ import annotations

import struct
import zlib

BGZF_HEADER_SIZE = 18
BGZF_TRAILER_SIZE = 8
STORED_BLOCK_HEADER = 5
MAX_BLOCK_SIZE = 65536
MAX_DEFLATE_STREAM = MAX_BLOCK_SIZE - BGZF_HEADER_SIZE - BGZF_TRAILER_SIZE
MAX_COMPRESSIBLE_PAYLOAD = MAX_DEFLATE_STREAM - STORED_BLOCK_HEADER
EMPTY_DEFLATE_STREAM = b"\x03\x00"
EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
DEFLATED, STORED, EMPTY = 0, 1, 2


def choose_framing(payload_len: int, deflate_len: int) -> int:
    if payload_len == 0:
        return EMPTY
    if deflate_len > MAX_DEFLATE_STREAM or deflate_len > payload_len + STORED_BLOCK_HEADER:
        return STORED
    return DEFLATED


def framed_size(payload_len: int, deflate_len: int) -> int:
    f = choose_framing(payload_len, deflate_len)
    body = deflate_len if f == DEFLATED else payload_len + STORED_BLOCK_HEADER if f == STORED else 2
    return BGZF_HEADER_SIZE + body + BGZF_TRAILER_SIZE


def header(block_size: int) -> bytes:
    return bytes([0x1F, 0x8B, 8, 4, 0, 0, 0, 0, 0, 0xFF]) + struct.pack("<HBBHH", 6, ord("B"), ord("C"), 2,
                                                                         block_size - 1)


def frame(payload: bytes, body: bytes) -> bytes:
    size = BGZF_HEADER_SIZE + len(body) + BGZF_TRAILER_SIZE
    return header(size) + body + struct.pack("<II", zlib.crc32(payload) & 0xFFFFFFFF, len(payload))


def stored_block(payload: bytes) -> bytes:
    """The exact framed bytes of a stored block (BFINAL 1, BTYPE 00, LEN, NLEN)."""
    n = len(payload)
    return frame(payload, bytes([1]) + struct.pack("<HH", n, n ^ 0xFFFF) + payload)


def empty_block() -> bytes:
    return frame(b"", EMPTY_DEFLATE_STREAM)


def members(stream: bytes):
    """(offset, block size, payload, body first byte) of each member, decoded
    by zlib's gzip reader (which checks CRC32 and ISIZE); BC must name the
    member's length."""
    out = []
    p = 0
    while p < len(stream):
        if len(stream) - p < BGZF_HEADER_SIZE:
            raise ValueError(f"truncated member at {p}")
        hdr = stream[p:p + BGZF_HEADER_SIZE]
        if hdr[:4] != b"\x1f\x8b\x08\x04" or hdr[10:16] != b"\x06\x00BC\x02\x00":
            raise ValueError(f"not a BGZF header at {p}")
        size = struct.unpack_from("<H", hdr, 16)[0] + 1
        block = stream[p:p + size]
        d = zlib.decompressobj(31)
        payload = d.decompress(block)
        if not d.eof or d.unused_data:
            raise ValueError(f"member at {p} does not end at its BC")
        crc, isize = struct.unpack_from("<II", block, size - 8)
        if crc != zlib.crc32(payload) & 0xFFFFFFFF or isize != len(payload):
            raise ValueError(f"trailer of the member at {p}")
        out.append((p, size, payload, block[BGZF_HEADER_SIZE]))
        p += size
    return out
