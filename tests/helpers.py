"""Shared test helpers: BGZF framing with zlib, BAM record building, a DEFLATE
bit writer for hand-made corrupt streams."""
import struct
import zlib

import numpy as np

EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def deflate_raw(data: bytes, level=6, strategy=zlib.Z_DEFAULT_STRATEGY, flushes=()):
    """Raw DEFLATE; `flushes` are byte positions after which a Z_FULL_FLUSH
    is emitted (an empty stored block), so streams can hold several blocks."""
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 8, strategy)
    out = []
    prev = 0
    for f in sorted(flushes):
        out.append(c.compress(data[prev:f]))
        out.append(c.flush(zlib.Z_FULL_FLUSH))
        prev = f
    out.append(c.compress(data[prev:]))
    out.append(c.flush(zlib.Z_FINISH))
    return b"".join(out)


def bgzf_member(payload: bytes, cdata: bytes = None, level=6, crc=None, isize=None) -> bytes:
    if cdata is None:
        cdata = deflate_raw(payload, level)
        if len(cdata) > 65495 + 15:  # noodles re-encodes at level 0 (deflate.rs:58-95)
            cdata = deflate_raw(payload, 0)
    bsize = 18 + len(cdata) + 8
    assert bsize <= 65536
    hdr = bytes([0x1F, 0x8B, 8, 4, 0, 0, 0, 0, 0, 0xFF]) + struct.pack("<HBBHH", 6, 66, 67, 2, bsize - 1)
    c = zlib.crc32(payload) if crc is None else crc
    n = len(payload) if isize is None else isize
    return hdr + cdata + struct.pack("<II", c & 0xFFFFFFFF, n)


def bgzf(data: bytes, cap=65495, level=6, eof=True) -> bytes:
    out = [bgzf_member(data[i:i + cap], level=level) for i in range(0, len(data), cap)]
    if eof:
        out.append(EOF_BLOCK)
    return b"".join(out)


BASES = b"=ACMGRSVTWYHKDBN"


def build_record(name: bytes, seq: bytes, quals=None, ref_id=0, pos=100, flag=99, mapq=60,
                 cigar=None, aux=b"", next_ref=0, next_pos=200, tlen=150) -> bytes:
    """A spec-shaped BAM record, block_size prefix included."""
    packed = bytearray((len(seq) + 1) // 2)
    for i, b in enumerate(seq):
        nib = BASES.index(bytes([b]))
        packed[i // 2] |= (nib << 4) if i % 2 == 0 else nib
    if cigar is None:
        cigar = [(len(seq) << 4) | 0] if seq else []
    q = bytes([0xFF] * len(seq)) if quals is None else bytes(quals)
    nm = name + b"\0"
    body = struct.pack("<iiBBHHHIiii", ref_id, pos, len(nm), mapq, 4680, len(cigar), flag, len(seq),
                       next_ref, next_pos, tlen)
    body += nm + b"".join(struct.pack("<I", c) for c in cigar) + bytes(packed) + q + aux
    return struct.pack("<I", len(body)) + body


def bam_header(refs=(("chr1", 1000),), text=b"@HD\tVN:1.6\n") -> bytes:
    out = b"BAM\1" + struct.pack("<i", len(text)) + text + struct.pack("<i", len(refs))
    for name, ln in refs:
        n = name.encode() + b"\0"
        out += struct.pack("<i", len(n)) + n + struct.pack("<i", ln)
    return out


def random_records(rng, n, min_len=0, max_len=300, name_prefix=b"read"):
    out = []
    for i in range(n):
        L = int(rng.integers(min_len, max_len + 1))
        seq = bytes(rng.choice(list(b"ACGT"), size=L).tolist()) if L else b""
        quals = rng.integers(2, 42, size=L, dtype=np.uint8).tobytes()
        aux = b"NMC" + bytes([int(rng.integers(0, 4))]) + b"RGZrg1\0"
        out.append(build_record(name_prefix + b"%06d" % i, seq, quals, pos=int(i * 10), aux=aux,
                                flag=int(rng.choice([0, 16]))))
    return out


class BitWriter:
    """LSB-first DEFLATE bit writer; Huffman codes are written MSB-first."""

    def __init__(self):
        self.bits = []

    def put(self, value: int, n: int):
        for k in range(n):
            self.bits.append((value >> k) & 1)

    def put_code(self, code: int, n: int):
        for k in range(n - 1, -1, -1):
            self.bits.append((code >> k) & 1)

    def bytes(self) -> bytes:
        out = bytearray((len(self.bits) + 7) // 8)
        for i, b in enumerate(self.bits):
            out[i // 8] |= b << (i % 8)
        return bytes(out)


def fixed_code(sym: int):
    """(code, length) of a literal/length symbol in the fixed Huffman code."""
    if sym < 144:
        return 0x30 + sym, 8
    if sym < 256:
        return 0x190 + sym - 144, 9
    if sym < 280:
        return sym - 256, 7
    return 0xC0 + sym - 280, 8


def corrupt_streams():
    """(name, payload, cap, expected status) for each device status code, built
    from RFC 1951 by hand."""
    cases = []
    good = deflate_raw(b"hello hello hello hello world" * 20, 6)
    cases.append(("truncated", good[: len(good) // 2], 1000, 1))
    w = BitWriter(); w.put(1, 1); w.put(3, 2)
    cases.append(("reserved block type", w.bytes() + b"\0\0", 100, 2))
    cases.append(("stored LEN/NLEN mismatch", bytes([0x01, 5, 0, 0, 0]) + b"abcde", 100, 3))
    cases.append(("output overflow", deflate_raw(b"x" * 500, 6), 100, 4))
    cases.append(("stored overflow", bytes([0x01, 5, 0, 0xFA, 0xFF]) + b"abcde", 3, 4))
    w = BitWriter(); w.put(1, 1); w.put(1, 2); w.put_code(*fixed_code(286)); w.put(0, 16)
    cases.append(("length code 286", w.bytes(), 100, 5))
    w = BitWriter(); w.put(1, 1); w.put(1, 2); w.put_code(*fixed_code(ord("a")))
    w.put_code(*fixed_code(257)); w.put_code(30, 5); w.put(0, 16)
    cases.append(("distance code 30", w.bytes(), 100, 6))
    w = BitWriter(); w.put(1, 1); w.put(1, 2); w.put_code(*fixed_code(257)); w.put_code(0, 5); w.put(0, 16)
    cases.append(("distance before output", w.bytes(), 100, 6))
    # dynamic: HLIT=0 (257), HDIST=0 (1), HCLEN=0 (4 code-length codes: 16,17,18,0)
    # code-length code lengths 1,1,0,0 -> complete (two 1-bit codes: 16 and 17)
    w = BitWriter(); w.put(1, 1); w.put(2, 2); w.put(0, 5); w.put(0, 5); w.put(0, 4)
    for L in (1, 1, 0, 0):
        w.put(L, 3)
    w.put_code(0, 1)  # symbol 16 first: repeat with no previous length
    w.put(0, 2); w.put(0, 32)
    cases.append(("repeat without a previous length", w.bytes(), 100, 8))
    w = BitWriter(); w.put(1, 1); w.put(2, 2); w.put(30, 5); w.put(0, 5); w.put(0, 4); w.put(0, 32)
    cases.append(("HLIT beyond 286", w.bytes(), 100, 9))
    # code-length code over-subscribed: three 1-bit codes
    w = BitWriter(); w.put(1, 1); w.put(2, 2); w.put(0, 5); w.put(0, 5); w.put(0, 4)
    for L in (1, 1, 1, 0):
        w.put(L, 3)
    w.put(0, 32)
    cases.append(("over-subscribed code-length code", w.bytes(), 100, 7))
    # all literal/length lengths zero via code 18 repeats -> lengths[256] == 0
    # code-length code: symbols 18 (len 1) and 0 (len 1) -> order 16,17,18,0 lengths 0,0,1,1
    w = BitWriter(); w.put(1, 1); w.put(2, 2); w.put(0, 5); w.put(0, 5); w.put(0, 4)
    for L in (0, 0, 1, 1):
        w.put(L, 3)
    # canonical: symbol 0 -> code 0, symbol 18 -> code 1
    w.put_code(1, 1); w.put(138 - 11, 7)   # 138 zeros
    w.put_code(1, 1); w.put(119 - 11, 7)   # 119 zeros -> 257
    w.put_code(0, 1)                        # distance length 0
    w.put(0, 32)
    cases.append(("no end-of-block code", w.bytes(), 100, 7))
    return cases
