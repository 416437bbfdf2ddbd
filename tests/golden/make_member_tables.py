"""Writes tests/golden/members.json: per-member (compressed offset, payload
length, ISIZE, CRC32) of every BGZF fixture, plus the inflated length and a
CRC of the whole inflated stream.

Derived with Python's struct + zlib only (independent of the C oracle): the
trailer fields are the writer's own (htslib/bcftools), and each member is
inflated with zlib (windowBits -15) and its CRC/ISIZE checked against them.
Run from the repository root: python tests/golden/make_member_tables.py
"""
import json
import os
import struct
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = ["htslib.bam", "htslib_multiblock.bam", "ont_ultralong.bam", "pacbio_hifi.bam",
            "kg_phase3.bcf", "giab_hg002.bcf", "illumina.fastq.gz", "ont_ultralong.fastq.gz",
            "pacbio_hifi.fastq.gz"]


def members(buf):
    pos = 0
    out = []
    while pos < len(buf):
        xlen = struct.unpack_from("<H", buf, pos + 10)[0]
        extra = buf[pos + 12:pos + 12 + xlen]
        bsize = None
        i = 0
        while i + 4 <= len(extra):
            si1, si2, slen = extra[i], extra[i + 1], struct.unpack_from("<H", extra, i + 2)[0]
            if si1 == 66 and si2 == 67 and slen == 2:
                bsize = struct.unpack_from("<H", extra, i + 4)[0] + 1
            i += 4 + slen
        crc, isize = struct.unpack_from("<II", buf, pos + bsize - 8)
        payload = buf[pos + 12 + xlen:pos + bsize - 8]
        data = zlib.decompress(payload, -15)
        assert len(data) == isize and zlib.crc32(data) == crc
        out.append({"offset": pos, "payload_len": len(payload), "isize": isize, "crc32": crc, "data": data})
        pos += bsize
    return out


def main():
    table = {}
    for name in FIXTURES:
        buf = open(os.path.join(HERE, name), "rb").read()
        ms = members(buf)
        whole = b"".join(m["data"] for m in ms)
        table[name] = {
            "members": [[m["offset"], m["payload_len"], m["isize"], m["crc32"]] for m in ms],
            "inflated_len": len(whole),
            "inflated_crc32": zlib.crc32(whole),
        }
    with open(os.path.join(HERE, "members.json"), "w") as f:
        json.dump(table, f, indent=1)


if __name__ == "__main__":
    main()
