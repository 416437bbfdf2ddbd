"""Synthetic BGZF/BAM inputs of the benchmark shapes (noodles framing).

Wraps ``csrc/synth.c`` (built into ``libfrsynth.so``). The generators are
deterministic for a given seed, whatever the thread count.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _build

_lib = None


class _BamParams(C.Structure):
    _fields_ = [("n_records", C.c_uint64), ("seed", C.c_uint64), ("read_mode", C.c_int32),
                ("len_fixed", C.c_uint32), ("len_min", C.c_uint32), ("len_max", C.c_uint32),
                ("ln_median", C.c_double), ("ln_sigma", C.c_double), ("len_cap", C.c_uint32),
                ("indel_p", C.c_double), ("aux_mode", C.c_int32), ("unaligned", C.c_int32),
                ("ref_len", C.c_uint32), ("threads", C.c_int32), ("cap_every", C.c_uint32)]


class _BgzfParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("level_mode", C.c_int32), ("level", C.c_int32),
                ("stored_frac", C.c_double), ("payload_cap", C.c_uint32), ("strategy", C.c_int32),
                ("eof_member", C.c_int32), ("threads", C.c_int32)]


def _l():
    global _lib
    if _lib is None:
        path = _build.synth_lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(path)
        _lib.synth_bam_stream.restype = C.c_int
        _lib.synth_bam_stream.argtypes = [C.POINTER(_BamParams), C.POINTER(C.c_void_p),
                                          C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        _lib.synth_bgzf.restype = C.c_int
        _lib.synth_bgzf.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(_BgzfParams),
                                    C.POINTER(C.c_void_p), C.POINTER(C.c_uint64)]
        _lib.synth_free.restype = None
        _lib.synth_free.argtypes = [C.c_void_p]
    return _lib


def _take(ptr: C.c_void_p, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint8)
    if n:
        C.memmove(out.ctypes.data, ptr.value, n)
    _l().synth_free(ptr)
    return out


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def bam_stream(n_records: int, seed: int, *, read_len=150, read_range=None, lognormal=None,
               indel_p=0.0, long_aux=False, unaligned=False, ref_len=1_000_000,
               threads=None, cap_every=0):
    """Uncompressed BAM stream (header + records). Returns (bytes array, records_start).
    With log-normal lengths, cap_every > 0 gives every cap_every-th record the
    length cap (tests of records that span many members)."""
    p = _BamParams()
    p.n_records = n_records
    p.seed = seed
    if lognormal is not None:
        p.read_mode = 2
        p.ln_median, p.ln_sigma, p.len_cap = lognormal
    elif read_range is not None:
        p.read_mode = 1
        p.len_min, p.len_max = read_range
    else:
        p.read_mode = 0
        p.len_fixed = read_len
    p.indel_p = indel_p
    p.aux_mode = 1 if long_aux else 0
    p.unaligned = 1 if unaligned else 0
    p.ref_len = ref_len
    p.threads = threads or default_threads()
    p.cap_every = cap_every
    ptr = C.c_void_p()
    n = C.c_uint64()
    rs = C.c_uint64()
    rc = _l().synth_bam_stream(C.byref(p), C.byref(ptr), C.byref(n), C.byref(rs))
    if rc != 0:
        raise RuntimeError(f"synth_bam_stream failed: {rc}")
    return _take(ptr, n.value), rs.value


def bgzf(data: np.ndarray, *, level=6, mixed_levels=False, stored_frac=0.0, seed=0,
         payload_cap=65495, eof=True, strategy=0, threads=None) -> np.ndarray:
    """Frames an uncompressed stream as BGZF members of `payload_cap` bytes each."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    b = _BgzfParams()
    b.seed = seed
    b.level_mode = 1 if mixed_levels else 0
    b.level = level
    b.stored_frac = stored_frac
    b.payload_cap = payload_cap
    b.strategy = strategy
    b.eof_member = 1 if eof else 0
    b.threads = threads or default_threads()
    ptr = C.c_void_p()
    n = C.c_uint64()
    rc = _l().synth_bgzf(data.ctypes.data_as(C.c_void_p), data.size, C.byref(b), C.byref(ptr),
                         C.byref(n))
    if rc != 0:
        raise RuntimeError(f"synth_bgzf failed: {rc}")
    return _take(ptr, n.value)


@dataclass
class SynthBam:
    compressed: np.ndarray
    inflated_len: int
    records_start: int
    n_records: int


CONFIG_SEED = 0x5EED0000


def config_bam(config: int, n_records=None, threads=None, cap_every=0) -> SynthBam:
    """BAM inputs of the benchmark configurations (1: short reads, 2: batched large,
    3: unaligned long reads); `n_records` scales a configuration down for tests,
    and for configuration 3 `cap_every` forces every cap_every-th read to the
    1 Mb cap. Configurations 1 and 2 share the generator: one 1 Mbp reference."""
    seed = CONFIG_SEED + config
    if config == 1:
        n = n_records or 10_000
        u, rs = bam_stream(n, seed, read_len=150, ref_len=1_000_000, threads=threads)
        c = bgzf(u, level=6, seed=seed, threads=threads)
    elif config == 2:
        n = n_records or 20_000_000
        u, rs = bam_stream(n, seed, read_range=(100, 250), indel_p=0.01, ref_len=1_000_000,
                           threads=threads)
        c = bgzf(u, mixed_levels=True, stored_frac=0.01, seed=seed, threads=threads)
    elif config == 3:
        n = n_records or 50_000
        u, rs = bam_stream(n, seed, lognormal=(20_000.0, 1.0, 1_000_000), long_aux=True,
                           unaligned=True, ref_len=1_000_000, threads=threads, cap_every=cap_every)
        c = bgzf(u, level=6, seed=seed, threads=threads)
    else:
        raise ValueError(f"no BAM configuration {config}")
    return SynthBam(c, int(u.size), rs, n)


@dataclass
class SynthFastq:
    compressed: np.ndarray
    inflated: np.ndarray
    n_records: int


def fastq_stream(n_records: int, seed: int, read_len: int = 150) -> np.ndarray:
    """Uncompressed FASTQ: names syn%010d, bases uniform ACGT, qualities
    round(N(35, 5)) clipped to [2, 41], Phred+33, an empty '+' line."""
    rng = np.random.default_rng(seed)
    name_len = 1 + 3 + 10 + 1
    width = name_len + (read_len + 1) + 2 + (read_len + 1)
    out = np.empty((n_records, width), dtype=np.uint8)
    out[:, 0] = ord("@")
    out[:, 1:4] = np.frombuffer(b"syn", dtype=np.uint8)
    idx = np.arange(n_records, dtype=np.int64)
    for d in range(10):
        out[:, 13 - d] = ord("0") + (idx // 10 ** d) % 10
    out[:, 14] = ord("\n")
    s0 = name_len
    out[:, s0:s0 + read_len] = np.frombuffer(b"ACGT", dtype=np.uint8)[
        rng.integers(0, 4, (n_records, read_len), dtype=np.uint8)]
    out[:, s0 + read_len] = ord("\n")
    p0 = s0 + read_len + 1
    out[:, p0] = ord("+")
    out[:, p0 + 1] = ord("\n")
    q0 = p0 + 2
    # round(N(35, 5)) clipped to [2, 41] by inverse CDF: a uniform 16-bit draw
    # indexes the distribution's quantiles at (k + 0.5) / 65536
    from scipy.stats import norm
    table = (np.clip(np.rint(norm.ppf((np.arange(65536) + 0.5) / 65536, 35.0, 5.0)), 2, 41) + 33).astype(np.uint8)
    out[:, q0:q0 + read_len] = table[rng.integers(0, 65536, (n_records, read_len), dtype=np.uint16)]
    out[:, q0 + read_len] = ord("\n")
    return out.reshape(-1)


def config_fastq(n_records=None, threads=None) -> SynthFastq:
    """FASTQ half of the mixed BCF + FASTQ configuration: 5,000,000 x 150 bp
    records, bgzip framing (65,280-byte payloads, level 6); `n_records` scales it
    down for tests."""
    n = n_records or 5_000_000
    u = fastq_stream(n, CONFIG_SEED + 4)
    c = bgzf(u, level=6, payload_cap=65280, seed=CONFIG_SEED + 4, threads=threads)
    return SynthFastq(c, u, n)


@dataclass
class SynthBcf:
    compressed: np.ndarray
    inflated: np.ndarray
    records_start: int
    n_records: int
    samples: int
    contigs: int


def bcf_stream(n_sites: int, n_samples: int, seed: int, contig_len: int = 50_000_000):
    """Uncompressed BCF: one contig, biallelic SNVs at sorted uniform positions,
    INFO AC and AN (int32) and AF (float), FILTER PASS, GT for every sample with
    alleles drawn from AF ~ Beta(0.1, 0.9). Returns (bytes, records_start)."""
    rng = np.random.default_rng(seed)
    names = "\t".join(f"S{i:05d}" for i in range(n_samples))
    text = ("##fileformat=VCFv4.2\n"
            '##FILTER=<ID=PASS,Description="All filters passed">\n'
            f"##contig=<ID=1,length={contig_len}>\n"
            '##INFO=<ID=AC,Number=A,Type=Integer,Description="Allele count">\n'
            '##INFO=<ID=AN,Number=1,Type=Integer,Description="Allele number">\n'
            '##INFO=<ID=AF,Number=A,Type=Float,Description="Allele frequency">\n'
            '##FORMAT=<ID=GT,Number=1,Type=String,Description="Genotype">\n'
            f"#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t{names}\n").encode() + b"\0"
    head = b"BCF\x02\x02" + len(text).to_bytes(4, "little") + text
    l_shared = 24 + 1 + 2 + 2 + 2 + 3 * 7
    l_indiv = 2 + 1 + 2 * n_samples
    width = 8 + l_shared + l_indiv
    rec = np.zeros((n_sites, width), dtype=np.uint8)

    def put(col, values, dtype):
        v = np.ascontiguousarray(values, dtype=dtype).view(np.uint8).reshape(n_sites, -1)
        rec[:, col:col + v.shape[1]] = v

    pos = np.sort(rng.integers(0, contig_len, n_sites)).astype(np.int32)
    af = rng.beta(0.1, 0.9, n_sites)
    hap = (rng.random((n_sites, 2 * n_samples)) < af[:, None]).astype(np.uint8)
    ac = hap.sum(axis=1).astype(np.int32)
    an = np.full(n_sites, 2 * n_samples, np.int32)
    bases = np.frombuffer(b"ACGT", np.uint8)
    ref = rng.integers(0, 4, n_sites)
    alt = (ref + rng.integers(1, 4, n_sites)) % 4
    put(0, np.full(n_sites, l_shared), np.uint32)
    put(4, np.full(n_sites, l_indiv), np.uint32)
    put(8, np.zeros(n_sites), np.int32)                       # CHROM 0
    put(12, pos, np.int32)
    put(16, np.ones(n_sites), np.int32)                       # rlen 1
    put(20, np.full(n_sites, 0x7F800001), np.uint32)          # QUAL missing
    put(24, np.full(n_sites, 3), np.uint16)                   # n_info
    put(26, np.full(n_sites, 2), np.uint16)                   # n_allele
    rec[:, 28:31] = np.frombuffer(n_samples.to_bytes(3, "little"), np.uint8)
    rec[:, 31] = 1                                            # n_fmt
    c = 32
    rec[:, c] = 0x07; c += 1                                  # ID missing
    rec[:, c] = 0x17; rec[:, c + 1] = bases[ref]; c += 2      # REF
    rec[:, c] = 0x17; rec[:, c + 1] = bases[alt]; c += 2      # ALT
    rec[:, c] = 0x11; rec[:, c + 1] = 0; c += 2               # FILTER PASS
    for key, typ, val in ((1, 0x13, ac), (2, 0x13, an), (3, 0x15, (ac / an).astype(np.float32))):
        rec[:, c] = 0x11; rec[:, c + 1] = key; rec[:, c + 2] = typ
        put(c + 3, val, np.float32 if typ == 0x15 else np.int32)
        c += 7
    assert c == 8 + l_shared
    rec[:, c] = 0x11; rec[:, c + 1] = 4; rec[:, c + 2] = 0x21; c += 3   # FORMAT GT, 2 x int8
    gt = ((hap.astype(np.uint8) + 1) << 1)                    # unphased allele index + 1
    rec[:, c:c + 2 * n_samples] = gt
    return np.concatenate([np.frombuffer(head, np.uint8), rec.reshape(-1)]), len(head)


def config_bcf(n_sites=None, n_samples=2504, threads=None) -> SynthBcf:
    """BCF half of the mixed BCF + FASTQ configuration: 200,000 biallelic sites
    x 2,504 diploid samples (GT only), bcftools framing (65,280-byte payloads,
    level 6); `n_sites` scales it down for tests."""
    n = n_sites or 200_000
    u, rs = bcf_stream(n, n_samples, CONFIG_SEED + 4)
    c = bgzf(u, level=6, payload_cap=65280, seed=CONFIG_SEED + 4, threads=threads)
    return SynthBcf(c, u, rs, n, n_samples, 1)
