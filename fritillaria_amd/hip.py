"""ctypes binding of ``include/frhip.h`` mirroring fritillaria's device API.

Names follow the reference crates: ``HipCodec`` (CudaCodec,
crates/fritillaria-cuda/src/lib.rs:152-294), ``DeviceInflateBatch``
(crates/fritillaria-core/src/device.rs:145-296), ``BamDecoder`` /
``DeviceRecordBatch`` (crates/fritillaria-cuda/src/bam.rs:103-187,
crates/fritillaria-bam/src/columnar/device.rs:135-282), ``DeviceBgzfReader``
(crates/fritillaria-bgzf/src/device_read.rs:47-280) and ``select_codec`` /
``Backend`` (crates/fritillaria/src/lib.rs:151-260).

There is no CPU fallback anywhere in this module: every call goes to the HIP
library, and a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _build

# ---------------------------------------------------------------------------
# errors (crates/fritillaria-core/src/error.rs:14-70)
# ---------------------------------------------------------------------------

OK, E_INVALID_BLOCK, E_CHECKSUM_MISMATCH, E_SIZE_MISMATCH, E_INFLATE, E_MALFORMED, \
    E_INVALID_DEVICE_BATCH, E_UNAVAILABLE, E_DEVICE, E_IO = range(10)


class FrError(Exception):
    kind = -1

    def __init__(self, err: "_Err"):
        self.offset = err.offset
        self.code = err.code
        self.expected = err.expected
        self.actual = err.actual
        self.format = err.format.decode(errors="replace")
        self.reason = err.reason.decode(errors="replace")
        super().__init__(self.describe())

    def describe(self) -> str:
        return f"{type(self).__name__} at {self.offset}: {self.reason}"


class InvalidBlock(FrError):
    kind = E_INVALID_BLOCK


class ChecksumMismatch(FrError):
    kind = E_CHECKSUM_MISMATCH


class SizeMismatch(FrError):
    kind = E_SIZE_MISMATCH


class InflateError(FrError):
    kind = E_INFLATE


class Malformed(FrError):
    kind = E_MALFORMED

    @property
    def position(self) -> int:
        return self.offset


class InvalidDeviceBatch(FrError):
    kind = E_INVALID_DEVICE_BATCH


class HipUnavailable(FrError):
    kind = E_UNAVAILABLE


class HipError(FrError):
    kind = E_DEVICE


class IoError(FrError):
    """Io: a malformed index (InvalidData) or a region it cannot answer (InvalidInput)."""
    kind = E_IO


_ERRORS = {c.kind: c for c in (InvalidBlock, ChecksumMismatch, SizeMismatch, InflateError, Malformed,
                               InvalidDeviceBatch, HipUnavailable, HipError, IoError)}


class _Err(C.Structure):
    _fields_ = [("kind", C.c_int32), ("code", C.c_uint32), ("offset", C.c_uint64),
                ("expected", C.c_uint64), ("actual", C.c_uint64), ("format", C.c_char * 16),
                ("reason", C.c_char * 200)]


class BlockSpanC(C.Structure):
    _fields_ = [("compressed_offset", C.c_uint64), ("payload_start", C.c_uint64),
                ("payload_len", C.c_uint64), ("isize", C.c_uint32), ("crc32", C.c_uint32)]


SPAN_DTYPE = np.dtype([("compressed_offset", "<u8"), ("payload_start", "<u8"),
                       ("payload_len", "<u8"), ("isize", "<u4"), ("crc32", "<u4")])


class InflateTimings(C.Structure):
    """InflateTimings (crates/fritillaria-cuda/src/backend.rs:136-177): per-phase
    host wall-clock in ms, phases separated by stream synchronisation; the
    timed calls add to it, so one object accumulates over batches."""
    _fields_ = [("plan_ms", C.c_double), ("upload_ms", C.c_double), ("kernel_ms", C.c_double),
                ("download_ms", C.c_double), ("verify_ms", C.c_double), ("blocks", C.c_uint64),
                ("compressed_bytes", C.c_uint64), ("inflated_bytes", C.c_uint64)]

    def total(self) -> float:
        return self.plan_ms + self.upload_ms + self.kernel_ms + self.download_ms + self.verify_ms


class DecodeTimings(C.Structure):
    """DecodeTimings (crates/fritillaria-cuda/src/bam.rs:55-101), same rules."""
    _fields_ = [("setup_ms", C.c_double), ("scan_ms", C.c_double), ("reconcile_ms", C.c_double),
                ("totals_ms", C.c_double), ("emit_ms", C.c_double), ("decode_ms", C.c_double),
                ("batches", C.c_uint64), ("blocks", C.c_uint64), ("records", C.c_uint64), ("bytes", C.c_uint64)]

    def total(self) -> float:
        return self.setup_ms + self.scan_ms + self.reconcile_ms + self.totals_ms + self.emit_ms + self.decode_ms


def _raise(rc: int, err: _Err):
    if rc != OK:
        raise _ERRORS.get(rc, FrError)(err)


# ---------------------------------------------------------------------------
# library
# ---------------------------------------------------------------------------

_lib = None
_vp = C.c_void_p
_u64 = C.c_uint64
_ep = C.POINTER(_Err)
# frhip_sink: int (*)(void *user, const uint8_t *p, uint64_t n)
_SINK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_uint8), C.c_uint64)


def lib():
    """Loads libfrhip.so. Raises if it has not been built: there is no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    path = _build.frhip_lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(path)
    sig = {
        "frhip_device_available": (C.c_int, []),
        "frhip_ctx_create": (C.c_int, [C.c_int, _vp, C.POINTER(_vp), _ep]),
        "frhip_ctx_destroy": (None, [_vp]),
        "frhip_ctx_ordinal": (C.c_int, [_vp]),
        "frhip_ctx_stream": (_vp, [_vp]),
        "frhip_ctx_set_bam_window": (C.c_int, [_vp, C.c_uint32]),
        "frhip_ctx_synchronize": (C.c_int, [_vp, _ep]),
        "frhip_ctx_set_inflate_schedule": (C.c_int, [_vp, C.c_int]),
        "frhip_ctx_set_two_phase_min_members": (C.c_int, [_vp, C.c_uint32]),
        "frhip_ctx_two_phase_min_members": (C.c_uint32, [_vp]),
        "frhip_ctx_compute_units": (C.c_uint32, [_vp]),
        "frhip_inflate_schedules": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int]),
        "frhip_codec_name": (C.c_char_p, []),
        "frhip_device_codec_name": (C.c_char_p, []),
        "frhip_buf_len": (_u64, [_vp]),
        "frhip_buf_ordinal": (C.c_int, [_vp]),
        "frhip_buf_device_ptr": (_vp, [_vp]),
        "frhip_buf_copy_to_host": (C.c_int, [_vp, _vp, _u64, _ep]),
        "frhip_buf_free": (None, [_vp]),
        "frhip_buf_alloc": (C.c_int, [_vp, _u64, C.POINTER(_vp), _ep]),
        "frhip_buf_copy_device": (C.c_int, [_vp, _vp, _vp, _ep]),
        "frhip_buf_copy_from_host": (C.c_int, [_vp, _u64, _vp, _u64, _ep]),
        "frhip_buf_ready_event": (_vp, [_vp]),
        "frhip_host_alloc": (C.c_int, [_u64, C.POINTER(_vp), _ep]),
        "frhip_host_free": (None, [_vp]),
        "frhip_pipeline_upload": (C.c_int, [_vp, _vp, _u64, _ep]),
        "frhip_event_wait": (C.c_int, [_vp, _vp, _ep]),
        "frhip_event_synchronize": (C.c_int, [_vp, _ep]),
        "frhip_inflate_batch_device_timed": (C.c_int, [_vp, _vp, _u64, _vp, _u64, C.POINTER(_vp), _vp, _vp, _ep]),
        "frhip_inflate_batch_timed": (C.c_int, [_vp, _vp, _u64, _vp, _u64, _vp, _u64, _vp, _vp, _ep]),
        "frhip_bam_decode_timed": (C.c_int, [_vp, _vp, _vp, _u64, _vp, _u64, C.POINTER(_vp), _vp, _ep]),
        "frhip_inflate_batch_device_noalloc": (C.c_int, [_vp, _vp, _u64, _vp, _u64, _vp, _u64, _vp, _ep]),
        "frhip_discover": (C.c_int, [_vp, _u64, _u64, _vp, _u64, C.POINTER(_u64), C.POINTER(_u64), _ep]),
        "frhip_inflate_batch_device": (C.c_int, [_vp, _vp, _u64, _vp, _u64, C.POINTER(_vp), _vp, _ep]),
        "frhip_inflate_batch": (C.c_int, [_vp, _vp, _u64, _vp, _u64, _vp, _u64, _vp, _ep]),
        "frhip_crc32_blocks": (C.c_int, [_vp, _vp, _vp, _vp, _u64, _vp, _ep]),
        "frhip_crc32_combine": (C.c_uint32, [C.c_uint32, C.c_uint32, _u64]),
        "frhip_bam_decode": (C.c_int, [_vp, _vp, _vp, _u64, _vp, _u64, C.POINTER(_vp), _ep]),
        "frhip_bam_batch_len": (_u64, [_vp]),
        "frhip_bam_batch_tail": (_u64, [_vp]),
        "frhip_bam_batch_fallbacks": (_u64, [_vp]),
        "frhip_bam_batch_column": (_vp, [_vp, C.c_int]),
        "frhip_bam_batch_free": (None, [_vp]),
        "frhip_bam_parse_header": (C.c_int, [_vp, _u64, C.POINTER(_u64), C.POINTER(C.c_uint32), _ep]),
        "frhip_bam_soa_build": (C.c_int, [_vp, _vp, _vp, C.POINTER(_vp), _ep]),
        "frhip_bam_soa_len": (_u64, [_vp]),
        "frhip_bam_soa_column": (_vp, [_vp, C.c_int]),
        "frhip_bam_soa_free": (None, [_vp]),
        "frhip_bcf_decode": (C.c_int, [_vp, _vp, _vp, _u64, _u64, C.c_uint32, C.c_uint32, C.POINTER(_vp), _ep]),
        "frhip_bcf_batch_len": (_u64, [_vp]),
        "frhip_bcf_batch_tail": (_u64, [_vp]),
        "frhip_bcf_batch_fallbacks": (_u64, [_vp]),
        "frhip_bcf_batch_column": (_vp, [_vp, C.c_int]),
        "frhip_bcf_batch_free": (None, [_vp]),
        "frhip_bcf_parse_header": (C.c_int, [_vp, _u64, C.POINTER(_u64), C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint32), _ep]),
        "frhip_fastq_decode": (C.c_int, [_vp, _vp, _u64, C.POINTER(_vp), _ep]),
        "frhip_text_scan": (C.c_int, [_vp, _vp, _u64, C.c_uint32, C.c_uint32, C.POINTER(_vp), _ep]),
        "frhip_text_batch_len": (_u64, [_vp]),
        "frhip_text_batch_headers": (_u64, [_vp]),
        "frhip_text_batch_tail": (_u64, [_vp]),
        "frhip_text_batch_column": (_vp, [_vp, C.c_int]),
        "frhip_text_batch_free": (None, [_vp]),
        "frhip_fasta_compact": (C.c_int, [_vp, _vp, C.POINTER(_vp), _ep]),
        "frhip_fasta_contigs": (_u64, [_vp]),
        "frhip_fasta_total_bases": (_u64, [_vp]),
        "frhip_fasta_reference_column": (_vp, [_vp, C.c_int]),
        "frhip_fasta_free": (None, [_vp]),
        "frhip_compress_batch_device": (C.c_int, [_vp, _vp, _vp, _u64, C.POINTER(_vp), _vp, _vp, _ep]),
        "frhip_framed_size": (_u64, [_u64, _u64]),
        "frhip_bgzf_writer_create": (C.c_int, [_vp, _SINK, _vp, _u64, _u64, C.POINTER(_vp), _ep]),
        "frhip_bgzf_writer_write_batch": (C.c_int, [_vp, _vp, _vp, _u64, _ep]),
        "frhip_bgzf_writer_write_all": (C.c_int, [_vp, _vp, _ep]),
        "frhip_bgzf_writer_finish": (C.c_int, [_vp, _ep]),
        "frhip_bgzf_writer_blocks_written": (_u64, [_vp]),
        "frhip_bgzf_writer_batches_run": (_u64, [_vp]),
        "frhip_bgzf_writer_compressed_bytes": (_u64, [_vp]),
        "frhip_bgzf_writer_uncompressed_bytes": (_u64, [_vp]),
        "frhip_bgzf_writer_virtual_position": (_u64, [_vp]),
        "frhip_bgzf_writer_free": (None, [_vp]),
        "frhip_bcf_fastq_decode": (C.c_int, [_vp, _vp, _u64, _vp, _u64, _u64, C.POINTER(_vp), _vp,
                                             C.POINTER(_u64), C.POINTER(_u64), C.POINTER(C.c_uint32),
                                             C.POINTER(C.c_uint32), C.POINTER(_vp), C.POINTER(_vp), _ep]),
        "frhip_fastq_batch_len": (_u64, [_vp]),
        "frhip_fastq_batch_tail": (_u64, [_vp]),
        "frhip_fastq_batch_survivors": (_u64, [_vp]),
        "frhip_fastq_batch_column": (_vp, [_vp, C.c_int]),
        "frhip_fastq_batch_free": (None, [_vp]),
        "frhip_reader_open": (C.c_int, [_vp, _vp, _u64, _u64, C.POINTER(_vp), _ep]),
        "frhip_reader_next_batch": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(_vp),
                                              C.POINTER(C.POINTER(_u64)), C.POINTER(C.POINTER(_u64)),
                                              C.POINTER(_u64), C.POINTER(_u64), _ep]),
        "frhip_reader_carry_from": (C.c_int, [_vp, _u64, _ep]),
        "frhip_reader_set_read_ahead": (C.c_int, [_vp, _u64, _u64, _ep]),
        "frhip_reader_free": (None, [_vp]),
        "frhip_pipeline_create": (C.c_int, [_vp, _vp, _u64, _vp, C.POINTER(_vp), _ep]),
        "frhip_pipeline_free": (None, [_vp]),
        "frhip_pipeline_run": (C.c_int, [_vp, _ep]),
        "frhip_pipeline_launch": (C.c_int, [_vp, _ep]),
        "frhip_pipeline_finish": (C.c_int, [_vp, _ep]),
        "frhip_pipeline_members": (_u64, [_vp]),
        "frhip_pipeline_inflated_len": (_u64, [_vp]),
        "frhip_pipeline_compressed_len": (_u64, [_vp]),
        "frhip_pipeline_records": (_u64, [_vp]),
        "frhip_pipeline_records_start": (_u64, [_vp]),
        "frhip_pipeline_tail": (_u64, [_vp]),
        "frhip_pipeline_fallbacks": (_u64, [_vp]),
        "frhip_pipeline_device_discovery": (C.c_int, [_vp]),
        "frhip_pipeline_inflated": (_vp, [_vp]),
        "frhip_pipeline_column": (_vp, [_vp, C.c_int]),
        "frhip_pipeline_soa_column": (_vp, [_vp, C.c_int]),
        "frhip_pipeline_member_offsets": (C.c_int, [_vp, _vp, _vp, _ep]),
        "frhip_pipeline_stage_ms": (C.c_int, [_vp, C.POINTER(C.c_double), C.c_int]),
        "frhip_server_create": (C.c_int, [_vp, _vp, _u64, C.POINTER(_vp), _ep]),
        "frhip_server_create_slots": (C.c_int, [_vp, _vp, _u64, C.c_int, C.POINTER(_vp), _ep]),
        "frhip_server_slots": (C.c_int, [_vp]),
        "frhip_server_launch": (C.c_int, [_vp, C.c_int, _vp, _u64, _ep]),
        "frhip_server_query": (C.c_int, [_vp, C.c_int]),
        "frhip_server_finish": (C.c_int, [_vp, C.c_int, _vp, _ep]),
        "frhip_server_batch_interval": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "frhip_server_slot_column": (_vp, [_vp, C.c_int, C.c_int]),
        "frhip_server_slot_inflated": (_vp, [_vp, C.c_int]),
        "frhip_server_free": (None, [_vp]),
        "frhip_server_members": (_u64, [_vp]),
        "frhip_server_member_table": (C.c_int, [_vp, _vp, _vp, _ep]),
        "frhip_server_batch_footprint": (_u64, [_vp, _vp, _u64]),
        "frhip_server_run": (C.c_int, [_vp, _vp, _u64, _vp, _ep]),
        "frhip_server_column": (_vp, [_vp, C.c_int]),
        "frhip_server_inflated": (_vp, [_vp]),
        "frhip_device_memory": (C.c_int, [_vp, C.POINTER(_u64), C.POINTER(_u64), _ep]),
        "frhip_server_launch_regions": (C.c_int, [_vp, C.c_int, _vp, _vp, _u64, _ep]),
        "frhip_server_finish_regions": (C.c_int, [_vp, C.c_int, _vp, _ep]),
        "frhip_server_slot_selection": (_vp, [_vp, C.c_int]),
        "frhip_index_read": (C.c_int, [_vp, _u64, C.POINTER(_vp), _ep]),
        "frhip_index_free": (None, [_vp]),
        "frhip_index_format": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "frhip_index_references": (C.c_uint32, [_vp]),
        "frhip_index_unplaced_unmapped": (C.c_int, [_vp, C.POINTER(_u64)]),
        "frhip_index_metadata": (C.c_int, [_vp, C.c_uint32, _vp]),
        "frhip_index_last_first_record_start": (C.c_int, [_vp, C.POINTER(_u64)]),
        "frhip_index_query": (C.c_int, [_vp, C.c_uint32, _u64, _u64, _vp, _u64, C.POINTER(_u64), _ep]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols() -> list:
    """Every function declared in include/*.h, for the load/export check."""
    import glob
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    names = []
    for p in sorted(glob.glob(os.path.join(os.path.dirname(here), "include", "*.h"))):
        txt = open(p).read()
        names += re.findall(r"\b(frhip_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def inflate_schedules() -> list:
    """(ring bytes, waves per workgroup, resident waves per CU, token form) per inflate variant."""
    ring = np.zeros(32, np.uint32)
    nw = np.zeros(32, np.uint32)
    wpc = np.zeros(32, np.uint32)
    tok = np.zeros(32, np.uint32)
    k = lib().frhip_inflate_schedules(_ptr(ring), _ptr(nw), _ptr(wpc), _ptr(tok), 32)
    return [(int(ring[i]), int(nw[i]), int(wpc[i]), int(tok[i])) for i in range(k)]


def device_is_available() -> bool:
    return bool(lib().frhip_device_available())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def _u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    return np.frombuffer(bytes(data), dtype=np.uint8)


# ---------------------------------------------------------------------------
# context, buffers
# ---------------------------------------------------------------------------


class HipContext:
    """CudaContext (crates/fritillaria-cuda/src/backend.rs:29-436)."""

    def __init__(self, ordinal: int = 0, stream: Optional[int] = None):
        h = _vp()
        e = _Err()
        rc = lib().frhip_ctx_create(ordinal, stream, C.byref(h), C.byref(e))
        _raise(rc, e)
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def ordinal(self) -> int:
        return lib().frhip_ctx_ordinal(self._h)

    @property
    def stream(self) -> int:
        return lib().frhip_ctx_stream(self._h) or 0

    def synchronize(self) -> None:
        e = _Err()
        _raise(lib().frhip_ctx_synchronize(self._h, C.byref(e)), e)

    def set_inflate_schedule(self, schedule: int) -> None:
        if lib().frhip_ctx_set_inflate_schedule(self._h, schedule) != OK:
            raise ValueError(f"no inflate schedule {schedule}")

    def set_two_phase_min_members(self, n: int) -> None:
        """Batches of fewer members than n run a two-phase schedule's one-wave form
        (default: the one-wave form's resident waves, 4096 on a 256-CU device)."""
        lib().frhip_ctx_set_two_phase_min_members(self._h, int(n))

    def two_phase_min_members(self) -> int:
        return int(lib().frhip_ctx_two_phase_min_members(self._h))

    def compute_units(self) -> int:
        return int(lib().frhip_ctx_compute_units(self._h))

    def alloc(self, nbytes: int) -> "DeviceBuffer":
        """A device buffer of nbytes (contents undefined)."""
        out = _vp()
        e = _Err()
        _raise(lib().frhip_buf_alloc(self.handle, nbytes, C.byref(out), C.byref(e)), e)
        return DeviceBuffer(out)

    def copy_device(self, dst: "DeviceBuffer", src: "DeviceBuffer") -> None:
        """Plain device-to-device copy kernel over buffers of equal length."""
        e = _Err()
        _raise(lib().frhip_buf_copy_device(self.handle, dst._h, src._h, C.byref(e)), e)

    def set_bam_window(self, window: int) -> None:
        lib().frhip_ctx_set_bam_window(self._h, window)

    def close(self):
        if getattr(self, "_h", None):
            lib().frhip_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Owning handle to device memory (DeviceBuffer, device.rs:82-143)."""

    def __init__(self, handle, owned: bool = True, keepalive=None):
        self._h = handle
        self._owned = owned
        self._keep = keepalive

    def byte_len(self) -> int:
        return int(lib().frhip_buf_len(self._h))

    def is_empty(self) -> bool:
        return self.byte_len() == 0

    def device_ordinal(self) -> int:
        return lib().frhip_buf_ordinal(self._h)

    def device_ptr(self) -> int:
        return lib().frhip_buf_device_ptr(self._h) or 0

    def copy_to_host(self, dst: np.ndarray) -> None:
        e = _Err()
        rc = lib().frhip_buf_copy_to_host(self._h, _ptr(dst), dst.nbytes, C.byref(e))
        _raise(rc, e)

    def to_vec(self) -> np.ndarray:
        out = np.empty(self.byte_len(), dtype=np.uint8)
        self.copy_to_host(out)
        return out

    def ready_event(self) -> Optional[int]:
        """The hipEvent_t recorded after the kernels that wrote this buffer
        (CudaAlloc::ready_event), or None."""
        return lib().frhip_buf_ready_event(self._h) or None

    def copy_from_host(self, src, offset: int = 0) -> None:
        a = _u8(src)
        e = _Err()
        _raise(lib().frhip_buf_copy_from_host(self._h, offset, _ptr(a), a.size, C.byref(e)), e)

    def free(self):
        if self._owned and self._h:
            lib().frhip_buf_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class InflateBatch:
    """Host output layout (InflateBatch, crates/fritillaria-core/src/codec.rs:53-157)."""
    data: np.ndarray
    offsets: np.ndarray

    def __len__(self):
        return len(self.offsets) - 1

    def block(self, i: int) -> np.ndarray:
        return self.data[int(self.offsets[i]):int(self.offsets[i + 1])]


class DeviceInflateBatch:
    """Inflated bytes that never left the device; offsets stay on the host."""

    def __init__(self, data: Optional[DeviceBuffer] = None, offsets=None, coffsets=None):
        self._data = data
        self._offsets = np.zeros(1, dtype=np.uint64) if offsets is None else np.asarray(offsets, np.uint64)
        self._coffsets = None if coffsets is None else np.asarray(coffsets, np.uint64)

    def data(self) -> Optional[DeviceBuffer]:
        return self._data

    def offsets(self) -> np.ndarray:
        return self._offsets

    def compressed_offsets(self) -> Optional[np.ndarray]:
        return self._coffsets

    def __len__(self) -> int:
        return len(self._offsets) - 1

    def is_empty(self) -> bool:
        return len(self) == 0

    def byte_len(self) -> int:
        return 0 if self._data is None else self._data.byte_len()

    def device_ordinal(self) -> Optional[int]:
        return None if self._data is None else self._data.device_ordinal()

    def block_range(self, i: int):
        return int(self._offsets[i]), int(self._offsets[i + 1])

    def to_host(self) -> InflateBatch:
        if self._data is None:
            return InflateBatch(np.zeros(0, np.uint8), np.zeros(1, np.uint64))
        return InflateBatch(self._data.to_vec(), self._offsets.copy())


def _spans_array(spans) -> np.ndarray:
    if isinstance(spans, np.ndarray) and spans.dtype == SPAN_DTYPE:
        return np.ascontiguousarray(spans)
    arr = np.zeros(len(spans), dtype=SPAN_DTYPE)
    for i, s in enumerate(spans):
        if isinstance(s, np.void):
            arr[i] = s
        else:
            arr[i] = (s.compressed_offset, s.payload_start, s.payload_len, s.isize, s.crc32)
    return arr


def discover(data, base: int = 0):
    """BlockDiscovery over host bytes: (spans, position). Raises InvalidBlock."""
    a = _u8(data)
    cap = a.size // 26 + 2
    spans = np.zeros(cap, dtype=SPAN_DTYPE)
    n = _u64()
    pos = _u64()
    e = _Err()
    rc = lib().frhip_discover(_ptr(a), a.size, base, _ptr(spans), cap, C.byref(n), C.byref(pos), C.byref(e))
    _raise(rc, e)
    return spans[:n.value].copy(), pos.value


class HipCodec:
    """BlockCodec + DeviceBlockCodec over the HIP inflate kernel."""

    def __init__(self, context: Optional[HipContext] = None, ordinal: int = 0):
        self.ctx = context or HipContext(ordinal)

    def name(self) -> str:
        return lib().frhip_codec_name().decode()

    def device_name(self) -> str:
        return lib().frhip_device_codec_name().decode()

    def device_ordinal(self) -> int:
        return self.ctx.ordinal

    def inflate_batch_device(self, batch, spans) -> DeviceInflateBatch:
        a = _u8(batch)
        sp = _spans_array(spans)
        offs = np.zeros(sp.size + 1, dtype=np.uint64)
        out = _vp()
        e = _Err()
        rc = lib().frhip_inflate_batch_device(self.ctx.handle, _ptr(a), a.size, _ptr(sp), sp.size,
                                              C.byref(out), _ptr(offs), C.byref(e))
        _raise(rc, e)
        return DeviceInflateBatch(DeviceBuffer(out), offs, sp["compressed_offset"].copy())

    @staticmethod
    def ready_event(batch: "DeviceInflateBatch") -> Optional[int]:
        """CudaCodec::ready_event (crates/fritillaria-cuda/src/lib.rs:207-224):
        the event marking completion of the inflate that produced `batch`; None
        for an empty batch."""
        d = batch.data()
        return None if d is None else d.ready_event()

    def inflate_batch_device_timed(self, batch, spans, timings: InflateTimings) -> DeviceInflateBatch:
        a = _u8(batch)
        sp = _spans_array(spans)
        offs = np.zeros(sp.size + 1, dtype=np.uint64)
        out = _vp()
        e = _Err()
        rc = lib().frhip_inflate_batch_device_timed(self.ctx.handle, _ptr(a), a.size, _ptr(sp), sp.size,
                                                    C.byref(out), _ptr(offs), C.byref(timings), C.byref(e))
        _raise(rc, e)
        return DeviceInflateBatch(DeviceBuffer(out), offs, sp["compressed_offset"].copy())

    def inflate_batch_timed(self, batch, spans, timings: InflateTimings) -> InflateBatch:
        a = _u8(batch)
        sp = _spans_array(spans)
        total = int(sp["isize"].astype(np.uint64).sum()) if sp.size else 0
        out = np.empty(max(total, 1), dtype=np.uint8)
        offs = np.zeros(sp.size + 1, dtype=np.uint64)
        e = _Err()
        rc = lib().frhip_inflate_batch_timed(self.ctx.handle, _ptr(a), a.size, _ptr(sp), sp.size, _ptr(out),
                                             out.size, _ptr(offs), C.byref(timings), C.byref(e))
        _raise(rc, e)
        return InflateBatch(out[:total], offs)

    def inflate_batch_device_noalloc(self, d_batch: "DeviceBuffer", batch_len: int, spans,
                                     d_out: "DeviceBuffer") -> np.ndarray:
        """frhip_inflate_batch_device_noalloc: inflates into the caller's d_out
        from the caller's d_batch (batch_len bytes, then 4096 zero bytes), with
        no device allocation; returns the member boundaries."""
        sp = _spans_array(spans)
        offs = np.zeros(sp.size + 1, dtype=np.uint64)
        e = _Err()
        rc = lib().frhip_inflate_batch_device_noalloc(self.ctx.handle, d_batch.device_ptr(), batch_len, _ptr(sp),
                                                      sp.size, d_out.device_ptr(), d_out.byte_len(), _ptr(offs),
                                                      C.byref(e))
        _raise(rc, e)
        return offs

    def inflate_batch(self, batch, spans) -> InflateBatch:
        a = _u8(batch)
        sp = _spans_array(spans)
        total = int(sp["isize"].astype(np.uint64).sum()) if sp.size else 0
        out = np.empty(max(total, 1), dtype=np.uint8)
        offs = np.zeros(sp.size + 1, dtype=np.uint64)
        e = _Err()
        rc = lib().frhip_inflate_batch(self.ctx.handle, _ptr(a), a.size, _ptr(sp), sp.size, _ptr(out),
                                       out.size, _ptr(offs), C.byref(e))
        _raise(rc, e)
        return InflateBatch(out[:total], offs)

    def crc32_blocks(self, data: DeviceBuffer, offsets, lengths) -> np.ndarray:
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        ln = np.ascontiguousarray(lengths, dtype=np.uint32)
        out = np.zeros(o.size, dtype=np.uint32)
        e = _Err()
        rc = lib().frhip_crc32_blocks(self.ctx.handle, data._h, _ptr(o), _ptr(ln), o.size, _ptr(out),
                                      C.byref(e))
        _raise(rc, e)
        return out


def crc32_combine(crc1: int, crc2: int, len2: int) -> int:
    return int(lib().frhip_crc32_combine(crc1, crc2, len2))


# ---------------------------------------------------------------------------
# BAM
# ---------------------------------------------------------------------------

# (name, dtype) in the order of crates/fritillaria-bam/src/columnar/device.rs:95-110
BAM_COLUMNS = [
    ("reference_sequence_id", np.int32), ("position", np.int32), ("flags", np.uint16),
    ("mapping_quality", np.uint8), ("sequence_len", np.uint32),
    ("mate_reference_sequence_id", np.int32), ("mate_position", np.int32),
    ("template_length", np.int32), ("record_offsets", np.uint64), ("cigar_start", np.uint32),
    ("sequence_start", np.uint32), ("quality_start", np.uint32), ("aux_start", np.uint32),
    ("record_end", np.uint32), ("virtual_offsets", np.uint64),
]
_COL_INDEX = {n: i for i, (n, _) in enumerate(BAM_COLUMNS)}


class DeviceRecordBatch:
    """BAM columns in device memory. Columns point into the inflated batch,
    which is kept alive by this object."""

    def __init__(self, handle, source: DeviceInflateBatch):
        self._h = handle
        self._source = source

    def __len__(self) -> int:
        return int(lib().frhip_bam_batch_len(self._h))

    def is_empty(self) -> bool:
        return len(self) == 0

    def tail(self) -> int:
        return int(lib().frhip_bam_batch_tail(self._h))

    def fallbacks(self) -> int:
        return int(lib().frhip_bam_batch_fallbacks(self._h))

    def column(self, name: str) -> Optional[DeviceBuffer]:
        h = lib().frhip_bam_batch_column(self._h, _COL_INDEX[name])
        return None if not h else DeviceBuffer(h, owned=False, keepalive=self)

    def device_ordinal(self) -> int:
        return self.column("reference_sequence_id").device_ordinal()

    def column_to_host(self, name: str) -> np.ndarray:
        buf = self.column(name)
        dt = dict(BAM_COLUMNS)[name]
        out = np.empty(buf.byte_len() // np.dtype(dt).itemsize, dtype=dt)
        buf.copy_to_host(out)
        return out

    def to_host(self) -> dict:
        """Every column (the RecordBatch columns plus field bounds) on the host."""
        cols = {}
        for name, _ in BAM_COLUMNS:
            if self.column(name) is not None:
                cols[name] = self.column_to_host(name)
        return cols

    def free(self):
        if getattr(self, "_h", None):
            lib().frhip_bam_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


AUX_ENTRY_DTYPE = np.dtype([("tag", "S2"), ("type", "u1"), ("subtype", "u1"), ("value_offset", "<u4"),
                            ("count", "<u4"), ("size", "<u4")])
SOA_COLUMNS = [("cigar_offsets", np.uint64), ("cigar_ops", np.uint32), ("sequence_offsets", np.uint64),
               ("sequence_packed", np.uint8), ("aux_offsets", np.uint64), ("aux_entries", AUX_ENTRY_DTYPE)]
_SOA_INDEX = {n: i for i, (n, _) in enumerate(SOA_COLUMNS)}


class BamSoa:
    """Resolved CIGAR ops, packed sequences and the aux tag index of a decoded
    BAM batch, as device columns (include/frhip.h, frhip_bam_soa_build)."""

    def __init__(self, handle, records: "DeviceRecordBatch"):
        self._h = handle
        self._records = records

    def __len__(self) -> int:
        return int(lib().frhip_bam_soa_len(self._h))

    def column(self, name: str) -> DeviceBuffer:
        return DeviceBuffer(lib().frhip_bam_soa_column(self._h, _SOA_INDEX[name]), owned=False, keepalive=self)

    def column_to_host(self, name: str) -> np.ndarray:
        buf = self.column(name)
        dt = np.dtype(dict(SOA_COLUMNS)[name])
        out = np.empty(buf.byte_len() // dt.itemsize, dtype=dt)
        buf.copy_to_host(out)
        return out

    def to_host(self) -> dict:
        return {name: self.column_to_host(name) for name, _ in SOA_COLUMNS}

    def free(self):
        if getattr(self, "_h", None):
            lib().frhip_bam_soa_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BamDecoder:
    """BamDecoder (crates/fritillaria-cuda/src/bam.rs:103-187)."""

    def __init__(self, context: Optional[HipContext] = None, ordinal: int = 0):
        self.ctx = context or HipContext(ordinal)

    def soa(self, batch: DeviceInflateBatch, records: DeviceRecordBatch) -> BamSoa:
        """Resolved CIGARs, packed sequences and the aux index of decoded records."""
        data = batch.data()
        out = _vp()
        e = _Err()
        rc = lib().frhip_bam_soa_build(self.ctx.handle, None if data is None else data._h, records._h, C.byref(out),
                                       C.byref(e))
        _raise(rc, e)
        return BamSoa(out, records)

    def decode(self, batch: DeviceInflateBatch, start: int) -> DeviceRecordBatch:
        return self.decode_timed(batch, start, None)

    def decode_timed(self, batch: DeviceInflateBatch, start: int,
                     timings: Optional[DecodeTimings]) -> DeviceRecordBatch:
        """BamDecoder::decode_timed (crates/fritillaria-cuda/src/bam.rs:161-187):
        decode, adding each phase's time to `timings` (None: untimed)."""
        data = batch.data()
        offs = np.ascontiguousarray(batch.offsets(), dtype=np.uint64)
        co = batch.compressed_offsets()
        co = None if co is None else np.ascontiguousarray(co, dtype=np.uint64)
        out = _vp()
        e = _Err()
        args = (self.ctx.handle, None if data is None else data._h, _ptr(offs), len(offs) - 1,
                None if co is None else _ptr(co), start, C.byref(out))
        if timings is None:
            rc = lib().frhip_bam_decode(*args, C.byref(e))
        else:
            rc = lib().frhip_bam_decode_timed(*args, C.byref(timings), C.byref(e))
        _raise(rc, e)
        return DeviceRecordBatch(out, batch)


# ---------------------------------------------------------------------------
# FASTQ
# ---------------------------------------------------------------------------

# (name, dtype) in the order of crates/fritillaria-fastq/src/columnar/device.rs:44-68
FASTQ_COLUMNS = [
    ("record_offsets", np.uint64), ("sequence_start", np.uint32), ("plus_start", np.uint32),
    ("quality_start", np.uint32), ("sequence_len", np.uint32), ("record_end", np.uint32),
]
_FQ_INDEX = {n: i for i, (n, _) in enumerate(FASTQ_COLUMNS)}


class FastqRecordBatch:
    """FASTQ columns in device memory (crates/fritillaria-fastq/src/columnar/device.rs:84-160).
    Bounds index the inflated batch, which this object keeps alive."""

    def __init__(self, handle, source: DeviceInflateBatch):
        self._h = handle
        self._source = source

    def __len__(self) -> int:
        return int(lib().frhip_fastq_batch_len(self._h))

    def is_empty(self) -> bool:
        return len(self) == 0

    def tail(self) -> int:
        return int(lib().frhip_fastq_batch_tail(self._h))

    def survivors(self) -> int:
        return int(lib().frhip_fastq_batch_survivors(self._h))

    def column(self, name: str) -> DeviceBuffer:
        return DeviceBuffer(lib().frhip_fastq_batch_column(self._h, _FQ_INDEX[name]), owned=False, keepalive=self)

    def column_to_host(self, name: str) -> np.ndarray:
        buf = self.column(name)
        dt = dict(FASTQ_COLUMNS)[name]
        out = np.empty(buf.byte_len() // np.dtype(dt).itemsize, dtype=dt)
        buf.copy_to_host(out)
        return out

    def to_host(self) -> dict:
        return {name: self.column_to_host(name) for name, _ in FASTQ_COLUMNS}

    def free(self):
        if getattr(self, "_h", None):
            lib().frhip_fastq_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SpeculativeScan:
    """crates/fritillaria-fastq/src/columnar/speculative.rs:81-93. `proof` is
    "tiled" when every line start that begins a record is a record of the walk
    (survivors == records), else "walked"; the offsets are the walk's either way."""

    def __init__(self, offsets, tail, survivors, proof):
        self.offsets = offsets
        self.tail = tail
        self.candidates = 0
        self.survivors = survivors
        self.proof = proof


class FastqScanner:
    """FastqScanner (crates/fritillaria-cuda/src/fastq.rs:48-135)."""

    def __init__(self, context: Optional[HipContext] = None, ordinal: int = 0):
        self.ctx = context or HipContext(ordinal)

    def decode(self, batch: DeviceInflateBatch, start: int) -> FastqRecordBatch:
        data = batch.data()
        out = _vp()
        e = _Err()
        rc = lib().frhip_fastq_decode(self.ctx.handle, None if data is None else data._h, start, C.byref(out),
                                      C.byref(e))
        _raise(rc, e)
        return FastqRecordBatch(out, batch)

    def scan(self, batch: DeviceInflateBatch, start: int) -> SpeculativeScan:
        d = self.decode(batch, start)
        offs = d.column_to_host("record_offsets")
        n, surv = len(d), d.survivors()
        return SpeculativeScan(offs, d.tail(), surv, "tiled" if surv == n else "walked")


# ---------------------------------------------------------------------------
# Text framing and FASTA compaction
# ---------------------------------------------------------------------------

class Dialect:
    """Which first byte makes a line a header
    (crates/fritillaria-text/src/columnar/batch.rs:3-35)."""

    def __init__(self, comment: bytes, secondary: Optional[bytes] = None):
        self.comment = comment[0]
        self.secondary = secondary[0] if secondary else None


Dialect.SAM = Dialect(b"@")
Dialect.VCF = Dialect(b"#")
Dialect.GFF = Dialect(b"#")
Dialect.BED = Dialect(b"#")

TEXT_COLUMNS = ["record_offsets", "record_ends", "header_offsets", "tabs", "field_starts"]


class TextRecordBatch:
    """Text columns in device memory (crates/fritillaria-text/src/columnar/device.rs:13-160),
    all u64 offsets into the inflated batch, which this object keeps alive."""

    def __init__(self, handle, source):
        self._h = handle
        self._source = source

    def __len__(self) -> int:
        return int(lib().frhip_text_batch_len(self._h))

    def is_empty(self) -> bool:
        return len(self) == 0

    def header_count(self) -> int:
        return int(lib().frhip_text_batch_headers(self._h))

    def tail(self) -> int:
        return int(lib().frhip_text_batch_tail(self._h))

    def column(self, name: str) -> DeviceBuffer:
        return DeviceBuffer(lib().frhip_text_batch_column(self._h, TEXT_COLUMNS.index(name)), owned=False,
                            keepalive=self)

    def column_to_host(self, name: str) -> np.ndarray:
        buf = self.column(name)
        out = np.empty(buf.byte_len() // 8, dtype=np.uint64)
        buf.copy_to_host(out)
        return out

    def to_host(self) -> dict:
        return {name: self.column_to_host(name) for name in TEXT_COLUMNS}

    def free(self):
        if getattr(self, "_h", None):
            lib().frhip_text_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class TextScanner:
    """TextScanner (crates/fritillaria-cuda/src/text.rs:10-66)."""

    def __init__(self, context: Optional[HipContext] = None, ordinal: int = 0):
        self.ctx = context or HipContext(ordinal)

    def scan(self, batch, start: int, dialect: Dialect) -> TextRecordBatch:
        """`batch` is a DeviceInflateBatch or a DeviceBuffer of text bytes."""
        data = batch.data() if isinstance(batch, DeviceInflateBatch) else batch
        out = _vp()
        e = _Err()
        sec = 0x100 if dialect.secondary is None else dialect.secondary
        _raise(lib().frhip_text_scan(self.ctx.handle, None if data is None else data._h, start, dialect.comment, sec,
                                     C.byref(out), C.byref(e)), e)
        return TextRecordBatch(out, batch)


FASTA_COLUMNS = [("sequence", np.uint8), ("sequence_offsets", np.uint64), ("sequence_lengths", np.uint64),
                 ("record_offsets", np.uint64)]


class DeviceReference:
    """A compacted FASTA reference in device memory
    (crates/fritillaria-fasta/src/columnar/device.rs:13-120)."""

    def __init__(self, handle, source):
        self._h = handle
        self._source = source

    def __len__(self) -> int:
        return int(lib().frhip_fasta_contigs(self._h))

    def total_bases(self) -> int:
        return int(lib().frhip_fasta_total_bases(self._h))

    def column(self, name: str) -> DeviceBuffer:
        idx = [n for n, _ in FASTA_COLUMNS].index(name)
        return DeviceBuffer(lib().frhip_fasta_reference_column(self._h, idx), owned=False, keepalive=self)

    def column_to_host(self, name: str) -> np.ndarray:
        buf = self.column(name)
        dt = dict(FASTA_COLUMNS)[name]
        out = np.empty(buf.byte_len() // np.dtype(dt).itemsize, dtype=dt)
        buf.copy_to_host(out)
        return out

    def free(self):
        if getattr(self, "_h", None):
            lib().frhip_fasta_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class FastaCompactor:
    """FastaCompactor (crates/fritillaria-cuda/src/fasta.rs:10-67)."""

    def __init__(self, context: Optional[HipContext] = None, ordinal: int = 0):
        self.ctx = context or HipContext(ordinal)

    def compact(self, batch) -> DeviceReference:
        data = batch.data() if isinstance(batch, DeviceInflateBatch) else batch
        out = _vp()
        e = _Err()
        _raise(lib().frhip_fasta_compact(self.ctx.handle, None if data is None else data._h, C.byref(out),
                                         C.byref(e)), e)
        return DeviceReference(out, batch)


# ---------------------------------------------------------------------------
# Writing BGZF from device payloads (include/frhip_write.h)
# ---------------------------------------------------------------------------

MAX_COMPRESSIBLE_PAYLOAD = 65505
DEFAULT_PAYLOAD_SIZE = 65280
DEFAULT_CHUNKS_PER_BATCH = 1024
FRAMING_DEFLATED, FRAMING_STORED, FRAMING_EMPTY = 0, 1, 2
EOF_BLOCK = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def framed_size(payload_len: int, deflate_len: int) -> int:
    """framed_size (crates/fritillaria-core/src/compress.rs:133-142)."""
    return int(lib().frhip_framed_size(payload_len, deflate_len))


class CompressedBatch:
    """Framed BGZF blocks back to back (crates/fritillaria-core/src/compress.rs:157-247),
    held in device memory; `offsets` (n + 1) are the block boundaries."""

    def __init__(self, device: DeviceBuffer, offsets: np.ndarray, framing: np.ndarray):
        self.device = device
        self._offsets = offsets
        self.framing = framing

    def __len__(self) -> int:
        return len(self._offsets) - 1

    def offsets(self) -> np.ndarray:
        return self._offsets

    def byte_len(self) -> int:
        return int(self._offsets[-1])

    def data(self) -> bytes:
        out = np.empty(self.byte_len(), np.uint8)
        if out.size:
            self.device.copy_to_host(out)
        return out.tobytes()

    def block(self, i: int) -> bytes:
        return self.data()[int(self._offsets[i]):int(self._offsets[i + 1])]


class HipCompressor:
    """DeviceBlockCompressor (crates/fritillaria-core/src/device.rs:350-372) on
    this library's own DEFLATE kernels."""

    def __init__(self, context: Optional[HipContext] = None, ordinal: int = 0):
        self.ctx = context or HipContext(ordinal)

    def name(self) -> str:
        return "hip"

    def device_ordinal(self) -> int:
        return self.ctx.ordinal

    def compress_batch_device(self, data: DeviceBuffer, bounds) -> CompressedBatch:
        b = np.ascontiguousarray(bounds, dtype=np.uint64)
        n = max(b.size - 1, 0)
        offs = np.zeros(n + 1, np.uint64)
        frm = np.zeros(max(n, 1), np.uint8)
        out = _vp()
        e = _Err()
        _raise(lib().frhip_compress_batch_device(self.ctx.handle, data._h if data is not None else None,
                                                 _ptr(b) if b.size else None, b.size, C.byref(out), _ptr(offs),
                                                 _ptr(frm), C.byref(e)), e)
        return CompressedBatch(DeviceBuffer(out), offs, frm[:n])


class DeviceBgzfWriter:
    """DeviceBgzfWriter (crates/fritillaria-bgzf/src/device_write.rs:70-259):
    writes BGZF to `sink` (an object with write(bytes), default an in-memory
    bytearray) from payloads in device memory. finish() writes the EOF block
    and returns the sink."""

    def __init__(self, sink=None, context: Optional[HipContext] = None, chunks_per_batch: int = 0,
                 payload_size: int = 0, ordinal: int = 0):
        self.ctx = context or HipContext(ordinal)
        self.sink = bytearray() if sink is None else sink
        self._error = None

        def write(_user, p, n):
            try:
                chunk = C.string_at(p, n)
                if isinstance(self.sink, bytearray):
                    self.sink += chunk
                else:
                    self.sink.write(chunk)
                return 0
            except Exception as exc:  # reported by the library as FRHIP_E_IO
                self._error = exc
                return 1

        self._cb = _SINK(write)
        h = _vp()
        e = _Err()
        _raise(lib().frhip_bgzf_writer_create(self.ctx.handle, self._cb, None, chunks_per_batch, payload_size,
                                              C.byref(h), C.byref(e)), e)
        self._h = h

    def write_batch(self, data: DeviceBuffer, bounds) -> None:
        b = np.ascontiguousarray(bounds, dtype=np.uint64)
        e = _Err()
        _raise(lib().frhip_bgzf_writer_write_batch(self._h, data._h if data is not None else None,
                                                   _ptr(b) if b.size else None, b.size, C.byref(e)), e)

    def write_all_device(self, data: DeviceBuffer) -> None:
        e = _Err()
        _raise(lib().frhip_bgzf_writer_write_all(self._h, data._h if data is not None else None, C.byref(e)), e)

    def blocks_written(self) -> int:
        return int(lib().frhip_bgzf_writer_blocks_written(self._h))

    def batches_run(self) -> int:
        return int(lib().frhip_bgzf_writer_batches_run(self._h))

    def compressed_bytes(self) -> int:
        return int(lib().frhip_bgzf_writer_compressed_bytes(self._h))

    def uncompressed_bytes(self) -> int:
        return int(lib().frhip_bgzf_writer_uncompressed_bytes(self._h))

    def virtual_position(self) -> int:
        return int(lib().frhip_bgzf_writer_virtual_position(self._h))

    def finish(self):
        e = _Err()
        _raise(lib().frhip_bgzf_writer_finish(self._h, C.byref(e)), e)
        return self.sink

    def close(self):
        if getattr(self, "_h", None):
            lib().frhip_bgzf_writer_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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
_BCF_INDEX = {n: i for i, (n, _) in enumerate(BCF_COLUMNS)}


class BcfRecordBatch:
    """BCF columns in device memory (crates/fritillaria-bcf/src/columnar/device.rs:100-200).
    Offsets index the inflated batch, which this object keeps alive."""

    def __init__(self, handle, source: DeviceInflateBatch):
        self._h = handle
        self._source = source

    def __len__(self) -> int:
        return int(lib().frhip_bcf_batch_len(self._h))

    def is_empty(self) -> bool:
        return len(self) == 0

    def tail(self) -> int:
        return int(lib().frhip_bcf_batch_tail(self._h))

    def fallbacks(self) -> int:
        return int(lib().frhip_bcf_batch_fallbacks(self._h))

    def column(self, name: str) -> DeviceBuffer:
        return DeviceBuffer(lib().frhip_bcf_batch_column(self._h, _BCF_INDEX[name]), owned=False, keepalive=self)

    def column_to_host(self, name: str) -> np.ndarray:
        buf = self.column(name)
        dt = dict(BCF_COLUMNS)[name]
        out = np.empty(buf.byte_len() // np.dtype(dt).itemsize, dtype=dt)
        buf.copy_to_host(out)
        return out

    def to_host(self) -> dict:
        return {name: self.column_to_host(name) for name, _ in BCF_COLUMNS}

    def free(self):
        if getattr(self, "_h", None):
            lib().frhip_bcf_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BcfScanner:
    """BcfScanner (crates/fritillaria-cuda/src/bcf.rs:42-143): `samples` and
    `contigs` come from the header (bcf_parse_header)."""

    def __init__(self, context: Optional[HipContext] = None, ordinal: int = 0):
        self.ctx = context or HipContext(ordinal)

    def decode(self, batch: DeviceInflateBatch, start: int, samples: int, contigs: int) -> BcfRecordBatch:
        data = batch.data()
        offs = np.ascontiguousarray(batch.offsets(), dtype=np.uint64)
        out = _vp()
        e = _Err()
        rc = lib().frhip_bcf_decode(self.ctx.handle, None if data is None else data._h, _ptr(offs), len(offs) - 1,
                                    start, samples, contigs, C.byref(out), C.byref(e))
        _raise(rc, e)
        return BcfRecordBatch(out, batch)

    def scan(self, batch: DeviceInflateBatch, start: int, samples: int, contigs: int):
        """(record offsets, tail)."""
        d = self.decode(batch, start, samples, contigs)
        return d.column_to_host("record_offsets"), d.tail()


class BcfFastqBatch:
    """A BCF stream and a bgzipped FASTQ stream decoded together
    (frhip_bcf_fastq_decode): `bcf` and `fastq` are the two scanners' column
    batches, `arena` the inflated bytes of both (the FASTQ part from
    `fastq_base`), which the column batches keep alive."""

    def __init__(self, arena, bcf, fastq, fastq_base, bcf_start, samples, contigs):
        self.arena = arena
        self.bcf = bcf
        self.fastq = fastq
        self.fastq_base = fastq_base
        self.bcf_start = bcf_start
        self.samples = samples
        self.contigs = contigs


def decode_bcf_fastq(context: "HipContext", bcf_compressed, bcf_spans, fastq_compressed,
                     fastq_spans) -> BcfFastqBatch:
    """One inflate launch over both streams' members, then the BCF and FASTQ
    boundary scans and columns (one BcfScanner::decode plus one
    FastqScanner::decode, crates/fritillaria-cuda/src/{bcf,fastq}.rs)."""
    both, sp, n_bcf = join_bcf_fastq(bcf_compressed, bcf_spans, fastq_compressed, fastq_spans)
    return decode_bcf_fastq_joined(context, both, sp, n_bcf)


def join_bcf_fastq(bcf_compressed, bcf_spans, fastq_compressed, fastq_spans):
    """(bytes, spans, n_bcf): the two BGZF streams back to back, with the FASTQ
    members' spans moved past the BCF bytes."""
    b = _u8(bcf_compressed)
    f = _u8(fastq_compressed)
    sb = _spans_array(bcf_spans).copy()
    sf = _spans_array(fastq_spans).copy()
    for k in ("compressed_offset", "payload_start"):
        sf[k] += np.uint64(b.size)
    return np.concatenate([b, f]), np.concatenate([sb, sf]), int(sb.size)


def decode_bcf_fastq_joined(context: "HipContext", both, spans, n_bcf: int) -> BcfFastqBatch:
    """decode_bcf_fastq over streams already joined by join_bcf_fastq."""
    a = _u8(both)
    sp = _spans_array(spans)
    offs = np.zeros(sp.size + 1, dtype=np.uint64)
    arena, bh, fh = _vp(), _vp(), _vp()
    base, start = _u64(), _u64()
    ns, nc = C.c_uint32(), C.c_uint32()
    e = _Err()
    rc = lib().frhip_bcf_fastq_decode(context.handle, _ptr(a), a.size, _ptr(sp), n_bcf, sp.size,
                                      C.byref(arena), _ptr(offs), C.byref(base), C.byref(start), C.byref(ns),
                                      C.byref(nc), C.byref(bh), C.byref(fh), C.byref(e))
    _raise(rc, e)
    src = DeviceInflateBatch(DeviceBuffer(arena), offs, sp["compressed_offset"].copy())
    return BcfFastqBatch(src, BcfRecordBatch(bh, src), FastqRecordBatch(fh, src), base.value, start.value,
                         ns.value, nc.value)


def bcf_parse_header(buf):
    """(records_start, samples, contigs) of a BCF header in host bytes."""
    a = _u8(buf)
    rs, ns, nc = _u64(), C.c_uint32(), C.c_uint32()
    e = _Err()
    _raise(lib().frhip_bcf_parse_header(_ptr(a), a.size, C.byref(rs), C.byref(ns), C.byref(nc), C.byref(e)), e)
    return rs.value, ns.value, nc.value


def bam_parse_header(buf):
    a = _u8(buf)
    rs = _u64()
    nref = C.c_uint32()
    e = _Err()
    rc = lib().frhip_bam_parse_header(_ptr(a), a.size, C.byref(rs), C.byref(nref), C.byref(e))
    _raise(rc, e)
    return rs.value, nref.value


# ---------------------------------------------------------------------------
# DeviceBgzfReader
# ---------------------------------------------------------------------------


@dataclass
class DeviceBatch:
    data: DeviceInflateBatch
    start: int


class DeviceBgzfReader:
    """DeviceBgzfReader over a BGZF stream in host memory."""

    def __init__(self, data, codec: HipCodec, blocks_per_batch: int = 256, read_ahead_members: int = 0,
                 read_ahead_bytes: int = 0):
        """read_ahead_*: bounds of the windows of members inflated in one
        launch (0: the defaults); the next window is inflated while batches
        are cut from the current one."""
        self._src = _u8(data)  # borrowed by the native reader: keep it alive
        self.codec = codec
        h = _vp()
        e = _Err()
        rc = lib().frhip_reader_open(codec.ctx.handle, _ptr(self._src), self._src.size,
                                     max(1, blocks_per_batch), C.byref(h), C.byref(e))
        _raise(rc, e)
        self._h = h
        if read_ahead_members or read_ahead_bytes:
            _raise(lib().frhip_reader_set_read_ahead(h, read_ahead_members, read_ahead_bytes, C.byref(e)), e)

    def codec_name(self) -> str:
        return self.codec.device_name()

    def next_batch(self) -> Optional[DeviceBatch]:
        has = C.c_int()
        out = _vp()
        offs = C.POINTER(_u64)()
        co = C.POINTER(_u64)()
        nb = _u64()
        start = _u64()
        e = _Err()
        rc = lib().frhip_reader_next_batch(self._h, C.byref(has), C.byref(out), C.byref(offs), C.byref(co),
                                           C.byref(nb), C.byref(start), C.byref(e))
        _raise(rc, e)
        if not has.value:
            return None
        n = nb.value
        offsets = np.ctypeslib.as_array(offs, shape=(n + 1,)).copy()
        coffsets = np.ctypeslib.as_array(co, shape=(n,)).copy()
        return DeviceBatch(DeviceInflateBatch(DeviceBuffer(out), offsets, coffsets), start.value)

    def carry_from(self, tail: int) -> None:
        e = _Err()
        rc = lib().frhip_reader_carry_from(self._h, tail, C.byref(e))
        _raise(rc, e)

    def close(self):
        if getattr(self, "_h", None):
            lib().frhip_reader_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# backend selection (crates/fritillaria/src/lib.rs:151-260)
# ---------------------------------------------------------------------------


class Backend(enum.Enum):
    AUTO = "auto"
    HIP = "hip"


def select_codec(backend: Backend = Backend.AUTO):
    """Returns (codec, resolved backend). An explicit HIP request is an assertion
    and raises when no usable device exists; AUTO resolves to HIP when a device is
    usable. This library ships no CPU codec (that is the host crate's CpuCodec),
    so AUTO on a machine without a device raises HipUnavailable rather than
    silently returning a different engine."""
    if backend in (Backend.AUTO, Backend.HIP):
        return HipCodec(), Backend.HIP
    raise ValueError(backend)


# ---------------------------------------------------------------------------
# batched, device-resident pipeline (include/frhip_pipeline.h)
# ---------------------------------------------------------------------------


class PipelineConfig(C.Structure):
    _fields_ = [("waves_per_cu", C.c_uint32), ("order", C.c_uint32), ("bam_window", C.c_uint32),
                ("use_graph", C.c_uint32), ("host_discovery", C.c_uint32), ("inflate_schedule", C.c_int32),
                ("soa", C.c_uint32)]


STAGES = ["discovery", "inflate", "verify", "speculate", "reconcile", "columns", "soa", "total"]


class PinnedBuffer:
    """Page-locked host memory (frhip_host_alloc) as a numpy uint8 array."""

    def __init__(self, nbytes: int):
        p = _vp()
        e = _Err()
        _raise(lib().frhip_host_alloc(nbytes, C.byref(p), C.byref(e)), e)
        self._p = p
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(nbytes, 1)).from_address(p.value))[:nbytes]

    def close(self):
        if getattr(self, "_p", None):
            self.array = None
            lib().frhip_host_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchedPipeline:
    """A whole BGZF BAM resident in HBM, decoded to columns in one enqueue."""

    def __init__(self, context: HipContext, file_bytes, *, waves_per_cu=0, order=0, bam_window=0,
                 use_graph=False, host_discovery=False, inflate_schedule=-1, soa=False):
        self.ctx = context
        self._file = _u8(file_bytes)
        cfg = PipelineConfig(waves_per_cu, order, bam_window, 1 if use_graph else 0, 1 if host_discovery else 0,
                             inflate_schedule, 1 if soa else 0)
        h = _vp()
        e = _Err()
        rc = lib().frhip_pipeline_create(context.handle, _ptr(self._file), self._file.size, C.byref(cfg),
                                         C.byref(h), C.byref(e))
        _raise(rc, e)
        self._h = h

    def run(self) -> None:
        e = _Err()
        _raise(lib().frhip_pipeline_run(self._h, C.byref(e)), e)

    def launch(self) -> None:
        e = _Err()
        _raise(lib().frhip_pipeline_launch(self._h, C.byref(e)), e)

    def upload(self, file_bytes) -> None:
        """Host-fed form: refills the resident file from host memory on the
        pipeline's copy stream (asynchronous; the next launch waits for it).
        `file_bytes` must stay alive until that launch."""
        a = _u8(file_bytes)
        self._upload_src = a
        e = _Err()
        _raise(lib().frhip_pipeline_upload(self._h, _ptr(a), a.size, C.byref(e)), e)

    def finish(self) -> None:
        e = _Err()
        _raise(lib().frhip_pipeline_finish(self._h, C.byref(e)), e)

    def members(self) -> int:
        return int(lib().frhip_pipeline_members(self._h))

    def inflated_len(self) -> int:
        return int(lib().frhip_pipeline_inflated_len(self._h))

    def compressed_len(self) -> int:
        return int(lib().frhip_pipeline_compressed_len(self._h))

    def records(self) -> int:
        return int(lib().frhip_pipeline_records(self._h))

    def records_start(self) -> int:
        return int(lib().frhip_pipeline_records_start(self._h))

    def tail(self) -> int:
        return int(lib().frhip_pipeline_tail(self._h))

    def fallbacks(self) -> int:
        return int(lib().frhip_pipeline_fallbacks(self._h))

    def device_discovery(self) -> bool:
        return bool(lib().frhip_pipeline_device_discovery(self._h))

    def inflated(self) -> DeviceBuffer:
        return DeviceBuffer(lib().frhip_pipeline_inflated(self._h), owned=False, keepalive=self)

    def column(self, name: str) -> DeviceBuffer:
        return DeviceBuffer(lib().frhip_pipeline_column(self._h, _COL_INDEX[name]), owned=False, keepalive=self)

    def column_to_host(self, name: str) -> np.ndarray:
        buf = self.column(name)
        dt = dict(BAM_COLUMNS)[name]
        out = np.empty(buf.byte_len() // np.dtype(dt).itemsize, dtype=dt)
        buf.copy_to_host(out)
        return out

    def columns_to_host(self) -> dict:
        return {name: self.column_to_host(name) for name, _ in BAM_COLUMNS}

    def soa_column(self, name: str) -> Optional[DeviceBuffer]:
        """SoA extras of the last run (pipelines created with soa=True)."""
        h = lib().frhip_pipeline_soa_column(self._h, _SOA_INDEX[name])
        return None if not h else DeviceBuffer(h, owned=False, keepalive=self)

    def soa_to_host(self) -> dict:
        out = {}
        for name, dt in SOA_COLUMNS:
            buf = self.soa_column(name)
            dt = np.dtype(dt)
            a = np.empty(buf.byte_len() // dt.itemsize, dtype=dt)
            buf.copy_to_host(a)
            out[name] = a
        return out

    def member_offsets(self):
        n = self.members()
        offs = np.zeros(n + 1, np.uint64)
        co = np.zeros(max(n, 1), np.uint64)
        e = _Err()
        _raise(lib().frhip_pipeline_member_offsets(self._h, _ptr(offs), _ptr(co), C.byref(e)), e)
        return offs, co[:n]

    def stage_ms(self) -> dict:
        out = (C.c_double * len(STAGES))()
        k = lib().frhip_pipeline_stage_ms(self._h, out, len(STAGES))
        return {STAGES[i]: out[i] for i in range(k)}

    def close(self):
        if getattr(self, "_h", None):
            lib().frhip_pipeline_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Batched virtual-offset range queries (include/frhip_serve.h)
# ---------------------------------------------------------------------------

QUERY_DTYPE = np.dtype([("vstart", "<u8"), ("vend", "<u8")])
QUERY_RESULT_DTYPE = np.dtype([("first", "<u8"), ("records", "<u8"), ("members", "<u8"),
                               ("inflated_bytes", "<u8"), ("status", "<i4"), ("reserved", "<u4"),
                               ("error_position", "<u8")])
TO_END = (1 << 64) - 1


def event_wait(event: int, stream: int) -> None:
    """Orders work later enqueued on `stream` after `event` (hipStreamWaitEvent)."""
    e = _Err()
    _raise(lib().frhip_event_wait(event, stream, C.byref(e)), e)


def event_synchronize(event: int) -> None:
    e = _Err()
    _raise(lib().frhip_event_synchronize(event, C.byref(e)), e)


def device_memory(ctx: HipContext):
    """(free, total) bytes of the context's device."""
    f, t = _u64(), _u64()
    e = _Err()
    _raise(lib().frhip_device_memory(ctx.handle, C.byref(f), C.byref(t), C.byref(e)), e)
    return f.value, t.value


CHUNK_DTYPE = np.dtype([("start", "<u8"), ("end", "<u8")])
REGION_DTYPE = np.dtype([("ref_id", "<i4"), ("reserved", "<u4"), ("start", "<u8"), ("end", "<u8")])
REGION_RESULT_DTYPE = np.dtype([("first", "<u8"), ("records", "<u8"), ("chunks", "<u8"), ("scanned", "<u8"),
                                ("members", "<u8"), ("inflated_bytes", "<u8"), ("status", "<i4"),
                                ("reserved", "<u4"), ("error_position", "<u8")])
INDEX_BAI, INDEX_CSI = 1, 2


class BinningIndex:
    """A BAI or CSI index read into host memory (include/frhip_index.h): the
    reference's bai::fs::read / csi::fs::read and BinningIndex::query
    (crates/fritillaria-csi/src/binning_index/index.rs:93-121)."""

    def __init__(self, data):
        d = _u8(data)
        h = _vp()
        e = _Err()
        _raise(lib().frhip_index_read(_ptr(d), d.size, C.byref(h), C.byref(e)), e)
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def format(self) -> str:
        return {INDEX_BAI: "bai", INDEX_CSI: "csi"}[lib().frhip_index_format(self._h, None, None)]

    @property
    def binning(self):
        """(min_shift, depth)."""
        a, b = C.c_uint32(), C.c_uint32()
        lib().frhip_index_format(self._h, C.byref(a), C.byref(b))
        return a.value, b.value

    def references(self) -> int:
        return int(lib().frhip_index_references(self._h))

    def unplaced_unmapped(self):
        n = _u64()
        return n.value if lib().frhip_index_unplaced_unmapped(self._h, C.byref(n)) else None

    def metadata(self, ref_id: int):
        out = np.zeros(4, np.uint64)
        return tuple(int(v) for v in out) if lib().frhip_index_metadata(self._h, ref_id, _ptr(out)) else None

    def last_first_record_start(self):
        v = _u64()
        return v.value if lib().frhip_index_last_first_record_start(self._h, C.byref(v)) else None

    def query(self, ref_id: int, start=None, end=None) -> np.ndarray:
        """Merged chunks (CHUNK_DTYPE) a query of [start, end] (1-based,
        inclusive; None = open) on reference ref_id reads."""
        n = _u64()
        e = _Err()
        _raise(lib().frhip_index_query(self._h, ref_id, start or 0, end or 0, None, 0, C.byref(n), C.byref(e)), e)
        out = np.zeros(n.value, CHUNK_DTYPE)
        if n.value:
            _raise(lib().frhip_index_query(self._h, ref_id, start or 0, end or 0, _ptr(out), n.value, C.byref(n),
                                           C.byref(e)), e)
        return out

    def close(self):
        if getattr(self, "_h", None):
            lib().frhip_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _regions(regions) -> np.ndarray:
    """(ref_id, start, end) triples, None for an open bound."""
    out = np.zeros(len(regions), REGION_DTYPE)
    for i, (rid, a, b) in enumerate(regions):
        out[i] = (rid, 0, a or 0, b or 0)
    return out


class QueryServer:
    """Serves [vstart, vend) virtual-offset ranges of a BAM resident in HBM, many
    queries per launch (noodles csi Query semantics,
    crates/fritillaria-csi/src/io/query.rs:66-82), and region queries through
    a BAI/CSI index (crates/fritillaria-bam/src/io/reader/query.rs:104-152)."""

    def __init__(self, context: HipContext, file_bytes, slots: int = 2):
        self.ctx = context
        self._file = _u8(file_bytes)
        h = _vp()
        e = _Err()
        _raise(lib().frhip_server_create_slots(context.handle, _ptr(self._file), self._file.size, slots,
                                               C.byref(h), C.byref(e)), e)
        self._h = h
        self._file = None
        self._launched = {}

    def slots(self) -> int:
        return int(lib().frhip_server_slots(self._h))

    def launch(self, slot: int, queries) -> None:
        """Enqueues the queries as one batch in `slot` and returns at once."""
        q = self._queries(queries)
        e = _Err()
        _raise(lib().frhip_server_launch(self._h, slot, _ptr(q), len(q), C.byref(e)), e)
        self._launched[slot] = len(q)

    def query(self, slot: int) -> bool:
        """True when the slot's batch is done on the device (or it has none)."""
        return bool(lib().frhip_server_query(self._h, slot))

    def finish(self, slot: int) -> np.ndarray:
        """Waits for the slot's batch; its QUERY_RESULT_DTYPE results."""
        res = np.zeros(self._launched.pop(slot, 0), QUERY_RESULT_DTYPE)
        e = _Err()
        _raise(lib().frhip_server_finish(self._h, slot, _ptr(res), C.byref(e)), e)
        return res

    def launch_regions(self, slot: int, index: BinningIndex, regions) -> None:
        """Maps the (ref_id, start, end) regions through `index` and enqueues
        their chunks as one batch in `slot`, with the overlap filter."""
        r = _regions(regions)
        e = _Err()
        _raise(lib().frhip_server_launch_regions(self._h, slot, index.handle, _ptr(r) if len(r) else None, len(r),
                                                 C.byref(e)), e)
        self._launched[slot] = len(r)

    def finish_regions(self, slot: int) -> np.ndarray:
        """Waits for the slot's region batch; its REGION_RESULT_DTYPE results."""
        res = np.zeros(self._launched.pop(slot, 0), REGION_RESULT_DTYPE)
        e = _Err()
        _raise(lib().frhip_server_finish_regions(self._h, slot, _ptr(res) if len(res) else None, C.byref(e)), e)
        return res

    def run_regions(self, index: BinningIndex, regions, slot: int = 0) -> np.ndarray:
        self.launch_regions(slot, index, regions)
        return self.finish_regions(slot)

    def selection(self, slot: int = 0) -> DeviceBuffer:
        """Kept rows (u64 indices into the slot's columns) of its last region batch."""
        return DeviceBuffer(lib().frhip_server_slot_selection(self._h, slot), owned=False, keepalive=self)

    def selection_to_host(self, slot: int = 0) -> np.ndarray:
        buf = self.selection(slot)
        out = np.empty(buf.byte_len() // 8, np.uint64)
        buf.copy_to_host(out)
        return out

    def batch_interval(self, slot: int):
        """(start, end) in ms from the server's creation of the slot's last batch on the device."""
        a, b = C.c_double(), C.c_double()
        lib().frhip_server_batch_interval(self._h, slot, C.byref(a), C.byref(b))
        return a.value, b.value

    def slot_column_to_host(self, slot: int, name: str) -> np.ndarray:
        buf = DeviceBuffer(lib().frhip_server_slot_column(self._h, slot, _COL_INDEX[name]), owned=False,
                           keepalive=self)
        dt = dict(BAM_COLUMNS)[name]
        out = np.empty(buf.byte_len() // np.dtype(dt).itemsize, dtype=dt)
        buf.copy_to_host(out)
        return out

    def members(self) -> int:
        return int(lib().frhip_server_members(self._h))

    def member_table(self):
        n = self.members()
        co = np.zeros(max(n, 1), np.uint64)
        up = np.zeros(n + 1, np.uint64)
        e = _Err()
        _raise(lib().frhip_server_member_table(self._h, _ptr(co), _ptr(up), C.byref(e)), e)
        return co[:n], up

    @staticmethod
    def _queries(queries) -> np.ndarray:
        q = np.zeros(len(queries), QUERY_DTYPE)
        for i, (a, b) in enumerate(queries):
            q[i] = (a, b)
        return q

    def footprint(self, queries) -> int:
        q = self._queries(queries)
        return int(lib().frhip_server_batch_footprint(self._h, _ptr(q), len(q)))

    def run(self, queries) -> np.ndarray:
        """Runs the (vstart, vend) queries as one batch; returns a QUERY_RESULT_DTYPE
        array. Columns of the batch are read with column()/column_to_host()."""
        q = self._queries(queries)
        res = np.zeros(len(q), QUERY_RESULT_DTYPE)
        e = _Err()
        _raise(lib().frhip_server_run(self._h, _ptr(q), len(q), _ptr(res), C.byref(e)), e)
        return res

    def column(self, name: str) -> DeviceBuffer:
        return DeviceBuffer(lib().frhip_server_column(self._h, _COL_INDEX[name]), owned=False, keepalive=self)

    def column_to_host(self, name: str) -> np.ndarray:
        buf = self.column(name)
        dt = dict(BAM_COLUMNS)[name]
        out = np.empty(buf.byte_len() // np.dtype(dt).itemsize, dtype=dt)
        buf.copy_to_host(out)
        return out

    def inflated(self) -> DeviceBuffer:
        return DeviceBuffer(lib().frhip_server_inflated(self._h), owned=False, keepalive=self)

    def close(self):
        if getattr(self, "_h", None):
            lib().frhip_server_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
