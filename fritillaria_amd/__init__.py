"""fritillaria_amd -- MI355X (gfx950) device path of fritillaria.

BGZF inflate + CRC32/ISIZE verification, BAM record boundaries and SoA columns,
FASTQ record boundaries and line-bound columns,
as hand-written HIP kernels behind the C ABI of ``include/frhip.h``. This
package is the binding; the native library is ``libfrhip.so``.
"""
from .hip import (  # noqa: F401
    BAM_COLUMNS, Backend, BamDecoder, ChecksumMismatch, DeviceBatch, DeviceBgzfReader, DeviceBuffer,
    DeviceInflateBatch, DeviceRecordBatch, FrError, HipCodec, HipContext, HipError, HipUnavailable,
    InflateBatch, InflateError, InvalidBlock, InvalidDeviceBatch, Malformed, SizeMismatch, SPAN_DTYPE,
    BatchedPipeline, FASTQ_COLUMNS, FastqRecordBatch, FastqScanner, SpeculativeScan, bam_parse_header,
    crc32_combine, device_is_available, discover, select_codec, QueryServer, QUERY_RESULT_DTYPE, TO_END,
    device_memory, BamSoa, SOA_COLUMNS, AUX_ENTRY_DTYPE, BCF_COLUMNS, BcfRecordBatch, BcfScanner, bcf_parse_header,
    BcfFastqBatch, decode_bcf_fastq, decode_bcf_fastq_joined, join_bcf_fastq,
    InflateTimings, DecodeTimings, event_wait, event_synchronize, PinnedBuffer,
    BinningIndex, IoError, CHUNK_DTYPE, REGION_DTYPE, REGION_RESULT_DTYPE,
    Dialect, TextScanner, TextRecordBatch, TEXT_COLUMNS, FastaCompactor, DeviceReference, FASTA_COLUMNS,
    HipCompressor, CompressedBatch, DeviceBgzfWriter, framed_size, MAX_COMPRESSIBLE_PAYLOAD, DEFAULT_PAYLOAD_SIZE,
    EOF_BLOCK,
)
