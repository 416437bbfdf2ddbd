/*
 * frhip.h -- C ABI of the MI355X (gfx950) device path of fritillaria.
 *
 * This is the boundary a `fritillaria-hip` crate binds in place of
 * `fritillaria-cuda`. Plain pointers and sizes only; every call is synchronous
 * on return unless stated otherwise, like the reference (each public call
 * synchronises before reading status: crates/fritillaria-cuda/src/backend.rs:463-582).
 *
 * Ownership follows crates/fritillaria-core/src/device.rs:1-40: a frhip_buf owns
 * device memory and frhip_buf_free releases it; decoded BAM columns point INTO
 * the inflated batch, which must outlive them.
 *
 * Return values: FRHIP_OK (0) or one of frhip_error_kind; details in *err when
 * err is not NULL.
 */
#ifndef FRHIP_H
#define FRHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* fritillaria_core::Error (crates/fritillaria-core/src/error.rs:14-70). */
typedef enum {
    FRHIP_OK = 0,
    FRHIP_E_INVALID_BLOCK = 1,        /* InvalidBlock{offset, reason} */
    FRHIP_E_CHECKSUM_MISMATCH = 2,    /* ChecksumMismatch{offset, expected, actual} */
    FRHIP_E_SIZE_MISMATCH = 3,        /* SizeMismatch{offset, expected, actual} */
    FRHIP_E_INFLATE = 4,              /* Inflate{offset, reason}; code = FR_ERR_* status */
    FRHIP_E_MALFORMED = 5,            /* Malformed{format, position(offset), reason} */
    FRHIP_E_INVALID_DEVICE_BATCH = 6, /* InvalidDeviceBatch{reason} */
    FRHIP_E_UNAVAILABLE = 7,          /* CudaUnavailable: no usable device */
    FRHIP_E_DEVICE = 8,               /* Cuda(String): a runtime call failed */
    FRHIP_E_IO = 9                    /* Io */
} frhip_error_kind;

typedef struct {
    int32_t kind;         /* frhip_error_kind */
    uint32_t code;        /* device status for FRHIP_E_INFLATE / FRHIP_E_MALFORMED */
    uint64_t offset;      /* compressed offset of the member, or position for Malformed */
    uint64_t expected;
    uint64_t actual;
    char format[16];      /* Malformed: "bam", "bgzf", "bcf", "fastq", "bam header" */
    char reason[200];
} frhip_error;

/* BlockSpan (crates/fritillaria-core/src/codec.rs:17-31); usize fields widened to u64. */
typedef struct {
    uint64_t compressed_offset;
    uint64_t payload_start;
    uint64_t payload_len;
    uint32_t isize;
    uint32_t crc32;
} frhip_block_span;

/* InflateTimings (crates/fritillaria-cuda/src/backend.rs:136-177): host
   wall-clock per phase with the phases separated by stream synchronisations,
   so the sum bounds the untimed call from above; the timed calls ADD to every
   field (accumulate over batches). */
typedef struct {
    double plan_ms;     /* bounds checks and the ISIZE prefix sum */
    double upload_ms;   /* compressed batch and member plan to the device */
    double kernel_ms;   /* inflate with the folded-in CRC32 */
    double download_ms; /* inflated bytes to the host (host path only) */
    double verify_ms;   /* status / ISIZE / CRC checks and the offsets */
    uint64_t blocks, compressed_bytes, inflated_bytes;
} frhip_inflate_timings;

/* DecodeTimings (crates/fritillaria-cuda/src/bam.rs:55-101), same rules. */
typedef struct {
    double setup_ms;     /* block starts uploaded, scratch ready */
    double scan_ms;      /* boundary speculation and the per-member exact walks */
    double reconcile_ms; /* the record chain across members */
    double totals_ms;    /* the one read-back of {records, tail, status} */
    double emit_ms;      /* record (and virtual) offsets */
    double decode_ms;    /* the field columns */
    uint64_t batches, blocks, records, bytes;
} frhip_decode_timings;

typedef struct frhip_ctx frhip_ctx;
typedef struct frhip_buf frhip_buf;
typedef struct frhip_bam_batch frhip_bam_batch;
typedef struct frhip_fastq_batch frhip_fastq_batch;
typedef struct frhip_bcf_batch frhip_bcf_batch;
typedef struct frhip_bam_soa frhip_bam_soa;
typedef struct frhip_reader frhip_reader;

/* ---- device and context ------------------------------------------------ */

/* device_is_available (crates/fritillaria-cuda/src/lib.rs:136-145): 1 if a gfx950 device can be used. */
int frhip_device_available(void);

/* CudaContext::new(ordinal) / with_context(ctx, stream) (backend.rs:390-436;
   lib.rs:194-201). `stream` is a hipStream_t owned by the caller, or NULL for a
   stream of the context's own. Kernels are precompiled code objects: no JIT. */
int frhip_ctx_create(int ordinal, void *stream, frhip_ctx **out, frhip_error *err);
/* Releases the creator's reference. Readers and pipelines hold their own, so the
   context outlives them whatever order they are freed in (the reference shares
   its context as Arc<CudaContext>). */
void frhip_ctx_destroy(frhip_ctx *ctx);
int frhip_ctx_ordinal(const frhip_ctx *ctx);
/* hipStream_t the context launches on. */
void *frhip_ctx_stream(const frhip_ctx *ctx);
/* Waits for all work on the context's device. */
int frhip_ctx_synchronize(frhip_ctx *ctx, frhip_error *err);
/* Bytes of each member whose offsets are tried as record-start candidates by the
   BAM boundary scan (default 2048). A tuning knob: results do not depend on it.
   No reference counterpart; the reference speculates on member starts only. */
int frhip_ctx_set_bam_window(frhip_ctx *ctx, uint32_t window);
/* Inflate kernel variant (output ring size per member x waves per workgroup),
   an index into the table frhip_inflate_schedules describes; the default is
   16 (two-phase, 4 KiB ring, whose small batches run the one-wave
   lane-speculative form). Results do not depend on it. No reference
   counterpart. */
int frhip_ctx_set_inflate_schedule(frhip_ctx *ctx, int schedule);
/* With a two-phase schedule (token form 3: one member per lane, then one wave
   per member), batches of fewer than min_members members run the one-wave
   lane-speculative form of the same ring size instead, which finishes a small
   batch sooner (default: the one-wave form's resident waves on the device,
   n_cu x 16 = 4096 on MI355X). Results do not depend on it. No reference
   counterpart. */
int frhip_ctx_set_two_phase_min_members(frhip_ctx *ctx, uint32_t min_members);
/* The current value of that threshold. No reference counterpart. */
uint32_t frhip_ctx_two_phase_min_members(const frhip_ctx *ctx);
/* Compute units of the context's device (per-CU throughput figures). No
   reference counterpart. */
uint32_t frhip_ctx_compute_units(const frhip_ctx *ctx);
/* Writes up to cap entries of (ring bytes, waves per workgroup, resident waves
   per CU, token form: 0 symbols straight into the ring, 1 token lists resolved
   64 at a time, 2 lane-speculative token lists, 3 two-phase); returns the
   number of schedules. Any array may be NULL. */
int frhip_inflate_schedules(uint32_t *ring, uint32_t *waves_per_group, uint32_t *waves_per_cu, uint32_t *tokens,
                            int cap);
/* Codec names of CudaCodec (lib.rs:152-294): host path and device path. */
const char *frhip_codec_name(void);        /* "hip" */
const char *frhip_device_codec_name(void); /* "hip-device" */

/* ---- device buffers: DeviceAlloc (crates/fritillaria-core/src/device.rs:51-75) ---- */

uint64_t frhip_buf_len(const frhip_buf *b);
int frhip_buf_ordinal(const frhip_buf *b);
void *frhip_buf_device_ptr(const frhip_buf *b);
/* copy_to_host: rejects dst_len != byte_len with FRHIP_E_INVALID_DEVICE_BATCH and synchronises. */
int frhip_buf_copy_to_host(const frhip_buf *b, void *dst, uint64_t dst_len, frhip_error *err);
void frhip_buf_free(frhip_buf *b);
/* CudaCodec::ready_event (crates/fritillaria-cuda/src/lib.rs:207-224): the
   hipEvent_t recorded on the producing context's stream after the kernels
   that wrote this buffer (inflated batches, BAM columns and SoA columns), for
   a consumer on its own stream to order against (frhip_event_wait) instead of
   synchronising the host. NULL for an empty buffer or one with no producer
   event (frhip_buf_alloc, pipeline and server views). Owned by the buffer. */
void *frhip_buf_ready_event(const frhip_buf *b);
/* hipStreamWaitEvent(stream, event): work enqueued on `stream` afterwards runs
   after the event's work. */
int frhip_event_wait(void *event, void *stream, frhip_error *err);
/* Blocks the host until the event's work is done. */
int frhip_event_synchronize(void *event, frhip_error *err);
/* No reference counterpart. A device buffer of len bytes (contents undefined),
   and a plain device-to-device copy kernel over two buffers of equal length
   (synchronous). The copy reads and writes every byte once, with 16-byte
   vector accesses, which makes it the calibration kernel for HBM byte counters. */
int frhip_buf_alloc(frhip_ctx *ctx, uint64_t len, frhip_buf **out, frhip_error *err);
int frhip_buf_copy_device(frhip_ctx *ctx, frhip_buf *dst, const frhip_buf *src, frhip_error *err);
/* No reference counterpart. Copies src[0..len) into the buffer at `offset`
   (synchronous); offset + len beyond byte_len is FRHIP_E_INVALID_DEVICE_BATCH. */
int frhip_buf_copy_from_host(frhip_buf *b, uint64_t offset, const void *src, uint64_t len, frhip_error *err);

/* No reference counterpart. Pinned (page-locked) host memory, which host to
   device copies read at full PCIe bandwidth. */
int frhip_host_alloc(uint64_t len, void **out, frhip_error *err);
void frhip_host_free(void *p);

/* ---- BGZF member discovery (crates/fritillaria-bgzf/src/discover.rs:51-113) ---- */

/* Host walk of member headers in buf[0..len), base = absolute offset of buf[0].
   Writes up to cap spans; *n = spans written; *position = resume point.
   Returns FRHIP_E_INVALID_BLOCK on a corrupt header (then *n spans precede it). */
int frhip_discover(const uint8_t *buf, uint64_t len, uint64_t base, frhip_block_span *spans,
                   uint64_t cap, uint64_t *n, uint64_t *position, frhip_error *err);

/* ---- inflate: BlockCodec / DeviceBlockCodec ----------------------------- */

/* DeviceBlockCodec::inflate_batch_device (crates/fritillaria-core/src/device.rs:306-324;
   crates/fritillaria-cuda/src/backend.rs:536-582). Uploads `batch`, inflates every
   span, verifies status -> ISIZE -> CRC32 in member order (backend.rs:221-256).
   On success *out_data owns sum(isize) bytes (NULL when that is 0: an empty batch
   owns no allocation) and out_offsets[0..n] holds the member boundaries.
   On failure nothing is returned. */
int frhip_inflate_batch_device(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                               const frhip_block_span *spans, uint64_t n, frhip_buf **out_data,
                               uint64_t *out_offsets, frhip_error *err);

/* BlockCodec::inflate_batch (crates/fritillaria-core/src/codec.rs:159-173): the same,
   then a download into out[0..out_cap). */
int frhip_inflate_batch(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                        const frhip_block_span *spans, uint64_t n, uint8_t *out, uint64_t out_cap,
                        uint64_t *out_offsets, frhip_error *err);

/* CudaContext::inflate_batch_device_timed / inflate_batch_timed
   (crates/fritillaria-cuda/src/backend.rs:463-582): the calls above, adding
   each phase's time to *t. */
int frhip_inflate_batch_device_timed(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                                     const frhip_block_span *spans, uint64_t n, frhip_buf **out_data,
                                     uint64_t *out_offsets, frhip_inflate_timings *t, frhip_error *err);
int frhip_inflate_batch_timed(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                              const frhip_block_span *spans, uint64_t n, uint8_t *out, uint64_t out_cap,
                              uint64_t *out_offsets, frhip_inflate_timings *t, frhip_error *err);

/* No counterpart in the reference: for callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc). The same inflate, verification order and error results as
   frhip_inflate_batch_device, but nothing is allocated: the caller passes the
   compressed batch in device memory, d_batch[0..batch_len) followed by 4096
   zero bytes (checked: FRHIP_E_INVALID_DEVICE_BATCH otherwise), and the output
   d_out[0..out_cap) with out_cap >= sum(isize); out_offsets[0..n] receives the
   member boundaries. The member plan and status words live in a fixed device
   array of the library, one per GPU, so n is at most
   FRHIP_NOALLOC_MAX_MEMBERS (larger batches: FRHIP_E_INVALID_DEVICE_BATCH).
   Calls that overlap on one GPU share that array and race. Synchronous. */
#define FRHIP_NOALLOC_MAX_MEMBERS 4096
int frhip_inflate_batch_device_noalloc(frhip_ctx *ctx, const void *d_batch, uint64_t batch_len,
                                       const frhip_block_span *spans, uint64_t n, void *d_out, uint64_t out_cap,
                                       uint64_t *out_offsets, frhip_error *err);

/* crc32_blocks (crates/fritillaria-cuda/src/backend.rs:737-810): CRC32 of byte
   ranges data[off[i] .. off[i]+len[i]) of a device buffer. */
int frhip_crc32_blocks(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *offsets,
                       const uint32_t *lengths, uint64_t n, uint32_t *crc_out, frhip_error *err);

/* zlib-compatible crc32_combine on the host, exported for tests of the CRC fold. */
uint32_t frhip_crc32_combine(uint32_t crc1, uint32_t crc2, uint64_t len2);

/* ---- BAM: BamDecoder::decode (crates/fritillaria-cuda/src/bam.rs:161-187, 324-574) ---- */

/* Column indices of DeviceRecordBatch, in the order of
   crates/fritillaria-bam/src/columnar/device.rs:95-110, plus the build's own. */
typedef enum {
    FRHIP_BAM_REFERENCE_SEQUENCE_ID = 0, /* i32 */
    FRHIP_BAM_POSITION = 1,              /* i32, 0-based */
    FRHIP_BAM_FLAGS = 2,                 /* u16 */
    FRHIP_BAM_MAPPING_QUALITY = 3,       /* u8 */
    FRHIP_BAM_SEQUENCE_LEN = 4,          /* u32 */
    FRHIP_BAM_MATE_REFERENCE_SEQUENCE_ID = 5, /* i32 */
    FRHIP_BAM_MATE_POSITION = 6,         /* i32 */
    FRHIP_BAM_TEMPLATE_LENGTH = 7,       /* i32 */
    FRHIP_BAM_RECORD_OFFSETS = 8,        /* u64, into the inflated batch */
    FRHIP_BAM_CIGAR_START = 9,           /* u32, relative to the record start */
    FRHIP_BAM_SEQUENCE_START = 10,       /* u32 */
    FRHIP_BAM_QUALITY_START = 11,        /* u32 */
    FRHIP_BAM_AUX_START = 12,            /* u32 */
    FRHIP_BAM_RECORD_END = 13,           /* u32 */
    FRHIP_BAM_VIRTUAL_OFFSETS = 14,      /* u64 (coffset << 16 | uoffset); only when coffsets were given */
    FRHIP_BAM_N_COLUMNS = 15
} frhip_bam_column;

/* Discovers record boundaries from `start` and decodes the columns.
   `data` is an inflated batch (NULL or empty: no records, tail = start),
   block_offsets its n_blocks+1 member boundaries. `coffsets` (n_blocks entries,
   may be NULL) are the members' compressed file offsets for the virtual-offset
   column. Malformed records yield FRHIP_E_MALFORMED with offset = the position
   of blocked.rs's reconcile. */
int frhip_bam_decode(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *block_offsets,
                     uint64_t n_blocks, const uint64_t *coffsets, uint64_t start,
                     frhip_bam_batch **out, frhip_error *err);
/* BamDecoder::decode_timed (crates/fritillaria-cuda/src/bam.rs:161-187): the
   same, adding each phase's time to *t. */
int frhip_bam_decode_timed(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *block_offsets,
                           uint64_t n_blocks, const uint64_t *coffsets, uint64_t start, frhip_bam_batch **out,
                           frhip_decode_timings *t, frhip_error *err);
uint64_t frhip_bam_batch_len(const frhip_bam_batch *b);
uint64_t frhip_bam_batch_tail(const frhip_bam_batch *b);
/* Members the true chain did not enter at a speculated candidate (a diagnostic,
   like BlockedScan::fallbacks, blocked.rs:80-88). */
uint64_t frhip_bam_batch_fallbacks(const frhip_bam_batch *b);
/* Column i, or NULL if absent. Owned by the batch. */
const frhip_buf *frhip_bam_batch_column(const frhip_bam_batch *b, int column);
void frhip_bam_batch_free(frhip_bam_batch *b);

/* Structure-of-arrays extras over decoded records. No device counterpart in the
   reference, which reads these per record on the host
   (crates/fritillaria-bam/src/columnar/record.rs:220-361, aux.rs):
   - CIGAR: the ops of Record::cigar_resolved, so a long-CIGAR placeholder
     ("<l_seq>S <n>N", two ops) is replaced by its CG:B:I array; record i owns
     ops [CIGAR_OFFSETS[i], CIGAR_OFFSETS[i+1]), u32 len << 4 | op each;
   - SEQUENCE: Record::sequence_packed (4-bit, high nibble first) back to back;
   - AUX: one 16-byte entry per aux field, in order: tag[2], type, B subtype
     (else 0), u32 value offset from the record start, u32 element count
     (scalars 1; Z/H string length without the NUL; B array count), u32 value
     bytes (Z/H with the NUL).
   Offset columns have n+1 u64 entries. A record with a malformed aux field, or
   with the placeholder and no CG:B:I, is FRHIP_E_MALFORMED at its offset. */
typedef enum {
    FRHIP_SOA_CIGAR_OFFSETS = 0,    /* u64, n + 1 */
    FRHIP_SOA_CIGAR_OPS = 1,        /* u32 */
    FRHIP_SOA_SEQUENCE_OFFSETS = 2, /* u64, n + 1 (bytes) */
    FRHIP_SOA_SEQUENCE_PACKED = 3,  /* u8 */
    FRHIP_SOA_AUX_OFFSETS = 4,      /* u64, n + 1 (entries) */
    FRHIP_SOA_AUX_ENTRIES = 5,      /* 16 bytes each */
    FRHIP_SOA_N_COLUMNS = 6
} frhip_soa_column;
int frhip_bam_soa_build(frhip_ctx *ctx, const frhip_buf *data, const frhip_bam_batch *records, frhip_bam_soa **out,
                  frhip_error *err);
uint64_t frhip_bam_soa_len(const frhip_bam_soa *s);
const frhip_buf *frhip_bam_soa_column(const frhip_bam_soa *s, int column);
void frhip_bam_soa_free(frhip_bam_soa *s);

/* BAM header (crates/fritillaria-bam/src/columnar/header.rs:43-103) on host bytes. */
int frhip_bam_parse_header(const uint8_t *buf, uint64_t len, uint64_t *records_start,
                           uint32_t *n_ref, frhip_error *err);

/* ---- FASTQ: FastqScanner::{scan, decode} (crates/fritillaria-cuda/src/fastq.rs:94-135, 469-560) ---- */

/* Column indices of the FASTQ DeviceRecordBatch, in the order of
   crates/fritillaria-fastq/src/columnar/device.rs:44-68. Bounds are relative
   to the record start; the definition line is record_offsets + 1 ..
   sequence_start - 1. */
typedef enum {
    FRHIP_FASTQ_RECORD_OFFSETS = 0, /* u64, into the inflated batch */
    FRHIP_FASTQ_SEQUENCE_START = 1, /* u32 */
    FRHIP_FASTQ_PLUS_START = 2,     /* u32 */
    FRHIP_FASTQ_QUALITY_START = 3,  /* u32 */
    FRHIP_FASTQ_SEQUENCE_LEN = 4,   /* u32, also the quality length */
    FRHIP_FASTQ_RECORD_END = 5,     /* u32 */
    FRHIP_FASTQ_N_COLUMNS = 6
} frhip_fastq_column;

/* Record boundaries from `start` and the line-bound columns of an inflated
   batch (NULL or empty, or start >= its length: no records, tail =
   min(start, length), as fastq.rs:469-479). The records are those of the
   serial walk of crates/fritillaria-fastq/src/columnar/record.rs:145-176:
   newlines between records are skipped and the first position that is not a
   whole record is the tail when it holds '@'; any other byte there is
   FRHIP_E_MALFORMED (format "fastq", offset = that position). */
int frhip_fastq_decode(frhip_ctx *ctx, const frhip_buf *data, uint64_t start, frhip_fastq_batch **out,
                       frhip_error *err);
uint64_t frhip_fastq_batch_len(const frhip_fastq_batch *b);
uint64_t frhip_fastq_batch_tail(const frhip_fastq_batch *b);
/* Line starts that begin a whole record (SpeculativeScan::survivors,
   crates/fritillaria-fastq/src/columnar/speculative.rs:86-92); equal to the
   record count exactly when the records tile the batch (Proof::Tiled). */
uint64_t frhip_fastq_batch_survivors(const frhip_fastq_batch *b);
const frhip_buf *frhip_fastq_batch_column(const frhip_fastq_batch *b, int column);
void frhip_fastq_batch_free(frhip_fastq_batch *b);

/* ---- Text framing: TextScanner::scan (crates/fritillaria-cuda/src/text.rs:49-66, 196-375) ---- */

typedef struct frhip_text_batch frhip_text_batch;

/* Column indices of the text DeviceRecordBatch
   (crates/fritillaria-text/src/columnar/device.rs:5-11); every column is u64.
   field_starts has len + 1 entries (none without records): record r's tabs
   are tabs[field_starts[r] .. field_starts[r + 1]). */
typedef enum {
    FRHIP_TEXT_RECORD_OFFSETS = 0,
    FRHIP_TEXT_RECORD_ENDS = 1,   /* the record's newline, excluded */
    FRHIP_TEXT_HEADER_OFFSETS = 2,
    FRHIP_TEXT_TABS = 3,
    FRHIP_TEXT_FIELD_STARTS = 4,
    FRHIP_TEXT_N_COLUMNS = 5
} frhip_text_column;

/* A dialect without a secondary header byte. */
#define FRHIP_TEXT_NO_SECONDARY 0x100u

/* Lines and fields of the inflated batch from `start`, as
   crates/fritillaria-text/src/columnar/batch.rs:56-86, 236-258 decodes them: a
   line starts at `start` or after a newline; blank lines are skipped; a line
   whose first byte is `comment` (or `secondary`, when below 256) is a header,
   any other a record; a final line without a newline is not decoded and its
   start is the tail (else the tail is the batch length). NULL or empty data,
   or start >= its length: no lines, tail = min(start, length), as
   text.rs:209-215. Quotes are data (no format here quotes a tab). */
int frhip_text_scan(frhip_ctx *ctx, const frhip_buf *data, uint64_t start, uint32_t comment, uint32_t secondary,
                    frhip_text_batch **out, frhip_error *err);
uint64_t frhip_text_batch_len(const frhip_text_batch *b);
uint64_t frhip_text_batch_headers(const frhip_text_batch *b);
uint64_t frhip_text_batch_tail(const frhip_text_batch *b);
const frhip_buf *frhip_text_batch_column(const frhip_text_batch *b, int column);
void frhip_text_batch_free(frhip_text_batch *b);

/* ---- FASTA: FastaCompactor::compact (crates/fritillaria-cuda/src/fasta.rs:45-67, 203-304) ---- */

typedef struct frhip_fasta_reference frhip_fasta_reference;

/* Columns of the DeviceReference (crates/fritillaria-fasta/src/columnar/device.rs:5-11). */
typedef enum {
    FRHIP_FASTA_SEQUENCE = 0,         /* u8: every contig's bases, back to back */
    FRHIP_FASTA_SEQUENCE_OFFSETS = 1, /* u64: each contig's first base in SEQUENCE */
    FRHIP_FASTA_SEQUENCE_LENGTHS = 2, /* u64 */
    FRHIP_FASTA_RECORD_OFFSETS = 3,   /* u64: the contig's '>' in the batch */
    FRHIP_FASTA_N_COLUMNS = 4
} frhip_fasta_column;

/* Contigs are '>' at a line start, from the first one on; a contig's
   sequence runs from after its definition line to the next contig, and its
   bases are those bytes without the newlines
   (crates/fritillaria-fasta/src/columnar/record.rs:58-126); a last '>' whose
   definition line has no newline is not a contig. NULL or empty data, or no
   '>' at a line start: no contigs. */
int frhip_fasta_compact(frhip_ctx *ctx, const frhip_buf *data, frhip_fasta_reference **out, frhip_error *err);
uint64_t frhip_fasta_contigs(const frhip_fasta_reference *r);
uint64_t frhip_fasta_total_bases(const frhip_fasta_reference *r);
const frhip_buf *frhip_fasta_reference_column(const frhip_fasta_reference *r, int column);
void frhip_fasta_free(frhip_fasta_reference *r);

/* ---- BCF: BcfScanner::{scan, decode} (crates/fritillaria-cuda/src/bcf.rs:94-143, 451-470) ---- */

/* Column indices of the BCF DeviceRecordBatch, in the order of
   crates/fritillaria-bcf/src/columnar/device.rs:59-98. Bounds are relative to
   the record start. */
typedef enum {
    FRHIP_BCF_CHROMOSOME_ID = 0,   /* i32 */
    FRHIP_BCF_POSITION = 1,        /* i32, 0-based */
    FRHIP_BCF_REFERENCE_SPAN = 2,  /* i32 (rlen) */
    FRHIP_BCF_QUALITY = 3,         /* u32, the float's bits */
    FRHIP_BCF_INFO_COUNT = 4,      /* u16 */
    FRHIP_BCF_ALLELE_COUNT = 5,    /* u16 */
    FRHIP_BCF_SAMPLE_COUNT = 6,    /* u32 */
    FRHIP_BCF_FORMAT_COUNT = 7,    /* u8 */
    FRHIP_BCF_RECORD_OFFSETS = 8,  /* u64, into the inflated batch */
    FRHIP_BCF_ALLELES_START = 9,   /* u32 */
    FRHIP_BCF_FILTERS_START = 10,  /* u32 */
    FRHIP_BCF_INFO_START = 11,     /* u32 */
    FRHIP_BCF_GENOTYPES_START = 12,/* u32 */
    FRHIP_BCF_RECORD_END = 13,     /* u32 */
    FRHIP_BCF_N_COLUMNS = 14
} frhip_bcf_column;

/* Record boundaries from `start` and the BCF columns of an inflated batch
   (NULL or empty: no records, tail = start). block_offsets are its n_blocks+1
   member boundaries; samples and contigs are the header's counts
   (frhip_bcf_parse_header). The records are those of scan_records
   (crates/fritillaria-bcf/src/columnar/record.rs:60-95): fewer than 8 bytes
   left or a record running past the batch is the tail; l_shared below 24 is
   FRHIP_E_MALFORMED at that record; a record whose ID, alleles or FILTER is not
   a valid typed value is FRHIP_E_MALFORMED at the record's offset (format "bcf"). */
int frhip_bcf_decode(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *block_offsets, uint64_t n_blocks,
                     uint64_t start, uint32_t samples, uint32_t contigs, frhip_bcf_batch **out, frhip_error *err);
uint64_t frhip_bcf_batch_len(const frhip_bcf_batch *b);
uint64_t frhip_bcf_batch_tail(const frhip_bcf_batch *b);
/* Members whose entry was not a speculated candidate (diagnostic). */
uint64_t frhip_bcf_batch_fallbacks(const frhip_bcf_batch *b);
const frhip_buf *frhip_bcf_batch_column(const frhip_bcf_batch *b, int column);
void frhip_bcf_batch_free(frhip_bcf_batch *b);
/* BCF header (crates/fritillaria-bcf/src/columnar/header.rs:43-140) on host
   bytes: where records start, samples (#CHROM columns after FORMAT) and contigs
   (##contig lines with an ID). */
int frhip_bcf_parse_header(const uint8_t *buf, uint64_t len, uint64_t *records_start, uint32_t *samples,
                           uint32_t *contigs, frhip_error *err);

/* A BCF stream and a bgzipped FASTQ stream decoded together. No single
   counterpart in the reference: it stands for one BcfScanner::decode
   (crates/fritillaria-cuda/src/bcf.rs:94-143) plus one FastqScanner::decode
   (crates/fritillaria-cuda/src/fastq.rs:94-135), each over its stream inflated
   as one batch (crates/fritillaria-cuda/src/backend.rs:536-582).
   batch[0..batch_len) holds both streams' compressed bytes; spans[0..n_bcf)
   are the BCF members and spans[n_bcf..n) the FASTQ members (BGZF members are
   independent, so the two files may simply be concatenated). All n members are
   inflated in ONE launch of the persistent inflate and verified in member
   order (status, then ISIZE, then CRC, as frhip_inflate_batch_device) into one
   arena: the BCF bytes at 0, the FASTQ bytes from *fastq_base (the BCF size
   rounded up to 4 KiB). offsets[n+1] receives each member's output end
   (offsets[0] = 0). The BCF header is parsed from the inflated bytes
   (*bcf_start, *samples, *contigs); then the BCF boundaries and columns from
   *bcf_start and the FASTQ boundaries and columns from the FASTQ part's first
   byte are found as by frhip_bcf_decode and frhip_fastq_decode, with their
   errors (positions relative to each part). On success the caller owns
   *arena, *bcf and *fastq; the columns index their part of the arena. */
int frhip_bcf_fastq_decode(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                           const frhip_block_span *spans, uint64_t n_bcf, uint64_t n,
                           frhip_buf **arena, uint64_t *offsets, uint64_t *fastq_base,
                           uint64_t *bcf_start, uint32_t *samples, uint32_t *contigs,
                           frhip_bcf_batch **bcf, frhip_fastq_batch **fastq, frhip_error *err);

/* ---- DeviceBgzfReader (crates/fritillaria-bgzf/src/device_read.rs:47-280) ---- */

/* A reader over a BGZF stream held in host memory [data, data+len) (borrowed for
   the reader's lifetime). blocks_per_batch budgets compressed bytes at 64 KiB per
   member, as with_blocks_per_batch (device_read.rs:113-116); 0 = 256. */
int frhip_reader_open(frhip_ctx *ctx, const uint8_t *data, uint64_t len, uint64_t blocks_per_batch,
                      frhip_reader **out, frhip_error *err);
/* next_batch: out_data, out_offsets and n_blocks describe the inflated batch
   (ownership of *out_data passes to the caller), *start = DeviceBatch::start,
   *coffsets points at the members' compressed offsets (valid until the next call).
   Returns FRHIP_OK with *has_batch = 0 at end of stream. */
int frhip_reader_next_batch(frhip_reader *r, int *has_batch, frhip_buf **out_data,
                            const uint64_t **out_offsets, const uint64_t **coffsets,
                            uint64_t *n_blocks, uint64_t *start, frhip_error *err);
/* carry_from (device_read.rs:155-182). */
int frhip_reader_carry_from(frhip_reader *r, uint64_t tail, frhip_error *err);
/* Bounds of the windows of members the reader inflates in one launch and
   cuts its batches from: at most `members` members and `bytes` compressed
   bytes each (0: max(32768, 16 x blocks_per_batch) members, 1 GiB). The next
   window is uploaded and inflated on the reader's own stream, by a helper
   thread, while batches are cut from the current one. Setting them discards windows already inflated. No
   reference counterpart: the reference inflates each batch on its own, and
   the batches and their errors are the same either way. */
int frhip_reader_set_read_ahead(frhip_reader *r, uint64_t members, uint64_t bytes, frhip_error *err);
void frhip_reader_free(frhip_reader *r);

#ifdef __cplusplus
}
#endif
#endif
