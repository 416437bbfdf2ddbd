/*
 * oracle.h -- test-only CPU restatement of fritillaria's device path.
 * Nothing in the product library includes this header or links the oracle.
 */
#ifndef FR_ORACLE_H
#define FR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device status codes of crates/fritillaria-cuda/kernels/inflate.cu:29-39. */
enum {
    OR_OK = 0,
    OR_ERR_TRUNCATED = 1,
    OR_ERR_BLOCK_TYPE = 2,
    OR_ERR_STORED_LEN = 3,
    OR_ERR_OUTPUT_OVERFLOW = 4,
    OR_ERR_BAD_CODE = 5,
    OR_ERR_BAD_DISTANCE = 6,
    OR_ERR_BAD_TABLE = 7,
    OR_ERR_BAD_LENGTH_REPEAT = 8,
    OR_ERR_BAD_COUNTS = 9
};

/* Header parse failures of crates/fritillaria-bgzf/src/block.rs:99-144, in check order. */
enum {
    OR_BLK_TRUNCATED_HEADER = 1,
    OR_BLK_BAD_MAGIC = 2,
    OR_BLK_BAD_METHOD = 3,
    OR_BLK_NO_FEXTRA = 4,
    OR_BLK_TRUNCATED_EXTRA = 5,
    OR_BLK_NO_BC = 6,
    OR_BLK_TOO_BIG = 7,
    OR_BLK_TOO_SMALL = 8,
    OR_BLK_ISIZE_TOO_BIG = 9
};

/* fritillaria_core::Error variants (crates/fritillaria-core/src/error.rs:14-70). */
enum {
    OR_E_OK = 0,
    OR_E_INVALID_BLOCK = 1,
    OR_E_CHECKSUM_MISMATCH = 2,
    OR_E_SIZE_MISMATCH = 3,
    OR_E_INFLATE = 4,
    OR_E_MALFORMED = 5,
    OR_E_INVALID_DEVICE_BATCH = 6
};

typedef struct {
    uint64_t compressed_offset;
    uint64_t payload_start;
    uint64_t payload_len;
    uint32_t isize;
    uint32_t crc32;
} or_span_t;

typedef struct {
    int32_t kind;
    uint32_t code;
    uint64_t offset;
    uint64_t expected;
    uint64_t actual;
} or_error_t;

uint32_t or_crc32(const uint8_t *p, uint64_t len);
int or_inflate_member(const uint8_t *in, uint32_t in_len, uint8_t *out, uint32_t out_cap,
                      uint32_t *produced);
int or_parse_header(const uint8_t *buf, uint64_t len, uint32_t *block_size, uint32_t *header_size);
uint64_t or_discover(const uint8_t *buf, uint64_t len, uint64_t base, or_span_t *spans,
                     uint64_t cap, uint64_t *position, int *err, uint64_t *err_offset);
int or_inflate_batch(const uint8_t *batch, uint64_t batch_len, const or_span_t *spans, uint64_t n,
                     uint8_t *out, uint64_t *offsets, or_error_t *e);
int or_bam_parse_header(const uint8_t *buf, uint64_t len, uint64_t *records_start, uint64_t *err_pos,
                        uint32_t *n_ref);
int or_bam_scan_records(const uint8_t *buf, uint64_t len, uint64_t start, uint64_t *offsets,
                        uint64_t cap, uint64_t *n, uint64_t *tail);
int or_bam_scan_blocked(const uint8_t *buf, uint64_t len, const uint64_t *block_starts, uint64_t nb,
                        uint64_t start, uint64_t *offsets, uint64_t cap, uint64_t *n,
                        uint64_t *tail, uint64_t *fallbacks);
void or_bam_decode_columns(const uint8_t *buf, const uint64_t *offsets, uint64_t n,
                           int32_t *ref_id, int32_t *pos, uint16_t *flags, uint8_t *mapq,
                           uint32_t *l_seq, int32_t *next_ref_id, int32_t *next_pos, int32_t *tlen,
                           uint64_t *rec_off, uint32_t *cigar_start, uint32_t *seq_start,
                           uint32_t *qual_start, uint32_t *aux_start, uint32_t *rec_end);
uint64_t or_virtual_offset(const uint64_t *block_offsets, const uint64_t *coffsets, uint64_t nb,
                           uint64_t p);
int or_fastq_bounds(const uint8_t *buf, uint64_t len, uint64_t pos, uint32_t b[5]);
int or_fastq_scan_records(const uint8_t *buf, uint64_t len, uint64_t start, uint64_t *offsets,
                          uint32_t *bounds, uint64_t cap, uint64_t *n, uint64_t *tail, uint64_t *err_pos);
uint64_t or_fastq_survivors(const uint8_t *buf, uint64_t len, uint64_t start);
int or_bcf_looks_like(const uint8_t *buf, uint64_t len, uint64_t pos, uint32_t samples, uint32_t contigs);
int or_bcf_scan_records(const uint8_t *buf, uint64_t len, uint64_t start, uint64_t *offsets, uint64_t cap,
                        uint64_t *n, uint64_t *tail, uint64_t *err_pos);
int or_bcf_decode_columns(const uint8_t *buf, const uint64_t *offsets, uint64_t n, uint64_t tail,
                          int32_t *chrom, int32_t *pos, int32_t *rlen, uint32_t *qual, uint16_t *n_info,
                          uint16_t *n_allele, uint32_t *n_sample, uint8_t *n_fmt, uint64_t *rec_off,
                          uint32_t *alleles_start, uint32_t *filters_start, uint32_t *info_start,
                          uint32_t *genotypes_start, uint32_t *rec_end, uint64_t *bad);
int or_bcf_validate(const uint8_t *buf, uint64_t len, uint64_t pos);

typedef struct {
    uint8_t tag[2];
    uint8_t type;
    uint8_t subtype;
    uint32_t value_offset;
    uint32_t count;
    uint32_t size;
} or_aux_entry_t;

int or_bam_soa(const uint8_t *buf, const uint64_t *offsets, uint64_t n, uint32_t *n_ops, uint32_t *seq_bytes,
               uint32_t *n_aux, const uint64_t *op_off, uint32_t *ops, const uint64_t *seq_off, uint8_t *seq,
               const uint64_t *aux_off, or_aux_entry_t *aux, uint64_t *bad);

#ifdef __cplusplus
}
#endif
#endif
