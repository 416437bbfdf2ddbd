/*
 * frhip_write.h -- BGZF written from payloads that stay on the device.
 *
 * Replaces:
 *   DeviceBlockCompressor::compress_batch_device  crates/fritillaria-core/src/device.rs:350-372
 *     (the reference's backend is nvCOMP plus crates/fritillaria-cuda/kernels/bgzf_frame.cu;
 *      here the DEFLATE compressor and the framing are this library's own kernels)
 *   choose_framing / framed_size                  crates/fritillaria-core/src/compress.rs:115-142
 *   DeviceBgzfWriter                              crates/fritillaria-bgzf/src/device_write.rs:70-259
 *
 * One block out per chunk in, in order: block boundaries are the caller's and
 * are never moved. A chunk holds at most FRHIP_MAX_COMPRESSIBLE_PAYLOAD bytes
 * (65536 - 18 - 8 - 5), so that storing it still frames to one 64 KiB block.
 * An empty chunk frames to exactly the 28-byte EOF marker's bytes. Output is
 * spec-valid BGZF (BC = block size - 1, CRC32 and ISIZE of the uncompressed
 * payload) whose blocks inflate back to their chunks; it is not byte-identical
 * to any other compressor's.
 */
#ifndef FRHIP_WRITE_H
#define FRHIP_WRITE_H
#include <stdint.h>

#include "frhip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FRHIP_MAX_COMPRESSIBLE_PAYLOAD 65505u
/* DEFAULT_PAYLOAD_SIZE (crates/fritillaria-bgzf/src/write.rs:19), htslib's choice */
#define FRHIP_DEFAULT_PAYLOAD_SIZE 65280u
/* DEFAULT_CHUNKS_PER_BATCH (crates/fritillaria-bgzf/src/device_write.rs:57) */
#define FRHIP_DEFAULT_CHUNKS_PER_BATCH 1024u

/* Framing of a block (crates/fritillaria-core/src/compress.rs:84-98). */
enum { FRHIP_FRAMING_DEFLATED = 0, FRHIP_FRAMING_STORED = 1, FRHIP_FRAMING_EMPTY = 2 };

/* Compresses each range bounds[i] .. bounds[i + 1] of `data` (n_bounds = n + 1
   entries, non-decreasing, the last within data) into one framed block. *out
   receives a device buffer of the n blocks back to back (no EOF block);
   block_offsets (host, n + 1 entries) their boundaries in it; framing (host,
   n entries, may be NULL) how each was framed. Bounds that are empty,
   decreasing or past the buffer, and a chunk above the maximum, are
   FRHIP_E_MALFORMED (format "bgzf"). */
int frhip_compress_batch_device(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *bounds, uint64_t n_bounds,
                                frhip_buf **out, uint64_t *block_offsets, uint8_t *framing, frhip_error *err);

/* The number of bytes a block takes once framed (framed_size). */
uint64_t frhip_framed_size(uint64_t payload_len, uint64_t deflate_len);

/* The writer's sink: write all n bytes of p; return 0, or nonzero to fail the
   write with FRHIP_E_IO. */
typedef int (*frhip_sink)(void *user, const uint8_t *p, uint64_t n);

typedef struct frhip_bgzf_writer frhip_bgzf_writer;

/* chunks_per_batch 0 = FRHIP_DEFAULT_CHUNKS_PER_BATCH; payload_size 0 =
   FRHIP_DEFAULT_PAYLOAD_SIZE (clamped to 1 .. FRHIP_MAX_COMPRESSIBLE_PAYLOAD). */
int frhip_bgzf_writer_create(frhip_ctx *ctx, frhip_sink sink, void *user, uint64_t chunks_per_batch,
                             uint64_t payload_size, frhip_bgzf_writer **out, frhip_error *err);
/* Compresses data at the caller's bounds (n_bounds = n + 1) and writes the
   blocks, compression batches being windows of at most chunks_per_batch
   chunks that overlap by one boundary, so the output does not depend on the
   batch size (device_write.rs:133-165). */
int frhip_bgzf_writer_write_batch(frhip_bgzf_writer *w, const frhip_buf *data, const uint64_t *bounds,
                                  uint64_t n_bounds, frhip_error *err);
/* Compresses the whole buffer chunked at the payload size (write_all_device:
   only for formats whose block boundaries carry nothing). */
int frhip_bgzf_writer_write_all(frhip_bgzf_writer *w, const frhip_buf *data, frhip_error *err);
/* Writes the EOF block; the writer takes no more data afterwards. */
int frhip_bgzf_writer_finish(frhip_bgzf_writer *w, frhip_error *err);
uint64_t frhip_bgzf_writer_blocks_written(const frhip_bgzf_writer *w);
uint64_t frhip_bgzf_writer_batches_run(const frhip_bgzf_writer *w);
uint64_t frhip_bgzf_writer_compressed_bytes(const frhip_bgzf_writer *w);
uint64_t frhip_bgzf_writer_uncompressed_bytes(const frhip_bgzf_writer *w);
/* Where the next block starts: (compressed_bytes << 16) | 0. */
uint64_t frhip_bgzf_writer_virtual_position(const frhip_bgzf_writer *w);
void frhip_bgzf_writer_free(frhip_bgzf_writer *w);

#ifdef __cplusplus
}
#endif
#endif
