/*
 * frhip_index.h -- BAI and CSI binning indexes: reading them and mapping a
 * region to the BGZF chunks a query reads.
 *
 * Replaces the index side of a region query in the reference:
 *   bai::fs::read / csi::fs::read        crates/fritillaria-bam/src/bai/io/reader/index.rs:9-37,
 *                                        crates/fritillaria-csi/src/io/reader/index.rs:72-120
 *   BinningIndex::query                  crates/fritillaria-csi/src/binning_index/index.rs:93-121
 *   (reg2bins, optimize_chunks,          crates/fritillaria-csi/src/binning_index/index/reference_sequence.rs:352-397,
 *    min_offset)                         crates/fritillaria-csi/src/binning_index.rs:102-160,
 *                                        .../reference_sequence/index/{linear,binned}_index.rs
 *   unplaced_unmapped_record_count,      crates/fritillaria-csi/src/binning_index/index.rs:89-91,123-129
 *   last_first_record_start_position
 * The chunks feed frhip_server_launch_regions (frhip_serve.h), which reads
 * them on the device and keeps the records that overlap the region.
 *
 * Errors are FRHIP_E_IO with a reason: the reference's io::Error of kind
 * InvalidData (a malformed index) or InvalidInput (a reference id the index
 * does not have, a region bound past the index's maximum position).
 */
#ifndef FRHIP_INDEX_H
#define FRHIP_INDEX_H
#include <stdint.h>

#include "frhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct frhip_index frhip_index;

/* A chunk: virtual offsets [start, end). */
typedef struct {
    uint64_t start;
    uint64_t end;
} frhip_chunk;

enum { FRHIP_INDEX_BAI = 1, FRHIP_INDEX_CSI = 2 };

/* Reads a BAI ("BAI\1") or CSI ("CSI\1") index held in memory. One
   deviation: the reference reads a CSI of any min_shift and depth from 0 to
   255, whose maximum position 2^(min_shift + 3 depth) - 1 then overflows its
   64-bit arithmetic in a query; such an index (min_shift + 3 depth > 62, or
   depth > 10) is refused here when read, as unsupported. */
int frhip_index_read(const uint8_t *data, uint64_t len, frhip_index **out, frhip_error *err);
void frhip_index_free(frhip_index *ix);
/* FRHIP_INDEX_BAI or FRHIP_INDEX_CSI, with the binning parameters (BAI: 14, 5). */
int frhip_index_format(const frhip_index *ix, uint32_t *min_shift, uint32_t *depth);
uint32_t frhip_index_references(const frhip_index *ix);
/* 1 and *n when the index records the unplaced, unmapped record count, else 0. */
int frhip_index_unplaced_unmapped(const frhip_index *ix, uint64_t *n);
/* 1 and the reference's metadata pseudo-bin (ref_beg, ref_end, n_mapped,
   n_unmapped) when present, else 0. */
int frhip_index_metadata(const frhip_index *ix, uint32_t ref_id, uint64_t out[4]);
/* 1 and the start of the first record in the last linear bin of the last
   reference that has one (BAI: last linear entry; CSI: greatest loffset),
   where the unmapped records after the placed ones begin; else 0. */
int frhip_index_last_first_record_start(const frhip_index *ix, uint64_t *vpos);
/* The merged chunks a query of reference ref_id over the 1-based closed
   interval [start, end] reads, in file order; start 0 = from 1, end 0 = to
   the index's maximum position (2^(min_shift + 3 depth) - 1). Writes at most
   cap chunks; *n is always the full count. */
int frhip_index_query(const frhip_index *ix, uint32_t ref_id, uint64_t start, uint64_t end, frhip_chunk *out,
                      uint64_t cap, uint64_t *n, frhip_error *err);

#ifdef __cplusplus
}
#endif
#endif
