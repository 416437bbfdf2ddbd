/*
 * frhip_serve.h -- virtual-offset range queries over a resident BGZF BAM,
 * served in batches.
 *
 * No counterpart in the reference, which serves one range at a time on the host
 * through noodles' csi::io::Query (crates/fritillaria-csi/src/io/query.rs:66-82)
 * over a BgzfReader. A query here is [vstart, vend): seek to vstart (a record
 * start's virtual offset, (coffset << 16) | uoffset) and serve records while
 * their virtual offset is below vend. A batch of queries is one launch per
 * stage: every member any query needs is inflated by one persistent kernel
 * (verified like the per-call path), each query is a separate record chain in
 * the same boundary launch, and all records land in one set of columns. A
 * query's records and columns are the same whether it runs alone or in a batch.
 */
#ifndef FRHIP_SERVE_H
#define FRHIP_SERVE_H
#include <stdint.h>

#include "frhip.h"
#include "frhip_index.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct frhip_server frhip_server;

typedef struct {
    uint64_t vstart; /* virtual offset of the first record served */
    uint64_t vend;   /* serve records whose virtual offset is below this; UINT64_MAX = to the end */
} frhip_query;

typedef struct {
    uint64_t first;          /* index of the query's first record in the batch columns */
    uint64_t records;        /* records served */
    uint64_t members;        /* BGZF members inflated for the query */
    uint64_t inflated_bytes; /* bytes those members inflate to */
    int32_t status;          /* FRHIP_OK, or the error kind of this query */
    uint32_t reserved;
    uint64_t error_position; /* for FRHIP_E_MALFORMED: position in the query's inflated bytes */
} frhip_query_result;

/* Uploads file[0..len) to HBM once and walks its member headers on the host
   (the same strict walk as frhip_discover). The server has `slots` batch slots
   (1..FRHIP_SERVER_MAX_SLOTS): each holds one batch on the device at a time,
   with its own stream (slot 0: the context's; the others: streams the server
   creates, default priority), scratch and columns, so batches in different
   slots run concurrently. frhip_server_create makes
   FRHIP_SERVER_DEFAULT_SLOTS. */
#define FRHIP_SERVER_DEFAULT_SLOTS 2
#define FRHIP_SERVER_MAX_SLOTS 4
int frhip_server_create(frhip_ctx *ctx, const uint8_t *file, uint64_t len, frhip_server **out, frhip_error *err);
int frhip_server_create_slots(frhip_ctx *ctx, const uint8_t *file, uint64_t len, int slots, frhip_server **out,
                              frhip_error *err);
int frhip_server_slots(const frhip_server *s);
void frhip_server_free(frhip_server *s);
uint64_t frhip_server_members(const frhip_server *s);
/* Compressed offset and inflated-prefix of every member (members and members+1 entries). */
int frhip_server_member_table(const frhip_server *s, uint64_t *coffsets, uint64_t *uprefix, frhip_error *err);
/* Device bytes a batch of these queries needs beyond the resident file (inflated
   bytes, columns at their worst-case record count, scan scratch): what an
   admission policy charges against free HBM. */
uint64_t frhip_server_batch_footprint(const frhip_server *s, const frhip_query *q, uint64_t n);
/* Runs the queries as one batch in slot 0, synchronously (launch, then
   finish). Per-query failures (a vstart or vend that names no member, a
   malformed record, a member that fails verification) are reported in
   res[i].status; the call itself fails only when the device does. */
int frhip_server_run(frhip_server *s, const frhip_query *q, uint64_t n, frhip_query_result *res, frhip_error *err);
/* Continuous batching. launch plans the n queries, uploads their member
   tables and enqueues the whole batch on the slot's stream, returning without
   waiting for the device (a slot with a batch not yet finished is
   FRHIP_E_INVALID_DEVICE_BATCH). query is 1 when the slot's batch is done on
   the device (or it has none), 0 while it runs. finish waits for it, fills
   res[0..n) as frhip_server_run does and frees the slot; the slot's columns
   then hold the batch's records. A query's results do not depend on which
   slot or batch it ran in. */
int frhip_server_launch(frhip_server *s, int slot, const frhip_query *q, uint64_t n, frhip_error *err);
int frhip_server_query(const frhip_server *s, int slot);
int frhip_server_finish(frhip_server *s, int slot, frhip_query_result *res, frhip_error *err);
/* Device interval of the slot's last finished batch, in ms from the server's
   creation (HIP events on the slot's stream around the batch's work). */
int frhip_server_batch_interval(const frhip_server *s, int slot, double *start_ms, double *end_ms);
/* Columns of the slot's last finished batch, in frhip_bam_column order (virtual
   offsets included), and the inflated bytes they index. Owned by the server;
   valid until the slot's next launch. frhip_server_column / _inflated are slot 0's. */
const frhip_buf *frhip_server_slot_column(const frhip_server *s, int slot, int column);
const frhip_buf *frhip_server_slot_inflated(const frhip_server *s, int slot);
const frhip_buf *frhip_server_column(const frhip_server *s, int column);
const frhip_buf *frhip_server_inflated(const frhip_server *s);
/* Region queries (crates/fritillaria-bam/src/io/reader.rs:373-390 with
   crates/fritillaria-bam/src/io/reader/query.rs:104-152): each region is
   mapped through a BAI/CSI index (frhip_index_query) to its merged chunks,
   every chunk becomes one range query of the batch, and after the decode a
   device pass keeps the records that overlap their region: same reference
   and, for a bounded region, an alignment [pos + 1, pos + span] (span from
   the raw CIGAR, at least 1) meeting the 1-based closed interval. The kept
   rows (indices into the slot's columns, u64, file order within a region)
   are one selection per batch; region i owns [first, first + records) of it.
   start/end 0 leave that bound open; both 0 keep every record of the
   reference, placed or not. */
typedef struct {
    int32_t ref_id;   /* reference sequence id */
    uint32_t reserved;
    uint64_t start;   /* 1-based, inclusive; 0 = from the start */
    uint64_t end;     /* 1-based, inclusive; 0 = to the end */
} frhip_region;

typedef struct {
    uint64_t first;          /* index of the region's first kept row in the selection */
    uint64_t records;        /* rows kept */
    uint64_t chunks;         /* chunks the index mapped the region to */
    uint64_t scanned;        /* records read from those chunks */
    uint64_t members;        /* BGZF members inflated for them */
    uint64_t inflated_bytes;
    int32_t status;          /* FRHIP_OK, or the error kind of this region */
    uint32_t reserved;
    /* FRHIP_E_MALFORMED: position in the failing chunk's inflated bytes, or,
       for a record failing the region check (reference id or position below
       -1, CIGAR op kind above 8), that record's ordinal among the scanned */
    uint64_t error_position;
} frhip_region_result;

/* Maps the regions through ix and launches them as one batch in the slot
   (see frhip_server_launch). A region the index rejects (FRHIP_E_IO: a
   reference id it does not have, a bound past its maximum position) fails
   alone. */
int frhip_server_launch_regions(frhip_server *s, int slot, const frhip_index *ix, const frhip_region *r, uint64_t n,
                                frhip_error *err);
/* Waits for the slot's region batch and fills res[0..n); the slot's columns
   then hold every record the chunks read, and its selection the kept rows. */
int frhip_server_finish_regions(frhip_server *s, int slot, frhip_region_result *res, frhip_error *err);
/* The selection of the slot's last finished region batch (u64 rows). Owned by
   the server; valid until the slot's next launch. */
const frhip_buf *frhip_server_slot_selection(const frhip_server *s, int slot);

/* hipMemGetInfo of the context's device. */
int frhip_device_memory(frhip_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes, frhip_error *err);

#ifdef __cplusplus
}
#endif
#endif
