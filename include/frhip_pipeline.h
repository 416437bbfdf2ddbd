/*
 * frhip_pipeline.h -- the batched, device-resident form of the BAM path.
 *
 * No counterpart in the reference, which only has the per-call drop-in
 * (crates/fritillaria-cuda/src/{backend,bam}.rs, driven batch by batch by
 * crates/fritillaria-bgzf/src/device_read.rs). It stands beside that drop-in:
 * a whole BGZF BAM already in HBM goes through member discovery, inflate with
 * CRC32/ISIZE verification, record boundaries and SoA columns in one enqueue
 * (optionally one hipGraph replay), with one small read-back at the end.
 * Results are the same bytes and columns the per-call path produces for the
 * whole file read as a single batch.
 */
#ifndef FRHIP_PIPELINE_H
#define FRHIP_PIPELINE_H
#include <stdint.h>

#include "frhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct frhip_pipeline frhip_pipeline;

typedef struct {
    uint32_t waves_per_cu;     /* persistent inflate grid = CUs x this; 0 = the schedule's resident waves */
    uint32_t order;            /* member order on the work queue: 0 file order, 1 largest payload first,
                                  2 smallest payload first, 3 members whose payload is no smaller than
                                  their ISIZE first, then the rest largest first within each run of
                                  1024 in file order */
    uint32_t bam_window;       /* speculation window of the boundary scan; 0 = 2048 */
    uint32_t use_graph;        /* 1: capture the whole run once and replay it as one hipGraph */
    uint32_t host_discovery;   /* 1: walk member headers on the host once at create; 0: sieve on the device every run */
    int32_t inflate_schedule;  /* inflate kernel variant (frhip_inflate_schedules); -1 = the context's */
    uint32_t soa;              /* 1: every run also builds the SoA extras of frhip_bam_soa_build (resolved
                                  CIGAR ops, packed 4-bit sequence, aux index), sized at create from one run */
} frhip_pipeline_config;

/* Uploads file[0..len) to HBM once and sizes every buffer for it. Runs member
   discovery once to learn the member count and inflated size; if the device
   sieve cannot prove its member chain the host walk is used instead. */
int frhip_pipeline_create(frhip_ctx *ctx, const uint8_t *file, uint64_t len, const frhip_pipeline_config *cfg,
                          frhip_pipeline **out, frhip_error *err);
void frhip_pipeline_free(frhip_pipeline *p);

/* One full pass over the resident file; synchronous, verifies everything
   (member chain, inflate status, ISIZE, CRC32, BAM header, record chain) and
   returns the first error in the reference's order. */
int frhip_pipeline_run(frhip_pipeline *p, frhip_error *err);
/* The same pass, enqueued without waiting; frhip_pipeline_finish waits and verifies. */
int frhip_pipeline_launch(frhip_pipeline *p, frhip_error *err);
int frhip_pipeline_finish(frhip_pipeline *p, frhip_error *err);

/* The host-fed form: copies file[0..len) into the pipeline's resident file
   buffer on a copy stream of the pipeline's own and returns at once; the copy
   waits for the last launch to be done with the old bytes, and the next
   launch waits for the copy, so a caller alternating two pipelines overlaps
   one file's upload with the other's decode. len must be the length the
   pipeline was created with and the file should have the same member table
   (the use: the next batch of a stream of equal batches, or the same file
   again from host memory); with device discovery, a member table whose
   inflated total differs makes the run fail verification. For full PCIe
   bandwidth the file should be in pinned memory (frhip_host_alloc). */
int frhip_pipeline_upload(frhip_pipeline *p, const uint8_t *file, uint64_t len, frhip_error *err);

uint64_t frhip_pipeline_members(const frhip_pipeline *p);
uint64_t frhip_pipeline_inflated_len(const frhip_pipeline *p);
uint64_t frhip_pipeline_compressed_len(const frhip_pipeline *p);
uint64_t frhip_pipeline_records(const frhip_pipeline *p);
uint64_t frhip_pipeline_records_start(const frhip_pipeline *p);
uint64_t frhip_pipeline_tail(const frhip_pipeline *p);
uint64_t frhip_pipeline_fallbacks(const frhip_pipeline *p);
/* 1 when member discovery runs on the device (the chain proof held). */
int frhip_pipeline_device_discovery(const frhip_pipeline *p);
/* The inflated file (all members back to back) and the BAM columns of the last
   run, in frhip_bam_column order; views owned by the pipeline, valid until the
   next run or free. */
const frhip_buf *frhip_pipeline_inflated(const frhip_pipeline *p);
const frhip_buf *frhip_pipeline_column(const frhip_pipeline *p, int column);
/* With cfg.soa: the SoA column of the last run, in frhip_soa_column order and
   with the layout frhip_bam_soa_build gives (offset columns records+1 entries);
   NULL without cfg.soa. Views owned by the pipeline. After a failed run every
   column view (these and frhip_pipeline_column's) is empty and
   frhip_pipeline_records is 0. */
const frhip_buf *frhip_pipeline_soa_column(const frhip_pipeline *p, int column);
/* Member boundaries (members+1 entries) and compressed offsets (members entries). */
int frhip_pipeline_member_offsets(const frhip_pipeline *p, uint64_t *offsets, uint64_t *coffsets, frhip_error *err);
/* Per-stage device times (ms) of the last run without a graph:
   [0] discovery [1] inflate [2] verify [3] boundary speculation [4] reconcile
   [5] offsets + columns [6] SoA extras [7] total. Returns the number of stages
   written. */
int frhip_pipeline_stage_ms(const frhip_pipeline *p, double *ms, int n);

#ifdef __cplusplus
}
#endif
#endif
