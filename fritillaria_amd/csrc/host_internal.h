// host_internal.h -- host-side types behind the opaque handles of frhip.h.
#ifndef FRHIP_HOST_INTERNAL_H
#define FRHIP_HOST_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <vector>

#include "../../include/frhip.h"

namespace frhip {

struct Scratch {
    void *ptr = nullptr;
    uint64_t cap = 0;
    void *get(uint64_t bytes, frhip_error *err, int *rc);
    void release();
};

struct PinnedScratch {
    void *ptr = nullptr;
    uint64_t cap = 0;
    void *get(uint64_t bytes, frhip_error *err, int *rc);
    void release();
};

// An event recorded after the kernels that produced a set of buffers, shared
// by them (CudaAlloc::ready_event, crates/fritillaria-cuda/src/lib.rs:207-224).
struct EventRef {
    hipEvent_t e = nullptr;
    std::atomic<int> refs{1};
};
// Records a new event on s; nullptr if that fails (the buffers then have none).
EventRef *event_record(hipStream_t s);
void event_release(EventRef *e);

// A device allocation shared by several buffers that view parts of it (a
// reader's inflated read-ahead and the batches cut from it); freed with the
// last reference.
struct BufOwner {
    void *ptr = nullptr;
    std::atomic<int> refs{1};
    int ordinal = 0;
    uint64_t cap = 0; // bytes of the allocation (0: not from dev_alloc)
};
void owner_release(BufOwner *o);
// Device blocks for buffers. Freed blocks are kept per device for reuse
// instead of released: hipFree waits for the whole device, so a free while
// another stream has work in flight (a reader's window inflating ahead) would
// stall on it. dev_alloc returns a block of at least n bytes (*cap: its size)
// on the current device; dev_free keeps it (cap 0: releases it); dev_flush
// releases a device's kept blocks.
void *dev_alloc(int ordinal, uint64_t n, uint64_t *cap, hipError_t *r);
void dev_free(int ordinal, void *p, uint64_t cap);
void dev_flush(int ordinal);

int set_err(frhip_error *e, int kind, uint64_t offset, const char *fmt, ...);
int hip_err(frhip_error *e, hipError_t r, const char *what);
const char *inflate_status_message(uint32_t status);
int discover(const uint8_t *buf, uint64_t len, uint64_t base, std::vector<frhip_block_span> &spans,
             uint64_t *position, frhip_error *err);
// Members [0, align_at) are placed back to back from 0; member align_at and
// those after it from the next 4 KiB boundary (align_at >= n: no gap).
// offsets[i + 1] is the end of member i's output.
// With t, the phases are separated by stream synchronisations and their host
// wall-clock times added to *t (InflateTimings, backend.rs:136-177).
int inflate_device(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len, const frhip_block_span *spans,
                   uint64_t n, frhip_buf **out, uint64_t *offsets, frhip_error *err, frhip_inflate_timings *t,
                   uint64_t align_at = ~0ull);
// The stream and scratch one inflate runs with: the context's, or a
// reader's own for the window it inflates ahead.
struct InflateRun {
    hipStream_t stream;
    Scratch *in, *meta, *status, *rec;
    PinnedScratch *pinned;
};
// inflate_device in two halves. inflate_launch plans, uploads and launches on
// run.stream without waiting: *out holds the output (its ready event
// recorded), out_off the members' output offsets, and *h_status (in
// run.pinned) receives status, produced and CRC per member once the stream
// passes its last copy. inflate_verify then checks them in the reference's
// order and returns the first failing member (err set), or n.
int inflate_launch(frhip_ctx *ctx, const InflateRun &run, const uint8_t *batch, uint64_t batch_len,
                   const frhip_block_span *spans, uint64_t n, frhip_buf **out, std::vector<uint64_t> &out_off,
                   const uint32_t **h_status, frhip_error *err, frhip_inflate_timings *t, uint64_t align_at = ~0ull);
uint64_t inflate_verify(const frhip_block_span *spans, uint64_t n, const uint32_t *h_status,
                        const std::vector<uint64_t> &out_off, uint64_t *offsets, frhip_error *err);
hipError_t launch_crc32_blocks(const uint8_t *data, const uint64_t *off, const uint32_t *len, uint32_t n,
                               uint32_t *crc, const uint32_t *x2n, hipStream_t s);

} // namespace frhip

struct frhip_buf {
    void *ptr;
    uint64_t len;
    int ordinal;
    frhip::EventRef *ready = nullptr; // owned reference, or none
    frhip::BufOwner *owner = nullptr; // set: ptr lies inside owner's allocation, which buf_free releases
    uint64_t cap = 0;                 // bytes of ptr's block from dev_alloc (0: none)
};

struct frhip_ctx {
    // Pipelines and readers keep a reference; the context is freed when the
    // last holder releases it (frhip_ctx_destroy releases the creator's).
    int refs = 1;
    int ordinal = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    uint32_t *x2n = nullptr;
    uint32_t bam_window = 2048;
    int inflate_schedule = 16; // index into INFLATE_SCHEDULES (DEFAULT_INFLATE_SCHEDULE)
    uint32_t two_phase_min = 4096; // smaller batches of a two-phase schedule run the one-wave form (set per device)
    frhip::Scratch in_scratch, meta_scratch, status_scratch, bam_scratch, bcf_scratch, soa_scratch, fq_small, fq_big,
        rec_scratch, tx_scratch, wr_scratch;
    frhip::PinnedScratch pinned;
    // device address of this GPU's no-allocation scratch array (noalloc.hip)
    void *noalloc_scratch = nullptr;
    hipError_t noalloc_rc = hipSuccess;
};

namespace frhip {
// Takes the device address of the no-allocation scratch for ctx's device.
void noalloc_bind(frhip_ctx *ctx);
frhip_buf *buf_alloc(frhip_ctx *ctx, uint64_t len, frhip_error *err);
void ctx_retain(frhip_ctx *ctx);
void ctx_release(frhip_ctx *ctx);
void buf_free(frhip_buf *b);
// Gives b a reference to e (e may be null).
void buf_attach_event(frhip_buf *b, EventRef *e);
// Host wall-clock in milliseconds, for the timed entry points.
double now_ms();
} // namespace frhip

#endif
