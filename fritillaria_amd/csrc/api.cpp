// api.cpp -- host side of the C ABI in include/frhip.h.
//
// Mirrors the planning, verification and ownership rules of
// crates/fritillaria-cuda/src/{backend,bam}.rs and the reader loop of
// crates/fritillaria-bgzf/src/device_read.rs, over HIP runtime calls and the
// kernels in inflate.hip / bam.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/frhip.h"
#include "crc_math.h"
#include "frhip_internal.h"
#include "host_internal.h"
#include "pipeline_internal.h"

using namespace frhip;

// ------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------

namespace frhip {

int set_err(frhip_error *e, int kind, uint64_t offset, const char *fmt, ...)
{
    if (e) {
        std::memset(e, 0, sizeof *e);
        e->kind = kind;
        e->offset = offset;
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(e->reason, sizeof e->reason, fmt, ap);
        va_end(ap);
    }
    return kind;
}

int hip_err(frhip_error *e, hipError_t r, const char *what)
{
    return set_err(e, FRHIP_E_DEVICE, 0, "%s: %s", what, hipGetErrorString(r));
}

const char *inflate_status_message(uint32_t status)
{
    // crates/fritillaria-cuda/src/backend.rs:66-79
    switch (status) {
    case 1: return "truncated deflate stream";
    case 2: return "reserved deflate block type";
    case 3: return "stored block LEN/NLEN mismatch";
    case 4: return "output overran the size declared by ISIZE";
    case 5: return "invalid Huffman code";
    case 6: return "invalid or out-of-range back-reference distance";
    case 7: return "malformed Huffman table";
    case 8: return "invalid code-length repeat";
    case 9: return "code counts exceed the DEFLATE maximum";
    default: return "unknown device error";
    }
}

// ------------------------------------------------------------------------
// buffers
// ------------------------------------------------------------------------

frhip_buf *buf_alloc(frhip_ctx *ctx, uint64_t len, frhip_error *err)
{
    frhip_buf *b = new (std::nothrow) frhip_buf;
    if (!b) {
        set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
        return nullptr;
    }
    b->len = len;
    b->ordinal = ctx->ordinal;
    b->ptr = nullptr;
    // zero-length allocations are padded to one byte with a logical length of 0
    // (crates/fritillaria-cuda/src/backend.rs:202-205,608-613)
    hipError_t r = hipSuccess;
    b->ptr = dev_alloc(ctx->ordinal, len ? len : 1, &b->cap, &r);
    if (!b->ptr) {
        delete b;
        hip_err(err, r, "hipMalloc");
        return nullptr;
    }
    return b;
}

void buf_free(frhip_buf *b)
{
    if (!b) return;
    event_release(b->ready);
    if (b->owner)
        owner_release(b->owner);
    else if (b->ptr)
        dev_free(b->ordinal, b->ptr, b->cap);
    delete b;
}

void owner_release(BufOwner *o)
{
    if (o && o->refs.fetch_sub(1) == 1) {
        if (o->ptr) dev_free(o->ordinal, o->ptr, o->cap);
        delete o;
    }
}

namespace {
// Blocks are sized in 64 KiB steps (4 KiB at least); a request takes the
// smallest kept block that holds it and is at most twice its size. At most
// BLOCK_KEEP_MAX bytes are kept per device (the largest go first beyond it).
constexpr uint64_t BLOCK_KEEP_MAX = 8ull << 30;
constexpr int KEEP_DEVICES = 64;
struct BlockKeep {
    std::mutex m;
    std::multimap<uint64_t, void *> blocks[KEEP_DEVICES];
    uint64_t bytes[KEEP_DEVICES] = {};
};
BlockKeep &block_keep()
{
    static BlockKeep k;
    return k;
}
uint64_t block_bytes(uint64_t n) { return n <= 4096 ? 4096 : (n + 65535) & ~65535ull; }
} // namespace

void *dev_alloc(int ordinal, uint64_t n, uint64_t *cap, hipError_t *r)
{
    const uint64_t want = block_bytes(n);
    *r = hipSuccess;
    if (ordinal >= 0 && ordinal < KEEP_DEVICES) {
        BlockKeep &k = block_keep();
        std::lock_guard<std::mutex> g(k.m);
        auto &bl = k.blocks[ordinal];
        auto it = bl.lower_bound(want);
        if (it != bl.end() && it->first <= 2 * want) {
            void *p = it->second;
            *cap = it->first;
            k.bytes[ordinal] -= it->first;
            bl.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    *r = hipMalloc(&p, want);
    if (*r != hipSuccess) {
        // the kept blocks may be what fills the device: release them, once
        (void)hipGetLastError();
        dev_flush(ordinal);
        *r = hipMalloc(&p, want);
    }
    if (*r != hipSuccess) {
        *cap = 0;
        return nullptr;
    }
    *cap = want;
    return p;
}

void dev_free(int ordinal, void *p, uint64_t cap)
{
    if (!p) return;
    if (cap == 0 || ordinal < 0 || ordinal >= KEEP_DEVICES) {
        (void)hipFree(p);
        return;
    }
    std::vector<void *> release;
    {
        BlockKeep &k = block_keep();
        std::lock_guard<std::mutex> g(k.m);
        auto &bl = k.blocks[ordinal];
        bl.emplace(cap, p);
        k.bytes[ordinal] += cap;
        while (k.bytes[ordinal] > BLOCK_KEEP_MAX && !bl.empty()) {
            auto last = std::prev(bl.end());
            release.push_back(last->second);
            k.bytes[ordinal] -= last->first;
            bl.erase(last);
        }
    }
    for (void *q : release) (void)hipFree(q);
}

void dev_flush(int ordinal)
{
    if (ordinal < 0 || ordinal >= KEEP_DEVICES) return;
    std::vector<void *> release;
    {
        BlockKeep &k = block_keep();
        std::lock_guard<std::mutex> g(k.m);
        for (auto &kv : k.blocks[ordinal]) release.push_back(kv.second);
        k.blocks[ordinal].clear();
        k.bytes[ordinal] = 0;
    }
    for (void *q : release) (void)hipFree(q);
}

EventRef *event_record(hipStream_t s)
{
    EventRef *e = new (std::nothrow) EventRef;
    if (!e) return nullptr;
    if (hipEventCreateWithFlags(&e->e, hipEventDisableTiming) != hipSuccess) {
        delete e;
        return nullptr;
    }
    if (hipEventRecord(e->e, s) != hipSuccess) {
        (void)hipEventDestroy(e->e);
        delete e;
        return nullptr;
    }
    return e;
}

void event_release(EventRef *e)
{
    if (e && e->refs.fetch_sub(1) == 1) {
        (void)hipEventDestroy(e->e);
        delete e;
    }
}

void buf_attach_event(frhip_buf *b, EventRef *e)
{
    if (!b || !e) return;
    e->refs.fetch_add(1);
    event_release(b->ready);
    b->ready = e;
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Grow-only device scratch owned by the context.
void *Scratch::get(uint64_t bytes, frhip_error *err, int *rc)
{
    *rc = FRHIP_OK;
    if (bytes <= cap && ptr) return ptr;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    uint64_t want = std::max<uint64_t>(bytes, 1 << 16);
    want = want + want / 4;
    hipError_t r = hipMalloc(&ptr, want);
    if (r != hipSuccess) {
        ptr = nullptr;
        *rc = hip_err(err, r, "hipMalloc(scratch)");
        return nullptr;
    }
    cap = want;
    return ptr;
}

void Scratch::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

void *PinnedScratch::get(uint64_t bytes, frhip_error *err, int *rc)
{
    *rc = FRHIP_OK;
    if (bytes <= cap && ptr) return ptr;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    uint64_t want = std::max<uint64_t>(bytes, 1 << 16);
    want = want + want / 4;
    hipError_t r = hipHostMalloc(&ptr, want, hipHostMallocDefault);
    if (r != hipSuccess) {
        ptr = nullptr;
        *rc = hip_err(err, r, "hipHostMalloc");
        return nullptr;
    }
    cap = want;
    return ptr;
}

void PinnedScratch::release()
{
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
}

} // namespace frhip

// ------------------------------------------------------------------------
// device and context
// ------------------------------------------------------------------------

extern "C" int frhip_device_available(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return 0;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 0;
    return std::strncmp(p.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

extern "C" int frhip_ctx_create(int ordinal, void *stream, frhip_ctx **out, frhip_error *err)
{
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return set_err(err, FRHIP_E_UNAVAILABLE, 0, "no HIP device is available");
    if (ordinal < 0 || ordinal >= n)
        return set_err(err, FRHIP_E_UNAVAILABLE, 0, "device %d does not exist (%d present)", ordinal, n);
    hipError_t r = hipSetDevice(ordinal);
    if (r != hipSuccess) return hip_err(err, r, "hipSetDevice");
    hipDeviceProp_t p;
    r = hipGetDeviceProperties(&p, ordinal);
    if (r != hipSuccess) return hip_err(err, r, "hipGetDeviceProperties");
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return set_err(err, FRHIP_E_UNAVAILABLE, 0, "device %d is %s, this build targets gfx950", ordinal,
                       p.gcnArchName);
    frhip_ctx *c = new (std::nothrow) frhip_ctx();
    if (!c) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    c->ordinal = ordinal;
    c->n_cu = p.multiProcessorCount;
    // a batch the one-wave form takes in one round of resident waves is done
    // sooner that way than one member per lane (measured crossover ~4,500
    // members of 65 KiB on 256 CUs)
    c->two_phase_min = (uint32_t)c->n_cu * inflate_waves_per_cu(15);
    if (stream) {
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    } else {
        r = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (r != hipSuccess) {
            delete c;
            return hip_err(err, r, "hipStreamCreate");
        }
        c->own_stream = true;
    }
    uint32_t x2n[32];
    fr_x2n_table(x2n);
    r = hipMalloc(&c->x2n, sizeof x2n);
    if (r == hipSuccess) r = hipMemcpy(c->x2n, x2n, sizeof x2n, hipMemcpyHostToDevice);
    if (r != hipSuccess) {
        frhip_ctx_destroy(c);
        return hip_err(err, r, "uploading CRC tables");
    }
    const char *w = std::getenv("FRHIP_BAM_WINDOW");
    if (w) c->bam_window = (uint32_t)std::strtoul(w, nullptr, 10);
    noalloc_bind(c);
    *out = c;
    return FRHIP_OK;
}

namespace frhip {

void ctx_retain(frhip_ctx *c) { __atomic_add_fetch(&c->refs, 1, __ATOMIC_ACQ_REL); }

void ctx_release(frhip_ctx *c)
{
    if (!c || __atomic_sub_fetch(&c->refs, 1, __ATOMIC_ACQ_REL) != 0) return;
    (void)hipSetDevice(c->ordinal);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dev_flush(c->ordinal);
    c->in_scratch.release();
    c->meta_scratch.release();
    c->status_scratch.release();
    c->bam_scratch.release();
    c->bcf_scratch.release();
    c->soa_scratch.release();
    c->fq_small.release();
    c->fq_big.release();
    c->rec_scratch.release();
    c->tx_scratch.release();
    c->wr_scratch.release();
    c->pinned.release();
    if (c->x2n) (void)hipFree(c->x2n);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

} // namespace frhip

extern "C" void frhip_ctx_destroy(frhip_ctx *c) { ctx_release(c); }

extern "C" int frhip_ctx_ordinal(const frhip_ctx *c) { return c->ordinal; }
extern "C" void *frhip_ctx_stream(const frhip_ctx *c) { return (void *)c->stream; }
extern "C" int frhip_ctx_synchronize(frhip_ctx *c, frhip_error *err)
{
    (void)hipSetDevice(c->ordinal);
    hipError_t r = hipDeviceSynchronize();
    return r == hipSuccess ? FRHIP_OK : hip_err(err, r, "hipDeviceSynchronize");
}
extern "C" int frhip_ctx_set_inflate_schedule(frhip_ctx *c, int schedule)
{
    if (schedule < 0 || schedule >= N_INFLATE_SCHEDULES) return FRHIP_E_INVALID_DEVICE_BATCH;
    c->inflate_schedule = schedule;
    return FRHIP_OK;
}
extern "C" int frhip_ctx_set_two_phase_min_members(frhip_ctx *c, uint32_t min_members)
{
    c->two_phase_min = min_members;
    return FRHIP_OK;
}
extern "C" uint32_t frhip_ctx_two_phase_min_members(const frhip_ctx *c) { return c->two_phase_min; }
extern "C" uint32_t frhip_ctx_compute_units(const frhip_ctx *c) { return (uint32_t)c->n_cu; }
extern "C" int frhip_inflate_schedules(uint32_t *ring, uint32_t *waves_per_group, uint32_t *waves_per_cu,
                                       uint32_t *tokens, int cap)
{
    int k = cap < N_INFLATE_SCHEDULES ? cap : N_INFLATE_SCHEDULES;
    for (int i = 0; i < k; ++i) {
        if (ring) ring[i] = INFLATE_SCHEDULES[i].ring;
        if (waves_per_group) waves_per_group[i] = INFLATE_SCHEDULES[i].waves_per_group;
        if (waves_per_cu) waves_per_cu[i] = inflate_waves_per_cu(i);
        if (tokens) tokens[i] = INFLATE_SCHEDULES[i].tokens;
    }
    return N_INFLATE_SCHEDULES;
}
extern "C" int frhip_ctx_set_bam_window(frhip_ctx *c, uint32_t window)
{
    c->bam_window = window < 1 ? 1 : window;
    return FRHIP_OK;
}
extern "C" const char *frhip_codec_name(void) { return "hip"; }
extern "C" const char *frhip_device_codec_name(void) { return "hip-device"; }

// ------------------------------------------------------------------------
// buffers
// ------------------------------------------------------------------------

extern "C" uint64_t frhip_buf_len(const frhip_buf *b) { return b ? b->len : 0; }
extern "C" int frhip_buf_ordinal(const frhip_buf *b) { return b ? b->ordinal : -1; }
extern "C" void *frhip_buf_device_ptr(const frhip_buf *b) { return b ? b->ptr : nullptr; }

extern "C" int frhip_buf_copy_to_host(const frhip_buf *b, void *dst, uint64_t dst_len, frhip_error *err)
{
    // crates/fritillaria-cuda/src/backend.rs:333-358
    if (!b) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "no buffer");
    if (dst_len != b->len)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0,
                       "copy_to_host destination is %llu bytes, the allocation is %llu",
                       (unsigned long long)dst_len, (unsigned long long)b->len);
    if (b->len == 0) return FRHIP_OK;
    (void)hipSetDevice(b->ordinal);
    hipError_t r = hipMemcpy(dst, b->ptr, b->len, hipMemcpyDeviceToHost);
    if (r != hipSuccess) return hip_err(err, r, "copy_to_host");
    return FRHIP_OK;
}

extern "C" void frhip_buf_free(frhip_buf *b) { buf_free(b); }

extern "C" void *frhip_buf_ready_event(const frhip_buf *b)
{
    return b && b->len && b->ready ? (void *)b->ready->e : nullptr;
}

extern "C" int frhip_event_wait(void *event, void *stream, frhip_error *err)
{
    if (!event) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "no event");
    hipError_t r = hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0);
    return r == hipSuccess ? FRHIP_OK : hip_err(err, r, "hipStreamWaitEvent");
}

extern "C" int frhip_event_synchronize(void *event, frhip_error *err)
{
    if (!event) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "no event");
    hipError_t r = hipEventSynchronize((hipEvent_t)event);
    return r == hipSuccess ? FRHIP_OK : hip_err(err, r, "hipEventSynchronize");
}

extern "C" int frhip_buf_alloc(frhip_ctx *ctx, uint64_t len, frhip_buf **out, frhip_error *err)
{
    (void)hipSetDevice(ctx->ordinal);
    *out = buf_alloc(ctx, len, err);
    return *out ? FRHIP_OK : (err ? err->kind : FRHIP_E_DEVICE);
}

extern "C" int frhip_buf_copy_device(frhip_ctx *ctx, frhip_buf *dst, const frhip_buf *src, frhip_error *err)
{
    if (!dst || !src || dst->len != src->len)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "copy between buffers of different lengths");
    if (dst->ordinal != ctx->ordinal || src->ordinal != ctx->ordinal)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "buffers are not on the context's device");
    (void)hipSetDevice(ctx->ordinal);
    hipError_t r = launch_copy((uint8_t *)dst->ptr, (const uint8_t *)src->ptr, src->len, ctx->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(ctx->stream);
    return r == hipSuccess ? FRHIP_OK : hip_err(err, r, "device copy");
}

extern "C" int frhip_host_alloc(uint64_t len, void **out, frhip_error *err)
{
    *out = nullptr;
    hipError_t r = hipHostMalloc(out, len ? len : 1, hipHostMallocDefault);
    if (r != hipSuccess) {
        *out = nullptr;
        return hip_err(err, r, "hipHostMalloc");
    }
    return FRHIP_OK;
}

extern "C" void frhip_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

extern "C" int frhip_buf_copy_from_host(frhip_buf *b, uint64_t offset, const void *src, uint64_t len,
                                        frhip_error *err)
{
    if (!b || offset > b->len || len > b->len - offset)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, offset, "copy past the end of the buffer");
    if (len == 0) return FRHIP_OK;
    (void)hipSetDevice(b->ordinal);
    hipError_t r = hipMemcpy((uint8_t *)b->ptr + offset, src, len, hipMemcpyHostToDevice);
    return r == hipSuccess ? FRHIP_OK : hip_err(err, r, "host to device copy");
}

// ------------------------------------------------------------------------
// discovery (host)
// ------------------------------------------------------------------------

namespace {

uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t rd32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// crates/fritillaria-bgzf/src/block.rs:99-164. Returns nullptr or the reason.
const char *parse_bgzf_header(const uint8_t *buf, uint64_t len, uint32_t *block_size, uint32_t *header_size)
{
    if (len < 12) return "truncated header";
    if (buf[0] != 0x1f || buf[1] != 0x8b) return "not a gzip member (bad magic)";
    if (buf[2] != 8) return "unsupported compression method";
    if ((buf[3] & 0x04) == 0) return "missing FEXTRA; not BGZF (plain gzip?)";
    uint32_t xlen = rd16(buf + 10);
    uint32_t hs = 12 + xlen;
    if (len < hs) return "truncated extra field";
    const uint8_t *x = buf + 12;
    uint32_t xl = xlen;
    bool found = false;
    uint32_t bsize = 0;
    while (xl >= 4) {
        uint32_t slen = rd16(x + 2);
        if (4 + slen > xl) break;
        if (x[0] == 'B' && x[1] == 'C') {
            if (slen != 2) break;
            bsize = rd16(x + 4);
            found = true;
            break;
        }
        x += 4 + slen;
        xl -= 4 + slen;
    }
    if (!found) return "missing BC subfield; not BGZF";
    uint32_t bs = bsize + 1;
    if (bs > 65536) return "block size exceeds 64 KiB";
    if (bs < hs + 8) return "block size smaller than its own framing";
    *block_size = bs;
    *header_size = hs;
    return nullptr;
}

} // namespace

namespace frhip {

// discover.rs:51-113
int discover(const uint8_t *buf, uint64_t len, uint64_t base, std::vector<frhip_block_span> &spans,
             uint64_t *position, frhip_error *err)
{
    uint64_t pos = 0;
    while (pos < len) {
        const uint8_t *rest = buf + pos;
        uint64_t rl = len - pos;
        if (rl < 12) break;
        uint64_t hl = 12 + (uint64_t)rd16(rest + 10);
        if (hl > rl) break;
        uint32_t bs = 0, hs = 0;
        const char *why = parse_bgzf_header(rest, rl, &bs, &hs);
        if (why) {
            *position = pos;
            return set_err(err, FRHIP_E_INVALID_BLOCK, base + pos, "%s", why);
        }
        if (bs > rl) break;
        uint32_t crc = rd32(rest + bs - 8);
        uint32_t isz = rd32(rest + bs - 4);
        if (isz > 65536) {
            *position = pos;
            return set_err(err, FRHIP_E_INVALID_BLOCK, base + pos, "ISIZE exceeds 64 KiB");
        }
        spans.push_back(frhip_block_span{base + pos, pos + hs, bs - hs - 8ull, isz, crc});
        pos += bs;
    }
    *position = pos;
    return FRHIP_OK;
}

} // namespace frhip

extern "C" int frhip_discover(const uint8_t *buf, uint64_t len, uint64_t base, frhip_block_span *spans,
                              uint64_t cap, uint64_t *n, uint64_t *position, frhip_error *err)
{
    std::vector<frhip_block_span> v;
    int rc = discover(buf, len, base, v, position, err);
    uint64_t k = std::min<uint64_t>(cap, v.size());
    if (spans && k) std::memcpy(spans, v.data(), k * sizeof(frhip_block_span));
    *n = k;
    return rc;
}

// ------------------------------------------------------------------------
// inflate
// ------------------------------------------------------------------------

namespace frhip {

// Plan + upload + launch on run's stream, without waiting. On success *out
// holds the output (its ready event recorded after the kernel), out_off the
// members' output offsets, and *h_status (in run's pinned buffer) receives
// status, produced and CRC per member once the stream passes the last copy.
int inflate_launch(frhip_ctx *ctx, const InflateRun &run, const uint8_t *batch, uint64_t batch_len,
                   const frhip_block_span *spans, uint64_t n, frhip_buf **out, std::vector<uint64_t> &out_off,
                   const uint32_t **h_status, frhip_error *err, frhip_inflate_timings *t, uint64_t align_at)
{
    *out = nullptr;
    *h_status = nullptr;
    (void)hipSetDevice(ctx->ordinal);
    double t0 = t ? now_ms() : 0;
    // ends a timed phase: wait for the stream, add the host time since t0
    auto phase = [&](double &acc) {
        if (!t) return hipSuccess;
        hipError_t r = hipStreamSynchronize(run.stream);
        const double t1 = now_ms();
        acc += t1 - t0;
        t0 = t1;
        return r;
    };
    // BatchPlan::new (backend.rs:95-133)
    std::vector<uint64_t> in_off(n);
    out_off.assign(n, 0);
    std::vector<uint32_t> caps(n), inl(n);
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const frhip_block_span &s = spans[i];
        if (s.isize > 65536)
            return set_err(err, FRHIP_E_INVALID_BLOCK, s.compressed_offset, "ISIZE exceeds 64 KiB");
        uint64_t end = s.payload_start + s.payload_len;
        if (end < s.payload_start || end > batch_len)
            return set_err(err, FRHIP_E_INVALID_BLOCK, s.compressed_offset,
                           "payload extends past the end of the batch");
        in_off[i] = s.payload_start;
        if (i == align_at) total = (total + 4095) & ~4095ull;
        out_off[i] = total;
        inl[i] = (uint32_t)s.payload_len;
        caps[i] = s.isize;
        total += s.isize;
    }
    if (t) {
        t->blocks += n;
        for (uint64_t i = 0; i < n; ++i) t->compressed_bytes += spans[i].payload_len;
        t->inflated_bytes += total;
        (void)phase(t->plan_ms);
    }
    frhip_buf *o = buf_alloc(ctx, total, err);
    if (!o) return err ? err->kind : FRHIP_E_DEVICE;
    if (n == 0) {
        *out = o;
        return FRHIP_OK;
    }
    int rc;
    // compressed bytes, padded: the kernel stages 1 KiB-aligned chunks past the end
    const uint64_t pad = 4096;
    uint8_t *d_in = (uint8_t *)run.in->get(batch_len + pad, err, &rc);
    if (!d_in) { buf_free(o); return rc; }
    // meta: in_off[n] u64, out_off[n] u64, in_len[n] u32, out_cap[n] u32; the
    // same pinned buffer later receives the status read-back (stream order)
    const uint64_t meta_bytes = n * 8 * 2 + n * 4 * 2;
    uint8_t *h_meta = (uint8_t *)run.pinned->get(n * 12 > meta_bytes ? n * 12 : meta_bytes, err, &rc);
    if (!h_meta) { buf_free(o); return rc; }
    std::memcpy(h_meta, in_off.data(), n * 8);
    std::memcpy(h_meta + n * 8, out_off.data(), n * 8);
    std::memcpy(h_meta + n * 16, inl.data(), n * 4);
    std::memcpy(h_meta + n * 20, caps.data(), n * 4);
    uint8_t *d_meta = (uint8_t *)run.meta->get(meta_bytes, err, &rc);
    if (!d_meta) { buf_free(o); return rc; }
    uint32_t *d_st = (uint32_t *)run.status->get(n * 12, err, &rc);
    if (!d_st) { buf_free(o); return rc; }
    hipError_t r = hipMemcpyAsync(d_in, batch, batch_len, hipMemcpyHostToDevice, run.stream);
    if (r == hipSuccess) r = hipMemsetAsync(d_in + batch_len, 0, pad, run.stream);
    if (r == hipSuccess) r = hipMemcpyAsync(d_meta, h_meta, meta_bytes, hipMemcpyHostToDevice, run.stream);
    if (r == hipSuccess && t) r = phase(t->upload_ms);
    if (r != hipSuccess) { buf_free(o); return hip_err(err, r, "uploading the batch"); }
    InflateArgs a{};
    a.in = d_in;
    a.in_off = (const uint64_t *)d_meta;
    a.out_off = (const uint64_t *)(d_meta + n * 8);
    a.in_len = (const uint32_t *)(d_meta + n * 16);
    a.out_cap = (const uint32_t *)(d_meta + n * 20);
    a.out = (uint8_t *)o->ptr;
    a.status = d_st;
    a.produced = d_st + n;
    a.crc = d_st + 2 * n;
    a.n = (uint32_t)n;
    a.queue = nullptr;
    a.order = nullptr;
    a.x2n = ctx->x2n;
    void *d_rec = nullptr;
    if (const uint64_t sb = inflate_scratch_bytes(ctx->inflate_schedule, total, n)) {
        d_rec = run.rec->get(sb, err, &rc);
        if (!d_rec) { buf_free(o); return rc; }
    }
    inflate_scratch_bind(a, d_rec, n);
    a.two_phase_min = ctx->two_phase_min;
    r = launch_inflate(a, (uint32_t)n, run.stream, ctx->inflate_schedule);
    if (r != hipSuccess) { buf_free(o); return hip_err(err, r, "launching inflate"); }
    // the consumer's ordering point: the inflated bytes are final here
    if (EventRef *ev = event_record(run.stream)) {
        buf_attach_event(o, ev);
        event_release(ev);
    }
    if (t) r = phase(t->kernel_ms);
    if (r == hipSuccess) r = hipMemcpyAsync(h_meta, d_st, n * 12, hipMemcpyDeviceToHost, run.stream);
    if (r != hipSuccess) { buf_free(o); return hip_err(err, r, "inflate"); }
    *out = o;
    *h_status = (const uint32_t *)h_meta;
    return FRHIP_OK;
}

// verify_and_offsets (backend.rs:221-256) over read-back statuses: status,
// then size, then CRC, member by member. Returns the index of the first
// member that fails (err set) or n; offsets[0 .. that index] are filled.
uint64_t inflate_verify(const frhip_block_span *spans, uint64_t n, const uint32_t *h_status,
                        const std::vector<uint64_t> &out_off, uint64_t *offsets, frhip_error *err)
{
    const uint32_t *st = h_status, *produced = h_status + n, *crc = h_status + 2 * n;
    offsets[0] = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (st[i] != 0) {
            set_err(err, FRHIP_E_INFLATE, spans[i].compressed_offset, "%s", inflate_status_message(st[i]));
            if (err) err->code = st[i];
            return i;
        }
        if (produced[i] != spans[i].isize) {
            set_err(err, FRHIP_E_SIZE_MISMATCH, spans[i].compressed_offset, "ISIZE says %u bytes, inflated %u",
                    spans[i].isize, produced[i]);
            if (err) { err->expected = spans[i].isize; err->actual = produced[i]; }
            return i;
        }
        if (crc[i] != spans[i].crc32) {
            set_err(err, FRHIP_E_CHECKSUM_MISMATCH, spans[i].compressed_offset, "expected crc32 %#010x, got %#010x",
                    spans[i].crc32, crc[i]);
            if (err) { err->expected = spans[i].crc32; err->actual = crc[i]; }
            return i;
        }
        offsets[i + 1] = out_off[i] + produced[i];
    }
    return n;
}

// Plan + upload + launch + verify on the context's stream. On success *out
// owns the output.
int inflate_device(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len, const frhip_block_span *spans,
                   uint64_t n, frhip_buf **out, uint64_t *offsets, frhip_error *err, frhip_inflate_timings *t,
                   uint64_t align_at)
{
    *out = nullptr;
    const InflateRun run{ctx->stream, &ctx->in_scratch, &ctx->meta_scratch, &ctx->status_scratch, &ctx->rec_scratch,
                         &ctx->pinned};
    std::vector<uint64_t> out_off;
    const uint32_t *h_st = nullptr;
    frhip_buf *o = nullptr;
    int rc = inflate_launch(ctx, run, batch, batch_len, spans, n, &o, out_off, &h_st, err, t, align_at);
    if (rc) return rc;
    if (n == 0) {
        offsets[0] = 0;
        *out = o;
        return FRHIP_OK;
    }
    const double t0 = t ? now_ms() : 0;
    hipError_t r = hipStreamSynchronize(ctx->stream);
    if (r != hipSuccess) { buf_free(o); return hip_err(err, r, "inflate"); }
    if (inflate_verify(spans, n, h_st, out_off, offsets, err) != n) {
        buf_free(o);
        return err ? err->kind : FRHIP_E_INFLATE;
    }
    if (t) t->verify_ms += now_ms() - t0;
    *out = o;
    return FRHIP_OK;
}

} // namespace frhip

extern "C" int frhip_inflate_batch_device(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                                          const frhip_block_span *spans, uint64_t n, frhip_buf **out_data,
                                          uint64_t *out_offsets, frhip_error *err)
{
    return inflate_device(ctx, batch, batch_len, spans, n, out_data, out_offsets, err, nullptr);
}

static int inflate_host(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len, const frhip_block_span *spans,
                        uint64_t n, uint8_t *out, uint64_t out_cap, uint64_t *out_offsets, frhip_inflate_timings *t,
                        frhip_error *err)
{
    frhip_buf *b = nullptr;
    int rc = inflate_device(ctx, batch, batch_len, spans, n, &b, out_offsets, err, t);
    if (rc) return rc;
    if (b->len > out_cap) {
        buf_free(b);
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "output buffer holds %llu bytes, batch needs %llu",
                       (unsigned long long)out_cap, (unsigned long long)b->len);
    }
    const double t0 = t ? now_ms() : 0;
    rc = frhip_buf_copy_to_host(b, out, b->len, err);
    if (t) t->download_ms += now_ms() - t0;
    buf_free(b);
    return rc;
}

extern "C" int frhip_inflate_batch(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                                   const frhip_block_span *spans, uint64_t n, uint8_t *out, uint64_t out_cap,
                                   uint64_t *out_offsets, frhip_error *err)
{
    return inflate_host(ctx, batch, batch_len, spans, n, out, out_cap, out_offsets, nullptr, err);
}

extern "C" int frhip_inflate_batch_device_timed(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                                                const frhip_block_span *spans, uint64_t n, frhip_buf **out_data,
                                                uint64_t *out_offsets, frhip_inflate_timings *t, frhip_error *err)
{
    return inflate_device(ctx, batch, batch_len, spans, n, out_data, out_offsets, err, t);
}

extern "C" int frhip_inflate_batch_timed(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                                         const frhip_block_span *spans, uint64_t n, uint8_t *out, uint64_t out_cap,
                                         uint64_t *out_offsets, frhip_inflate_timings *t, frhip_error *err)
{
    return inflate_host(ctx, batch, batch_len, spans, n, out, out_cap, out_offsets, t, err);
}

extern "C" uint32_t frhip_crc32_combine(uint32_t crc1, uint32_t crc2, uint64_t len2)
{
    uint32_t x2n[32];
    fr_x2n_table(x2n);
    return fr_mulmod(crc1, fr_xpow8(x2n, len2)) ^ crc2;
}

extern "C" int frhip_crc32_blocks(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *offsets,
                                  const uint32_t *lengths, uint64_t n, uint32_t *crc_out, frhip_error *err)
{
    if (n == 0) return FRHIP_OK;
    (void)hipSetDevice(ctx->ordinal);
    for (uint64_t i = 0; i < n; ++i)
        if (offsets[i] + lengths[i] > data->len)
            return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "range %llu lies outside the buffer",
                           (unsigned long long)i);
    int rc;
    uint8_t *d = (uint8_t *)ctx->meta_scratch.get(n * 16, err, &rc);
    if (!d) return rc;
    std::vector<uint8_t> h(n * 12);
    std::memcpy(h.data(), offsets, n * 8);
    std::memcpy(h.data() + n * 8, lengths, n * 4);
    hipError_t r = hipMemcpyAsync(d, h.data(), n * 12, hipMemcpyHostToDevice, ctx->stream);
    uint32_t *d_crc = (uint32_t *)ctx->status_scratch.get(n * 4, err, &rc);
    if (!d_crc) return rc;
    if (r == hipSuccess)
        r = launch_crc32_blocks((const uint8_t *)data->ptr, (const uint64_t *)d, (const uint32_t *)(d + n * 8),
                                (uint32_t)n, d_crc, ctx->x2n, ctx->stream);
    if (r == hipSuccess) r = hipMemcpyAsync(crc_out, d_crc, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(ctx->stream);
    if (r != hipSuccess) return hip_err(err, r, "crc32_blocks");
    return FRHIP_OK;
}
