// writer_host.cpp -- host side of device BGZF compression and the streaming writer.
//
// frhip_compress_batch_device is DeviceBlockCompressor::compress_batch_device
// (crates/fritillaria-core/src/device.rs:350-372): CRC32 of every chunk (the
// range CRC kernel), one compressor workgroup per chunk at a time writing
// each framed block into a 64 KiB slot, the block sizes back to the host for
// the layout, then one gather into a dense stream. frhip_bgzf_writer is
// DeviceBgzfWriter (crates/fritillaria-bgzf/src/device_write.rs:70-259).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/frhip.h"
#include "../../include/frhip_write.h"
#include "deflate_internal.h"
#include "host_internal.h"
#include "pipeline_internal.h"

using namespace frhip;

namespace {

// The 28-byte EOF block: an empty fixed-Huffman block, CRC 0, ISIZE 0.
const uint8_t EOF_BLOCK[28] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 'B', 'C', 2, 0, 0x1b, 0,
                               3,    0,    0, 0, 0, 0, 0, 0, 0, 0};

int malformed(frhip_error *err, uint64_t position, const char *fmt, unsigned long long a = 0,
              unsigned long long b = 0)
{
    int k = set_err(err, FRHIP_E_MALFORMED, position, fmt, a, b);
    if (err) std::strcpy(err->format, "bgzf");
    return k;
}

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

int compress(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *bounds, uint64_t nb1, frhip_buf **out,
             uint64_t *offs, uint8_t *framing_out, uint64_t position, frhip_error *err)
{
    *out = nullptr;
    if (nb1 == 0) return malformed(err, position, "bounds must have at least one entry");
    const uint64_t len = data ? data->len : 0;
    if (bounds[nb1 - 1] > len)
        return malformed(err, position, "bounds reach byte %llu of %llu device bytes",
                         (unsigned long long)bounds[nb1 - 1], (unsigned long long)len);
    for (uint64_t i = 0; i + 1 < nb1; ++i) {
        if (bounds[i + 1] < bounds[i]) return malformed(err, position, "bounds decrease at entry %llu", i + 1);
        if (bounds[i + 1] - bounds[i] > FRHIP_MAX_COMPRESSIBLE_PAYLOAD)
            return malformed(err, position, "chunk of %llu bytes exceeds the %llu-byte maximum",
                             (unsigned long long)(bounds[i + 1] - bounds[i]),
                             (unsigned long long)FRHIP_MAX_COMPRESSIBLE_PAYLOAD);
    }
    const uint64_t n = nb1 - 1;
    offs[0] = 0;
    if (n == 0) {
        *out = buf_alloc(ctx, 0, err);
        return *out ? FRHIP_OK : (err ? err->kind : FRHIP_E_DEVICE);
    }
    if (data->ordinal != ctx->ordinal)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "batch is on device %d but this compressor is on %d",
                       data->ordinal, ctx->ordinal);
    if (n >= (1ull << 31)) return malformed(err, position, "too many chunks in one batch");
    (void)hipSetDevice(ctx->ordinal);
    hipStream_t st = ctx->stream;
    const uint32_t wgs = (uint32_t)std::min<uint64_t>(n, (uint64_t)(ctx->n_cu > 0 ? ctx->n_cu : 256));
    uint64_t p = 0;
    auto take = [&](uint64_t b) { uint64_t o = p; p = align_up(p + b, 256); return o; };
    const uint64_t o_bounds = take(nb1 * 8), o_coff = take(n * 8), o_clen = take(n * 4), o_crc = take(n * 4),
                   o_size = take(n * 4), o_frm = take(n), o_q = take(8), o_offs = take(n * 8),
                   o_wg = take(wgs * DEFLATE_WG_SCRATCH), o_slots = take(n * DEFLATE_SLOT);
    int rc;
    uint8_t *s = (uint8_t *)ctx->wr_scratch.get(p, err, &rc);
    if (!s) return rc;
    // host side of the plan: bounds, and each chunk's range for its CRC
    std::vector<uint8_t> h(o_clen + n * 4);
    std::memcpy(h.data() + o_bounds, bounds, nb1 * 8);
    uint64_t *coff = (uint64_t *)(h.data() + o_coff);
    uint32_t *clen = (uint32_t *)(h.data() + o_clen);
    for (uint64_t i = 0; i < n; ++i) {
        coff[i] = bounds[i];
        clen[i] = (uint32_t)(bounds[i + 1] - bounds[i]);
    }
    hipError_t r = hipMemcpyAsync(s, h.data(), h.size(), hipMemcpyHostToDevice, st);
    DeflateArgs a{};
    a.in = (const uint8_t *)data->ptr;
    a.bounds = (const uint64_t *)(s + o_bounds);
    a.crc = (const uint32_t *)(s + o_crc);
    a.n = (uint32_t)n;
    a.queue = (uint32_t *)(s + o_q);
    a.slots = s + o_slots;
    a.block_size = (uint32_t *)(s + o_size);
    a.framing = s + o_frm;
    a.dist = (uint16_t *)(s + o_wg);
    a.syms = (uint32_t *)(s + o_wg + (uint64_t)wgs * 65536 * 2);
    if (r == hipSuccess)
        r = launch_crc32_blocks(a.in, (const uint64_t *)(s + o_coff), (const uint32_t *)(s + o_clen), (uint32_t)n,
                                (uint32_t *)(s + o_crc), ctx->x2n, st);
    if (r == hipSuccess) r = launch_fill_u64((uint64_t *)a.queue, 0, 1, st);
    if (r == hipSuccess) r = launch_deflate(a, wgs, st);
    std::vector<uint32_t> sizes(n);
    std::vector<uint8_t> frm(n);
    if (r == hipSuccess) r = hipMemcpyAsync(sizes.data(), a.block_size, n * 4, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipMemcpyAsync(frm.data(), a.framing, n, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return hip_err(err, r, "compressing the batch");
    for (uint64_t i = 0; i < n; ++i) offs[i + 1] = offs[i] + sizes[i];
    if (framing_out) std::memcpy(framing_out, frm.data(), n);
    frhip_buf *o = buf_alloc(ctx, offs[n], err);
    if (!o) return err ? err->kind : FRHIP_E_DEVICE;
    r = hipMemcpyAsync(s + o_offs, offs, n * 8, hipMemcpyHostToDevice, st);
    if (r == hipSuccess)
        r = launch_gather_blocks(a.slots, a.block_size, (const uint64_t *)(s + o_offs), (uint32_t)n, (uint8_t *)o->ptr,
                                 st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) {
        buf_free(o);
        return hip_err(err, r, "framing the batch");
    }
    *out = o;
    return FRHIP_OK;
}

} // namespace

extern "C" int frhip_compress_batch_device(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *bounds,
                                           uint64_t n_bounds, frhip_buf **out, uint64_t *block_offsets,
                                           uint8_t *framing, frhip_error *err)
{
    return compress(ctx, data, bounds, n_bounds, out, block_offsets, framing, 0, err);
}

extern "C" uint64_t frhip_framed_size(uint64_t payload_len, uint64_t deflate_len)
{
    // choose_framing (crates/fritillaria-core/src/compress.rs:115-127)
    uint64_t body;
    if (payload_len == 0) body = 2;
    else if (deflate_len > DEFLATE_MAX_STREAM || deflate_len > payload_len + 5) body = payload_len + 5;
    else body = deflate_len;
    return 18 + body + 8;
}

struct frhip_bgzf_writer {
    frhip_ctx *ctx = nullptr;
    frhip_sink sink = nullptr;
    void *user = nullptr;
    uint64_t chunks_per_batch = FRHIP_DEFAULT_CHUNKS_PER_BATCH;
    uint64_t payload_size = FRHIP_DEFAULT_PAYLOAD_SIZE;
    bool finished = false;
    uint64_t blocks = 0, batches = 0, compressed = 0, uncompressed = 0;
    std::vector<uint8_t> host;
    std::vector<uint64_t> offs;
};

extern "C" int frhip_bgzf_writer_create(frhip_ctx *ctx, frhip_sink sink, void *user, uint64_t chunks_per_batch,
                                        uint64_t payload_size, frhip_bgzf_writer **out, frhip_error *err)
{
    *out = nullptr;
    if (!sink) return set_err(err, FRHIP_E_IO, 0, "no sink");
    frhip_bgzf_writer *w = new (std::nothrow) frhip_bgzf_writer;
    if (!w) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    w->ctx = ctx;
    ctx_retain(ctx);
    w->sink = sink;
    w->user = user;
    if (chunks_per_batch) w->chunks_per_batch = chunks_per_batch;
    if (payload_size) w->payload_size = std::min<uint64_t>(payload_size, FRHIP_MAX_COMPRESSIBLE_PAYLOAD);
    *out = w;
    return FRHIP_OK;
}

extern "C" int frhip_bgzf_writer_write_batch(frhip_bgzf_writer *w, const frhip_buf *data, const uint64_t *bounds,
                                             uint64_t n_bounds, frhip_error *err)
{
    if (w->finished) return set_err(err, FRHIP_E_IO, w->compressed, "the writer has been finished");
    if (n_bounds == 0) return malformed(err, w->compressed, "bounds must have at least one entry");
    const uint64_t len = data ? data->len : 0;
    if (bounds[n_bounds - 1] > len)
        return malformed(err, w->compressed, "bounds reach byte %llu of %llu device bytes",
                         (unsigned long long)bounds[n_bounds - 1], (unsigned long long)len);
    // windows of the boundary list overlapping by one, so a batch boundary is
    // always a chunk boundary
    uint64_t start = 0;
    while (start + 1 < n_bounds) {
        const uint64_t end = std::min(start + w->chunks_per_batch + 1, n_bounds);
        const uint64_t nb1 = end - start;
        w->offs.assign(nb1, 0);
        frhip_buf *o = nullptr;
        int rc = compress(w->ctx, data, bounds + start, nb1, &o, w->offs.data(), nullptr, w->compressed, err);
        if (rc != FRHIP_OK) return rc;
        const uint64_t bytes = w->offs[nb1 - 1];
        w->host.resize(bytes);
        if (bytes) {
            (void)hipSetDevice(w->ctx->ordinal);
            hipError_t r = hipMemcpy(w->host.data(), o->ptr, bytes, hipMemcpyDeviceToHost);
            if (r != hipSuccess) {
                buf_free(o);
                return hip_err(err, r, "reading the compressed batch");
            }
        }
        buf_free(o);
        if (bytes && w->sink(w->user, w->host.data(), bytes) != 0)
            return set_err(err, FRHIP_E_IO, w->compressed, "the sink refused %llu bytes", (unsigned long long)bytes);
        w->uncompressed += bounds[end - 1] - bounds[start];
        w->batches += 1;
        w->blocks += nb1 - 1;
        w->compressed += bytes;
        start = end - 1;
    }
    return FRHIP_OK;
}

extern "C" int frhip_bgzf_writer_write_all(frhip_bgzf_writer *w, const frhip_buf *data, frhip_error *err)
{
    const uint64_t len = data ? data->len : 0;
    std::vector<uint64_t> b;
    for (uint64_t x = 0; x < len; x += w->payload_size) b.push_back(x);
    b.push_back(len);
    return frhip_bgzf_writer_write_batch(w, data, b.data(), b.size(), err);
}

extern "C" int frhip_bgzf_writer_finish(frhip_bgzf_writer *w, frhip_error *err)
{
    if (w->finished) return set_err(err, FRHIP_E_IO, w->compressed, "the writer has been finished");
    w->finished = true;
    if (w->sink(w->user, EOF_BLOCK, sizeof EOF_BLOCK) != 0)
        return set_err(err, FRHIP_E_IO, w->compressed, "the sink refused the EOF block");
    return FRHIP_OK;
}

extern "C" uint64_t frhip_bgzf_writer_blocks_written(const frhip_bgzf_writer *w) { return w->blocks; }
extern "C" uint64_t frhip_bgzf_writer_batches_run(const frhip_bgzf_writer *w) { return w->batches; }
extern "C" uint64_t frhip_bgzf_writer_compressed_bytes(const frhip_bgzf_writer *w) { return w->compressed; }
extern "C" uint64_t frhip_bgzf_writer_uncompressed_bytes(const frhip_bgzf_writer *w) { return w->uncompressed; }
extern "C" uint64_t frhip_bgzf_writer_virtual_position(const frhip_bgzf_writer *w)
{
    return w->compressed < (1ull << 48) ? w->compressed << 16 : 0;
}

extern "C" void frhip_bgzf_writer_free(frhip_bgzf_writer *w)
{
    if (!w) return;
    if (w->ctx) ctx_release(w->ctx);
    delete w;
}
