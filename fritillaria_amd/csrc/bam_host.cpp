// bam_host.cpp -- BamDecoder::decode and DeviceBgzfReader behind the C ABI.
//
// decode follows crates/fritillaria-cuda/src/bam.rs:324-574: an empty batch
// gives no records and tail = start; otherwise boundaries on the device, one
// host read of {records, tail, status} to size the columns (bam.rs:447-468),
// then offsets and columns on the device. The reader follows
// crates/fritillaria-bgzf/src/device_read.rs:155-280 line for line in
// behaviour: compressed-member carry, window growth, truncation and
// "stream ends inside a record" errors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "../../include/frhip.h"
#include "frhip_internal.h"
#include "host_internal.h"
#include "bam_layout.h"
#include "pipeline_internal.h"
#include "soa_internal.h"

using namespace frhip;

static const int BAM_WIDTHS[FRHIP_BAM_N_COLUMNS] = {4, 4, 2, 1, 4, 4, 4, 4, 8, 4, 4, 4, 4, 4, 8};

struct frhip_bam_batch {
    uint64_t n = 0;
    uint64_t tail = 0;
    uint64_t fallbacks = 0;
    frhip_buf *cols[FRHIP_BAM_N_COLUMNS] = {};
};

namespace {

int empty_batch(frhip_ctx *ctx, uint64_t start, bool with_voff, frhip_bam_batch **out, frhip_error *err)
{
    frhip_bam_batch *b = new (std::nothrow) frhip_bam_batch;
    if (!b) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    b->tail = start;
    for (int i = 0; i < FRHIP_BAM_N_COLUMNS; ++i) {
        if (i == FRHIP_BAM_VIRTUAL_OFFSETS && !with_voff) continue;
        b->cols[i] = buf_alloc(ctx, 0, err);
        if (!b->cols[i]) {
            frhip_bam_batch_free(b);
            return err ? err->kind : FRHIP_E_DEVICE;
        }
    }
    *out = b;
    return FRHIP_OK;
}

} // namespace

namespace {

// The decode with, when t is given, stream synchronisations between its
// phases and each phase's host time added to *t (bam.rs:324-574 timings).
int bam_decode(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *block_offsets, uint64_t nb,
               const uint64_t *coffsets, uint64_t start, frhip_bam_batch **out, frhip_decode_timings *t,
               frhip_error *err)
{
    *out = nullptr;
    double t0 = t ? now_ms() : 0;
    auto phase = [&](double &acc) {
        if (!t) return hipSuccess;
        hipError_t r = hipStreamSynchronize(ctx->stream);
        const double t1 = now_ms();
        acc += t1 - t0;
        t0 = t1;
        return r;
    };
    if (t) ++t->batches;
    if (!data || data->len == 0) return empty_batch(ctx, start, coffsets != nullptr, out, err);
    if (t) {
        t->blocks += nb;
        t->bytes += data->len;
    }
    if (data->ordinal != ctx->ordinal)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "batch is on device %d but this decoder is on %d",
                       data->ordinal, ctx->ordinal);
    const uint64_t len = data->len;
    if (nb == 0 || block_offsets[0] != 0 || block_offsets[nb] != len)
        return set_err(err, FRHIP_E_MALFORMED, 0, "block starts must span the buffer from 0 to its length");
    for (uint64_t i = 0; i < nb; ++i)
        if (block_offsets[i + 1] < block_offsets[i] || block_offsets[i + 1] - block_offsets[i] > 65536)
            return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "member %llu has an invalid extent",
                           (unsigned long long)i);
    (void)hipSetDevice(ctx->ordinal);
    BamScratchLayout L = bam_layout(nb);
    int rc;
    uint8_t *s = (uint8_t *)ctx->bam_scratch.get(L.bytes, err, &rc);
    if (!s) return rc;
    hipError_t r = hipMemcpyAsync(s + L.o_bs, block_offsets, (nb + 1) * 8, hipMemcpyHostToDevice, ctx->stream);
    if (r == hipSuccess && coffsets)
        r = hipMemcpyAsync(s + L.o_coff, coffsets, nb * 8, hipMemcpyHostToDevice, ctx->stream);
    if (r == hipSuccess) r = phase(t ? t->setup_ms : t0);
    if (r != hipSuccess) return hip_err(err, r, "uploading block starts");

    BamSpecArgs sa{};
    sa.buf = (const uint8_t *)data->ptr;
    sa.len = len;
    sa.block_starts = (const uint64_t *)(s + L.o_bs);
    sa.nb = (uint32_t)nb;
    sa.window = ctx->bam_window;
    sa.spec_ok = s + L.o_ok;
    sa.n_surv = s + L.o_ns;
    sa.surv_pos = (uint64_t *)(s + L.o_spos);
    sa.surv_cnt = (uint32_t *)(s + L.o_scnt);
    sa.land = (uint64_t *)(s + L.o_land);
    sa.exitp = (uint64_t *)(s + L.o_exit);
    sa.cnt_tail = (uint32_t *)(s + L.o_ctail);
    sa.complete = s + L.o_comp;
    sa.rel_offs = (uint16_t *)(s + L.o_rel);
    r = launch_bam_spec(sa, ctx->stream);
    if (r == hipSuccess) r = phase(t ? t->scan_ms : t0);
    if (r != hipSuccess) return hip_err(err, r, "launching bam_spec");

    BamReconcileArgs ra{};
    ra.buf = sa.buf;
    ra.len = len;
    ra.block_starts = sa.block_starts;
    ra.nb = (uint32_t)nb;
    ra.start = start;
    ra.spec_ok = sa.spec_ok;
    ra.n_surv = sa.n_surv;
    ra.surv_pos = sa.surv_pos;
    ra.surv_cnt = sa.surv_cnt;
    ra.exitp = sa.exitp;
    ra.cnt_tail = sa.cnt_tail;
    ra.complete = sa.complete;
    ra.entry = (uint64_t *)(s + L.o_entry);
    ra.count = (uint32_t *)(s + L.o_count);
    ra.first_index = (uint64_t *)(s + L.o_first);
    ra.mode = s + L.o_mode;
    ra.totals = (uint64_t *)(s + L.o_totals);
    ra.fast_scan = (uint64_t *)(s + L.o_fscan);
    ra.fast_state = (uint32_t *)(s + L.o_fstate);
    r = launch_bam_reconcile(ra, ctx->stream);
    if (r == hipSuccess) r = phase(t ? t->reconcile_ms : t0);
    if (r != hipSuccess) return hip_err(err, r, "launching bam_reconcile");
    uint64_t totals[4];
    r = hipMemcpyAsync(totals, ra.totals, sizeof totals, hipMemcpyDeviceToHost, ctx->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(ctx->stream);
    if (r == hipSuccess) r = phase(t ? t->totals_ms : t0);
    if (r != hipSuccess) return hip_err(err, r, "bam boundaries");
    if (totals[2] != 0) {
        int k = set_err(err, FRHIP_E_MALFORMED, totals[1], "malformed record reached from a confirmed boundary");
        if (err) {
            err->code = (uint32_t)totals[2];
            std::strcpy(err->format, "bam");
        }
        return k;
    }
    const uint64_t n = totals[0];
    frhip_bam_batch *b = new (std::nothrow) frhip_bam_batch;
    if (!b) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    b->n = n;
    b->tail = totals[1];
    b->fallbacks = totals[3];
    for (int i = 0; i < FRHIP_BAM_N_COLUMNS; ++i) {
        if (i == FRHIP_BAM_VIRTUAL_OFFSETS && !coffsets) continue;
        b->cols[i] = buf_alloc(ctx, n * BAM_WIDTHS[i], err);
        if (!b->cols[i]) {
            frhip_bam_batch_free(b);
            return err ? err->kind : FRHIP_E_DEVICE;
        }
    }
    BamEmitArgs ea{};
    ea.buf = sa.buf;
    ea.block_starts = sa.block_starts;
    ea.coffsets = coffsets ? (const uint64_t *)(s + L.o_coff) : nullptr;
    ea.nb = (uint32_t)nb;
    ea.entry = ra.entry;
    ea.count = ra.count;
    ea.first_index = ra.first_index;
    ea.mode = ra.mode;
    ea.surv_cnt = sa.surv_cnt;
    ea.rel_offs = sa.rel_offs;
    ea.record_offsets = (uint64_t *)b->cols[FRHIP_BAM_RECORD_OFFSETS]->ptr;
    ea.virtual_offsets = coffsets ? (uint64_t *)b->cols[FRHIP_BAM_VIRTUAL_OFFSETS]->ptr : nullptr;
    r = n ? launch_bam_emit(ea, ctx->stream) : hipSuccess;
    if (r == hipSuccess) r = phase(t ? t->emit_ms : t0);
    if (r != hipSuccess) {
        frhip_bam_batch_free(b);
        return hip_err(err, r, "launching bam_emit");
    }
    BamDecodeArgs da{};
    da.buf = sa.buf;
    da.len = len;
    da.n = n;
    da.n_dev = nullptr;
    da.c.ref_id = (int32_t *)b->cols[0]->ptr;
    da.c.pos = (int32_t *)b->cols[1]->ptr;
    da.c.flags = (uint16_t *)b->cols[2]->ptr;
    da.c.mapq = (uint8_t *)b->cols[3]->ptr;
    da.c.l_seq = (uint32_t *)b->cols[4]->ptr;
    da.c.next_ref_id = (int32_t *)b->cols[5]->ptr;
    da.c.next_pos = (int32_t *)b->cols[6]->ptr;
    da.c.tlen = (int32_t *)b->cols[7]->ptr;
    da.c.record_offsets = (uint64_t *)b->cols[8]->ptr;
    da.c.cigar_start = (uint32_t *)b->cols[9]->ptr;
    da.c.seq_start = (uint32_t *)b->cols[10]->ptr;
    da.c.qual_start = (uint32_t *)b->cols[11]->ptr;
    da.c.aux_start = (uint32_t *)b->cols[12]->ptr;
    da.c.record_end = (uint32_t *)b->cols[13]->ptr;
    r = launch_bam_decode(da, n, ctx->stream);
    // the consumer's ordering point for every column of the batch
    if (r == hipSuccess)
        if (EventRef *ev = event_record(ctx->stream)) {
            for (frhip_buf *c : b->cols) buf_attach_event(c, ev);
            event_release(ev);
        }
    if (r == hipSuccess) r = hipStreamSynchronize(ctx->stream);
    if (r == hipSuccess) r = phase(t ? t->decode_ms : t0);
    if (r != hipSuccess) {
        frhip_bam_batch_free(b);
        return hip_err(err, r, "bam decode");
    }
    if (t) t->records += n;
    *out = b;
    return FRHIP_OK;
}

} // namespace

extern "C" int frhip_bam_decode(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *block_offsets,
                                uint64_t nb, const uint64_t *coffsets, uint64_t start, frhip_bam_batch **out,
                                frhip_error *err)
{
    return bam_decode(ctx, data, block_offsets, nb, coffsets, start, out, nullptr, err);
}

extern "C" int frhip_bam_decode_timed(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *block_offsets,
                                      uint64_t nb, const uint64_t *coffsets, uint64_t start, frhip_bam_batch **out,
                                      frhip_decode_timings *t, frhip_error *err)
{
    return bam_decode(ctx, data, block_offsets, nb, coffsets, start, out, t, err);
}

extern "C" uint64_t frhip_bam_batch_len(const frhip_bam_batch *b) { return b->n; }
extern "C" uint64_t frhip_bam_batch_tail(const frhip_bam_batch *b) { return b->tail; }
extern "C" uint64_t frhip_bam_batch_fallbacks(const frhip_bam_batch *b) { return b->fallbacks; }
extern "C" const frhip_buf *frhip_bam_batch_column(const frhip_bam_batch *b, int column)
{
    if (!b || column < 0 || column >= FRHIP_BAM_N_COLUMNS) return nullptr;
    return b->cols[column];
}
extern "C" void frhip_bam_batch_free(frhip_bam_batch *b)
{
    if (!b) return;
    for (auto *c : b->cols) buf_free(c);
    delete b;
}

// crates/fritillaria-bam/src/columnar/header.rs:43-103
extern "C" int frhip_bam_parse_header(const uint8_t *buf, uint64_t len, uint64_t *records_start, uint32_t *n_ref,
                                      frhip_error *err)
{
    auto mal = [&](uint64_t pos, const char *why) {
        int k = set_err(err, FRHIP_E_MALFORMED, pos, "%s", why);
        if (err) std::strcpy(err->format, "bam header");
        return k;
    };
    auto rd32 = [&](uint64_t p) -> int32_t {
        return (int32_t)((uint32_t)buf[p] | ((uint32_t)buf[p + 1] << 8) | ((uint32_t)buf[p + 2] << 16) |
                         ((uint32_t)buf[p + 3] << 24));
    };
    if (len < 12) return mal(0, "truncated header");
    if (std::memcmp(buf, "BAM\1", 4) != 0) return mal(0, "bad magic; expected BAM\\1");
    int32_t l_text = rd32(4);
    if (l_text < 0) return mal(4, "negative l_text");
    uint64_t text_end = 8 + (uint64_t)l_text;
    if (text_end > len) return mal(4, "l_text runs past the end of the buffer");
    if (text_end + 4 > len) return mal(text_end, "truncated n_ref");
    int32_t nref = rd32(text_end);
    if (nref < 0) return mal(text_end, "negative n_ref");
    uint64_t pos = text_end + 4;
    for (int32_t i = 0; i < nref; ++i) {
        if (pos + 4 > len) return mal(pos, "truncated l_name");
        int32_t l_name = rd32(pos);
        if (l_name < 0) return mal(pos, "negative l_name");
        pos += 4;
        if (pos + (uint64_t)l_name > len) return mal(pos, "truncated name");
        pos += (uint64_t)l_name;
        if (pos + 4 > len) return mal(pos, "truncated l_ref");
        pos += 4;
    }
    *records_start = pos;
    *n_ref = (uint32_t)nref;
    return FRHIP_OK;
}

// ------------------------------------------------------------------------
// DeviceBgzfReader
// ------------------------------------------------------------------------

struct frhip_reader {
    frhip_ctx *ctx = nullptr;
    const uint8_t *src = nullptr;
    uint64_t src_len = 0, src_pos = 0;
    uint64_t blocks_per_batch = 256;
    // The current batch's compressed bytes are src[batch_offset, batch_offset
    // + comp_len); the carry (a trailing partial member, or the members from
    // the one holding a carried record on) is the last carry_len of them, and
    // always ends at src_pos. Both are views of the caller's buffer, which
    // outlives the reader: nothing is copied on the host.
    uint64_t comp_len = 0;
    uint64_t batch_offset = 0;
    uint64_t carry_len = 0;
    uint64_t carry_start = 0;
    std::vector<frhip_block_span> spans;
    std::vector<uint64_t> offsets, coffsets;
    bool inner_exhausted = false;
    // Read-ahead: the members [cache_coff[0], ...) inflated by one launch
    // (cache_uoff: their output offsets, n + 1), which the next batches are
    // cut from as views instead of inflating each batch on its own.
    BufOwner *cache = nullptr;
    EventRef *cache_ready = nullptr;
    std::vector<uint64_t> cache_coff, cache_uoff;
    uint64_t cache_cend = 0; // compressed end of the last cached member
    // Window bounds (0: the defaults below).
    uint64_t ra_members = 0, ra_bytes = 0;
    // The next window, inflating on the reader's own stream with its own
    // scratch while batches are cut from the current one; it becomes the
    // cache when a batch needs it.
    hipStream_t ahead_stream = nullptr;
    Scratch ahead_in, ahead_meta, ahead_status, ahead_rec;
    PinnedScratch ahead_pinned;
    // The window's upload and launch run on a helper thread: an upload from
    // pageable memory holds up the thread that issues it until it is staged.
    // The thread writes ahead_data, ahead_out_off, ahead_status_h and
    // ahead_rc; they are read only after it is joined.
    std::thread ahead_thread;
    int ahead_rc = 0;
    bool ahead_pending = false;
    uint64_t ahead_c0 = 0;
    frhip_buf *ahead_data = nullptr;
    const uint32_t *ahead_status_h = nullptr;
    std::vector<frhip_block_span> ahead_spans;
    std::vector<uint64_t> ahead_out_off;
};

namespace {

// Members one read-ahead launch inflates: enough for the two-phase form (one
// member per lane) to fill the device, and many batches' worth.
uint64_t read_ahead_members(const frhip_reader *r)
{
    return r->ra_members ? r->ra_members : std::max<uint64_t>(32768, 16 * r->blocks_per_batch);
}
constexpr uint64_t READ_AHEAD_MAX_BYTES = 1ull << 30; // compressed bytes per read-ahead launch
uint64_t read_ahead_bytes(const frhip_reader *r) { return r->ra_bytes ? r->ra_bytes : READ_AHEAD_MAX_BYTES; }

void drop_cache(frhip_reader *r)
{
    owner_release(r->cache);
    event_release(r->cache_ready);
    r->cache = nullptr;
    r->cache_ready = nullptr;
    r->cache_coff.clear();
    r->cache_uoff.clear();
    r->cache_cend = 0;
}

// Waits for the window inflating ahead, if any, and discards it.
void drop_ahead(frhip_reader *r)
{
    if (!r->ahead_pending) return;
    if (r->ahead_thread.joinable()) r->ahead_thread.join();
    (void)hipStreamSynchronize(r->ahead_stream);
    buf_free(r->ahead_data);
    r->ahead_data = nullptr;
    r->ahead_pending = false;
}

// The current batch (r->spans) as a view into the read-ahead, or nullptr when
// the read-ahead does not hold all of its members.
frhip_buf *cache_view(frhip_reader *r, frhip_error *err)
{
    const uint64_t k = r->spans.size();
    if (!r->cache || k == 0) return nullptr;
    auto it = std::lower_bound(r->cache_coff.begin(), r->cache_coff.end(), r->spans[0].compressed_offset);
    if (it == r->cache_coff.end() || *it != r->spans[0].compressed_offset) return nullptr;
    const uint64_t i0 = (uint64_t)(it - r->cache_coff.begin());
    if (i0 + k > r->cache_coff.size() || r->cache_coff[i0 + k - 1] != r->spans[k - 1].compressed_offset) return nullptr;
    frhip_buf *b = new (std::nothrow) frhip_buf;
    if (!b) {
        set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
        return nullptr;
    }
    b->ptr = (uint8_t *)r->cache->ptr + r->cache_uoff[i0];
    b->len = r->cache_uoff[i0 + k] - r->cache_uoff[i0];
    b->ordinal = r->ctx->ordinal;
    r->cache->refs.fetch_add(1);
    b->owner = r->cache;
    buf_attach_event(b, r->cache_ready);
    for (uint64_t j = 0; j <= k; ++j) r->offsets[j] = r->cache_uoff[i0 + j] - r->cache_uoff[i0];
    return b;
}

// The whole members from compressed offset c0 on that one window holds
// (discovery stops at a partial or bad member).
void window_members(const frhip_reader *r, uint64_t c0, std::vector<frhip_block_span> &ra)
{
    ra.clear();
    if (c0 >= r->src_len) return;
    const uint64_t avail = std::min<uint64_t>(r->src_len - c0, read_ahead_bytes(r));
    uint64_t consumed = 0;
    frhip_error e{};
    (void)discover(r->src + c0, avail, c0, ra, &consumed, &e);
    if (ra.size() > read_ahead_members(r)) ra.resize(read_ahead_members(r));
}

// Compressed bytes of the window ra read from c0 (spans' payload offsets are
// relative to c0).
uint64_t window_bytes(const std::vector<frhip_block_span> &ra)
{
    return ra.empty() ? 0 : ra.back().payload_start + ra.back().payload_len + 8;
}

// A finished window becomes the cache: its members before `good`, the first
// that failed verification (good == ra.size(): none failed), so that the
// batch holding a bad member is inflated on its own and reports the failure
// as it always has. Returns whether every member passed.
bool adopt_window(frhip_reader *r, frhip_buf *data, uint64_t c0, const std::vector<frhip_block_span> &ra,
                  uint64_t good, const uint64_t *offs)
{
    drop_cache(r);
    if (good == 0) {
        buf_free(data);
        return false;
    }
    r->cache = new (std::nothrow) BufOwner;
    if (!r->cache) {
        buf_free(data);
        return false;
    }
    r->cache->ptr = data->ptr;
    r->cache->ordinal = data->ordinal;
    r->cache->cap = data->cap;
    r->cache_ready = data->ready;
    data->ptr = nullptr;
    data->ready = nullptr;
    buf_free(data);
    r->cache_uoff.assign(offs, offs + good + 1);
    r->cache_coff.resize(good);
    for (uint64_t i = 0; i < good; ++i) r->cache_coff[i] = ra[i].compressed_offset;
    r->cache_cend = c0 + ra[good - 1].payload_start + ra[good - 1].payload_len + 8;
    return good == ra.size();
}

// Launches the window after the cache on the reader's stream. It starts
// `overlap` compressed bytes before the cache's end, so that a batch that
// straddles that end, with its carry, lies wholly in the new window; a batch
// that does not (a carry longer than that) is inflated by a refill instead.
void launch_ahead(frhip_reader *r)
{
    if (r->ahead_pending || !r->cache || r->cache_coff.empty() || r->cache_cend >= r->src_len) return;
    const uint64_t span = r->cache_cend - r->cache_coff[0];
    const uint64_t overlap = std::min<uint64_t>(2 * r->blocks_per_batch * 65536ull + (8ull << 20), span / 2);
    // the first cached member at or past cache_cend - overlap (the last one
    // when that point lies inside it)
    auto it = std::lower_bound(r->cache_coff.begin(), r->cache_coff.end(), r->cache_cend - overlap);
    const uint64_t from = it == r->cache_coff.end() ? r->cache_coff.back() : *it;
    window_members(r, from, r->ahead_spans);
    // a window that would not reach past the cache adds nothing
    if (r->ahead_spans.empty() || from + window_bytes(r->ahead_spans) <= r->cache_cend) return;
    if (!r->ahead_stream && hipStreamCreateWithFlags(&r->ahead_stream, hipStreamNonBlocking) != hipSuccess) {
        r->ahead_stream = nullptr;
        return;
    }
    r->ahead_c0 = from;
    r->ahead_data = nullptr;
    r->ahead_rc = FRHIP_E_DEVICE;
    r->ahead_pending = true;
    auto work = [r, from]() {
        const InflateRun run{r->ahead_stream, &r->ahead_in, &r->ahead_meta, &r->ahead_status, &r->ahead_rec,
                             &r->ahead_pinned};
        frhip_error e{};
        frhip_buf *d = nullptr;
        const int rc = inflate_launch(r->ctx, run, r->src + from, window_bytes(r->ahead_spans), r->ahead_spans.data(),
                                      r->ahead_spans.size(), &d, r->ahead_out_off, &r->ahead_status_h, &e, nullptr);
        r->ahead_data = rc == FRHIP_OK ? d : nullptr;
        r->ahead_rc = rc;
    };
    try {
        r->ahead_thread = std::thread(work);
    } catch (...) {
        // no thread to be had (std::system_error): no window ahead; the
        // next batch past the cache refills it
        r->ahead_pending = false;
    }
}

// Makes the window inflated ahead the cache when it holds every member of
// the current batch (r->spans): waits for it, verifies it and launches the
// next. Returns whether it was taken.
bool take_ahead(frhip_reader *r)
{
    const uint64_t k = r->spans.size();
    if (!r->ahead_pending || k == 0) return false;
    auto holds = [&](uint64_t coff) {
        auto it = std::lower_bound(r->ahead_spans.begin(), r->ahead_spans.end(), coff,
                                   [](const frhip_block_span &s, uint64_t v) { return s.compressed_offset < v; });
        return it != r->ahead_spans.end() && it->compressed_offset == coff;
    };
    if (!holds(r->spans[0].compressed_offset) || !holds(r->spans[k - 1].compressed_offset)) return false;
    r->ahead_thread.join();
    r->ahead_pending = false;
    frhip_buf *data = r->ahead_data;
    r->ahead_data = nullptr;
    if (r->ahead_rc != FRHIP_OK) {
        (void)hipStreamSynchronize(r->ahead_stream);
        buf_free(data);
        return false;
    }
    if (hipStreamSynchronize(r->ahead_stream) != hipSuccess) {
        buf_free(data);
        return false;
    }
    std::vector<uint64_t> offs(r->ahead_spans.size() + 1);
    frhip_error e{};
    const uint64_t good = inflate_verify(r->ahead_spans.data(), r->ahead_spans.size(), r->ahead_status_h,
                                         r->ahead_out_off, offs.data(), &e);
    if (adopt_window(r, data, r->ahead_c0, r->ahead_spans, good, offs.data())) launch_ahead(r);
    return true;
}

// Inflates a window from the batch's first member on, on the context's
// stream, waiting for it; then launches the next one ahead.
void refill_cache(frhip_reader *r)
{
    drop_cache(r);
    if (r->spans.empty()) return;
    const uint64_t c0 = r->spans[0].compressed_offset;
    std::vector<frhip_block_span> ra;
    window_members(r, c0, ra);
    if (ra.size() <= r->spans.size()) return;
    frhip_ctx *ctx = r->ctx;
    const InflateRun run{ctx->stream, &ctx->in_scratch, &ctx->meta_scratch, &ctx->status_scratch, &ctx->rec_scratch,
                         &ctx->pinned};
    frhip_buf *data = nullptr;
    std::vector<uint64_t> out_off;
    const uint32_t *hst = nullptr;
    frhip_error e{};
    if (inflate_launch(ctx, run, r->src + c0, window_bytes(ra), ra.data(), ra.size(), &data, out_off, &hst, &e,
                       nullptr) != FRHIP_OK)
        return;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        buf_free(data);
        return;
    }
    std::vector<uint64_t> offs(ra.size() + 1);
    const uint64_t good = inflate_verify(ra.data(), ra.size(), hst, out_off, offs.data(), &e);
    if (adopt_window(r, data, c0, ra, good, offs.data())) launch_ahead(r);
}

} // namespace

extern "C" int frhip_reader_open(frhip_ctx *ctx, const uint8_t *data, uint64_t len, uint64_t blocks_per_batch,
                                 frhip_reader **out, frhip_error *err)
{
    frhip_reader *r = new (std::nothrow) frhip_reader;
    if (!r) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    r->ctx = ctx;
    ctx_retain(ctx);
    r->src = data;
    r->src_len = len;
    r->blocks_per_batch = blocks_per_batch ? blocks_per_batch : 256;
    *out = r;
    return FRHIP_OK;
}

extern "C" void frhip_reader_free(frhip_reader *r)
{
    if (!r) return;
    drop_ahead(r);
    drop_cache(r);
    if (r->ahead_stream) (void)hipStreamDestroy(r->ahead_stream);
    r->ahead_in.release();
    r->ahead_meta.release();
    r->ahead_status.release();
    r->ahead_rec.release();
    r->ahead_pinned.release();
    ctx_release(r->ctx);
    delete r;
}

extern "C" int frhip_reader_set_read_ahead(frhip_reader *r, uint64_t members, uint64_t bytes, frhip_error *err)
{
    (void)err;
    drop_ahead(r);
    drop_cache(r);
    r->ra_members = members;
    r->ra_bytes = bytes;
    return FRHIP_OK;
}

extern "C" int frhip_reader_carry_from(frhip_reader *r, uint64_t tail, frhip_error *err)
{
    // device_read.rs:155-182
    uint64_t total = r->offsets.empty() ? 0 : r->offsets.back();
    if (tail > total)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "carry offset %llu is past the batch's %llu bytes",
                       (unsigned long long)tail, (unsigned long long)total);
    if (tail == total) {
        r->carry_start = 0;
        return FRHIP_OK;
    }
    uint64_t block = (uint64_t)(std::upper_bound(r->offsets.begin(), r->offsets.end(), tail) - r->offsets.begin()) - 1;
    uint64_t block_start = r->spans[block].compressed_offset - r->batch_offset;
    r->carry_len = r->comp_len - block_start;
    r->carry_start = tail - r->offsets[block];
    return FRHIP_OK;
}

extern "C" int frhip_reader_next_batch(frhip_reader *r, int *has_batch, frhip_buf **out_data,
                                       const uint64_t **out_offsets, const uint64_t **coffsets, uint64_t *n_blocks,
                                       uint64_t *start, frhip_error *err)
{
    *has_batch = 0;
    *out_data = nullptr;
    for (;;) {
        if (r->inner_exhausted && r->carry_len == 0) return FRHIP_OK;
        if (r->inner_exhausted && r->comp_len != 0 && r->carry_len == r->comp_len) {
            int k = set_err(err, FRHIP_E_MALFORMED, r->batch_offset,
                            "stream ends inside a record: the trailing bytes are not a complete record");
            if (err) std::strcpy(err->format, "bgzf");
            return k;
        }
        // the carry becomes the front of the next batch, which continues it
        r->batch_offset = r->src_pos - r->carry_len;
        r->comp_len = r->carry_len;
        r->carry_len = 0;
        const uint64_t want = r->blocks_per_batch * 65536ull;
        const uint64_t remaining = r->src_len - r->src_pos;
        const uint64_t take = std::min(want, remaining);
        r->comp_len += take;
        r->src_pos += take;
        if (remaining < want) r->inner_exhausted = true;
        const uint8_t *comp = r->src + r->batch_offset;

        r->spans.clear();
        uint64_t consumed = 0;
        int rc = discover(comp, r->comp_len, r->batch_offset, r->spans, &consumed, err);
        if (rc) return rc;
        r->carry_len = r->comp_len - consumed;
        if (r->inner_exhausted && r->carry_len != 0) {
            int k = set_err(err, FRHIP_E_MALFORMED, r->batch_offset + consumed,
                            "truncated BGZF: %llu trailing bytes are not a complete block",
                            (unsigned long long)r->carry_len);
            if (err) std::strcpy(err->format, "bgzf");
            return k;
        }
        if (r->spans.empty()) {
            if (r->inner_exhausted) return FRHIP_OK;
            continue;
        }
        r->offsets.assign(r->spans.size() + 1, 0);
        frhip_buf *data = cache_view(r, err);
        if (!data && take_ahead(r)) data = cache_view(r, err);
        if (!data) {
            refill_cache(r);
            data = cache_view(r, err);
        }
        if (!data) {
            rc = inflate_device(r->ctx, comp, r->comp_len, r->spans.data(), r->spans.size(), &data,
                                r->offsets.data(), err, nullptr);
            if (rc) return rc;
        }
        if (data->len == 0) {
            buf_free(data);
            if (r->inner_exhausted) return FRHIP_OK;
            continue;
        }
        r->coffsets.resize(r->spans.size());
        for (size_t i = 0; i < r->spans.size(); ++i) r->coffsets[i] = r->spans[i].compressed_offset;
        *has_batch = 1;
        *out_data = data;
        *out_offsets = r->offsets.data();
        *coffsets = r->coffsets.data();
        *n_blocks = r->spans.size();
        *start = r->carry_start;
        r->carry_start = 0;
        return FRHIP_OK;
    }
}

// ---- structure-of-arrays extras (resolved CIGAR, packed sequence, aux index) ----

static const int SOA_WIDTHS[FRHIP_SOA_N_COLUMNS] = {8, 4, 8, 1, 8, 16};

struct frhip_bam_soa {
    uint64_t n = 0;
    frhip_buf *cols[FRHIP_SOA_N_COLUMNS] = {};
};

extern "C" void frhip_bam_soa_free(frhip_bam_soa *s)
{
    if (!s) return;
    for (auto *c : s->cols) buf_free(c);
    delete s;
}
extern "C" uint64_t frhip_bam_soa_len(const frhip_bam_soa *s) { return s->n; }
extern "C" const frhip_buf *frhip_bam_soa_column(const frhip_bam_soa *s, int column)
{
    if (!s || column < 0 || column >= FRHIP_SOA_N_COLUMNS) return nullptr;
    return s->cols[column];
}

extern "C" int frhip_bam_soa_build(frhip_ctx *ctx, const frhip_buf *data, const frhip_bam_batch *records,
                             frhip_bam_soa **out, frhip_error *err)
{
    *out = nullptr;
    frhip_bam_soa *o = new (std::nothrow) frhip_bam_soa;
    if (!o) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    const uint64_t n = records ? records->n : 0;
    o->n = n;
    (void)hipSetDevice(ctx->ordinal);
    hipStream_t st = ctx->stream;
    auto fail = [&](int rc) {
        frhip_bam_soa_free(o);
        return rc;
    };
    if (n == 0 || !data) {
        const uint64_t lens[FRHIP_SOA_N_COLUMNS] = {8, 0, 8, 0, 8, 0};
        for (int c = 0; c < FRHIP_SOA_N_COLUMNS; ++c) {
            o->cols[c] = buf_alloc(ctx, lens[c], err);
            if (!o->cols[c]) return fail(err ? err->kind : FRHIP_E_DEVICE);
            if (lens[c]) {
                hipError_t r = hipMemsetAsync(o->cols[c]->ptr, 0, 8, st);
                if (r == hipSuccess) r = hipStreamSynchronize(st);
                if (r != hipSuccess) return fail(hip_err(err, r, "empty offsets"));
            }
        }
        *out = o;
        return FRHIP_OK;
    }
    // the three offset columns are outputs; per-record counts and tile sums live in scratch
    for (int c : {FRHIP_SOA_CIGAR_OFFSETS, FRHIP_SOA_SEQUENCE_OFFSETS, FRHIP_SOA_AUX_OFFSETS}) {
        o->cols[c] = buf_alloc(ctx, (n + 1) * 8, err);
        if (!o->cols[c]) return fail(err ? err->kind : FRHIP_E_DEVICE);
    }
    int rc;
    const SoaScratchLayout L = soa_scratch_layout(n);
    uint8_t *s = (uint8_t *)ctx->soa_scratch.get(L.bytes, err, &rc);
    if (!s) return fail(rc);
    SoaArgs a{};
    soa_bind_scratch(a, s, L);
    a.buf = (const uint8_t *)data->ptr;
    a.n = n;
    a.cap_rec = n;
    a.cap_ops = a.cap_seq = a.cap_aux = 0;
    soa_bind_columns(a, (const uint64_t *)records->cols[FRHIP_BAM_RECORD_OFFSETS]->ptr,
                     (const uint32_t *)records->cols[FRHIP_BAM_SEQUENCE_LEN]->ptr,
                     (const uint32_t *)records->cols[FRHIP_BAM_CIGAR_START]->ptr,
                     (const uint32_t *)records->cols[FRHIP_BAM_SEQUENCE_START]->ptr,
                     (const uint32_t *)records->cols[FRHIP_BAM_AUX_START]->ptr,
                     (const uint32_t *)records->cols[FRHIP_BAM_RECORD_END]->ptr);
    a.op_off = (uint64_t *)o->cols[FRHIP_SOA_CIGAR_OFFSETS]->ptr;
    a.seq_off = (uint64_t *)o->cols[FRHIP_SOA_SEQUENCE_OFFSETS]->ptr;
    a.aux_off = (uint64_t *)o->cols[FRHIP_SOA_AUX_OFFSETS]->ptr;
    hipError_t r = launch_fill_u64((uint64_t *)a.first_bad, ~0ull, 1, st);
    if (r == hipSuccess) r = launch_soa_count(a, st);
    if (r == hipSuccess) r = launch_soa_scan(a, st);
    uint64_t tot[SOA_N_RESULT + 1];
    if (r == hipSuccess) r = hipMemcpyAsync(tot, a.result, sizeof(tot), hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return fail(hip_err(err, r, "SoA counts"));
    const uint64_t bad = tot[SOA_N_RESULT];
    if (bad != ~0ull) {
        uint64_t at = 0;
        (void)hipMemcpy(&at, a.record_offsets + bad, 8, hipMemcpyDeviceToHost);
        int k = set_err(err, FRHIP_E_MALFORMED, at,
                        "record %llu: malformed aux field, or a long-CIGAR placeholder without CG:B:I",
                        (unsigned long long)bad);
        if (err) std::strcpy(err->format, "bam");
        return fail(k);
    }
    const uint64_t sizes[3] = {tot[SOA_TOT_OPS] * 4, tot[SOA_TOT_SEQ], tot[SOA_TOT_AUX] * 16};
    const int data_cols[3] = {FRHIP_SOA_CIGAR_OPS, FRHIP_SOA_SEQUENCE_PACKED, FRHIP_SOA_AUX_ENTRIES};
    for (int k = 0; k < 3; ++k) {
        o->cols[data_cols[k]] = buf_alloc(ctx, sizes[k], err);
        if (!o->cols[data_cols[k]]) return fail(err ? err->kind : FRHIP_E_DEVICE);
    }
    a.ops = (uint32_t *)o->cols[FRHIP_SOA_CIGAR_OPS]->ptr;
    a.seq = (uint8_t *)o->cols[FRHIP_SOA_SEQUENCE_PACKED]->ptr;
    a.aux = (AuxEntry *)o->cols[FRHIP_SOA_AUX_ENTRIES]->ptr;
    a.cap_ops = tot[SOA_TOT_OPS];
    a.cap_seq = tot[SOA_TOT_SEQ];
    a.cap_aux = tot[SOA_TOT_AUX];
    r = launch_soa_fill(a, st);
    if (r == hipSuccess)
        if (EventRef *ev = event_record(st)) {
            for (frhip_buf *c : o->cols) buf_attach_event(c, ev);
            event_release(ev);
        }
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return fail(hip_err(err, r, "SoA fill"));
    *out = o;
    return FRHIP_OK;
}
