// serve_host.cpp -- batched virtual-offset range queries over a resident BAM.
//
// A batch of queries becomes one set of launches: the members every query
// needs go on one persistent inflate work queue, each query's inflated bytes
// are a separate stretch of one arena, and the boundary scan runs all the
// record chains in one launch (a segment per query, one workgroup each, see
// bam_reconcile_kernel). Columns are shared; a query owns the record range
// [first, first + records). Query semantics are noodles' csi Query
// (crates/fritillaria-csi/src/io/query.rs:66-82): from vstart, while the
// record's virtual offset is below vend.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/frhip.h"
#include "../../include/frhip_index.h"
#include "../../include/frhip_serve.h"
#include "bam_layout.h"
#include "frhip_internal.h"
#include "host_internal.h"
#include "pipeline_internal.h"
#include "region_internal.h"

using namespace frhip;

namespace {

constexpr int N_COLS = FRHIP_BAM_N_COLUMNS;
const int WIDTHS[N_COLS] = {4, 4, 2, 1, 4, 4, 4, 4, 8, 4, 4, 4, 4, 4, 8};
constexpr uint64_t NONE = ~0ull;

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

// Where a query lands in the file's member table.
struct Plan {
    int32_t status = FRHIP_OK;
    uint64_t m0 = 0, m1 = 0;  // members [m0, m1) are inflated
    uint64_t start = 0;       // first record: offset in the query's inflated bytes
    uint64_t end = 0;         // records start below this offset
    uint64_t bytes = 0;       // inflated bytes of [m0, m1)
};

} // namespace

// One batch in flight: its own stream, scratch, columns and host staging, so
// that several batches can be on the device at once.
struct ServeSlot {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    Scratch meta, work, arena, recs, cols[N_COLS];
    PinnedScratch pin;
    hipEvent_t ev0 = nullptr, ev1 = nullptr; // the batch's first and last device work
    // the batch on the device (launched, not yet finished)
    bool busy = false, timed = false;
    std::vector<Plan> plans;
    std::vector<uint64_t> seg_of, base_of;
    uint64_t nq = 0, nb = 0, ns = 0, total = 0;
    uint8_t *hb = nullptr; // read-back staging: member status, segment totals and bases
    uint8_t *du = nullptr;
    uint8_t *dc[N_COLS] = {};
    // region batches: the regions, the chunk queries [q0[r], q0[r + 1]) of
    // region r, an index failure per region, the filter's scratch and rows
    bool regions = false;
    std::vector<RegionDev> reg;
    std::vector<uint64_t> reg_q0;
    std::vector<int32_t> reg_status;
    std::vector<uint32_t> q_region;
    std::vector<frhip_query> reg_queries;
    std::vector<frhip_query_result> reg_qres;
    Scratch rscratch, sel;
    uint64_t *d_sel = nullptr;
    // results of the last finished batch
    frhip_buf col_view[N_COLS] = {};
    frhip_buf arena_view = {};
    frhip_buf sel_view = {};
    uint64_t last_records = 0;
    double t0_ms = 0, t1_ms = 0;
};

struct frhip_server {
    frhip_ctx *ctx = nullptr;
    uint8_t *d_file = nullptr;
    uint64_t file_len = 0;
    uint64_t n = 0;
    std::vector<uint64_t> coff, in_off, uprefix;
    std::vector<uint32_t> in_len, isize, crc;
    std::vector<ServeSlot> slots;
    hipEvent_t ev_ref = nullptr; // the origin of the batches' device intervals

    // Member holding compressed offset c as its start, or NONE.
    uint64_t member_at(uint64_t c) const
    {
        auto it = std::lower_bound(coff.begin(), coff.end(), c);
        if (it == coff.end() || *it != c) return NONE;
        return (uint64_t)(it - coff.begin());
    }

    Plan plan(const frhip_query &q) const
    {
        Plan p;
        const uint64_t c0 = q.vstart >> 16, u0 = q.vstart & 0xFFFF;
        const uint64_t m0 = member_at(c0);
        if (m0 == NONE || u0 > isize[m0]) {
            p.status = FRHIP_E_INVALID_BLOCK;
            return p;
        }
        uint64_t m_end = n, u_end = 0;
        if (q.vend != UINT64_MAX && (q.vend >> 16) < file_len) {
            m_end = member_at(q.vend >> 16);
            u_end = q.vend & 0xFFFF;
            if (m_end == NONE || u_end > isize[m_end]) {
                p.status = FRHIP_E_INVALID_BLOCK;
                return p;
            }
        }
        p.m0 = m0;
        p.m1 = u_end ? m_end + 1 : m_end;
        if (p.m1 < p.m0) p.m1 = p.m0;
        p.bytes = uprefix[p.m1] - uprefix[p.m0];
        p.start = u0;
        const uint64_t e = uprefix[m_end] + u_end;
        p.end = e > uprefix[m0] + u0 ? e - uprefix[m0] : u0;
        if (p.end > p.bytes) p.end = p.bytes;
        if (p.start > p.bytes) p.start = p.bytes;
        return p;
    }
};

namespace {

// The per-member and per-segment arrays of one batch, packed into one upload.
struct MetaLayout {
    uint64_t o_in_off, o_in_len, o_isize, o_crc, o_coff, o_out_off, o_until, o_len, o_seg, o_mb, o_start, o_slen,
        o_reg, o_sreg, bytes;
};

MetaLayout meta_layout(uint64_t nb, uint64_t ns, uint64_t nr = 0)
{
    MetaLayout L{};
    uint64_t p = 0;
    auto take = [&](uint64_t b) { uint64_t o = p; p = align_up(p + b, 256); return o; };
    L.o_in_off = take(nb * 8);
    L.o_in_len = take(nb * 4);
    L.o_isize = take(nb * 4);
    L.o_crc = take(nb * 4);
    L.o_coff = take(nb * 8);
    L.o_out_off = take((nb + 1) * 8);
    L.o_until = take(nb * 8);
    L.o_len = take(nb * 8);
    L.o_seg = take(nb * 4);
    L.o_mb = take((ns + 1) * 4);
    L.o_start = take(ns * 8);
    L.o_slen = take(ns * 8);
    L.o_reg = take(nr * sizeof(RegionDev));
    L.o_sreg = take(nr ? ns * 4 : 0);
    L.bytes = p;
    return L;
}

// What one batch reads back: member status words, segment totals and bases,
// and for a region batch each segment's first selected row and first row
// failing the region check.
struct BackLayout {
    uint64_t o_tot, o_base, o_first, o_bad, bytes;
};

BackLayout back_layout(uint64_t nb, uint64_t ns, bool regions)
{
    BackLayout B{};
    B.o_tot = align_up(nb * 12, 256);
    B.o_base = B.o_tot + ns * 32;
    B.o_first = B.o_base + (ns + 1) * 8;
    B.o_bad = B.o_first + (regions ? (ns + 1) * 8 : 0);
    B.bytes = B.o_bad + (regions ? ns * 8 : 0);
    return B;
}

// Device scratch: inflate status, the work queue, the boundary scan, segment totals.
struct WorkLayout {
    BamScratchLayout bam;
    uint64_t o_st, o_queue, o_bam, o_tot, o_base, bytes;
};

WorkLayout work_layout(uint64_t nb, uint64_t ns)
{
    WorkLayout W{};
    W.bam = bam_layout(nb);
    uint64_t p = 0;
    auto take = [&](uint64_t b) { uint64_t o = p; p = align_up(p + b, 256); return o; };
    W.o_st = take(nb * 12);
    W.o_queue = take(8);
    W.o_bam = take(W.bam.bytes);
    W.o_tot = take(ns * 32);
    W.o_base = take((ns + 1) * 8);
    W.bytes = p;
    return W;
}

uint64_t records_cap(uint64_t bytes) { return bytes / 36 + 1; }

} // namespace

extern "C" int frhip_server_create(frhip_ctx *ctx, const uint8_t *file, uint64_t len, frhip_server **out,
                                   frhip_error *err)
{
    return frhip_server_create_slots(ctx, file, len, FRHIP_SERVER_DEFAULT_SLOTS, out, err);
}

extern "C" int frhip_server_create_slots(frhip_ctx *ctx, const uint8_t *file, uint64_t len, int slots,
                                         frhip_server **out, frhip_error *err)
{
    *out = nullptr;
    if (slots < 1 || slots > FRHIP_SERVER_MAX_SLOTS)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "a server has 1 to %d slots", FRHIP_SERVER_MAX_SLOTS);
    std::vector<frhip_block_span> spans;
    uint64_t position = 0;
    int rc = discover(file, len, 0, spans, &position, err);
    if (rc != FRHIP_OK) return rc;
    frhip_server *s = new (std::nothrow) frhip_server;
    if (!s) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    s->ctx = ctx;
    ctx_retain(ctx);
    s->file_len = len;
    s->n = spans.size();
    s->coff.resize(s->n);
    s->in_off.resize(s->n);
    s->in_len.resize(s->n);
    s->isize.resize(s->n);
    s->crc.resize(s->n);
    s->uprefix.assign(s->n + 1, 0);
    for (uint64_t i = 0; i < s->n; ++i) {
        s->coff[i] = spans[i].compressed_offset;
        s->in_off[i] = spans[i].payload_start;
        s->in_len[i] = (uint32_t)spans[i].payload_len;
        s->isize[i] = spans[i].isize;
        s->crc[i] = spans[i].crc32;
        s->uprefix[i + 1] = s->uprefix[i] + spans[i].isize;
    }
    (void)hipSetDevice(ctx->ordinal);
    // the bit reader loads whole 512-byte chunks: 4 KiB of zeros past the end
    hipError_t r = hipMalloc(&s->d_file, len + 4096);
    if (r == hipSuccess) r = hipMemcpy(s->d_file, file, len, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemset(s->d_file + len, 0, 4096);
    // slot 0 runs on the context's stream, the others on streams of their own
    // (default priority)
    s->slots.resize(slots);
    for (int k = 0; r == hipSuccess && k < slots; ++k) {
        ServeSlot &sl = s->slots[k];
        if (k == 0) {
            sl.stream = ctx->stream;
        } else {
            r = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
            sl.own_stream = r == hipSuccess;
        }
        if (r == hipSuccess) r = hipEventCreate(&sl.ev0);
        if (r == hipSuccess) r = hipEventCreate(&sl.ev1);
        for (int c = 0; c < N_COLS; ++c) sl.col_view[c] = frhip_buf{nullptr, 0, ctx->ordinal};
        sl.arena_view = frhip_buf{nullptr, 0, ctx->ordinal};
        sl.sel_view = frhip_buf{nullptr, 0, ctx->ordinal};
    }
    if (r == hipSuccess) r = hipEventCreate(&s->ev_ref);
    if (r == hipSuccess) r = hipEventRecord(s->ev_ref, ctx->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(ctx->stream);
    if (r != hipSuccess) {
        frhip_server_free(s);
        return hip_err(err, r, "uploading the file");
    }
    *out = s;
    return FRHIP_OK;
}

extern "C" void frhip_server_free(frhip_server *s)
{
    if (!s) return;
    if (s->ctx) {
        (void)hipSetDevice(s->ctx->ordinal);
        (void)hipStreamSynchronize(s->ctx->stream);
    }
    for (ServeSlot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        sl.meta.release();
        sl.work.release();
        sl.arena.release();
        sl.recs.release();
        for (auto &c : sl.cols) c.release();
        sl.rscratch.release();
        sl.sel.release();
        sl.pin.release();
        if (sl.ev0) (void)hipEventDestroy(sl.ev0);
        if (sl.ev1) (void)hipEventDestroy(sl.ev1);
        if (sl.own_stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->ev_ref) (void)hipEventDestroy(s->ev_ref);
    if (s->d_file) (void)hipFree(s->d_file);
    if (s->ctx) ctx_release(s->ctx);
    delete s;
}

extern "C" uint64_t frhip_server_members(const frhip_server *s) { return s->n; }

extern "C" int frhip_server_member_table(const frhip_server *s, uint64_t *coffsets, uint64_t *uprefix,
                                         frhip_error *err)
{
    (void)err;
    if (coffsets) std::memcpy(coffsets, s->coff.data(), s->n * 8);
    if (uprefix) std::memcpy(uprefix, s->uprefix.data(), (s->n + 1) * 8);
    return FRHIP_OK;
}

extern "C" uint64_t frhip_server_batch_footprint(const frhip_server *s, const frhip_query *q, uint64_t n)
{
    uint64_t bytes = 0, members = 0;
    for (uint64_t i = 0; i < n; ++i) {
        Plan p = s->plan(q[i]);
        if (p.status != FRHIP_OK) continue;
        bytes += p.bytes;
        members += p.m1 - p.m0;
    }
    uint64_t per_rec = 0;
    for (int c = 0; c < N_COLS; ++c) per_rec += WIDTHS[c];
    return bytes + records_cap(bytes) * per_rec + meta_layout(members, n).bytes + work_layout(members, n).bytes +
           inflate_scratch_bytes(s->ctx->inflate_schedule, bytes, members);
}

extern "C" int frhip_server_slots(const frhip_server *s) { return (int)s->slots.size(); }

namespace {

// Plans and enqueues one batch of queries in a slot. A region batch (the
// slot's reg / q_region filled by the caller) also runs the region filter
// after the decode and reads back its per-segment results.
int launch_batch(frhip_server *s, int slot, const frhip_query *q, uint64_t nq, bool regions, frhip_error *err)
{
    if (slot < 0 || slot >= (int)s->slots.size())
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "no slot %d", slot);
    ServeSlot &sl = s->slots[slot];
    if (sl.busy) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "slot %d has a batch in flight", slot);
    frhip_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->ordinal);
    hipStream_t st = sl.stream;
    // plan every query; queries with members become segments
    std::vector<Plan> &plans = sl.plans;
    std::vector<uint64_t> &seg_of = sl.seg_of, &base_of = sl.base_of;
    plans.assign(nq, Plan{});
    seg_of.assign(nq, NONE);
    base_of.assign(nq, 0);
    uint64_t nb = 0, ns = 0, total = 0;
    for (uint64_t i = 0; i < nq; ++i) {
        plans[i] = s->plan(q[i]);
        if (plans[i].status != FRHIP_OK || plans[i].m1 == plans[i].m0) continue;
        seg_of[i] = ns++;
        base_of[i] = total;
        nb += plans[i].m1 - plans[i].m0;
        total += plans[i].bytes;
    }
    sl.nq = nq;
    sl.nb = nb;
    sl.ns = ns;
    sl.total = total;
    sl.last_records = 0;
    for (int c = 0; c < N_COLS; ++c) sl.col_view[c] = frhip_buf{nullptr, 0, ctx->ordinal};
    sl.arena_view = frhip_buf{nullptr, 0, ctx->ordinal};
    sl.sel_view = frhip_buf{nullptr, 0, ctx->ordinal};
    sl.regions = regions;
    sl.timed = false;
    sl.busy = true;
    sl.t0_ms = sl.t1_ms = 0;
    if (ns == 0) return FRHIP_OK;

    // host side of the member and segment arrays
    const uint64_t nr = regions ? sl.reg.size() : 0;
    const MetaLayout M = meta_layout(nb, ns, nr);
    const WorkLayout W = work_layout(nb, ns);
    const BackLayout B = back_layout(nb, ns, regions);
    const uint64_t back = B.bytes;
    int rc;
    // a failure before the batch is on the stream leaves the slot free
    auto fail = [&](int k) {
        sl.busy = false;
        return k;
    };
    uint8_t *h = (uint8_t *)sl.pin.get(std::max(M.bytes, back), err, &rc);
    if (!h) return fail(rc);
    {
        uint64_t *in_off = (uint64_t *)(h + M.o_in_off), *coff = (uint64_t *)(h + M.o_coff);
        uint64_t *out_off = (uint64_t *)(h + M.o_out_off), *until = (uint64_t *)(h + M.o_until);
        uint64_t *blen = (uint64_t *)(h + M.o_len), *sstart = (uint64_t *)(h + M.o_start);
        uint64_t *slen = (uint64_t *)(h + M.o_slen);
        uint32_t *in_len = (uint32_t *)(h + M.o_in_len), *isz = (uint32_t *)(h + M.o_isize);
        uint32_t *crc = (uint32_t *)(h + M.o_crc), *seg = (uint32_t *)(h + M.o_seg), *mb = (uint32_t *)(h + M.o_mb);
        uint32_t *sreg = (uint32_t *)(h + M.o_sreg);
        if (nr) std::memcpy(h + M.o_reg, sl.reg.data(), nr * sizeof(RegionDev));
        uint64_t j = 0;
        for (uint64_t i = 0; i < nq; ++i) {
            if (seg_of[i] == NONE) continue;
            const Plan &p = plans[i];
            const uint64_t sq = seg_of[i], base = base_of[i];
            mb[sq] = (uint32_t)j;
            if (nr) sreg[sq] = sl.q_region[i];
            sstart[sq] = base + p.start;
            slen[sq] = base + p.bytes;
            for (uint64_t m = p.m0; m < p.m1; ++m, ++j) {
                in_off[j] = s->in_off[m];
                in_len[j] = s->in_len[m];
                isz[j] = s->isize[m];
                crc[j] = s->crc[m];
                coff[j] = s->coff[m];
                out_off[j] = base + (s->uprefix[m] - s->uprefix[p.m0]);
                until[j] = std::min(base + (s->uprefix[m + 1] - s->uprefix[p.m0]), base + p.end);
                blen[j] = base + p.bytes;
                seg[j] = (uint32_t)sq;
            }
        }
        mb[ns] = (uint32_t)j;
        out_off[nb] = total;
    }
    uint8_t *dm = (uint8_t *)sl.meta.get(M.bytes, err, &rc);
    if (!dm) return fail(rc);
    uint8_t *dw = (uint8_t *)sl.work.get(W.bytes, err, &rc);
    if (!dw) return fail(rc);
    uint8_t *du = (uint8_t *)sl.arena.get(total + 16, err, &rc);
    if (!du) return fail(rc);
    const uint64_t cap = records_cap(total);
    uint8_t **dc = sl.dc;
    for (int c = 0; c < N_COLS; ++c) {
        dc[c] = (uint8_t *)sl.cols[c].get(cap * WIDTHS[c], err, &rc);
        if (!dc[c]) return fail(rc);
    }
    sl.du = du;
    RegionFilterArgs fa{};
    RegionScratchLayout RL{};
    if (regions) {
        RL = region_scratch_layout(cap, ns);
        uint8_t *rs = (uint8_t *)sl.rscratch.get(RL.bytes, err, &rc);
        if (!rs) return fail(rc);
        sl.d_sel = (uint64_t *)sl.sel.get(cap * 8, err, &rc);
        if (!sl.d_sel) return fail(rc);
        region_bind_scratch(fa, rs, RL);
        fa.sel = sl.d_sel;
    }
    hipError_t r = hipEventRecord(sl.ev0, st);
    if (r == hipSuccess) r = hipMemcpyAsync(dm, h, M.bytes, hipMemcpyHostToDevice, st);
    uint32_t *d_st = (uint32_t *)(dw + W.o_st);
    if (r == hipSuccess) r = launch_fill_u64((uint64_t *)(dw + W.o_queue), 0, 1, st);

    // inflate every member of the batch on one work queue
    InflateArgs a{};
    a.in = s->d_file;
    a.in_off = (const uint64_t *)(dm + M.o_in_off);
    a.in_len = (const uint32_t *)(dm + M.o_in_len);
    a.out = du;
    a.out_off = (const uint64_t *)(dm + M.o_out_off);
    a.out_cap = (const uint32_t *)(dm + M.o_isize);
    a.status = d_st;
    a.produced = d_st + nb;
    a.crc = d_st + 2 * nb;
    a.n = (uint32_t)nb;
    a.queue = (uint32_t *)(dw + W.o_queue);
    a.order = nullptr;
    a.x2n = ctx->x2n;
    void *d_rec = nullptr;
    if (const uint64_t sb = inflate_scratch_bytes(ctx->inflate_schedule, total, nb)) {
        d_rec = sl.recs.get(sb, err, &rc);
        if (!d_rec) return fail(rc);
    }
    inflate_scratch_bind(a, d_rec, nb);
    a.two_phase_min = ctx->two_phase_min;
    if (r == hipSuccess) {
        const int sched = ctx->inflate_schedule;
        uint64_t waves = (uint64_t)ctx->n_cu * inflate_waves_per_cu(sched);
        if (waves > nb) waves = nb;
        r = launch_inflate(a, (uint32_t)waves, st, sched);
    }

    // boundaries: one chain per query
    uint8_t *b = dw + W.o_bam;
    const BamScratchLayout &L = W.bam;
    BamSpecArgs sa{};
    sa.buf = du;
    sa.len = total;
    sa.block_starts = a.out_off;
    sa.nb = (uint32_t)nb;
    sa.window = ctx->bam_window;
    sa.spec_ok = b + L.o_ok;
    sa.n_surv = b + L.o_ns;
    sa.surv_pos = (uint64_t *)(b + L.o_spos);
    sa.surv_cnt = (uint32_t *)(b + L.o_scnt);
    sa.land = (uint64_t *)(b + L.o_land);
    sa.exitp = (uint64_t *)(b + L.o_exit);
    sa.cnt_tail = (uint32_t *)(b + L.o_ctail);
    sa.complete = b + L.o_comp;
    sa.rel_offs = (uint16_t *)(b + L.o_rel);
    sa.mem_until = (const uint64_t *)(dm + M.o_until);
    sa.mem_len = (const uint64_t *)(dm + M.o_len);
    if (r == hipSuccess) r = launch_bam_spec(sa, st);
    BamReconcileArgs ra{};
    ra.buf = du;
    ra.len = total;
    ra.block_starts = a.out_off;
    ra.nb = (uint32_t)nb;
    ra.spec_ok = sa.spec_ok;
    ra.n_surv = sa.n_surv;
    ra.surv_pos = sa.surv_pos;
    ra.surv_cnt = sa.surv_cnt;
    ra.exitp = sa.exitp;
    ra.cnt_tail = sa.cnt_tail;
    ra.complete = sa.complete;
    ra.entry = (uint64_t *)(b + L.o_entry);
    ra.count = (uint32_t *)(b + L.o_count);
    ra.first_index = (uint64_t *)(b + L.o_first);
    ra.mode = b + L.o_mode;
    ra.seg_mb = (const uint32_t *)(dm + M.o_mb);
    ra.seg_start = (const uint64_t *)(dm + M.o_start);
    ra.seg_len = (const uint64_t *)(dm + M.o_slen);
    ra.seg_totals = (uint64_t *)(dw + W.o_tot);
    ra.mem_until = sa.mem_until;
    if (r == hipSuccess) r = launch_bam_reconcile_segments(ra, (uint32_t)ns, st);
    uint64_t *seg_base = (uint64_t *)(dw + W.o_base);
    if (r == hipSuccess) r = launch_bam_segment_bases(ra.seg_totals, (uint32_t)ns, seg_base, st);
    BamEmitArgs ea{};
    ea.buf = du;
    ea.block_starts = a.out_off;
    ea.coffsets = (const uint64_t *)(dm + M.o_coff);
    ea.nb = (uint32_t)nb;
    ea.entry = ra.entry;
    ea.count = ra.count;
    ea.first_index = ra.first_index;
    ea.mode = ra.mode;
    ea.surv_cnt = sa.surv_cnt;
    ea.rel_offs = sa.rel_offs;
    ea.record_offsets = (uint64_t *)dc[FRHIP_BAM_RECORD_OFFSETS];
    ea.virtual_offsets = (uint64_t *)dc[FRHIP_BAM_VIRTUAL_OFFSETS];
    ea.mem_seg = (const uint32_t *)(dm + M.o_seg);
    ea.seg_base = seg_base;
    if (r == hipSuccess) r = launch_bam_emit(ea, st);
    BamDecodeArgs da{};
    da.buf = du;
    da.len = total;
    da.n = 0;
    da.n_dev = seg_base + ns;
    da.c.ref_id = (int32_t *)dc[0];
    da.c.pos = (int32_t *)dc[1];
    da.c.flags = (uint16_t *)dc[2];
    da.c.mapq = (uint8_t *)dc[3];
    da.c.l_seq = (uint32_t *)dc[4];
    da.c.next_ref_id = (int32_t *)dc[5];
    da.c.next_pos = (int32_t *)dc[6];
    da.c.tlen = (int32_t *)dc[7];
    da.c.record_offsets = (uint64_t *)dc[8];
    da.c.cigar_start = (uint32_t *)dc[9];
    da.c.seq_start = (uint32_t *)dc[10];
    da.c.qual_start = (uint32_t *)dc[11];
    da.c.aux_start = (uint32_t *)dc[12];
    da.c.record_end = (uint32_t *)dc[13];
    if (r == hipSuccess) r = launch_bam_decode(da, cap, st);
    if (regions) {
        fa.buf = du;
        fa.record_offsets = da.c.record_offsets;
        fa.ref_id = da.c.ref_id;
        fa.pos = da.c.pos;
        fa.cigar_start = da.c.cigar_start;
        fa.seq_start = da.c.seq_start;
        fa.seg_base = seg_base;
        fa.ns = (uint32_t)ns;
        fa.seg_region = (const uint32_t *)(dm + M.o_sreg);
        fa.regions = (const RegionDev *)(dm + M.o_reg);
        if (r == hipSuccess) r = launch_fill_u64((uint64_t *)fa.seg_bad, ~0ull, ns, st);
        if (r == hipSuccess) r = launch_region_filter(fa, cap, st);
    }

    // one read-back: member status, segment totals and bases (and the
    // region filter's per-segment results)
    uint8_t *hb = h; // the upload has been consumed once the stream reaches this copy
    if (r == hipSuccess) r = hipMemcpyAsync(hb, d_st, nb * 12, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipMemcpyAsync(hb + B.o_tot, ra.seg_totals, ns * 32, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipMemcpyAsync(hb + B.o_base, seg_base, (ns + 1) * 8, hipMemcpyDeviceToHost, st);
    if (regions) {
        if (r == hipSuccess) r = hipMemcpyAsync(hb + B.o_first, fa.sel_first, (ns + 1) * 8, hipMemcpyDeviceToHost, st);
        if (r == hipSuccess) r = hipMemcpyAsync(hb + B.o_bad, fa.seg_bad, ns * 8, hipMemcpyDeviceToHost, st);
    }
    if (r == hipSuccess) r = hipEventRecord(sl.ev1, st);
    if (r != hipSuccess) {
        // whatever was enqueued finishes before the slot is used again
        (void)hipStreamSynchronize(st);
        return fail(hip_err(err, r, "serving batch"));
    }
    sl.hb = hb;
    sl.timed = true;
    return FRHIP_OK;
}

} // namespace

extern "C" int frhip_server_launch(frhip_server *s, int slot, const frhip_query *q, uint64_t nq, frhip_error *err)
{
    return launch_batch(s, slot, q, nq, false, err);
}

extern "C" int frhip_server_query(const frhip_server *s, int slot)
{
    if (slot < 0 || slot >= (int)s->slots.size()) return 1;
    const ServeSlot &sl = s->slots[slot];
    if (!sl.busy || !sl.timed) return 1;
    return hipEventQuery(sl.ev1) == hipSuccess ? 1 : 0;
}

extern "C" int frhip_server_finish(frhip_server *s, int slot, frhip_query_result *res, frhip_error *err)
{
    if (slot < 0 || slot >= (int)s->slots.size())
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "no slot %d", slot);
    ServeSlot &sl = s->slots[slot];
    if (!sl.busy) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "slot %d has no batch", slot);
    frhip_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->ordinal);
    const uint64_t nq = sl.nq, nb = sl.nb, ns = sl.ns, total = sl.total;
    const std::vector<Plan> &plans = sl.plans;
    const std::vector<uint64_t> &seg_of = sl.seg_of, &base_of = sl.base_of;
    for (uint64_t i = 0; i < nq; ++i) {
        std::memset(&res[i], 0, sizeof res[i]);
        res[i].status = plans[i].status;
        if (seg_of[i] == NONE) continue;
        res[i].members = plans[i].m1 - plans[i].m0;
        res[i].inflated_bytes = plans[i].bytes;
    }
    sl.busy = false;
    if (ns == 0) return FRHIP_OK;
    hipError_t r = hipEventSynchronize(sl.ev1);
    if (r != hipSuccess) return hip_err(err, r, "serving batch");
    float a0 = 0, a1 = 0;
    (void)hipEventElapsedTime(&a0, s->ev_ref, sl.ev0);
    (void)hipEventElapsedTime(&a1, s->ev_ref, sl.ev1);
    sl.t0_ms = a0;
    sl.t1_ms = a1;
    const uint8_t *hb = sl.hb;
    const BackLayout B = back_layout(nb, ns, sl.regions);
    const uint32_t *status = (const uint32_t *)hb, *produced = status + nb, *crcs = status + 2 * nb;
    const uint64_t *tot = (const uint64_t *)(hb + B.o_tot), *base = (const uint64_t *)(hb + B.o_base);

    uint64_t j = 0;
    for (uint64_t i = 0; i < nq; ++i) {
        if (seg_of[i] == NONE) continue;
        const Plan &p = plans[i];
        const uint64_t sq = seg_of[i];
        // member verification in reference order: status, size, CRC
        for (uint64_t m = p.m0; m < p.m1; ++m, ++j) {
            if (res[i].status != FRHIP_OK) continue;
            if (status[j] != 0) res[i].status = FRHIP_E_INFLATE;
            else if (produced[j] != s->isize[m]) res[i].status = FRHIP_E_SIZE_MISMATCH;
            else if (crcs[j] != s->crc[m]) res[i].status = FRHIP_E_CHECKSUM_MISMATCH;
        }
        if (res[i].status != FRHIP_OK) continue;
        res[i].first = base[sq];
        res[i].records = tot[4 * sq + 0];
        const uint64_t tail = tot[4 * sq + 1] - base_of[i];
        if (tot[4 * sq + 2] != 0) {
            res[i].status = FRHIP_E_MALFORMED;
            res[i].error_position = tail;
        } else if (tail < p.end) {
            // the inflated members end inside a record that starts before vend
            res[i].status = FRHIP_E_MALFORMED;
            res[i].error_position = tail;
        }
    }
    sl.last_records = base[ns];
    for (int c = 0; c < N_COLS; ++c) sl.col_view[c] = frhip_buf{sl.dc[c], sl.last_records * WIDTHS[c], ctx->ordinal};
    sl.arena_view = frhip_buf{sl.du, total, ctx->ordinal};
    return FRHIP_OK;
}

extern "C" int frhip_server_run(frhip_server *s, const frhip_query *q, uint64_t nq, frhip_query_result *res,
                                frhip_error *err)
{
    const int rc = frhip_server_launch(s, 0, q, nq, err);
    if (rc != FRHIP_OK) return rc;
    return frhip_server_finish(s, 0, res, err);
}

extern "C" int frhip_server_batch_interval(const frhip_server *s, int slot, double *start_ms, double *end_ms)
{
    if (slot < 0 || slot >= (int)s->slots.size()) return FRHIP_E_INVALID_DEVICE_BATCH;
    const ServeSlot &sl = s->slots[slot];
    if (start_ms) *start_ms = sl.t0_ms;
    if (end_ms) *end_ms = sl.t1_ms;
    return FRHIP_OK;
}

extern "C" const frhip_buf *frhip_server_slot_column(const frhip_server *s, int slot, int column)
{
    if (!s || slot < 0 || slot >= (int)s->slots.size() || column < 0 || column >= N_COLS) return nullptr;
    return &s->slots[slot].col_view[column];
}

extern "C" const frhip_buf *frhip_server_slot_inflated(const frhip_server *s, int slot)
{
    if (!s || slot < 0 || slot >= (int)s->slots.size()) return nullptr;
    return &s->slots[slot].arena_view;
}

extern "C" const frhip_buf *frhip_server_column(const frhip_server *s, int column)
{
    return frhip_server_slot_column(s, 0, column);
}

extern "C" const frhip_buf *frhip_server_inflated(const frhip_server *s) { return frhip_server_slot_inflated(s, 0); }

extern "C" int frhip_server_launch_regions(frhip_server *s, int slot, const frhip_index *ix, const frhip_region *r,
                                           uint64_t n, frhip_error *err)
{
    if (slot < 0 || slot >= (int)s->slots.size())
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "no slot %d", slot);
    ServeSlot &sl = s->slots[slot];
    if (sl.busy) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "slot %d has a batch in flight", slot);
    if (!ix) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "no index");
    sl.reg.assign(n, RegionDev{});
    sl.reg_q0.assign(n + 1, 0);
    sl.reg_status.assign(n, FRHIP_OK);
    sl.q_region.clear();
    sl.reg_queries.clear();
    std::vector<frhip_chunk> chunks;
    for (uint64_t k = 0; k < n; ++k) {
        sl.reg_q0[k] = sl.reg_queries.size();
        RegionDev &d = sl.reg[k];
        d.ref_id = r[k].ref_id;
        d.bounded = (r[k].start || r[k].end) ? 1u : 0u;
        d.start = r[k].start ? r[k].start : 1;
        d.end = r[k].end ? r[k].end : (uint64_t)INT64_MAX;
        if (r[k].ref_id < 0) {
            sl.reg_status[k] = FRHIP_E_IO;
            continue;
        }
        uint64_t nc = 0;
        frhip_error e2{};
        int rc = frhip_index_query(ix, (uint32_t)r[k].ref_id, r[k].start, r[k].end, nullptr, 0, &nc, &e2);
        if (rc == FRHIP_OK && nc) {
            chunks.resize(nc);
            rc = frhip_index_query(ix, (uint32_t)r[k].ref_id, r[k].start, r[k].end, chunks.data(), nc, &nc, &e2);
        }
        if (rc != FRHIP_OK) {
            sl.reg_status[k] = rc;
            continue;
        }
        for (uint64_t c = 0; c < nc; ++c) {
            sl.reg_queries.push_back(frhip_query{chunks[c].start, chunks[c].end});
            sl.q_region.push_back((uint32_t)k);
        }
    }
    sl.reg_q0[n] = sl.reg_queries.size();
    return launch_batch(s, slot, sl.reg_queries.data(), sl.reg_queries.size(), true, err);
}

extern "C" int frhip_server_finish_regions(frhip_server *s, int slot, frhip_region_result *res, frhip_error *err)
{
    if (slot < 0 || slot >= (int)s->slots.size())
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "no slot %d", slot);
    ServeSlot &sl = s->slots[slot];
    if (!sl.busy) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "slot %d has no batch", slot);
    if (!sl.regions) return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "slot %d holds no region batch", slot);
    const uint64_t nq = sl.reg_queries.size(), nr = sl.reg.size();
    sl.reg_qres.assign(nq, frhip_query_result{});
    const int rc = frhip_server_finish(s, slot, sl.reg_qres.data(), err);
    if (rc != FRHIP_OK) return rc;
    const uint64_t ns = sl.ns;
    const BackLayout B = back_layout(sl.nb, ns, true);
    const uint64_t *first = (const uint64_t *)(sl.hb + B.o_first), *bad = (const uint64_t *)(sl.hb + B.o_bad);
    const uint64_t *base = (const uint64_t *)(sl.hb + B.o_base);
    for (uint64_t k = 0; k < nr; ++k) {
        frhip_region_result &o = res[k];
        std::memset(&o, 0, sizeof o);
        o.status = sl.reg_status[k];
        o.chunks = sl.reg_q0[k + 1] - sl.reg_q0[k];
        if (o.status != FRHIP_OK) continue;
        uint64_t s_lo = NONE, s_hi = 0;
        for (uint64_t qi = sl.reg_q0[k]; qi < sl.reg_q0[k + 1]; ++qi) {
            const frhip_query_result &qr = sl.reg_qres[qi];
            const uint64_t before = o.scanned;
            o.members += qr.members;
            o.inflated_bytes += qr.inflated_bytes;
            o.scanned += qr.records;
            if (o.status == FRHIP_OK && qr.status != FRHIP_OK) {
                o.status = qr.status;
                o.error_position = qr.error_position;
            }
            const uint64_t sq = sl.seg_of[qi];
            if (sq == NONE) continue;
            if (s_lo == NONE) s_lo = sq;
            s_hi = sq + 1;
            if (o.status == FRHIP_OK && bad[sq] != ~0ull) {
                o.status = FRHIP_E_MALFORMED;
                o.error_position = before + (bad[sq] - base[sq]);
            }
        }
        if (o.status != FRHIP_OK || s_lo == NONE) continue;
        o.first = first[s_lo];
        o.records = first[s_hi] - first[s_lo];
    }
    sl.sel_view = frhip_buf{sl.d_sel, ns ? first[ns] * 8 : 0, s->ctx->ordinal};
    return FRHIP_OK;
}

extern "C" const frhip_buf *frhip_server_slot_selection(const frhip_server *s, int slot)
{
    if (!s || slot < 0 || slot >= (int)s->slots.size()) return nullptr;
    return &s->slots[slot].sel_view;
}

extern "C" int frhip_device_memory(frhip_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes, frhip_error *err)
{
    (void)hipSetDevice(ctx->ordinal);
    size_t f = 0, t = 0;
    hipError_t r = hipMemGetInfo(&f, &t);
    if (r != hipSuccess) return hip_err(err, r, "hipMemGetInfo");
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return FRHIP_OK;
}
