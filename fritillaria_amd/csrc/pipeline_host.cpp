// pipeline_host.cpp -- the batched, device-resident BAM path (frhip_pipeline.h).
//
// One run enqueues, on the context's stream: member discovery (sieve, scan,
// ordered compaction, chain proof, spans, ISIZE prefix sum), the persistent
// work-queue inflate, batch verification, the BAM header walk, the boundary
// speculation / reconcile / emit, and the column decode. Nothing is read back
// until the end, so the whole run can be captured once as a hipGraph and
// replayed. Buffers are sized at create from one discovery pass: every run over
// the same resident file reproduces the same member table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/frhip_pipeline.h"
#include "bam_layout.h"
#include "frhip_internal.h"
#include "host_internal.h"
#include "pipeline_internal.h"
#include "soa_internal.h"

using namespace frhip;

namespace {

const int WIDTHS[FRHIP_BAM_N_COLUMNS] = {4, 4, 2, 1, 4, 4, 4, 4, 8, 4, 4, 4, 4, 4, 8};
const int SOA_WIDTHS[FRHIP_SOA_N_COLUMNS] = {8, 4, 8, 1, 8, 16};

// Small block read back once per run.
struct ResultBlock {
    uint64_t chain_bad;      // member-chain proof failures (device discovery)
    uint64_t first_bad;      // first member failing verification, or ~0
    uint64_t hdr_status;     // BAM header walk
    uint64_t records_start;
    uint64_t totals[4];      // records, tail, status, fallbacks
    uint64_t soa[SOA_N_RESULT + 1]; // SoA totals, overflow, first malformed record
};

enum { EV_START, EV_DISC, EV_INFL, EV_VERIFY, EV_SPEC, EV_REC, EV_COLS, EV_END, N_EV };
constexpr int N_STAGES = N_EV; // the stages between events, then the total

} // namespace

struct frhip_pipeline {
    frhip_ctx *ctx = nullptr;
    frhip_pipeline_config cfg{};
    uint64_t file_len = 0, n = 0, U = 0, cap_rec = 0, n_tiles = 0, cand_cap = 0;
    bool device_discovery = false;
    bool sized = false;          // U and the buffers sized from it are final
    uint8_t *d_file = nullptr;
    // re-uploads of the file (frhip_pipeline_upload): a copy stream, the
    // copy's event the next launch waits for, the last launch's end event
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_uploaded = nullptr, ev_done = nullptr;
    bool upload_pending = false, launched = false;
    // discovery
    uint32_t *d_tile_counts = nullptr;
    uint64_t *d_tile_base = nullptr;
    uint64_t *d_cand = nullptr;
    uint32_t *d_cand_bs = nullptr, *d_cand_hs = nullptr;
    PipeMembers m{};
    uint64_t *d_n = nullptr;     // member count (device)
    uint64_t *d_ulen = nullptr;  // inflated length (device)
    // inflate
    uint8_t *d_u = nullptr;
    uint32_t *d_st = nullptr;    // status, produced, crc: 3n
    uint64_t *d_queue = nullptr; // low 32 bits are the work-queue counter
    uint32_t *d_order = nullptr;
    int sched = 0;               // inflate schedule, fixed at creation
    void *d_rec = nullptr;       // two-phase schedules: match records
    // bam
    uint8_t *d_bam = nullptr;
    BamScratchLayout L{};
    void *d_cols[FRHIP_BAM_N_COLUMNS] = {};
    ResultBlock *d_res = nullptr;
    ResultBlock res{};
    ResultBlock *h_res = nullptr; // pinned
    // SoA extras (cfg.soa): sized from the records of a run at create
    bool soa = false;
    uint8_t *d_soa_scratch = nullptr;
    void *d_soa_cols[FRHIP_SOA_N_COLUMNS] = {};
    uint64_t soa_cap_rec = 0, soa_cap[3] = {};
    // views
    frhip_buf u_view{};
    frhip_buf col_view[FRHIP_BAM_N_COLUMNS] = {};
    frhip_buf soa_view[FRHIP_SOA_N_COLUMNS] = {};
    // graph and timing
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[N_EV] = {};
    bool timed = false;
    bool last_ok = false;
    double stage_ms[N_STAGES] = {};
    std::vector<void *> allocs;
};

namespace {

int dalloc(frhip_pipeline *p, void **out, uint64_t bytes, frhip_error *err, const char *what)
{
    hipError_t r = hipMalloc(out, bytes ? bytes : 1);
    if (r != hipSuccess) {
        *out = nullptr;
        return hip_err(err, r, what);
    }
    p->allocs.push_back(*out);
    return FRHIP_OK;
}

// Enqueues discovery: sieve count, tile scan, ordered write, chain proof and
// spans, ISIZE scan. The member arrays are written for `cap` candidates at most.
hipError_t enqueue_discovery(frhip_pipeline *p, hipStream_t s)
{
    hipError_t r = launch_fill_u64((uint64_t *)&p->d_res->chain_bad, 0, 1, s);
    if (r == hipSuccess) r = launch_sieve_count(p->d_file, p->file_len, p->d_tile_counts, s);
    if (r == hipSuccess) r = launch_scan_u32_u64(p->d_tile_counts, p->d_tile_base, nullptr, p->n_tiles, s);
    if (r == hipSuccess)
        r = launch_sieve_write(p->d_file, p->file_len, p->d_tile_base, p->d_cand, p->d_cand_bs, p->d_cand_hs,
                               p->cand_cap, s);
    if (r == hipSuccess)
        r = launch_spans(p->d_file, p->file_len, p->d_cand, p->d_cand_bs, p->d_cand_hs, p->d_tile_base + p->n_tiles,
                         p->cand_cap, p->m, (uint64_t *)&p->d_res->chain_bad, s);
    if (r == hipSuccess)
        r = launch_scan_u32_u64(p->m.isize, p->m.out_off, p->d_tile_base + p->n_tiles, 0, s);
    if (r == hipSuccess && p->sized)
        r = launch_guard_total(p->m.out_off, p->m.isize, p->n, p->U, (uint64_t *)&p->d_res->chain_bad, s);
    return r;
}

SoaArgs soa_args(frhip_pipeline *p)
{
    SoaArgs a{};
    soa_bind_scratch(a, p->d_soa_scratch, soa_scratch_layout(p->soa_cap_rec));
    a.result = p->d_res->soa;
    a.first_bad = (unsigned long long *)&p->d_res->soa[SOA_N_RESULT];
    a.buf = p->d_u;
    a.n = p->soa_cap_rec;
    a.n_dev = &p->d_res->totals[0];
    a.cap_rec = p->soa_cap_rec;
    soa_bind_columns(a, (const uint64_t *)p->d_cols[FRHIP_BAM_RECORD_OFFSETS],
                     (const uint32_t *)p->d_cols[FRHIP_BAM_SEQUENCE_LEN],
                     (const uint32_t *)p->d_cols[FRHIP_BAM_CIGAR_START],
                     (const uint32_t *)p->d_cols[FRHIP_BAM_SEQUENCE_START],
                     (const uint32_t *)p->d_cols[FRHIP_BAM_AUX_START],
                     (const uint32_t *)p->d_cols[FRHIP_BAM_RECORD_END]);
    a.cap_ops = p->soa_cap[0];
    a.cap_seq = p->soa_cap[1];
    a.cap_aux = p->soa_cap[2];
    a.op_off = (uint64_t *)p->d_soa_cols[FRHIP_SOA_CIGAR_OFFSETS];
    a.seq_off = (uint64_t *)p->d_soa_cols[FRHIP_SOA_SEQUENCE_OFFSETS];
    a.aux_off = (uint64_t *)p->d_soa_cols[FRHIP_SOA_AUX_OFFSETS];
    a.ops = (uint32_t *)p->d_soa_cols[FRHIP_SOA_CIGAR_OPS];
    a.seq = (uint8_t *)p->d_soa_cols[FRHIP_SOA_SEQUENCE_PACKED];
    a.aux = (AuxEntry *)p->d_soa_cols[FRHIP_SOA_AUX_ENTRIES];
    return a;
}

// The SoA counts and their prefix; with `fill`, the offsets and data too.
hipError_t enqueue_soa(frhip_pipeline *p, hipStream_t s, bool fill)
{
    const SoaArgs a = soa_args(p);
    hipError_t r = launch_fill_u64((uint64_t *)a.first_bad, ~0ull, 1, s);
    if (r == hipSuccess) r = launch_soa_count(a, s);
    if (r == hipSuccess) r = launch_soa_scan(a, s);
    if (r == hipSuccess && fill) r = launch_soa_fill(a, s);
    return r;
}

hipError_t enqueue_run(frhip_pipeline *p, hipStream_t s, bool timed)
{
    hipError_t r = hipSuccess;
    auto rec = [&](int e) {
        if (timed && r == hipSuccess) r = hipEventRecord(p->ev[e], s);
    };
    rec(EV_START);
    if (p->device_discovery) r = enqueue_discovery(p, s);
    if (r == hipSuccess) r = launch_fill_u64(p->d_queue, 0, 1, s);
    if (r == hipSuccess) r = launch_fill_u64((uint64_t *)&p->d_res->first_bad, ~0ull, 1, s);
    rec(EV_DISC);
    if (r == hipSuccess) {
        InflateArgs a{};
        a.in = p->d_file;
        a.in_off = p->m.in_off;
        a.in_len = p->m.in_len;
        a.out = p->d_u;
        a.out_off = p->m.out_off;
        a.out_cap = p->m.isize;
        a.status = p->d_st;
        a.produced = p->d_st + p->n;
        a.crc = p->d_st + 2 * p->n;
        a.n = (uint32_t)p->n;
        a.queue = (uint32_t *)p->d_queue;
        a.order = p->cfg.order ? p->d_order : nullptr;
        a.x2n = p->ctx->x2n;
        const int sched = p->sched;
        inflate_scratch_bind(a, p->d_rec, p->n);
        a.two_phase_min = p->ctx->two_phase_min;
        uint32_t wpc = p->cfg.waves_per_cu ? p->cfg.waves_per_cu : inflate_waves_per_cu(sched);
        uint64_t waves = (uint64_t)p->ctx->n_cu * wpc;
        if (waves > p->n) waves = p->n;
        if (waves == 0) waves = 1;
        r = launch_inflate(a, (uint32_t)waves, s, sched);
    }
    rec(EV_INFL);
    if (r == hipSuccess)
        r = launch_verify(p->d_st, p->d_st + p->n, p->d_st + 2 * p->n, p->m, p->d_n, p->n,
                          (unsigned long long *)&p->d_res->first_bad, s);
    if (r == hipSuccess) r = launch_bam_header(p->d_u, p->d_ulen, &p->d_res->records_start, &p->d_res->hdr_status, s);
    rec(EV_VERIFY);
    uint8_t *b = p->d_bam;
    const BamScratchLayout &L = p->L;
    BamSpecArgs sa{};
    sa.buf = p->d_u;
    sa.len = p->U;
    sa.block_starts = p->m.out_off;
    sa.nb = (uint32_t)p->n;
    sa.window = p->cfg.bam_window ? p->cfg.bam_window : p->ctx->bam_window;
    sa.spec_ok = b + L.o_ok;
    sa.n_surv = b + L.o_ns;
    sa.surv_pos = (uint64_t *)(b + L.o_spos);
    sa.surv_cnt = (uint32_t *)(b + L.o_scnt);
    sa.land = (uint64_t *)(b + L.o_land);
    sa.exitp = (uint64_t *)(b + L.o_exit);
    sa.cnt_tail = (uint32_t *)(b + L.o_ctail);
    sa.complete = b + L.o_comp;
    sa.rel_offs = (uint16_t *)(b + L.o_rel);
    if (r == hipSuccess) r = launch_bam_spec(sa, s);
    rec(EV_SPEC);
    BamReconcileArgs ra{};
    ra.buf = p->d_u;
    ra.len = p->U;
    ra.block_starts = p->m.out_off;
    ra.nb = (uint32_t)p->n;
    ra.start = 0;
    ra.start_dev = &p->d_res->records_start;
    ra.len_dev = nullptr;
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
    ra.totals = p->d_res->totals;
    ra.fast_scan = (uint64_t *)(b + L.o_fscan);
    ra.fast_state = (uint32_t *)(b + L.o_fstate);
    if (r == hipSuccess) r = launch_bam_reconcile(ra, s);
    rec(EV_REC);
    BamEmitArgs ea{};
    ea.buf = p->d_u;
    ea.block_starts = p->m.out_off;
    ea.coffsets = p->m.coff;
    ea.nb = (uint32_t)p->n;
    ea.entry = ra.entry;
    ea.count = ra.count;
    ea.first_index = ra.first_index;
    ea.mode = ra.mode;
    ea.surv_cnt = sa.surv_cnt;
    ea.rel_offs = sa.rel_offs;
    ea.record_offsets = (uint64_t *)p->d_cols[FRHIP_BAM_RECORD_OFFSETS];
    ea.virtual_offsets = (uint64_t *)p->d_cols[FRHIP_BAM_VIRTUAL_OFFSETS];
    if (r == hipSuccess) r = launch_bam_emit(ea, s);
    BamDecodeArgs da{};
    da.buf = p->d_u;
    da.len = p->U;
    da.n = 0;
    da.n_dev = &p->d_res->totals[0];
    da.c.ref_id = (int32_t *)p->d_cols[0];
    da.c.pos = (int32_t *)p->d_cols[1];
    da.c.flags = (uint16_t *)p->d_cols[2];
    da.c.mapq = (uint8_t *)p->d_cols[3];
    da.c.l_seq = (uint32_t *)p->d_cols[4];
    da.c.next_ref_id = (int32_t *)p->d_cols[5];
    da.c.next_pos = (int32_t *)p->d_cols[6];
    da.c.tlen = (int32_t *)p->d_cols[7];
    da.c.record_offsets = (uint64_t *)p->d_cols[8];
    da.c.cigar_start = (uint32_t *)p->d_cols[9];
    da.c.seq_start = (uint32_t *)p->d_cols[10];
    da.c.qual_start = (uint32_t *)p->d_cols[11];
    da.c.aux_start = (uint32_t *)p->d_cols[12];
    da.c.record_end = (uint32_t *)p->d_cols[13];
    if (r == hipSuccess) r = launch_bam_decode(da, p->cap_rec, s);
    rec(EV_COLS);
    if (r == hipSuccess && p->soa && p->d_soa_scratch) r = enqueue_soa(p, s, true);
    rec(EV_END);
    return r;
}

void clear_views(frhip_pipeline *p)
{
    for (auto &v : p->col_view) v.len = 0;
    for (auto &v : p->soa_view) v.len = 0;
}

int check_result_inner(frhip_pipeline *p, frhip_error *err)
{
    const ResultBlock &R = p->res;
    if (p->device_discovery && R.chain_bad)
        return set_err(err, FRHIP_E_INVALID_BLOCK, 0, "device member discovery lost the member chain");
    if (R.first_bad != ~0ull) {
        uint64_t i = R.first_bad;
        uint32_t st[3], isz, crc;
        uint64_t coff;
        hipError_t r = hipMemcpy(&st[0], p->d_st + i, 4, hipMemcpyDeviceToHost);
        if (r == hipSuccess) r = hipMemcpy(&st[1], p->d_st + p->n + i, 4, hipMemcpyDeviceToHost);
        if (r == hipSuccess) r = hipMemcpy(&st[2], p->d_st + 2 * p->n + i, 4, hipMemcpyDeviceToHost);
        if (r == hipSuccess) r = hipMemcpy(&isz, p->m.isize + i, 4, hipMemcpyDeviceToHost);
        if (r == hipSuccess) r = hipMemcpy(&crc, p->m.crc_expect + i, 4, hipMemcpyDeviceToHost);
        if (r == hipSuccess) r = hipMemcpy(&coff, p->m.coff + i, 8, hipMemcpyDeviceToHost);
        if (r != hipSuccess) return hip_err(err, r, "reading the failing member");
        if (st[0] != 0) {
            int k = set_err(err, FRHIP_E_INFLATE, coff, "%s", inflate_status_message(st[0]));
            if (err) err->code = st[0];
            return k;
        }
        if (st[1] != isz) {
            int k = set_err(err, FRHIP_E_SIZE_MISMATCH, coff, "ISIZE says %u bytes, inflated %u", isz, st[1]);
            if (err) { err->expected = isz; err->actual = st[1]; }
            return k;
        }
        int k = set_err(err, FRHIP_E_CHECKSUM_MISMATCH, coff, "expected crc32 %#010x, got %#010x", crc, st[2]);
        if (err) { err->expected = crc; err->actual = st[2]; }
        return k;
    }
    if (R.hdr_status) {
        int k = set_err(err, FRHIP_E_MALFORMED, 0, "BAM header does not parse");
        if (err) std::strcpy(err->format, "bam header");
        return k;
    }
    if (R.totals[2]) {
        int k = set_err(err, FRHIP_E_MALFORMED, R.totals[1], "malformed record reached from a confirmed boundary");
        if (err) {
            err->code = (uint32_t)R.totals[2];
            std::strcpy(err->format, "bam");
        }
        return k;
    }
    if (p->soa) {
        const uint64_t *S = R.soa;
        if (S[SOA_N_RESULT] != ~0ull) {
            uint64_t at = 0;
            (void)hipMemcpy(&at, (const uint64_t *)p->d_cols[FRHIP_BAM_RECORD_OFFSETS] + S[SOA_N_RESULT], 8,
                            hipMemcpyDeviceToHost);
            int k = set_err(err, FRHIP_E_MALFORMED, at,
                            "record %llu: malformed aux field, or a long-CIGAR placeholder without CG:B:I",
                            (unsigned long long)S[SOA_N_RESULT]);
            if (err) std::strcpy(err->format, "bam");
            return k;
        }
        if (S[SOA_OVERFLOW] || S[SOA_TOT_OPS] > p->soa_cap[0] || S[SOA_TOT_SEQ] > p->soa_cap[1] ||
            S[SOA_TOT_AUX] > p->soa_cap[2])
            return set_err(err, FRHIP_E_DEVICE, 0, "SoA columns outgrew the sizes found at create");
    }
    const uint64_t n = R.totals[0];
    for (int c = 0; c < FRHIP_BAM_N_COLUMNS; ++c) p->col_view[c].len = n * WIDTHS[c];
    if (p->soa) {
        const uint64_t lens[FRHIP_SOA_N_COLUMNS] = {(n + 1) * 8, R.soa[SOA_TOT_OPS] * 4, (n + 1) * 8,
                                                    R.soa[SOA_TOT_SEQ], (n + 1) * 8, R.soa[SOA_TOT_AUX] * 16};
        for (int c = 0; c < FRHIP_SOA_N_COLUMNS; ++c) p->soa_view[c].len = lens[c];
    }
    return FRHIP_OK;
}

// A failed run leaves every column view empty (the reference leaves a failed
// batch's output empty: crates/fritillaria-cuda/tests/device_resident.rs:208-236).
int check_result(frhip_pipeline *p, frhip_error *err)
{
    clear_views(p);
    const int rc = check_result_inner(p, err);
    p->last_ok = rc == FRHIP_OK;
    return rc;
}

} // namespace

extern "C" int frhip_pipeline_create(frhip_ctx *ctx, const uint8_t *file, uint64_t len,
                                     const frhip_pipeline_config *cfg, frhip_pipeline **out, frhip_error *err)
{
    *out = nullptr;
    (void)hipSetDevice(ctx->ordinal);
    frhip_pipeline *p = new (std::nothrow) frhip_pipeline;
    if (!p) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    p->ctx = ctx;
    ctx_retain(ctx);
    p->cfg.inflate_schedule = -1;
    if (cfg) p->cfg = *cfg;
    p->file_len = len;
    hipStream_t s = ctx->stream;
    int rc;
    auto fail = [&](int k) {
        frhip_pipeline_free(p);
        return k;
    };
    if ((rc = dalloc(p, (void **)&p->d_file, len + 4096, err, "file buffer"))) return fail(rc);
    hipError_t r = hipMemcpy(p->d_file, file, len, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemset(p->d_file + len, 0, 4096);
    if (r != hipSuccess) return fail(hip_err(err, r, "uploading the file"));
    if ((rc = dalloc(p, (void **)&p->d_res, sizeof(ResultBlock), err, "result block"))) return fail(rc);
    r = hipHostMalloc((void **)&p->h_res, sizeof(ResultBlock), hipHostMallocDefault);
    if (r != hipSuccess) return fail(hip_err(err, r, "hipHostMalloc"));

    // discovery once, to size everything
    std::vector<frhip_block_span> host_spans;
    p->n_tiles = sieve_tiles(len);
    bool try_device = !p->cfg.host_discovery && len > 0;
    if (try_device) {
        p->cand_cap = len / 28 + 2;
        if ((rc = dalloc(p, (void **)&p->d_tile_counts, p->n_tiles * 4 + 4, err, "tile counts"))) return fail(rc);
        if ((rc = dalloc(p, (void **)&p->d_tile_base, (p->n_tiles + 1) * 8, err, "tile bases"))) return fail(rc);
        r = launch_sieve_count(p->d_file, len, p->d_tile_counts, s);
        if (r == hipSuccess) r = launch_scan_u32_u64(p->d_tile_counts, p->d_tile_base, nullptr, p->n_tiles, s);
        uint64_t n_cand = 0;
        if (r == hipSuccess)
            r = hipMemcpyAsync(&n_cand, p->d_tile_base + p->n_tiles, 8, hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipStreamSynchronize(s);
        if (r != hipSuccess) return fail(hip_err(err, r, "member sieve"));
        p->n = n_cand;
        p->cand_cap = n_cand;
    }
    if (!try_device) {
        uint64_t pos = 0;
        rc = discover(file, len, 0, host_spans, &pos, err);
        if (rc) return fail(rc);
        if (pos != len)
            return fail(set_err(err, FRHIP_E_INVALID_BLOCK, pos, "truncated BGZF: trailing bytes are not a member"));
        p->n = host_spans.size();
    }
    const uint64_t n = p->n;
    uint64_t nn = n ? n : 1;
    if ((rc = dalloc(p, (void **)&p->m.coff, nn * 8, err, "members"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->m.in_off, nn * 8, err, "members"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->m.in_len, nn * 4, err, "members"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->m.isize, nn * 4, err, "members"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->m.crc_expect, nn * 4, err, "members"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->m.out_off, (nn + 1) * 8, err, "members"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->d_n, 8, err, "count"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->d_ulen, 8, err, "count"))) return fail(rc);
    if (try_device) {
        if ((rc = dalloc(p, (void **)&p->d_cand, nn * 8, err, "candidates"))) return fail(rc);
        if ((rc = dalloc(p, (void **)&p->d_cand_bs, nn * 4, err, "candidates"))) return fail(rc);
        if ((rc = dalloc(p, (void **)&p->d_cand_hs, nn * 4, err, "candidates"))) return fail(rc);
        r = enqueue_discovery(p, s);
        uint64_t bad = 0;
        if (r == hipSuccess) r = hipMemcpyAsync(&bad, &p->d_res->chain_bad, 8, hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipStreamSynchronize(s);
        if (r != hipSuccess) return fail(hip_err(err, r, "member discovery"));
        p->device_discovery = bad == 0;
        if (!p->device_discovery) {
            // the sieve's candidates are not exactly the member chain: walk it
            uint64_t pos = 0;
            host_spans.clear();
            rc = discover(file, len, 0, host_spans, &pos, err);
            if (rc) return fail(rc);
            if (pos != len)
                return fail(set_err(err, FRHIP_E_INVALID_BLOCK, pos, "truncated BGZF: trailing bytes are not a member"));
            if (host_spans.size() > p->cand_cap)
                return fail(set_err(err, FRHIP_E_DEVICE, 0, "member table larger than sized"));
            p->n = host_spans.size();
        }
    }
    if (!p->device_discovery) {
        const uint64_t k = p->n;
        std::vector<uint64_t> coff(k), in_off(k), out_off(k + 1);
        std::vector<uint32_t> in_len(k), isz(k), crc(k);
        out_off[0] = 0;
        for (uint64_t i = 0; i < k; ++i) {
            coff[i] = host_spans[i].compressed_offset;
            in_off[i] = host_spans[i].payload_start;
            in_len[i] = (uint32_t)host_spans[i].payload_len;
            isz[i] = host_spans[i].isize;
            crc[i] = host_spans[i].crc32;
            out_off[i + 1] = out_off[i] + isz[i];
        }
        r = hipMemcpy(p->m.coff, coff.data(), k * 8, hipMemcpyHostToDevice);
        if (r == hipSuccess) r = hipMemcpy(p->m.in_off, in_off.data(), k * 8, hipMemcpyHostToDevice);
        if (r == hipSuccess) r = hipMemcpy(p->m.in_len, in_len.data(), k * 4, hipMemcpyHostToDevice);
        if (r == hipSuccess) r = hipMemcpy(p->m.isize, isz.data(), k * 4, hipMemcpyHostToDevice);
        if (r == hipSuccess) r = hipMemcpy(p->m.crc_expect, crc.data(), k * 4, hipMemcpyHostToDevice);
        if (r == hipSuccess) r = hipMemcpy(p->m.out_off, out_off.data(), (k + 1) * 8, hipMemcpyHostToDevice);
        if (r != hipSuccess) return fail(hip_err(err, r, "uploading members"));
    }
    r = hipMemcpy(&p->U, p->m.out_off + p->n, 8, hipMemcpyDeviceToHost);
    p->sized = true;
    if (r == hipSuccess) r = hipMemcpy(p->d_n, &p->n, 8, hipMemcpyHostToDevice);
    if (r == hipSuccess) r = hipMemcpy(p->d_ulen, &p->U, 8, hipMemcpyHostToDevice);
    if (r != hipSuccess) return fail(hip_err(err, r, "member table"));
    if (p->n > 0xFFFFFFFFull) return fail(set_err(err, FRHIP_E_DEVICE, 0, "too many members for one pipeline"));

    // inflate buffers
    if ((rc = dalloc(p, (void **)&p->d_u, p->U, err, "inflated output"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->d_st, nn * 12, err, "member status"))) return fail(rc);
    if ((rc = dalloc(p, (void **)&p->d_queue, 8, err, "work queue"))) return fail(rc);
    p->sched = p->cfg.inflate_schedule >= 0 && p->cfg.inflate_schedule < N_INFLATE_SCHEDULES ? p->cfg.inflate_schedule
                                                                                            : ctx->inflate_schedule;
    if (const uint64_t sb = inflate_scratch_bytes(p->sched, p->U, nn))
        if ((rc = dalloc(p, &p->d_rec, sb, err, "match records"))) return fail(rc);
    if (p->cfg.order) {
        // largest compressed payload first: the long members start early
        std::vector<uint32_t> in_len(p->n);
        r = hipMemcpy(in_len.data(), p->m.in_len, p->n * 4, hipMemcpyDeviceToHost);
        if (r != hipSuccess) return fail(hip_err(err, r, "member lengths"));
        std::vector<uint32_t> order(p->n);
        std::iota(order.begin(), order.end(), 0u);
        if (p->cfg.order == 2) {
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return in_len[a] < in_len[b]; });
        } else if (p->cfg.order == 3) {
            // members whose payload is no smaller than their ISIZE (stored
            // blocks: a lane copies them, which would hold up a wave of
            // decoding lanes) first, together; then the rest largest first
            // within each run of 1024 file-order members: a wave's members
            // are alike and still come from one stretch of the file
            std::vector<uint32_t> isz(p->n);
            r = hipMemcpy(isz.data(), p->m.isize, p->n * 4, hipMemcpyDeviceToHost);
            if (r != hipSuccess) return fail(hip_err(err, r, "member sizes"));
            const auto mid = std::stable_partition(order.begin(), order.end(),
                                                   [&](uint32_t a) { return in_len[a] >= isz[a]; });
            for (auto i = mid; i < order.end(); i += std::min<std::ptrdiff_t>(1024, order.end() - i)) {
                const auto e = i + std::min<std::ptrdiff_t>(1024, order.end() - i);
                std::stable_sort(i, e, [&](uint32_t a, uint32_t b) { return in_len[a] > in_len[b]; });
            }
        } else {
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return in_len[a] > in_len[b]; });
        }
        if ((rc = dalloc(p, (void **)&p->d_order, nn * 4, err, "member order"))) return fail(rc);
        r = hipMemcpy(p->d_order, order.data(), p->n * 4, hipMemcpyHostToDevice);
        if (r != hipSuccess) return fail(hip_err(err, r, "member order"));
    }
    // BAM buffers: at most one record per 36 inflated bytes
    p->L = bam_layout(nn);
    if ((rc = dalloc(p, (void **)&p->d_bam, p->L.bytes, err, "boundary scan"))) return fail(rc);
    p->cap_rec = p->U / 36 + 1;
    for (int c = 0; c < FRHIP_BAM_N_COLUMNS; ++c) {
        if ((rc = dalloc(p, &p->d_cols[c], p->cap_rec * WIDTHS[c], err, "columns"))) return fail(rc);
        p->col_view[c] = frhip_buf{p->d_cols[c], 0, ctx->ordinal};
    }
    p->u_view = frhip_buf{p->d_u, p->U, ctx->ordinal};
    for (auto &e : p->ev) {
        r = hipEventCreate(&e);
        if (r != hipSuccess) return fail(hip_err(err, r, "hipEventCreate"));
    }
    p->soa = p->cfg.soa != 0;
    if (p->soa && p->n > 0) {
        // One run learns the record count, and the SoA counts of those records
        // the totals the three data columns are sized for; the resident file
        // gives the same counts on every later run.
        r = enqueue_run(p, s, false);
        if (r == hipSuccess) r = hipMemcpyAsync(p->h_res, p->d_res, sizeof(ResultBlock), hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipStreamSynchronize(s);
        if (r != hipSuccess) return fail(hip_err(err, r, "sizing run"));
        p->soa_cap_rec = std::min<uint64_t>(p->h_res->totals[0], p->cap_rec);
        if ((rc = dalloc(p, (void **)&p->d_soa_scratch, soa_scratch_layout(p->soa_cap_rec).bytes, err,
                         "SoA scratch")))
            return fail(rc);
        for (int c : {FRHIP_SOA_CIGAR_OFFSETS, FRHIP_SOA_SEQUENCE_OFFSETS, FRHIP_SOA_AUX_OFFSETS})
            if ((rc = dalloc(p, &p->d_soa_cols[c], (p->soa_cap_rec + 1) * 8, err, "SoA offsets"))) return fail(rc);
        r = enqueue_soa(p, s, false);
        if (r == hipSuccess) r = hipMemcpyAsync(p->h_res, p->d_res, sizeof(ResultBlock), hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipStreamSynchronize(s);
        if (r != hipSuccess) return fail(hip_err(err, r, "sizing the SoA columns"));
        const uint64_t *S = p->h_res->soa;
        p->soa_cap[0] = S[SOA_TOT_OPS];
        p->soa_cap[1] = S[SOA_TOT_SEQ];
        p->soa_cap[2] = S[SOA_TOT_AUX];
        const int data_cols[3] = {FRHIP_SOA_CIGAR_OPS, FRHIP_SOA_SEQUENCE_PACKED, FRHIP_SOA_AUX_ENTRIES};
        for (int k = 0; k < 3; ++k)
            if ((rc = dalloc(p, &p->d_soa_cols[data_cols[k]], p->soa_cap[k] * SOA_WIDTHS[data_cols[k]], err,
                             "SoA columns")))
                return fail(rc);
        for (int c = 0; c < FRHIP_SOA_N_COLUMNS; ++c) p->soa_view[c] = frhip_buf{p->d_soa_cols[c], 0, ctx->ordinal};
    }
    if (p->cfg.use_graph && p->n > 0) {
        r = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (r == hipSuccess) {
            hipError_t r2 = enqueue_run(p, s, false);
            r = hipStreamEndCapture(s, &p->graph);
            if (r2 != hipSuccess) r = r2;
        }
        if (r == hipSuccess) r = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
        if (r != hipSuccess) return fail(hip_err(err, r, "capturing the pipeline graph"));
    }
    *out = p;
    return FRHIP_OK;
}

extern "C" void frhip_pipeline_free(frhip_pipeline *p)
{
    if (!p) return;
    (void)hipSetDevice(p->ctx->ordinal);
    (void)hipStreamSynchronize(p->ctx->stream);
    if (p->copy_stream) (void)hipStreamSynchronize(p->copy_stream);
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->ev_uploaded) (void)hipEventDestroy(p->ev_uploaded);
    if (p->ev_done) (void)hipEventDestroy(p->ev_done);
    if (p->copy_stream) (void)hipStreamDestroy(p->copy_stream);
    for (auto e : p->ev)
        if (e) (void)hipEventDestroy(e);
    for (void *a : p->allocs) (void)hipFree(a);
    if (p->h_res) (void)hipHostFree(p->h_res);
    ctx_release(p->ctx);
    delete p;
}

extern "C" int frhip_pipeline_launch(frhip_pipeline *p, frhip_error *err)
{
    (void)hipSetDevice(p->ctx->ordinal);
    hipStream_t s = p->ctx->stream;
    if (p->n == 0) return FRHIP_OK;
    hipError_t r = hipSuccess;
    if (p->upload_pending) {
        r = hipStreamWaitEvent(s, p->ev_uploaded, 0);
        if (r != hipSuccess) return hip_err(err, r, "waiting for the upload");
        p->upload_pending = false;
    }
    if (p->exec) {
        r = hipGraphLaunch(p->exec, s);
        p->timed = false;
    } else {
        r = enqueue_run(p, s, true);
        p->timed = true;
    }
    if (r == hipSuccess) r = hipMemcpyAsync(p->h_res, p->d_res, sizeof(ResultBlock), hipMemcpyDeviceToHost, s);
    if (r == hipSuccess && p->ev_done) {
        r = hipEventRecord(p->ev_done, s);
        p->launched = true;
    }
    if (r != hipSuccess) return hip_err(err, r, "launching the pipeline");
    return FRHIP_OK;
}

extern "C" int frhip_pipeline_upload(frhip_pipeline *p, const uint8_t *file, uint64_t len, frhip_error *err)
{
    if (len != p->file_len)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "file of %llu bytes, the pipeline was built for %llu",
                       (unsigned long long)len, (unsigned long long)p->file_len);
    (void)hipSetDevice(p->ctx->ordinal);
    hipError_t r = hipSuccess;
    if (!p->copy_stream) {
        // runs launched before there was an end event to wait for
        r = hipStreamSynchronize(p->ctx->stream);
        if (r == hipSuccess) r = hipStreamCreateWithFlags(&p->copy_stream, hipStreamNonBlocking);
        if (r == hipSuccess) r = hipEventCreateWithFlags(&p->ev_uploaded, hipEventDisableTiming);
        if (r == hipSuccess) r = hipEventCreateWithFlags(&p->ev_done, hipEventDisableTiming);
        if (r != hipSuccess) return hip_err(err, r, "creating the upload stream");
    }
    // the last run has read the old bytes before the copy overwrites them
    if (p->launched) r = hipStreamWaitEvent(p->copy_stream, p->ev_done, 0);
    const uint64_t chunk = 64ull << 20;
    for (uint64_t o = 0; r == hipSuccess && o < len; o += chunk)
        r = hipMemcpyAsync(p->d_file + o, file + o, std::min(chunk, len - o), hipMemcpyHostToDevice,
                           p->copy_stream);
    if (r == hipSuccess) r = hipEventRecord(p->ev_uploaded, p->copy_stream);
    if (r != hipSuccess) return hip_err(err, r, "uploading the file");
    p->upload_pending = true;
    return FRHIP_OK;
}

extern "C" int frhip_pipeline_finish(frhip_pipeline *p, frhip_error *err)
{
    if (p->n == 0) {
        p->res = ResultBlock{};
        p->res.first_bad = ~0ull;
        p->last_ok = false;
        clear_views(p);
        return set_err(err, FRHIP_E_MALFORMED, 0, "no BGZF members");
    }
    hipError_t r = hipStreamSynchronize(p->ctx->stream);
    if (r != hipSuccess) return hip_err(err, r, "pipeline");
    p->res = *p->h_res;
    if (p->timed) {
        float ms = 0;
        for (int i = 0; i + 1 < N_EV; ++i) {
            (void)hipEventElapsedTime(&ms, p->ev[i], p->ev[i + 1]);
            p->stage_ms[i] = ms;
        }
        (void)hipEventElapsedTime(&ms, p->ev[EV_START], p->ev[EV_END]);
        p->stage_ms[N_STAGES - 1] = ms;
    }
    return check_result(p, err);
}

extern "C" int frhip_pipeline_run(frhip_pipeline *p, frhip_error *err)
{
    int rc = frhip_pipeline_launch(p, err);
    if (rc) return rc;
    return frhip_pipeline_finish(p, err);
}

extern "C" uint64_t frhip_pipeline_members(const frhip_pipeline *p) { return p->n; }
extern "C" uint64_t frhip_pipeline_inflated_len(const frhip_pipeline *p) { return p->U; }
extern "C" uint64_t frhip_pipeline_compressed_len(const frhip_pipeline *p) { return p->file_len; }
extern "C" uint64_t frhip_pipeline_records(const frhip_pipeline *p) { return p->last_ok ? p->res.totals[0] : 0; }
extern "C" uint64_t frhip_pipeline_records_start(const frhip_pipeline *p) { return p->res.records_start; }
extern "C" uint64_t frhip_pipeline_tail(const frhip_pipeline *p) { return p->res.totals[1]; }
extern "C" uint64_t frhip_pipeline_fallbacks(const frhip_pipeline *p) { return p->res.totals[3]; }
extern "C" int frhip_pipeline_device_discovery(const frhip_pipeline *p) { return p->device_discovery ? 1 : 0; }
extern "C" const frhip_buf *frhip_pipeline_inflated(const frhip_pipeline *p) { return &p->u_view; }
extern "C" const frhip_buf *frhip_pipeline_column(const frhip_pipeline *p, int column)
{
    if (column < 0 || column >= FRHIP_BAM_N_COLUMNS) return nullptr;
    return &p->col_view[column];
}

extern "C" const frhip_buf *frhip_pipeline_soa_column(const frhip_pipeline *p, int column)
{
    if (!p->soa || column < 0 || column >= FRHIP_SOA_N_COLUMNS) return nullptr;
    return &p->soa_view[column];
}

extern "C" int frhip_pipeline_member_offsets(const frhip_pipeline *p, uint64_t *offsets, uint64_t *coffsets,
                                             frhip_error *err)
{
    hipError_t r = hipSuccess;
    if (offsets) r = hipMemcpy(offsets, p->m.out_off, (p->n + 1) * 8, hipMemcpyDeviceToHost);
    if (r == hipSuccess && coffsets && p->n) r = hipMemcpy(coffsets, p->m.coff, p->n * 8, hipMemcpyDeviceToHost);
    if (r != hipSuccess) return hip_err(err, r, "member offsets");
    return FRHIP_OK;
}

extern "C" int frhip_pipeline_stage_ms(const frhip_pipeline *p, double *ms, int n)
{
    int k = n < N_STAGES ? n : N_STAGES;
    for (int i = 0; i < k; ++i) ms[i] = p->timed ? p->stage_ms[i] : -1.0;
    return k;
}
