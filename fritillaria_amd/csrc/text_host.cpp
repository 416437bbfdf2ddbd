// text_host.cpp -- host side of text framing and FASTA compaction.
//
// TextScanner::scan (crates/fritillaria-cuda/src/text.rs:196-375) and
// FastaCompactor::compact (crates/fritillaria-cuda/src/fasta.rs:203-304) over
// an inflated batch already on the device. Each reads back only a few
// totals (to size its columns); the layout work is on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>

#include "../../include/frhip.h"
#include "host_internal.h"
#include "pipeline_internal.h"
#include "text_internal.h"

using namespace frhip;

struct frhip_text_batch {
    uint64_t n = 0, headers = 0, tail = 0;
    frhip_buf *cols[FRHIP_TEXT_N_COLUMNS] = {};
};

struct frhip_fasta_reference {
    uint64_t contigs = 0, total = 0;
    frhip_buf *cols[FRHIP_FASTA_N_COLUMNS] = {};
};

namespace {

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

int check_device(frhip_ctx *ctx, const frhip_buf *data, frhip_error *err)
{
    if (data->ordinal != ctx->ordinal)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "batch is on device %d but this scanner is on %d",
                       data->ordinal, ctx->ordinal);
    return FRHIP_OK;
}

int new_text(frhip_ctx *ctx, const uint64_t bytes[FRHIP_TEXT_N_COLUMNS], frhip_text_batch **out, frhip_error *err)
{
    frhip_text_batch *b = new (std::nothrow) frhip_text_batch;
    if (!b) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    for (int i = 0; i < FRHIP_TEXT_N_COLUMNS; ++i) {
        b->cols[i] = buf_alloc(ctx, bytes[i], err);
        if (!b->cols[i]) {
            frhip_text_batch_free(b);
            return err ? err->kind : FRHIP_E_DEVICE;
        }
    }
    *out = b;
    return FRHIP_OK;
}

int new_fasta(frhip_ctx *ctx, const uint64_t bytes[FRHIP_FASTA_N_COLUMNS], frhip_fasta_reference **out,
              frhip_error *err)
{
    frhip_fasta_reference *r = new (std::nothrow) frhip_fasta_reference;
    if (!r) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    for (int i = 0; i < FRHIP_FASTA_N_COLUMNS; ++i) {
        r->cols[i] = buf_alloc(ctx, bytes[i], err);
        if (!r->cols[i]) {
            frhip_fasta_free(r);
            return err ? err->kind : FRHIP_E_DEVICE;
        }
    }
    *out = r;
    return FRHIP_OK;
}

} // namespace

extern "C" int frhip_text_scan(frhip_ctx *ctx, const frhip_buf *data, uint64_t start, uint32_t comment,
                               uint32_t secondary, frhip_text_batch **out, frhip_error *err)
{
    *out = nullptr;
    const uint64_t len = data ? data->len : 0;
    if (!data || start >= len) {
        const uint64_t zero[FRHIP_TEXT_N_COLUMNS] = {};
        int rc = new_text(ctx, zero, out, err);
        if (rc == FRHIP_OK) (*out)->tail = std::min(start, len);
        return rc;
    }
    if (int rc = check_device(ctx, data, err)) return rc;
    (void)hipSetDevice(ctx->ordinal);
    hipStream_t st = ctx->stream;
    TextArgs a{};
    a.buf = (const uint8_t *)data->ptr;
    a.len = len;
    a.start = start;
    a.comment = comment;
    a.secondary = secondary;
    a.first_tile = start / TX_TILE;
    a.n_tiles = tx_tiles(len) - a.first_tile;
    if (a.n_tiles >= (1ull << 31))
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "batch of %llu bytes is too large",
                       (unsigned long long)len);
    const uint64_t T = a.n_tiles;
    uint64_t p = 0;
    auto take = [&](uint64_t b) { uint64_t o = p; p = align_up(p + b, 256); return o; };
    const uint64_t o_rec = take(T * 4), o_hdr = take(T * 4), o_tr = take(T * 4), o_th = take(T * 4),
                   o_lk = take(T), o_ls = take(T * 8), o_cy = take(T), o_rb = take((T + 1) * 8),
                   o_hb = take((T + 1) * 8), o_tb = take((T + 1) * 8), o_tot = take(TX_T_N * 8);
    const uint64_t G = text_groups(T);
    const uint64_t o_g64 = take(G * 8 * 5), o_g32 = take(G * 4 * 2);
    int rc;
    uint8_t *s = (uint8_t *)ctx->tx_scratch.get(p, err, &rc);
    if (!s) return rc;
    a.t_rec = (uint32_t *)(s + o_rec);
    a.t_hdr = (uint32_t *)(s + o_hdr);
    a.t_tabs_rec = (uint32_t *)(s + o_tr);
    a.t_tabs_head = (uint32_t *)(s + o_th);
    a.t_last_kind = s + o_lk;
    a.t_last_ls = (uint64_t *)(s + o_ls);
    a.t_carry = s + o_cy;
    a.rec_base = (uint64_t *)(s + o_rb);
    a.hdr_base = (uint64_t *)(s + o_hb);
    a.tab_base = (uint64_t *)(s + o_tb);
    a.totals = (uint64_t *)(s + o_tot);
    a.g_rec = (uint64_t *)(s + o_g64);
    a.g_hdr = a.g_rec + G;
    a.g_known = a.g_hdr + G;
    a.g_pending = a.g_known + G;
    a.g_last_ls = a.g_pending + G;
    a.g_last_kind = (uint32_t *)(s + o_g32);
    a.g_carry = a.g_last_kind + G;
    hipError_t r = launch_text_count(a, st);
    if (r == hipSuccess) r = launch_text_tiles(a, st);
    uint64_t tot[TX_T_N] = {};
    uint8_t last = 0;
    if (r == hipSuccess) r = hipMemcpyAsync(tot, a.totals, sizeof tot, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipMemcpyAsync(&last, a.buf + len - 1, 1, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return hip_err(err, r, "text line count");
    const uint64_t R = tot[TX_T_RECORDS], H = tot[TX_T_HEADERS], NT = tot[TX_T_TABS];
    const uint64_t bytes[FRHIP_TEXT_N_COLUMNS] = {R * 8, R * 8, H * 8, NT * 8, (R + 1) * 8};
    frhip_text_batch *b = nullptr;
    rc = new_text(ctx, bytes, &b, err);
    if (rc != FRHIP_OK) return rc;
    a.record_offsets = (uint64_t *)b->cols[FRHIP_TEXT_RECORD_OFFSETS]->ptr;
    a.record_ends = (uint64_t *)b->cols[FRHIP_TEXT_RECORD_ENDS]->ptr;
    a.header_offsets = (uint64_t *)b->cols[FRHIP_TEXT_HEADER_OFFSETS]->ptr;
    a.tabs = (uint64_t *)b->cols[FRHIP_TEXT_TABS]->ptr;
    a.field_starts = (uint64_t *)b->cols[FRHIP_TEXT_FIELD_STARTS]->ptr;
    r = launch_text_write(a, st);
    if (r == hipSuccess) r = launch_fill_u64(a.field_starts + R, NT, 1, st);
    // a final line without a newline is not decoded: it is the tail
    uint64_t n = R, headers = H, tabs = NT, tail = len;
    const bool partial = last != '\n' && tot[TX_T_LAST_LS] != ~0ull;
    uint64_t fs = NT;
    if (partial) {
        tail = tot[TX_T_LAST_LS];
        if (tot[TX_T_LAST_KIND] == TX_RECORD) {
            n = R - 1;
            if (r == hipSuccess) r = hipMemcpyAsync(&fs, a.field_starts + n, 8, hipMemcpyDeviceToHost, st);
        } else {
            headers = H - 1;
        }
    }
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) {
        frhip_text_batch_free(b);
        return hip_err(err, r, "text fields");
    }
    if (partial && tot[TX_T_LAST_KIND] == TX_RECORD) tabs = fs;
    b->n = n;
    b->headers = headers;
    b->tail = tail;
    b->cols[FRHIP_TEXT_RECORD_OFFSETS]->len = n * 8;
    b->cols[FRHIP_TEXT_RECORD_ENDS]->len = n * 8;
    b->cols[FRHIP_TEXT_HEADER_OFFSETS]->len = headers * 8;
    b->cols[FRHIP_TEXT_TABS]->len = tabs * 8;
    b->cols[FRHIP_TEXT_FIELD_STARTS]->len = n ? (n + 1) * 8 : 0;
    *out = b;
    return FRHIP_OK;
}

extern "C" uint64_t frhip_text_batch_len(const frhip_text_batch *b) { return b->n; }
extern "C" uint64_t frhip_text_batch_headers(const frhip_text_batch *b) { return b->headers; }
extern "C" uint64_t frhip_text_batch_tail(const frhip_text_batch *b) { return b->tail; }
extern "C" const frhip_buf *frhip_text_batch_column(const frhip_text_batch *b, int column)
{
    if (!b || column < 0 || column >= FRHIP_TEXT_N_COLUMNS) return nullptr;
    return b->cols[column];
}
extern "C" void frhip_text_batch_free(frhip_text_batch *b)
{
    if (!b) return;
    for (auto *c : b->cols) buf_free(c);
    delete b;
}

extern "C" int frhip_fasta_compact(frhip_ctx *ctx, const frhip_buf *data, frhip_fasta_reference **out,
                                   frhip_error *err)
{
    *out = nullptr;
    const uint64_t zero[FRHIP_FASTA_N_COLUMNS] = {};
    const uint64_t len = data ? data->len : 0;
    if (len == 0) return new_fasta(ctx, zero, out, err);
    if (int rc = check_device(ctx, data, err)) return rc;
    (void)hipSetDevice(ctx->ordinal);
    hipStream_t st = ctx->stream;
    FastaArgs a{};
    a.buf = (const uint8_t *)data->ptr;
    a.len = len;
    a.n_tiles = tx_tiles(len);
    if (a.n_tiles >= (1ull << 31))
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "batch of %llu bytes is too large",
                       (unsigned long long)len);
    const uint64_t T = a.n_tiles;
    uint64_t p = 0;
    auto take = [&](uint64_t b) { uint64_t o = p; p = align_up(p + b, 256); return o; };
    const uint64_t o_gt = take(T * 4), o_nl = take(T * 4), o_gb = take((T + 1) * 8), o_nb = take((T + 1) * 8),
                   o_res = take(FA_T_N * 8);
    int rc;
    uint8_t *s = (uint8_t *)ctx->tx_scratch.get(p, err, &rc);
    if (!s) return rc;
    a.t_gt = (uint32_t *)(s + o_gt);
    a.t_nl = (uint32_t *)(s + o_nl);
    a.gt_base = (uint64_t *)(s + o_gb);
    a.nl_base = (uint64_t *)(s + o_nb);
    a.result = (uint64_t *)(s + o_res);
    hipError_t r = launch_fasta_count(a, st);
    if (r == hipSuccess) r = launch_scan_u32_u64(a.t_gt, a.gt_base, nullptr, T, st);
    if (r == hipSuccess) r = launch_scan_u32_u64(a.t_nl, a.nl_base, nullptr, T, st);
    uint64_t n_gt = 0;
    if (r == hipSuccess) r = hipMemcpyAsync(&n_gt, a.gt_base + T, 8, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return hip_err(err, r, "fasta contig count");
    if (n_gt == 0) return new_fasta(ctx, zero, out, err);
    a.n_gt = n_gt;
    // per contig: starts (the record offsets), sequence lengths and offsets
    // are columns; sequence starts and non-newline prefixes are scratch
    frhip_buf *per[3] = {};
    frhip_fasta_reference *ref = nullptr;
    const uint64_t bytes[FRHIP_FASTA_N_COLUMNS] = {0, (n_gt + 1) * 8, n_gt * 8, n_gt * 8};
    rc = new_fasta(ctx, bytes, &ref, err);
    if (rc != FRHIP_OK) return rc;
    per[0] = buf_alloc(ctx, n_gt * 16, err);
    if (!per[0]) {
        frhip_fasta_free(ref);
        return err ? err->kind : FRHIP_E_DEVICE;
    }
    a.starts = (uint64_t *)ref->cols[FRHIP_FASTA_RECORD_OFFSETS]->ptr;
    a.seq_len = (uint64_t *)ref->cols[FRHIP_FASTA_SEQUENCE_LENGTHS]->ptr;
    a.seq_off = (uint64_t *)ref->cols[FRHIP_FASTA_SEQUENCE_OFFSETS]->ptr;
    a.seq_start = (uint64_t *)per[0]->ptr;
    a.nonnl0 = a.seq_start + n_gt;
    r = launch_fasta_starts(a, st);
    if (r == hipSuccess) r = launch_fasta_contigs(a, st);
    if (r == hipSuccess) r = launch_scan_u64_u64(a.seq_len, a.seq_off, a.result + FA_T_VALID, 0, st);
    uint64_t n_valid = 0, total = 0;
    if (r == hipSuccess) r = hipMemcpyAsync(&n_valid, a.result + FA_T_VALID, 8, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r == hipSuccess && n_valid > n_gt) r = hipErrorInvalidValue;
    if (r == hipSuccess) r = hipMemcpyAsync(&total, a.seq_off + n_valid, 8, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) {
        buf_free(per[0]);
        frhip_fasta_free(ref);
        return hip_err(err, r, "fasta layout");
    }
    buf_free(ref->cols[FRHIP_FASTA_SEQUENCE]);
    ref->cols[FRHIP_FASTA_SEQUENCE] = buf_alloc(ctx, total, err);
    if (!ref->cols[FRHIP_FASTA_SEQUENCE]) {
        buf_free(per[0]);
        frhip_fasta_free(ref);
        return err ? err->kind : FRHIP_E_DEVICE;
    }
    a.out = (uint8_t *)ref->cols[FRHIP_FASTA_SEQUENCE]->ptr;
    if (total) r = launch_fasta_compact(a, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    buf_free(per[0]);
    if (r != hipSuccess) {
        frhip_fasta_free(ref);
        return hip_err(err, r, "fasta compaction");
    }
    ref->contigs = n_valid;
    ref->total = total;
    ref->cols[FRHIP_FASTA_SEQUENCE]->len = total;
    ref->cols[FRHIP_FASTA_SEQUENCE_OFFSETS]->len = n_valid * 8;
    ref->cols[FRHIP_FASTA_SEQUENCE_LENGTHS]->len = n_valid * 8;
    ref->cols[FRHIP_FASTA_RECORD_OFFSETS]->len = n_valid * 8;
    *out = ref;
    return FRHIP_OK;
}

extern "C" uint64_t frhip_fasta_contigs(const frhip_fasta_reference *r) { return r->contigs; }
extern "C" uint64_t frhip_fasta_total_bases(const frhip_fasta_reference *r) { return r->total; }
extern "C" const frhip_buf *frhip_fasta_reference_column(const frhip_fasta_reference *r, int column)
{
    if (!r || column < 0 || column >= FRHIP_FASTA_N_COLUMNS) return nullptr;
    return r->cols[column];
}
extern "C" void frhip_fasta_free(frhip_fasta_reference *r)
{
    if (!r) return;
    for (auto *c : r->cols) buf_free(c);
    delete r;
}
