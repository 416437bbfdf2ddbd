// fastq_host.cpp -- FastqScanner::decode behind the C ABI.
//
// Follows crates/fritillaria-cuda/src/fastq.rs:469-560 at the boundary: an
// empty batch or a start at or past its end gives no records and
// tail = min(start, len); otherwise the boundaries and the six columns of
// crates/fritillaria-fastq/src/columnar/device.rs:44-68 stay on the device.
// Two small reads come back: the newline count (to size the line index) and
// {records, tail, status, survivors}.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "../../include/frhip.h"
#include "fastq_internal.h"
#include "host_internal.h"

using namespace frhip;

static const int FQ_WIDTHS[FRHIP_FASTQ_N_COLUMNS] = {8, 4, 4, 4, 4, 4};

struct frhip_fastq_batch {
    uint64_t n = 0;
    uint64_t tail = 0;
    uint64_t survivors = 0;
    frhip_buf *cols[FRHIP_FASTQ_N_COLUMNS] = {};
};

namespace {

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

int new_batch(frhip_ctx *ctx, uint64_t cap, frhip_fastq_batch **out, frhip_error *err)
{
    frhip_fastq_batch *b = new (std::nothrow) frhip_fastq_batch;
    if (!b) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    for (int i = 0; i < FRHIP_FASTQ_N_COLUMNS; ++i) {
        b->cols[i] = buf_alloc(ctx, cap * FQ_WIDTHS[i], err);
        if (!b->cols[i]) {
            frhip_fastq_batch_free(b);
            return err ? err->kind : FRHIP_E_DEVICE;
        }
    }
    *out = b;
    return FRHIP_OK;
}

} // namespace

extern "C" int frhip_fastq_decode(frhip_ctx *ctx, const frhip_buf *data, uint64_t start, frhip_fastq_batch **out,
                                  frhip_error *err)
{
    *out = nullptr;
    const uint64_t len = data ? data->len : 0;
    if (!data || start >= len) {
        int rc = new_batch(ctx, 0, out, err);
        if (rc == FRHIP_OK) (*out)->tail = start < len ? start : len;
        return rc;
    }
    if (data->ordinal != ctx->ordinal)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "batch is on device %d but this scanner is on %d",
                       data->ordinal, ctx->ordinal);
    (void)hipSetDevice(ctx->ordinal);
    hipStream_t st = ctx->stream;

    FastqArgs a{};
    a.buf = (const uint8_t *)data->ptr;
    a.start = start;
    a.len = len;
    a.first_tile = start / FQ_TILE;
    a.n_tiles = fastq_tiles(start, len);
    // small scratch: tile counts, tile base (+1), totals
    const uint64_t o_cnt = 0;
    const uint64_t o_base = align_up(o_cnt + a.n_tiles * 4, 256);
    const uint64_t o_tot = align_up(o_base + (a.n_tiles + 1) * 8, 256);
    const uint64_t small = o_tot + FQ_T_N * 8;
    int rc;
    uint8_t *s = (uint8_t *)ctx->fq_small.get(small, err, &rc);
    if (!s) return rc;
    a.tile_counts = (uint32_t *)(s + o_cnt);
    a.tile_base = (uint64_t *)(s + o_base);
    a.totals = (uint64_t *)(s + o_tot);
    hipError_t r = launch_fill_u64(a.totals, 0, FQ_T_N, st);
    if (r == hipSuccess) r = launch_fastq_index(a, st);
    uint64_t n_nl = 0;
    uint8_t last = 0;
    if (r == hipSuccess) r = hipMemcpyAsync(&n_nl, a.tile_base + a.n_tiles, 8, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipMemcpyAsync(&last, a.buf + len - 1, 1, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return hip_err(err, r, "fastq newline index");

    a.n_nl = n_nl;
    // a batch that does not end in a newline ends inside one more line, which
    // the serial walk also reaches
    a.n_lines = n_nl + (last != '\n' ? 1 : 0);
    if (a.n_lines >= (1ull << 32))
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "batch has %llu lines; at most 2^32 - 1 are supported",
                       (unsigned long long)a.n_lines);
    a.n_lines_pad = align_up(a.n_lines + 4, FQ_PAD);
    a.n_seg = a.n_lines_pad / FQ_SEG;
    const uint64_t o_nl = 0;
    const uint64_t o_type = align_up(o_nl + n_nl * 8, 256);
    const uint64_t o_f = align_up(o_type + a.n_lines_pad, 256);
    const uint64_t o_state = align_up(o_f + a.n_seg * 8, 256);
    const uint64_t o_sbase = align_up(o_state + a.n_seg, 256);
    const uint64_t big = o_sbase + a.n_seg * 8;
    uint8_t *g = (uint8_t *)ctx->fq_big.get(big, err, &rc);
    if (!g) return rc;
    a.nl = (uint64_t *)(g + o_nl);
    a.type = g + o_type;
    a.seg_f = (uint64_t *)(g + o_f);
    a.seg_state = g + o_state;
    a.seg_base = (uint64_t *)(g + o_sbase);
    // every record takes four lines
    const uint64_t cap = a.n_lines / 4 + 1;
    frhip_fastq_batch *b = nullptr;
    rc = new_batch(ctx, cap, &b, err);
    if (rc != FRHIP_OK) return rc;
    a.cap = cap;
    a.record_offsets = (uint64_t *)b->cols[FRHIP_FASTQ_RECORD_OFFSETS]->ptr;
    a.sequence_start = (uint32_t *)b->cols[FRHIP_FASTQ_SEQUENCE_START]->ptr;
    a.plus_start = (uint32_t *)b->cols[FRHIP_FASTQ_PLUS_START]->ptr;
    a.quality_start = (uint32_t *)b->cols[FRHIP_FASTQ_QUALITY_START]->ptr;
    a.sequence_len = (uint32_t *)b->cols[FRHIP_FASTQ_SEQUENCE_LEN]->ptr;
    a.record_end = (uint32_t *)b->cols[FRHIP_FASTQ_RECORD_END]->ptr;

    r = launch_fastq_lines(a, st);
    if (r == hipSuccess) r = launch_fastq_emit(a, st);
    uint64_t tot[FQ_T_N];
    if (r == hipSuccess) r = hipMemcpyAsync(tot, a.totals, sizeof tot, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) {
        frhip_fastq_batch_free(b);
        return hip_err(err, r, "fastq boundaries");
    }
    if (tot[FQ_T_STATUS] != 0) {
        const uint64_t pos = tot[FQ_T_TAIL];
        uint8_t ch = 0;
        if (hipMemcpyAsync(&ch, a.buf + pos, 1, hipMemcpyDeviceToHost, st) == hipSuccess)
            (void)hipStreamSynchronize(st);
        frhip_fastq_batch_free(b);
        int k = set_err(err, FRHIP_E_MALFORMED, pos, "expected a record to start with '@', found 0x%02x", ch);
        if (err) std::strcpy(err->format, "fastq");
        return k;
    }
    if (tot[FQ_T_RECORDS] > cap) {
        frhip_fastq_batch_free(b);
        return set_err(err, FRHIP_E_DEVICE, 0, "fastq walk produced %llu records for a capacity of %llu",
                       (unsigned long long)tot[FQ_T_RECORDS], (unsigned long long)cap);
    }
    b->n = tot[FQ_T_RECORDS];
    b->tail = tot[FQ_T_TAIL];
    b->survivors = tot[FQ_T_SURVIVORS];
    for (int i = 0; i < FRHIP_FASTQ_N_COLUMNS; ++i) b->cols[i]->len = b->n * FQ_WIDTHS[i];
    *out = b;
    return FRHIP_OK;
}

extern "C" uint64_t frhip_fastq_batch_len(const frhip_fastq_batch *b) { return b->n; }
extern "C" uint64_t frhip_fastq_batch_tail(const frhip_fastq_batch *b) { return b->tail; }
extern "C" uint64_t frhip_fastq_batch_survivors(const frhip_fastq_batch *b) { return b->survivors; }
extern "C" const frhip_buf *frhip_fastq_batch_column(const frhip_fastq_batch *b, int column)
{
    if (!b || column < 0 || column >= FRHIP_FASTQ_N_COLUMNS) return nullptr;
    return b->cols[column];
}
extern "C" void frhip_fastq_batch_free(frhip_fastq_batch *b)
{
    if (!b) return;
    for (int i = 0; i < FRHIP_FASTQ_N_COLUMNS; ++i) buf_free(b->cols[i]);
    delete b;
}
