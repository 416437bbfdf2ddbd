// bcf_host.cpp -- BcfScanner::decode behind the C ABI.
//
// Follows crates/fritillaria-cuda/src/bcf.rs:451-470 at the boundary: an empty
// batch gives no records and tail = start; otherwise record boundaries, the
// site-core columns and the typed-value bounds of
// crates/fritillaria-bcf/src/columnar/device.rs:59-98 stay on the device.
// Boundaries are the serial scan of crates/fritillaria-bcf/src/columnar/record.rs
// (scan_records), found by the same member-window speculation and one-wave
// reconcile as BAM with BCF's record step; every byte offset of a member is a
// candidate (BCF writers pack members full, so member starts are almost never
// record starts), filtered by looks_like_a_record with the header's sample and
// contig counts. One small read comes back to size the columns, one at the end.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "../../include/frhip.h"
#include "bam_layout.h"
#include "frhip_internal.h"
#include "host_internal.h"
#include "pipeline_internal.h"

using namespace frhip;

static const int BCF_WIDTHS[FRHIP_BCF_N_COLUMNS] = {4, 4, 4, 4, 2, 2, 4, 1, 8, 4, 4, 4, 4, 4};

struct frhip_bcf_batch {
    uint64_t n = 0;
    uint64_t tail = 0;
    uint64_t fallbacks = 0;
    frhip_buf *cols[FRHIP_BCF_N_COLUMNS] = {};
};

namespace {

int new_batch(frhip_ctx *ctx, uint64_t n, frhip_bcf_batch **out, frhip_error *err)
{
    frhip_bcf_batch *b = new (std::nothrow) frhip_bcf_batch;
    if (!b) return set_err(err, FRHIP_E_DEVICE, 0, "host allocation failed");
    b->n = n;
    for (int i = 0; i < FRHIP_BCF_N_COLUMNS; ++i) {
        b->cols[i] = buf_alloc(ctx, n * BCF_WIDTHS[i], err);
        if (!b->cols[i]) {
            frhip_bcf_batch_free(b);
            return err ? err->kind : FRHIP_E_DEVICE;
        }
    }
    *out = b;
    return FRHIP_OK;
}

int malformed(frhip_error *err, uint64_t pos, const char *why)
{
    int k = set_err(err, FRHIP_E_MALFORMED, pos, "%s", why);
    if (err) std::strcpy(err->format, "bcf");
    return k;
}

} // namespace

extern "C" int frhip_bcf_decode(frhip_ctx *ctx, const frhip_buf *data, const uint64_t *block_offsets, uint64_t nb,
                                uint64_t start, uint32_t samples, uint32_t contigs, frhip_bcf_batch **out,
                                frhip_error *err)
{
    *out = nullptr;
    if (!data || data->len == 0) {
        int rc = new_batch(ctx, 0, out, err);
        if (rc == FRHIP_OK) (*out)->tail = start;
        return rc;
    }
    if (data->ordinal != ctx->ordinal)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "batch is on device %d but this scanner is on %d",
                       data->ordinal, ctx->ordinal);
    const uint64_t len = data->len;
    if (nb == 0 || block_offsets[0] != 0 || block_offsets[nb] != len)
        return set_err(err, FRHIP_E_MALFORMED, 0, "block starts must span the buffer from 0 to its length");
    for (uint64_t i = 0; i < nb; ++i)
        if (block_offsets[i + 1] < block_offsets[i] || block_offsets[i + 1] - block_offsets[i] > 65536)
            return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "member %llu has an invalid extent",
                           (unsigned long long)i);
    (void)hipSetDevice(ctx->ordinal);
    hipStream_t st = ctx->stream;
    const BamScratchLayout L = bam_layout(nb);
    int rc;
    uint8_t *s = (uint8_t *)ctx->bcf_scratch.get(L.bytes + 256, err, &rc);
    if (!s) return rc;
    unsigned long long *d_bad = (unsigned long long *)(s + L.bytes);
    hipError_t r = hipMemcpyAsync(s + L.o_bs, block_offsets, (nb + 1) * 8, hipMemcpyHostToDevice, st);
    if (r != hipSuccess) return hip_err(err, r, "uploading block starts");

    BamSpecArgs sa{};
    sa.format = CHAIN_BCF;
    sa.samples = samples;
    sa.contigs = contigs;
    sa.buf = (const uint8_t *)data->ptr;
    sa.len = len;
    sa.block_starts = (const uint64_t *)(s + L.o_bs);
    sa.nb = (uint32_t)nb;
    sa.window = 65536;
    sa.spec_ok = s + L.o_ok;
    sa.n_surv = s + L.o_ns;
    sa.surv_pos = (uint64_t *)(s + L.o_spos);
    sa.surv_cnt = (uint32_t *)(s + L.o_scnt);
    sa.land = (uint64_t *)(s + L.o_land);
    sa.exitp = (uint64_t *)(s + L.o_exit);
    sa.cnt_tail = (uint32_t *)(s + L.o_ctail);
    sa.complete = s + L.o_comp;
    sa.rel_offs = (uint16_t *)(s + L.o_rel);
    r = launch_bam_spec(sa, st);
    if (r != hipSuccess) return hip_err(err, r, "launching the BCF candidate scan");
    BamReconcileArgs ra{};
    ra.format = CHAIN_BCF;
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
    r = launch_bam_reconcile(ra, st);
    uint64_t totals[4];
    if (r == hipSuccess) r = hipMemcpyAsync(totals, ra.totals, sizeof totals, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return hip_err(err, r, "BCF boundaries");
    if (totals[2] != 0) return malformed(err, totals[1], "l_shared smaller than the 24-byte site core");
    const uint64_t n = totals[0];
    frhip_bcf_batch *b = nullptr;
    if ((rc = new_batch(ctx, n, &b, err)) != FRHIP_OK) return rc;
    b->tail = totals[1];
    b->fallbacks = totals[3];
    BamEmitArgs ea{};
    ea.format = CHAIN_BCF;
    ea.buf = sa.buf;
    ea.block_starts = sa.block_starts;
    ea.nb = (uint32_t)nb;
    ea.entry = ra.entry;
    ea.count = ra.count;
    ea.first_index = ra.first_index;
    ea.mode = ra.mode;
    ea.surv_cnt = sa.surv_cnt;
    ea.rel_offs = sa.rel_offs;
    ea.record_offsets = (uint64_t *)b->cols[FRHIP_BCF_RECORD_OFFSETS]->ptr;
    r = n ? launch_bam_emit(ea, st) : hipSuccess;
    if (r == hipSuccess) r = launch_fill_u64((uint64_t *)d_bad, ~0ull, 1, st);
    BcfDecodeArgs da{};
    da.buf = sa.buf;
    da.n = n;
    da.tail = b->tail;
    da.c.chrom = (int32_t *)b->cols[0]->ptr;
    da.c.pos = (int32_t *)b->cols[1]->ptr;
    da.c.rlen = (int32_t *)b->cols[2]->ptr;
    da.c.qual = (uint32_t *)b->cols[3]->ptr;
    da.c.n_info = (uint16_t *)b->cols[4]->ptr;
    da.c.n_allele = (uint16_t *)b->cols[5]->ptr;
    da.c.n_sample = (uint32_t *)b->cols[6]->ptr;
    da.c.n_fmt = (uint8_t *)b->cols[7]->ptr;
    da.c.record_offsets = (uint64_t *)b->cols[8]->ptr;
    da.c.alleles_start = (uint32_t *)b->cols[9]->ptr;
    da.c.filters_start = (uint32_t *)b->cols[10]->ptr;
    da.c.info_start = (uint32_t *)b->cols[11]->ptr;
    da.c.genotypes_start = (uint32_t *)b->cols[12]->ptr;
    da.c.record_end = (uint32_t *)b->cols[13]->ptr;
    da.first_bad = d_bad;
    if (r == hipSuccess) r = launch_bcf_decode(da, st);
    unsigned long long bad = ~0ull;
    if (r == hipSuccess) r = hipMemcpyAsync(&bad, d_bad, 8, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) {
        frhip_bcf_batch_free(b);
        return hip_err(err, r, "BCF decode");
    }
    if (bad != ~0ull) {
        uint64_t at = 0;
        (void)hipMemcpy(&at, (uint64_t *)b->cols[8]->ptr + bad, 8, hipMemcpyDeviceToHost);
        frhip_bcf_batch_free(b);
        return malformed(err, at, "ID, an allele or FILTER is not a valid typed value");
    }
    *out = b;
    return FRHIP_OK;
}

extern "C" uint64_t frhip_bcf_batch_len(const frhip_bcf_batch *b) { return b->n; }
extern "C" uint64_t frhip_bcf_batch_tail(const frhip_bcf_batch *b) { return b->tail; }
extern "C" uint64_t frhip_bcf_batch_fallbacks(const frhip_bcf_batch *b) { return b->fallbacks; }
extern "C" const frhip_buf *frhip_bcf_batch_column(const frhip_bcf_batch *b, int column)
{
    if (!b || column < 0 || column >= FRHIP_BCF_N_COLUMNS) return nullptr;
    return b->cols[column];
}
extern "C" void frhip_bcf_batch_free(frhip_bcf_batch *b)
{
    if (!b) return;
    for (auto *c : b->cols) buf_free(c);
    delete b;
}

// crates/fritillaria-bcf/src/columnar/header.rs:43-140, the counts the scanner needs.
extern "C" int frhip_bcf_parse_header(const uint8_t *buf, uint64_t len, uint64_t *records_start, uint32_t *samples,
                                      uint32_t *contigs, frhip_error *err)
{
    auto mal = [&](uint64_t pos, const char *why) {
        int k = set_err(err, FRHIP_E_MALFORMED, pos, "%s", why);
        if (err) std::strcpy(err->format, "bcf header");
        return k;
    };
    if (len < 9) return mal(0, "truncated header");
    if (std::memcmp(buf, "BCF", 3) != 0) return mal(0, "bad magic; expected BCF");
    if (buf[3] != 2) return mal(3, "unsupported major version; expected 2");
    const uint64_t l_text = (uint64_t)buf[5] | ((uint64_t)buf[6] << 8) | ((uint64_t)buf[7] << 16) |
                            ((uint64_t)buf[8] << 24);
    if (9 + l_text > len) return mal(5, "l_text runs past the end of the buffer");
    const uint8_t *t = buf + 9, *e = buf + 9 + l_text;
    while (e > t && e[-1] == 0) --e;
    uint32_t n_contig = 0, n_sample = 0;
    for (const uint8_t *line = t; line < e;) {
        const uint8_t *nl = (const uint8_t *)std::memchr(line, '\n', (size_t)(e - line));
        const uint8_t *end = nl ? nl : e;
        uint64_t ll = (uint64_t)(end - line);
        if (ll && line[ll - 1] == '\r') --ll;
        if (ll >= 6 && std::memcmp(line, "#CHROM", 6) == 0) {
            uint32_t tabs = 0;
            for (uint64_t i = 0; i < ll; ++i) tabs += line[i] == '\t';
            n_sample = tabs >= 9 ? tabs - 8 : 0;
        } else if (ll >= 10 && std::memcmp(line, "##contig=<", 10) == 0) {
            for (uint64_t i = 9; i + 3 < ll; ++i)
                if ((line[i] == '<' || line[i] == ',') && std::memcmp(line + i + 1, "ID=", 3) == 0) {
                    ++n_contig;
                    break;
                }
        }
        line = end + 1;
    }
    *records_start = 9 + l_text;
    *samples = n_sample;
    *contigs = n_contig;
    return FRHIP_OK;
}
