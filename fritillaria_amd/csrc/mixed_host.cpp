// mixed_host.cpp -- a BCF stream and a bgzipped FASTQ stream decoded in one call.
//
// The reference decodes each format batch by batch through its own scanner
// (crates/fritillaria-cuda/src/bcf.rs:94-143, crates/fritillaria-cuda/src/fastq.rs:94-135),
// each batch inflated by its own codec call (crates/fritillaria-cuda/src/backend.rs:536-582).
// Here both streams' members go on one persistent inflate launch (one work
// queue, uneven members of both formats mixed) into one arena, verified in
// member order as the per-call path verifies; then the BCF boundary scan with
// its site-core columns and the FASTQ boundary scan with its columns run on the
// two parts of the arena. The results are those of frhip_bcf_decode and
// frhip_fastq_decode on each stream inflated alone.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/frhip.h"
#include "host_internal.h"

using namespace frhip;

extern "C" int frhip_bcf_fastq_decode(frhip_ctx *ctx, const uint8_t *batch, uint64_t batch_len,
                                      const frhip_block_span *spans, uint64_t n_bcf, uint64_t n,
                                      frhip_buf **arena, uint64_t *offsets, uint64_t *fastq_base,
                                      uint64_t *bcf_start, uint32_t *samples, uint32_t *contigs,
                                      frhip_bcf_batch **bcf, frhip_fastq_batch **fastq, frhip_error *err)
{
    *arena = nullptr;
    *bcf = nullptr;
    *fastq = nullptr;
    if (n_bcf == 0 || n_bcf > n)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0,
                       "the BCF part needs at least one member and at most the batch's %llu", (unsigned long long)n);
    frhip_buf *a = nullptr;
    int rc = inflate_device(ctx, batch, batch_len, spans, n, &a, offsets, err, nullptr, n_bcf);
    if (rc != FRHIP_OK) return rc;
    const uint64_t bcf_len = offsets[n_bcf];
    const uint64_t base = n_bcf < n ? (bcf_len + 4095) & ~4095ull : bcf_len;
    const uint64_t fq_len = a->len - base;

    // the BCF header: magic, l_text, then the text (crates/fritillaria-bcf/src/columnar/header.rs)
    std::vector<uint8_t> head(bcf_len < 9 ? bcf_len : 9);
    hipError_t r = hipMemcpyAsync(head.data(), a->ptr, head.size(), hipMemcpyDeviceToHost, ctx->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(ctx->stream);
    if (r == hipSuccess && head.size() == 9) {
        const uint64_t l_text = (uint64_t)head[5] | ((uint64_t)head[6] << 8) | ((uint64_t)head[7] << 16) |
                                ((uint64_t)head[8] << 24);
        const uint64_t want = 9 + l_text < bcf_len ? 9 + l_text : bcf_len;
        head.resize(want);
        r = hipMemcpyAsync(head.data(), a->ptr, want, hipMemcpyDeviceToHost, ctx->stream);
        if (r == hipSuccess) r = hipStreamSynchronize(ctx->stream);
    }
    if (r != hipSuccess) {
        buf_free(a);
        return hip_err(err, r, "reading the BCF header");
    }
    rc = frhip_bcf_parse_header(head.data(), head.size(), bcf_start, samples, contigs, err);
    if (rc != FRHIP_OK) {
        buf_free(a);
        return rc;
    }

    frhip_buf bview{a->ptr, bcf_len, a->ordinal};
    rc = frhip_bcf_decode(ctx, &bview, offsets, n_bcf, *bcf_start, *samples, *contigs, bcf, err);
    if (rc != FRHIP_OK) {
        buf_free(a);
        return rc;
    }
    frhip_buf fview{(uint8_t *)a->ptr + base, fq_len, a->ordinal};
    rc = frhip_fastq_decode(ctx, &fview, 0, fastq, err);
    if (rc != FRHIP_OK) {
        frhip_bcf_batch_free(*bcf);
        *bcf = nullptr;
        buf_free(a);
        return rc;
    }
    *fastq_base = base;
    *arena = a;
    return FRHIP_OK;
}
