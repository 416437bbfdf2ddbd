// noalloc.hip -- the inflate entry point that allocates nothing.
//
// frhip_inflate_batch_device_noalloc is frhip_inflate_batch_device for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc): the caller passes the
// compressed batch and the output already in device memory, and the member
// plan, the per-member status words and the verification result live in the
// file-scope array `scratch` below instead of in library-allocated buffers.
// It runs the same inflate kernels and applies the same verification order
// (status, then ISIZE, then CRC32, first failing member in member order).
//
// Each GPU holds its own copy of `scratch` until the process ends. Its device
// address is taken once per device with hipGetSymbolAddress (at context
// creation, with the device current) and only that address is handed to
// kernels and copies. Every call on one GPU shares the array, so calls that
// overlap on one GPU race; the array keeps the last call's contents, and what
// must start at zero is cleared by a fill kernel each call.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "frhip_internal.h"
#include "host_internal.h"
#include "pipeline_internal.h"

namespace frhip {
namespace {

// Scratch layout for at most FRHIP_NOALLOC_MAX_MEMBERS members.
constexpr uint64_t NA_MAX = FRHIP_NOALLOC_MAX_MEMBERS;
constexpr uint64_t NA_PAD = 4096; // zero bytes the caller leaves after the batch
constexpr uint64_t O_IN_OFF = 0;                     // u64[NA_MAX] payload start
constexpr uint64_t O_OUT_OFF = O_IN_OFF + 8 * NA_MAX; // u64[NA_MAX] output offset
constexpr uint64_t O_IN_LEN = O_OUT_OFF + 8 * NA_MAX; // u32[NA_MAX] payload length
constexpr uint64_t O_ISIZE = O_IN_LEN + 4 * NA_MAX;   // u32[NA_MAX] ISIZE (output capacity)
constexpr uint64_t O_CRC = O_ISIZE + 4 * NA_MAX;      // u32[NA_MAX] trailer CRC32
constexpr uint64_t META_BYTES = O_CRC + 4 * NA_MAX;   // what the host uploads
constexpr uint64_t O_STATUS = META_BYTES;             // u32[3 * NA_MAX] status, produced, crc
constexpr uint64_t O_RESULT = O_STATUS + 12 * NA_MAX; // NaResult
constexpr uint64_t SCRATCH_BYTES = O_RESULT + 256;
static_assert(O_RESULT % 256 == 0, "result block alignment");

struct NaResult {
    unsigned long long first_bad; // first failing member, or ~0
    unsigned long long pad_bad;   // nonzero: a padding byte after the batch is not zero
    uint32_t status, produced, crc, reserved;
};

__device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// One workgroup: the first member failing verification (status, ISIZE, CRC
// against the plan), the status words of that member, and the padding check.
__global__ __launch_bounds__(1024) void na_verify_kernel(uint8_t *s, uint32_t n, const uint8_t *batch,
                                                         uint64_t batch_len)
{
    __shared__ unsigned long long first;
    __shared__ uint32_t pad;
    const uint32_t *st = (const uint32_t *)(s + O_STATUS);
    const uint32_t *isz = (const uint32_t *)(s + O_ISIZE);
    const uint32_t *crc = (const uint32_t *)(s + O_CRC);
    if (threadIdx.x == 0) {
        first = ~0ull;
        pad = 0;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
        if (st[i] != 0 || st[NA_MAX + i] != isz[i] || st[2 * NA_MAX + i] != crc[i]) atomicMin(&first, (unsigned long long)i);
    uint32_t any = 0;
    for (uint32_t j = threadIdx.x; j < NA_PAD / 4; j += blockDim.x) {
        const uint8_t *p = batch + batch_len + 4 * j;
        any |= (uint32_t)p[0] | p[1] | p[2] | p[3];
    }
    if (any) atomicOr(&pad, 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        NaResult *r = (NaResult *)(s + O_RESULT);
        r->first_bad = first;
        r->pad_bad = pad;
        if (first != ~0ull) {
            r->status = st[first];
            r->produced = st[NA_MAX + first];
            r->crc = st[2 * NA_MAX + first];
        }
    }
}

constexpr int NA_DEVICES = 64;
std::mutex na_mu;
void *na_ptr[NA_DEVICES];
hipError_t na_rc[NA_DEVICES];
bool na_done[NA_DEVICES];

// Host staging of the plan: per thread, so calls from threads driving
// different GPUs do not share it.
thread_local uint8_t na_meta[META_BYTES];

} // namespace

// Once per device, with that device current: the device address of this
// GPU's copy of `scratch`.
void noalloc_bind(frhip_ctx *ctx)
{
    const int d = ctx->ordinal;
    if (d < 0 || d >= NA_DEVICES) {
        ctx->noalloc_scratch = nullptr;
        ctx->noalloc_rc = hipErrorInvalidDevice;
        return;
    }
    std::lock_guard<std::mutex> g(na_mu);
    if (!na_done[d]) {
        void *p = nullptr;
        na_rc[d] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        na_ptr[d] = na_rc[d] == hipSuccess ? p : nullptr;
        na_done[d] = true;
    }
    ctx->noalloc_scratch = na_ptr[d];
    ctx->noalloc_rc = na_rc[d];
}

} // namespace frhip

using namespace frhip;

extern "C" int frhip_inflate_batch_device_noalloc(frhip_ctx *ctx, const void *d_batch, uint64_t batch_len,
                                                  const frhip_block_span *spans, uint64_t n, void *d_out,
                                                  uint64_t out_cap, uint64_t *out_offsets, frhip_error *err)
{
    if (!ctx->noalloc_scratch)
        return hip_err(err, ctx->noalloc_rc != hipSuccess ? ctx->noalloc_rc : hipErrorInvalidSymbol,
                       "device address of the no-allocation scratch");
    if (n > NA_MAX)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0,
                       "batch of %llu members exceeds the no-allocation maximum of %llu",
                       (unsigned long long)n, (unsigned long long)NA_MAX);
    (void)hipSetDevice(ctx->ordinal);
    uint8_t *s = (uint8_t *)ctx->noalloc_scratch;
    // BatchPlan::new (crates/fritillaria-cuda/src/backend.rs:95-133) into the host staging
    uint64_t *in_off = (uint64_t *)(na_meta + O_IN_OFF), *out_off = (uint64_t *)(na_meta + O_OUT_OFF);
    uint32_t *in_len = (uint32_t *)(na_meta + O_IN_LEN), *isz = (uint32_t *)(na_meta + O_ISIZE),
             *crc = (uint32_t *)(na_meta + O_CRC);
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const frhip_block_span &sp = spans[i];
        if (sp.isize > 65536) return set_err(err, FRHIP_E_INVALID_BLOCK, sp.compressed_offset, "ISIZE exceeds 64 KiB");
        const uint64_t end = sp.payload_start + sp.payload_len;
        if (end < sp.payload_start || end > batch_len)
            return set_err(err, FRHIP_E_INVALID_BLOCK, sp.compressed_offset, "payload extends past the end of the batch");
        in_off[i] = sp.payload_start;
        out_off[i] = total;
        in_len[i] = (uint32_t)sp.payload_len;
        isz[i] = sp.isize;
        crc[i] = sp.crc32;
        total += sp.isize;
    }
    if (total > out_cap)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "output of %llu bytes does not fit %llu",
                       (unsigned long long)total, (unsigned long long)out_cap);
    out_offsets[0] = 0;
    if (n == 0) return FRHIP_OK;
    if (!d_batch || (total && !d_out))
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, 0, "null device buffer");
    hipStream_t st = ctx->stream;
    // the plan's arrays sit at the same offsets in the staging and the scratch
    hipError_t r = hipSuccess;
    const uint64_t parts[5][2] = {{O_IN_OFF, 8}, {O_OUT_OFF, 8}, {O_IN_LEN, 4}, {O_ISIZE, 4}, {O_CRC, 4}};
    for (const auto &pt : parts)
        if (r == hipSuccess) r = hipMemcpyAsync(s + pt[0], na_meta + pt[0], n * pt[1], hipMemcpyHostToDevice, st);
    if (r != hipSuccess) return hip_err(err, r, "uploading the member plan");
    InflateArgs a{};
    a.in = (const uint8_t *)d_batch;
    a.in_off = (const uint64_t *)(s + O_IN_OFF);
    a.out_off = (const uint64_t *)(s + O_OUT_OFF);
    a.in_len = (const uint32_t *)(s + O_IN_LEN);
    a.out_cap = (const uint32_t *)(s + O_ISIZE);
    a.out = (uint8_t *)d_out;
    a.status = (uint32_t *)(s + O_STATUS);
    a.produced = a.status + NA_MAX;
    a.crc = a.status + 2 * NA_MAX;
    a.n = (uint32_t)n;
    a.queue = nullptr;
    a.order = nullptr;
    a.x2n = ctx->x2n;
    // one wave per member: the two-phase forms keep match records in scratch
    // sized by the output, so a two-phase schedule runs its one-wave form
    inflate_scratch_bind(a, nullptr, n);
    a.two_phase_min = 0xFFFFFFFFu;
    r = launch_inflate(a, (uint32_t)n, st, ctx->inflate_schedule);
    if (r == hipSuccess) {
        hipLaunchKernelGGL(na_verify_kernel, dim3(1), dim3(1024), 0, st, s, (uint32_t)n, (const uint8_t *)d_batch,
                           batch_len);
        r = hipGetLastError();
    }
    NaResult res;
    if (r == hipSuccess) r = hipMemcpyAsync(&res, s + O_RESULT, sizeof res, hipMemcpyDeviceToHost, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (r != hipSuccess) return hip_err(err, r, "inflate");
    if (res.pad_bad)
        return set_err(err, FRHIP_E_INVALID_DEVICE_BATCH, batch_len,
                       "the 4096 bytes after the batch must be zero");
    if (res.first_bad != ~0ull) {
        const frhip_block_span &sp = spans[res.first_bad];
        if (res.status != 0) {
            int k = set_err(err, FRHIP_E_INFLATE, sp.compressed_offset, "%s", inflate_status_message(res.status));
            if (err) err->code = res.status;
            return k;
        }
        if (res.produced != sp.isize) {
            int k = set_err(err, FRHIP_E_SIZE_MISMATCH, sp.compressed_offset, "ISIZE says %u bytes, inflated %u",
                            sp.isize, res.produced);
            if (err) { err->expected = sp.isize; err->actual = res.produced; }
            return k;
        }
        int k = set_err(err, FRHIP_E_CHECKSUM_MISMATCH, sp.compressed_offset, "expected crc32 %#010x, got %#010x",
                        sp.crc32, res.crc);
        if (err) { err->expected = sp.crc32; err->actual = res.crc; }
        return k;
    }
    for (uint64_t i = 0; i < n; ++i) out_offsets[i + 1] = out_off[i] + isz[i];
    return FRHIP_OK;
}
