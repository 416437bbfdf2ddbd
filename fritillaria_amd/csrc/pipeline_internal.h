// pipeline_internal.h -- device arrays and launchers of the batched path.
#ifndef FRHIP_PIPELINE_INTERNAL_H
#define FRHIP_PIPELINE_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frhip {

// Per-member arrays (structure of arrays), sized for the file's member count.
struct PipeMembers {
    uint64_t *coff;       // compressed offset of the member
    uint64_t *in_off;     // payload start (absolute in the file)
    uint32_t *in_len;     // payload length
    uint32_t *isize;      // ISIZE
    uint32_t *crc_expect; // CRC32 from the trailer
    uint64_t *out_off;    // exclusive prefix sum of ISIZE, n + 1 entries
};

uint64_t sieve_tiles(uint64_t len);
hipError_t launch_sieve_count(const uint8_t *f, uint64_t len, uint32_t *counts, hipStream_t s);
hipError_t launch_sieve_write(const uint8_t *f, uint64_t len, const uint64_t *tile_base, uint64_t *cand,
                              uint32_t *cand_bs, uint32_t *cand_hs, uint64_t cap, hipStream_t s);
hipError_t launch_spans(const uint8_t *f, uint64_t len, const uint64_t *cand, const uint32_t *cand_bs,
                        const uint32_t *cand_hs, const uint64_t *n_cand, uint64_t cap, const PipeMembers &m,
                        uint64_t *bad, hipStream_t s);
hipError_t launch_scan_u32_u64(const uint32_t *in, uint64_t *out, const uint64_t *n_dev, uint64_t n_host,
                               hipStream_t s);
hipError_t launch_scan_u64_u64(const uint64_t *in, uint64_t *out, const uint64_t *n_dev, uint64_t n_host,
                               hipStream_t s);
hipError_t launch_verify(const uint32_t *status, const uint32_t *produced, const uint32_t *crc, const PipeMembers &m,
                         const uint64_t *n_dev, uint64_t cap, unsigned long long *first_bad, hipStream_t s);
hipError_t launch_bam_header(const uint8_t *u, const uint64_t *u_len, uint64_t *records_start, uint64_t *status,
                             hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);
// out_off[n] != expect: every member's extent emptied, *bad set (see pipeline.hip)
hipError_t launch_guard_total(uint64_t *out_off, uint32_t *isize, uint64_t n, uint64_t expect, uint64_t *bad,
                              hipStream_t s);
// dst and src 16-byte aligned (device allocations are)
hipError_t launch_copy(uint8_t *dst, const uint8_t *src, uint64_t n, hipStream_t s);

} // namespace frhip
#endif
