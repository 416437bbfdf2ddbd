// deflate_internal.h -- device side of BGZF compression.
#ifndef FRHIP_DEFLATE_INTERNAL_H
#define FRHIP_DEFLATE_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frhip {

// Largest chunk: 65536 - 18 (header) - 8 (trailer) - 5 (a stored block's
// header), so that storing it still frames to at most one 64 KiB block
// (crates/fritillaria-core/src/compress.rs:47-65).
constexpr uint32_t DEFLATE_MAX_PAYLOAD = 65505;
constexpr uint32_t DEFLATE_MAX_STREAM = 65536 - 18 - 8;
constexpr uint32_t DEFLATE_SLOT = 65536; // one framed block at most

// Framing of a block (crates/fritillaria-core/src/compress.rs:84-127).
enum { FRAMING_DEFLATED = 0, FRAMING_STORED = 1, FRAMING_EMPTY = 2 };

struct DeflateArgs {
    const uint8_t *in;       // the payloads
    const uint64_t *bounds;  // n + 1: chunk i is [bounds[i], bounds[i + 1])
    const uint32_t *crc;     // CRC32 of each chunk
    uint32_t n;
    uint32_t *queue;         // work counter (zeroed before the launch)
    uint8_t *slots;          // n x DEFLATE_SLOT: each chunk's framed block
    uint32_t *block_size;    // n: framed bytes
    uint8_t *framing;        // n: FRAMING_*
    // per resident workgroup: match distances (u16 x 65536) and symbols (u32 x 65536)
    uint16_t *dist;
    uint32_t *syms;
};

// Bytes of per-workgroup scratch, and the workgroups a launch keeps resident.
constexpr uint64_t DEFLATE_WG_SCRATCH = 65536ull * 2 + 65536ull * 4;
hipError_t launch_deflate(const DeflateArgs &a, uint32_t workgroups, hipStream_t s);
// out[offsets[i] ..) = slots[i * DEFLATE_SLOT ..) for block_size[i] bytes.
hipError_t launch_gather_blocks(const uint8_t *slots, const uint32_t *block_size, const uint64_t *offsets, uint32_t n,
                                uint8_t *out, hipStream_t s);

} // namespace frhip
#endif
