// soa_internal.h -- device arrays and launchers of the BAM SoA extras.
#ifndef FRHIP_SOA_INTERNAL_H
#define FRHIP_SOA_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frhip {

// One aux field: tag, type byte, B subtype (0 otherwise), value offset from
// the record start, element count (1 for scalars; Z/H: string length without
// the NUL; B: array count) and value bytes (Z/H: with the NUL).
struct AuxEntry {
    uint8_t tag[2];
    uint8_t type;
    uint8_t subtype;
    uint32_t value_offset;
    uint32_t count;
    uint32_t size;
};
static_assert(sizeof(AuxEntry) == 16, "aux index entry");

struct SoaArgs {
    const uint8_t *buf;
    const uint64_t *record_offsets;
    uint64_t n;
    uint32_t *n_ops, *seq_bytes, *n_aux, *cg_at; // count pass, per record
    unsigned long long *first_bad;              // lowest failing record (init ~0)
    const uint64_t *op_off, *seq_off, *aux_off;  // n + 1 each (exclusive prefix, total last)
    uint32_t *ops;
    uint8_t *seq;
    AuxEntry *aux;
};

hipError_t launch_soa_count(const SoaArgs &a, hipStream_t s);
hipError_t launch_soa_fill(const SoaArgs &a, hipStream_t s);

} // namespace frhip
#endif
