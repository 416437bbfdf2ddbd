// bam_layout.h -- one device allocation holding every per-member array of the
// BAM boundary scan, carved at 256-byte alignment.
#ifndef FRHIP_BAM_LAYOUT_H
#define FRHIP_BAM_LAYOUT_H
#include <stdint.h>

#include "frhip_internal.h"

namespace frhip {

struct BamScratchLayout {
    uint64_t bytes;
    uint64_t o_bs, o_coff, o_ok, o_ns, o_spos, o_scnt, o_land, o_exit, o_ctail, o_comp, o_rel, o_entry, o_count,
        o_first, o_mode, o_totals, o_fscan, o_fstate;
};

inline uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

inline BamScratchLayout bam_layout(uint64_t nb)
{
    BamScratchLayout L{};
    uint64_t p = 0;
    auto take = [&](uint64_t bytes) { uint64_t o = p; p = align256(p + bytes); return o; };
    L.o_bs = take((nb + 1) * 8);
    L.o_coff = take(nb * 8);
    L.o_ok = take(nb);
    L.o_ns = take(nb);
    L.o_spos = take(nb * BAM_K * 8);
    L.o_scnt = take(nb * BAM_K * 4);
    L.o_land = take(nb * 8);
    L.o_exit = take(nb * 8);
    L.o_ctail = take(nb * 4);
    L.o_comp = take(nb);
    L.o_rel = take(nb * (uint64_t)BAM_CAPB * 2);
    L.o_entry = take(nb * 8);
    L.o_count = take(nb * 4);
    L.o_first = take(nb * 8);
    L.o_mode = take(nb);
    L.o_totals = take(4 * 8);
    L.o_fscan = take((nb + 1) * 8);
    L.o_fstate = take(4 * 4);
    L.bytes = p;
    // the offsets slab takes 16-byte stores: its base must be 16-byte aligned
    // within an allocation that is itself at least 256-byte aligned
    static_assert(BAM_CAPB % 8 == 0, "member slabs of u16 offsets must be a multiple of 16 bytes");
    return L;
}

} // namespace frhip
#endif
