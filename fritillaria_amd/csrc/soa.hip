// soa.hip -- structure-of-arrays extras over decoded BAM records: the resolved
// CIGAR ops, the packed 4-bit sequences and an aux tag index.
//
// The reference keeps these behind per-record accessors on the host
// (crates/fritillaria-bam/src/columnar/record.rs:220-361, aux.rs): cigar_raw /
// cigar_resolved (the long-CIGAR placeholder "<l_seq>S <n>N" replaced by the
// CG:B:I array), sequence_packed, and the aux field walk of aux.rs. Here they
// become device columns: a count pass (one thread per record), three prefix
// sums, and a fill pass that copies ops and sequence bytes and writes one
// 16-byte index entry per aux field.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "soa_internal.h"

namespace frhip {
namespace {

__device__ __forceinline__ uint32_t ld32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ uint32_t ld16(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }

__device__ __forceinline__ bool is_hex(uint8_t c)
{
    return (c >= '0' && c <= '9') || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F');
}

// One aux field at p of [.., end) (aux.rs read_value / read_array): fills e and
// returns the next field's offset, or 0 when the field is malformed.
__device__ uint64_t aux_field(const uint8_t *r, uint64_t p, uint64_t end, AuxEntry &e)
{
    if (p + 3 > end) return 0;
    const uint8_t ty = r[p + 2];
    uint64_t v = p + 3, size, count = 1;
    uint8_t sub = 0;
    switch (ty) {
    case 'A': case 'c': case 'C': size = 1; break;
    case 's': case 'S': size = 2; break;
    case 'i': case 'I': case 'f': size = 4; break;
    case 'Z': case 'H': {
        uint64_t q = v;
        while (q < end && r[q] != 0) ++q;
        if (q >= end) return 0;
        count = q - v;
        size = count + 1;
        if (ty == 'H') {
            if (count & 1) return 0;
            for (uint64_t k = v; k < q; ++k)
                if (!is_hex(r[k])) return 0;
        }
        break;
    }
    case 'B': {
        if (v + 5 > end) return 0;
        sub = r[v];
        const uint32_t w = sub == 'c' || sub == 'C' ? 1 : sub == 's' || sub == 'S' ? 2
                         : sub == 'i' || sub == 'I' || sub == 'f' ? 4 : 0;
        if (!w) return 0;
        count = ld32(r + v + 1);
        v += 5;
        size = count * w;
        break;
    }
    default: return 0;
    }
    if (v + size > end) return 0;
    e.tag[0] = r[p];
    e.tag[1] = r[p + 1];
    e.type = ty;
    e.subtype = sub;
    e.value_offset = (uint32_t)v;
    e.count = (uint32_t)count;
    e.size = (uint32_t)size;
    return v + size;
}

struct Shape {
    uint64_t c0, s0, a0, end;
    uint32_t nc, ls;
};

__device__ __forceinline__ Shape shape(const uint8_t *r)
{
    Shape s;
    s.end = 4 + (uint64_t)ld32(r);
    s.nc = ld16(r + 16);
    s.ls = ld32(r + 20);
    s.c0 = 36 + (uint64_t)r[12];
    s.s0 = s.c0 + 4ull * s.nc;
    s.a0 = s.s0 + (s.ls + 1) / 2 + s.ls;
    return s;
}

// Record::has_long_cigar_placeholder (record.rs:316-327)
__device__ __forceinline__ bool placeholder(const uint8_t *r, const Shape &s)
{
    if (s.nc != 2) return false;
    const uint32_t o0 = ld32(r + s.c0), o1 = ld32(r + s.c0 + 4);
    return (o0 & 15) == 4 && (o0 >> 4) == s.ls && (o1 & 15) == 3;
}

__global__ __launch_bounds__(256) void soa_count_kernel(SoaArgs a)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint8_t *r = a.buf + a.record_offsets[i];
        const Shape s = shape(r);
        uint32_t k = 0, cg = 0, cg_count = 0, cg_ok = 0;
        bool ok = true;
        for (uint64_t p = s.a0; p < s.end;) {
            AuxEntry e;
            const uint64_t nx = aux_field(r, p, s.end, e);
            if (!nx) {
                ok = false;
                break;
            }
            if (!cg && e.tag[0] == 'C' && e.tag[1] == 'G') {
                cg = e.value_offset;
                cg_count = e.count;
                cg_ok = e.type == 'B' && e.subtype == 'I';
            }
            ++k;
            p = nx;
        }
        const bool ph = placeholder(r, s);
        if (ph && (!cg || !cg_ok)) ok = false;
        a.n_aux[i] = k;
        a.n_ops[i] = ph ? cg_count : s.nc;
        a.cg_at[i] = ph ? cg : 0;
        a.seq_bytes[i] = (s.ls + 1) / 2;
        if (!ok) atomicMin(a.first_bad, (unsigned long long)i);
    }
}

__global__ __launch_bounds__(256) void soa_fill_kernel(SoaArgs a)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint8_t *r = a.buf + a.record_offsets[i];
        const Shape s = shape(r);
        // CIGAR: inline ops, or the CG array the count pass found
        const uint64_t src = a.cg_at[i] ? a.cg_at[i] : s.c0;
        uint32_t *ops = a.ops + a.op_off[i];
        const uint32_t nops = (uint32_t)(a.op_off[i + 1] - a.op_off[i]);
        for (uint32_t j = 0; j < nops; ++j) ops[j] = ld32(r + src + 4ull * j);
        uint8_t *seq = a.seq + a.seq_off[i];
        const uint32_t nb = (s.ls + 1) / 2;
        for (uint32_t j = 0; j < nb; ++j) seq[j] = r[s.s0 + j];
        AuxEntry *aux = a.aux + a.aux_off[i];
        uint32_t k = 0;
        for (uint64_t p = s.a0; p < s.end;) {
            AuxEntry e;
            const uint64_t nx = aux_field(r, p, s.end, e);
            if (!nx) break;
            aux[k++] = e;
            p = nx;
        }
    }
}

uint32_t grid_for(uint64_t n)
{
    uint64_t g = (n + 255) / 256;
    if (g > 8192) g = 8192;
    return g ? (uint32_t)g : 1;
}

} // namespace

hipError_t launch_soa_count(const SoaArgs &a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(soa_count_kernel, dim3(grid_for(a.n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_soa_fill(const SoaArgs &a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(soa_fill_kernel, dim3(grid_for(a.n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace frhip
