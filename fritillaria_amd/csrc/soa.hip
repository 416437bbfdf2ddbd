// soa.hip -- structure-of-arrays extras over decoded BAM records: the resolved
// CIGAR ops, the packed 4-bit sequences and an aux tag index.
//
// The reference keeps these behind per-record accessors on the host
// (crates/fritillaria-bam/src/columnar/record.rs:220-361, aux.rs): cigar_raw /
// cigar_resolved (the long-CIGAR placeholder "<l_seq>S <n>N" replaced by the
// CG:B:I array), sequence_packed, and the aux field walk of aux.rs. Here they
// become device columns in three launches, each sized by the record count
// alone so that the batched form captures them in its graph:
//
//   count  one thread per record, in tiles of 256: op and aux-field counts,
//          the CG:B:I source of a placeholder, and each tile's three sums;
//   scan   one workgroup: exclusive prefix of the tile sums and the totals;
//   fill   the same tiles again: a workgroup prefix of its records' counts
//          plus the tile's base gives every offset; each thread copies its
//          record's ops and writes its 16-byte aux entries, then each
//          half-wave copies its wave's packed sequences one record at a time,
//          consecutive lanes on consecutive aligned words (source words
//          funnel-shifted), so a store touches one or two cache lines
//          instead of one per lane.
//
// A record's field bounds come from the decoded columns (cigar_start,
// seq_start, aux_start, record_end, l_seq: coalesced 4-byte loads), not from
// its header again.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "soa_internal.h"

namespace frhip {
namespace {

static_assert(SOA_TILE == 256, "the workgroup scans below assume 4 waves of 64");

// The aligned word at or below p, and the next one only when p is not
// aligned: neither reaches past the last byte of [p, p + 4).
__device__ __forceinline__ uint32_t ldu32(const uint8_t *p)
{
    const uintptr_t a = (uintptr_t)p;
    const uint32_t sh = (uint32_t)(a & 3);
    const uint32_t *w = (const uint32_t *)(a - sh);
    const uint32_t lo = w[0];
    if (!sh) return lo;
    return __builtin_amdgcn_alignbyte(w[1], lo, sh);
}

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
        // the first NUL, four bytes per step while a whole word remains:
        // (w - 0x01010101) & ~w & 0x80808080 has its lowest set bit in the
        // first zero byte (bits above it may be set by the borrow)
        uint64_t q = v;
        for (; q + 4 <= end; q += 4) {
            const uint32_t w = ldu32(r + q);
            const uint32_t z = (w - 0x01010101u) & ~w & 0x80808080u;
            if (z) {
                q += __builtin_ctz(z) >> 3;
                break;
            }
        }
        if (q + 4 > end)
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
        count = ldu32(r + v + 1);
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

// Record::has_long_cigar_placeholder (record.rs:316-327): two ops, the first
// <l_seq>S, the second an N.
__device__ __forceinline__ bool placeholder(const uint8_t *r, uint32_t c0, uint32_t nc, uint32_t ls)
{
    if (nc != 2) return false;
    const uint32_t o0 = ldu32(r + c0), o1 = ldu32(r + c0 + 4);
    return (o0 & 15) == 4 && (o0 >> 4) == ls && (o1 & 15) == 3;
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v, uint32_t lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t u = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += u;
    }
    return v;
}

__device__ __forceinline__ uint64_t record_count(const SoaArgs &a)
{
    const uint64_t n = a.n_dev ? *a.n_dev : a.n;
    return n < a.cap_rec ? n : a.cap_rec;
}

__global__ __launch_bounds__(256) void soa_count_kernel(SoaArgs a)
{
    __shared__ uint64_t red[3][4];
    const uint64_t n = record_count(a);
    const uint64_t tiles = soa_tiles(n);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t i = t * SOA_TILE + threadIdx.x;
        uint32_t ops = 0, sb = 0, k = 0;
        if (i < n) {
            const uint8_t *r = a.buf + a.record_offsets[i];
            const uint32_t c0 = a.cigar_start[i], s0 = a.seq_start[i], a0 = a.aux_start[i], end = a.record_end[i];
            const uint32_t ls = a.l_seq[i];
            const uint32_t nc = (s0 - c0) >> 2;
            uint32_t cg = 0, cg_count = 0, cg_ok = 0, bnd = 0;
            bool ok = true;
            for (uint64_t p = a0; p < end;) {
                if (k >= 1 && k <= 4) {
                    if (p - a0 < 256) bnd |= (uint32_t)(p - a0) << (8 * (k - 1));
                    else bnd = SOA_WALK;
                }
                AuxEntry e;
                const uint64_t nx = aux_field(r, p, end, e);
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
            const bool ph = placeholder(r, c0, nc, ls);
            if (ph && !(cg && cg_ok)) ok = false;
            // a malformed record copies nothing: its counts stay inside the record
            ops = ph ? (ok ? cg_count : 0) : nc;
            sb = (ls + 1) / 2;
            a.n_ops[i] = ops;
            a.n_aux[i] = k;
            a.cg_at[i] = ph && ok ? cg : 0;
            a.aux_bounds[i] = k <= 5 ? bnd : SOA_WALK;
            if (!ok) atomicMin(a.first_bad, (unsigned long long)i);
        }
        const uint64_t s0 = wave_sum(ops), s1 = wave_sum(sb), s2 = wave_sum(k);
        if (lane == 0) {
            red[0][wave] = s0;
            red[1][wave] = s1;
            red[2][wave] = s2;
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            const uint32_t c = threadIdx.x;
            a.tile_sums[3 * t + c] = red[c][0] + red[c][1] + red[c][2] + red[c][3];
        }
        __syncthreads();
    }
}

// Exclusive prefix of the per-tile sums by one workgroup of 1024 threads, four
// tiles per thread per step; the totals and the record overflow go to result.
__global__ __launch_bounds__(1024) void soa_scan_kernel(SoaArgs a)
{
    __shared__ uint64_t wsum[3][16];
    __shared__ uint64_t carry[3];
    const uint64_t n_all = a.n_dev ? *a.n_dev : a.n;
    const uint64_t n = record_count(a);
    const uint64_t tiles = soa_tiles(n);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x < 3) carry[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t base = 0; base < tiles; base += 1024 * 4) {
        const uint64_t t0 = base + (uint64_t)threadIdx.x * 4;
        uint64_t v[4][3], s[3] = {0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                v[j][c] = t0 + j < tiles ? a.tile_sums[3 * (t0 + j) + c] : 0;
                s[c] += v[j][c];
            }
        uint64_t incl[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) incl[c] = wave_incl_scan(s[c], lane);
        if (lane == 63)
#pragma unroll
            for (int c = 0; c < 3; ++c) wsum[c][wave] = incl[c];
        __syncthreads();
        uint64_t run[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            uint64_t pre = carry[c];
            for (uint32_t w = 0; w < wave; ++w) pre += wsum[c][w];
            run[c] = pre + incl[c] - s[c];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (t0 + j < tiles) a.tile_base[3 * (t0 + j) + c] = run[c];
                run[c] += v[j][c];
            }
        __syncthreads();
        if (threadIdx.x == 1023)
#pragma unroll
            for (int c = 0; c < 3; ++c) carry[c] = run[c];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.result[SOA_TOT_OPS] = carry[0];
        a.result[SOA_TOT_SEQ] = carry[1];
        a.result[SOA_TOT_AUX] = carry[2];
        a.result[SOA_OVERFLOW] = n_all > a.cap_rec ? 1 : 0;
    }
}

// dst[0, len) = src[0, len): head bytes up to dst's 4-byte alignment, aligned
// 4-byte stores of funnel-shifted source words, tail bytes. No source word
// read lies past the last source byte's word.
__device__ __forceinline__ void copy_bytes(uint8_t *dst, const uint8_t *src, uint32_t len)
{
    uint32_t head = (uint32_t)((4 - ((uintptr_t)dst & 3)) & 3);
    if (head > len) head = len;
    for (uint32_t j = 0; j < head; ++j) dst[j] = src[j];
    const uint32_t nw = (len - head) >> 2;
    if (nw) {
        const uintptr_t sa = (uintptr_t)(src + head);
        const uint32_t sh = (uint32_t)(sa & 3);
        const uint32_t *sw = (const uint32_t *)(sa - sh);
        uint32_t *dw = (uint32_t *)(dst + head);
        if (sh == 0) {
            for (uint32_t k = 0; k < nw; ++k) dw[k] = sw[k];
        } else {
            uint32_t lo = sw[0];
            for (uint32_t k = 0; k < nw; ++k) {
                const uint32_t hi = sw[k + 1];
                dw[k] = __builtin_amdgcn_alignbyte(hi, lo, sh);
                lo = hi;
            }
        }
    }
    for (uint32_t j = head + 4 * nw; j < len; ++j) dst[j] = src[j];
}

// The 16 bytes at p (any alignment): one 16-byte load from the aligned word
// at or below p, plus the next word when p is unaligned (it holds the last
// wanted byte, so it never crosses a page), funnel-shifted.
typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4))) __attribute__((aligned(4)));
__device__ __forceinline__ uint4 ld16u(const uint8_t *p)
{
    const uintptr_t ad = (uintptr_t)p;
    const uint32_t sh = (uint32_t)(ad & 3) * 8;
    const uint32_t *w = (const uint32_t *)(ad & ~(uintptr_t)3);
    const u32x4_a4 v = *reinterpret_cast<const u32x4_a4 *>(w);
    const uint32_t w4 = sh ? w[4] : 0u;
    uint4 o;
    o.x = __builtin_amdgcn_alignbit(v.y, v.x, sh);
    o.y = __builtin_amdgcn_alignbit(v.z, v.y, sh);
    o.z = __builtin_amdgcn_alignbit(v.w, v.z, sh);
    o.w = __builtin_amdgcn_alignbit(w4, v.w, sh);
    return o;
}

// Bytes j = at .. at + 3 of a chunk: from a where j < k, else from b.
__device__ __forceinline__ uint32_t merge4(uint32_t a, uint32_t b, uint32_t k, uint32_t at)
{
    const uint32_t n = k > at ? k - at : 0; // bytes of this word from a
    const uint32_t m = n >= 4 ? ~0u : (1u << (8 * n)) - 1;
    return (a & m) | (b & ~m);
}

__device__ __forceinline__ void store_entry(AuxEntry *at, const AuxEntry &e)
{
    uint4 v;
    v.x = (uint32_t)e.tag[0] | (uint32_t)e.tag[1] << 8 | (uint32_t)e.type << 16 | (uint32_t)e.subtype << 24;
    v.y = e.value_offset;
    v.z = e.count;
    v.w = e.size;
    *(uint4 *)at = v;
}

__global__ __launch_bounds__(256) void soa_fill_kernel(SoaArgs a)
{
    __shared__ uint64_t wsum[3][4];
    __shared__ uint64_t s_rel[SOA_TILE + 1]; // each record's sequence offset in the tile's stretch, and its end
    __shared__ uint64_t s_src[SOA_TILE];     // each record's packed sequence
    const uint64_t n_all = a.n_dev ? *a.n_dev : a.n;
    if (n_all > a.cap_rec) return;
    const uint64_t n = n_all;
    const uint64_t *R = a.result;
    const bool fits = R[SOA_TOT_OPS] <= a.cap_ops && R[SOA_TOT_SEQ] <= a.cap_seq && R[SOA_TOT_AUX] <= a.cap_aux;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.op_off[n] = R[SOA_TOT_OPS];
        a.seq_off[n] = R[SOA_TOT_SEQ];
        a.aux_off[n] = R[SOA_TOT_AUX];
    }
    const uint64_t tiles = soa_tiles(n);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t i = t * SOA_TILE + threadIdx.x;
        const bool live = i < n;
        uint32_t ops = 0, sb = 0, k = 0;
        if (live) {
            ops = a.n_ops[i];
            sb = (a.l_seq[i] + 1) / 2;
            k = a.n_aux[i];
        }
        const uint64_t i0 = wave_incl_scan(ops, lane), i1 = wave_incl_scan(sb, lane), i2 = wave_incl_scan(k, lane);
        if (lane == 63) {
            wsum[0][wave] = i0;
            wsum[1][wave] = i1;
            wsum[2][wave] = i2;
        }
        __syncthreads();
        uint64_t o0 = a.tile_base[3 * t] + i0 - ops, o1 = a.tile_base[3 * t + 1] + i1 - sb,
                 o2 = a.tile_base[3 * t + 2] + i2 - k;
        for (uint32_t w = 0; w < wave; ++w) {
            o0 += wsum[0][w];
            o1 += wsum[1][w];
            o2 += wsum[2][w];
        }
        __syncthreads();
        uint64_t seq_src = 0;
        if (live) {
            a.op_off[i] = o0;
            a.seq_off[i] = o1;
            a.aux_off[i] = o2;
        }
        if (live && fits) {
            const uint8_t *r = a.buf + a.record_offsets[i];
            // CIGAR: the inline ops, or the CG:B:I array the count pass found
            const uint32_t cg = a.cg_at[i];
            const uint32_t src = cg ? cg : a.cigar_start[i];
            copy_bytes((uint8_t *)(a.ops + o0), r + src, 4 * ops);
            const uint32_t end = a.record_end[i], a0 = a.aux_start[i], bnd = a.aux_bounds[i];
            AuxEntry *aux = a.aux + o2;
            if (bnd != SOA_WALK) {
                // field f spans [a0 + b(f), a0 + b(f + 1)) with b(0) = 0 and
                // the last field ending at the record end: each head (tag,
                // type, first value byte) is one word, B counts another
                uint32_t st[5], hd[5], bc[5];
#pragma unroll
                for (uint32_t f = 0; f < 5; ++f) {
                    st[f] = f == 0 ? a0 : a0 + ((bnd >> (8 * (f - 1))) & 0xFF);
                    if (f < k) hd[f] = ldu32(r + st[f]);
                }
#pragma unroll
                for (uint32_t f = 0; f < 5; ++f)
                    if (f < k && (hd[f] >> 16 & 0xFF) == 'B') bc[f] = ldu32(r + st[f] + 4);
#pragma unroll
                for (uint32_t f = 0; f < 5; ++f) {
                    if (f >= k) break;
                    const uint32_t fe = f + 1 < k ? st[f + 1] : end;
                    const uint32_t ty = hd[f] >> 16 & 0xFF;
                    uint4 v;
                    if (ty == 'B') {
                        v.x = hd[f];
                        v.y = st[f] + 8;
                        v.z = bc[f];
                    } else {
                        v.x = hd[f] & 0x00FFFFFFu;
                        v.y = st[f] + 3;
                        v.z = ty == 'Z' || ty == 'H' ? fe - st[f] - 4 : 1;
                    }
                    v.w = fe - v.y;
                    *(uint4 *)(aux + f) = v;
                }
            } else {
                uint32_t m = 0;
                for (uint64_t p = a0; p < end && m < k;) {
                    AuxEntry e;
                    const uint64_t nx = aux_field(r, p, end, e);
                    if (!nx) break;
                    store_entry(aux + m++, e);
                    p = nx;
                }
            }
            seq_src = (uint64_t)(uintptr_t)(r + a.seq_start[i]);
        }
        // Packed sequences: the tile's records' sequences are one stretch
        // [T0, T1) of the column. The workgroup writes it in 16-byte chunks
        // aligned in memory (coalesced stores), each lane finding its chunk's
        // record by a binary search over the tile's offsets in LDS: a chunk
        // inside one record is one unaligned 16-byte load, a chunk across one
        // record boundary two loads merged by byte, anything else (the tile's
        // ends, records shorter than a chunk) byte by byte.
        if (live) {
            s_rel[threadIdx.x] = o1 - a.tile_base[3 * t + 1];
            s_src[threadIdx.x] = seq_src;
        }
        const uint32_t nt = n - t * SOA_TILE < SOA_TILE ? (uint32_t)(n - t * SOA_TILE) : SOA_TILE;
        if (live && threadIdx.x == nt - 1) s_rel[nt] = o1 + sb - a.tile_base[3 * t + 1];
        __syncthreads();
        if (fits) {
            const uint64_t T0 = a.tile_base[3 * t + 1], T1 = T0 + s_rel[nt];
            const uint32_t mis = (uint32_t)((uintptr_t)a.seq & 15);
            // chunk g holds column bytes [16 g - mis, 16 g - mis + 16)
            const uint64_t g0 = (T0 + mis) / 16, g1 = (T1 + mis + 15) / 16;
            for (uint64_t g = g0 + threadIdx.x; g < g1; g += SOA_TILE) {
                const uint64_t c0 = 16 * g - mis; // wraps for the column's first chunk: clamped below
                const uint64_t lo = g == 0 || c0 < T0 ? T0 : c0;
                const uint64_t hi = c0 + 16 < T1 ? c0 + 16 : T1;
                const uint64_t xlo = lo - T0, xhi = hi - T0;
                if (xlo >= xhi) continue; // a tile whose records hold no sequence
                // r: the last record whose sequence starts at or before xlo
                uint32_t lo_i = 0, hi_i = nt; // s_rel[lo_i] <= xlo < s_rel[hi_i]
                while (hi_i - lo_i > 1) {
                    const uint32_t mid = (lo_i + hi_i) / 2;
                    if (s_rel[mid] <= xlo) lo_i = mid;
                    else hi_i = mid;
                }
                uint32_t r = lo_i;
                while (s_rel[r + 1] <= xlo) ++r; // past zero-length records
                const bool full = g != 0 && c0 >= T0 && hi == c0 + 16;
                const uint8_t *src_r = (const uint8_t *)(uintptr_t)s_src[r] + (xlo - s_rel[r]);
                if (full && s_rel[r + 1] >= xhi) {
                    *reinterpret_cast<uint4 *>(a.seq + lo) = ld16u(src_r);
                } else if (full && r + 1 < nt && s_rel[r + 1] < xhi && s_rel[r + 2] >= xhi) {
                    const uint32_t k = (uint32_t)(s_rel[r + 1] - xlo); // bytes of record r, 1..15
                    const uint4 v1 = ld16u(src_r);
                    const uint4 v2 = ld16u((const uint8_t *)(uintptr_t)s_src[r + 1] - k);
                    uint4 v;
                    v.x = merge4(v1.x, v2.x, k, 0);
                    v.y = merge4(v1.y, v2.y, k, 4);
                    v.z = merge4(v1.z, v2.z, k, 8);
                    v.w = merge4(v1.w, v2.w, k, 12);
                    *reinterpret_cast<uint4 *>(a.seq + lo) = v;
                } else {
                    for (uint64_t x = xlo; x < xhi; ++x) {
                        while (s_rel[r + 1] <= x) ++r;
                        a.seq[T0 + x] = ((const uint8_t *)(uintptr_t)s_src[r])[x - s_rel[r]];
                    }
                }
            }
        }
        __syncthreads(); // before the next tile's offsets replace these
    }
}

uint32_t grid_for_tiles(uint64_t n)
{
    uint64_t g = soa_tiles(n);
    if (g > 8192) g = 8192;
    return g ? (uint32_t)g : 1;
}

} // namespace

hipError_t launch_soa_count(const SoaArgs &a, hipStream_t s)
{
    const uint64_t n = a.n < a.cap_rec ? a.n : a.cap_rec;
    if (n == 0 && !a.n_dev) return hipSuccess;
    hipLaunchKernelGGL(soa_count_kernel, dim3(grid_for_tiles(n)), dim3(SOA_TILE), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_soa_scan(const SoaArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(soa_scan_kernel, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_soa_fill(const SoaArgs &a, hipStream_t s)
{
    const uint64_t n = a.n < a.cap_rec ? a.n : a.cap_rec;
    hipLaunchKernelGGL(soa_fill_kernel, dim3(grid_for_tiles(n)), dim3(SOA_TILE), 0, s, a);
    return hipGetLastError();
}

} // namespace frhip
