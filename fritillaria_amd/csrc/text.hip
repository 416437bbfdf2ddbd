// text.hip -- line and field framing of tab-delimited text (SAM, VCF, BED,
// GFF, GTF) on the device.
//
// The decode is crates/fritillaria-text/src/columnar/batch.rs:56-86 and
// :236-258: lines from `start`; a line starts at `start` or after a newline,
// blank lines are skipped; a line without a newline after it ends the scan and
// is the tail; a line whose first byte is the dialect's comment (or secondary)
// byte is a header, any other a record; a record ends at its newline
// (excluded), and its tabs, in order, are the field table, with field_starts[r]
// the number of record tabs before record r.
//
// The reference's kernels (crates/fritillaria-cuda/kernels/text_scan.cu)
// append line starts with an atomic counter, sort them on the host, then walk
// each record with one thread. Here every pass is over bytes in tiles, so
// the work does not depend on line length (a 2,504-sample VCF line is 20 KB,
// a BED line 30 bytes) and nothing returns to the host but five totals:
//   1. count   per tile: record / header line starts, the tabs whose line is
//              known inside the tile, the tabs before the tile's first line
//              start, and the kind of its last line start;
//   2. tiles   one workgroup: the kind of the line open at each tile's start
//              (the last line start before it), and exclusive prefixes of
//              records, headers and record tabs;
//   3. write   per tile: each line start, line end and record tab at its rank.
// Traffic: the batch is read twice; 8 bytes are written per record offset,
// record end, header offset, tab and field start.
#include <hip/hip_runtime.h>

#include "text_internal.h"

namespace frhip {
namespace {

constexpr uint32_t THREADS = 256;
constexpr uint32_t PER = TX_TILE / THREADS; // 16 bytes per thread
static_assert(PER == 16, "a thread takes one 16-byte chunk");

__device__ inline uint32_t lane_id() { return threadIdx.x & 63; }

template <typename T> __device__ inline T wave_incl_sum(T v)
{
    for (int o = 1; o < 64; o <<= 1) {
        T x = __shfl_up(v, o, 64);
        if ((int)lane_id() >= o) v += x;
    }
    return v;
}

template <typename T> __device__ inline T wave_incl_max(T v)
{
    for (int o = 1; o < 64; o <<= 1) {
        T x = __shfl_up(v, o, 64);
        if ((int)lane_id() >= o && x > v) v = x;
    }
    return v;
}

// Exclusive prefix sum over the workgroup (NW waves); *total receives the sum.
template <uint32_t NW, typename T> __device__ inline T block_excl_sum(T v, T *lds, T *total)
{
    const T inc = wave_incl_sum(v);
    const uint32_t w = threadIdx.x / 64;
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    T before = 0, all = 0;
    for (uint32_t k = 0; k < NW; ++k) {
        if (k < w) before += lds[k];
        all += lds[k];
    }
    __syncthreads();
    if (total) *total = all;
    return before + inc - v;
}

// Maximum over the threads before this one (0 when none); *total: over all.
template <uint32_t NW, typename T> __device__ inline T block_excl_max(T v, T *lds, T *total)
{
    const T inc = wave_incl_max(v);
    T ex = __shfl_up(inc, 1, 64);
    if (lane_id() == 0) ex = 0;
    const uint32_t w = threadIdx.x / 64;
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    T before = 0, all = 0;
    for (uint32_t k = 0; k < NW; ++k) {
        if (k < w && lds[k] > before) before = lds[k];
        if (lds[k] > all) all = lds[k];
    }
    __syncthreads();
    if (total) *total = all;
    return ex > before ? ex : before;
}

// One thread's 16 bytes [base, base + 16) and the byte before them.
struct Chunk {
    uint8_t b[PER];
    uint8_t prev;
    uint64_t base;
};

__device__ inline void load_chunk(const uint8_t *buf, uint64_t len, uint64_t base, Chunk &c)
{
    c.base = base;
    c.prev = base > 0 && base - 1 < len ? buf[base - 1] : 0;
    if (base + PER <= len && ((uintptr_t)(buf + base) & 15) == 0) {
        const uint4 v = *(const uint4 *)(buf + base);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) c.b[k] = (uint8_t)(w[k / 4] >> (8 * (k % 4)));
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) c.b[k] = base + k < len ? buf[base + k] : 0;
    }
}

// Per-thread summary of a chunk: line starts by kind, the kind and position of
// the last one, record tabs after the first line start, and tabs before it.
struct Summary {
    uint32_t rec, hdr, tabs_rec, tabs_head;
    uint32_t last_kind;
    uint64_t last_ls;
};

__device__ inline uint32_t kind_of(uint8_t c, uint32_t comment, uint32_t secondary)
{
    return ((uint32_t)c == comment || (secondary < 256u && (uint32_t)c == secondary)) ? TX_HEADER : TX_RECORD;
}

__device__ inline Summary summarize(const TextArgs &a, const Chunk &c)
{
    Summary s{0, 0, 0, 0, TX_NONE, ~0ull};
    uint32_t cur = TX_NONE;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint64_t p = c.base + j;
        if (p < a.start || p >= a.len) continue;
        const uint8_t ch = c.b[j];
        const uint8_t pc = j ? c.b[j - 1] : c.prev;
        if ((p == a.start || pc == '\n') && ch != '\n') {
            cur = kind_of(ch, a.comment, a.secondary);
            if (cur == TX_RECORD) ++s.rec;
            else ++s.hdr;
            s.last_kind = cur;
            s.last_ls = p;
        }
        if (ch == '\t') {
            if (cur == TX_RECORD) ++s.tabs_rec;
            else if (cur == TX_NONE) ++s.tabs_head;
        }
    }
    return s;
}

__global__ __launch_bounds__(THREADS) void text_count_kernel(TextArgs a)
{
    __shared__ uint32_t lds32[THREADS / 64];
    __shared__ uint64_t lds64[THREADS / 64];
    const uint64_t t = blockIdx.x;
    const uint64_t base = (a.first_tile + t) * TX_TILE + (uint64_t)threadIdx.x * PER;
    Chunk c;
    load_chunk(a.buf, a.len, base, c);
    Summary s = summarize(a, c);
    // kind of the line open at this thread's first byte, from earlier threads
    const uint32_t key = s.last_kind ? ((threadIdx.x + 1) << 2 | s.last_kind) : 0u;
    uint32_t kall = 0;
    const uint32_t kin = block_excl_max<THREADS / 64>(key, lds32, &kall) & 3u;
    uint32_t head = 0, rec_tabs = s.tabs_rec;
    if (kin == TX_RECORD) rec_tabs += s.tabs_head;
    else if (kin == TX_NONE) head = s.tabs_head;
    uint32_t tot_rec, tot_hdr, tot_tabs, tot_head;
    block_excl_sum<THREADS / 64>(s.rec, lds32, &tot_rec);
    block_excl_sum<THREADS / 64>(s.hdr, lds32, &tot_hdr);
    block_excl_sum<THREADS / 64>(rec_tabs, lds32, &tot_tabs);
    block_excl_sum<THREADS / 64>(head, lds32, &tot_head);
    const uint64_t ls = s.last_ls == ~0ull ? 0ull : s.last_ls + 1;
    uint64_t ls_all = 0;
    block_excl_max<THREADS / 64>(ls, lds64, &ls_all);
    if (threadIdx.x == 0) {
        a.t_rec[t] = tot_rec;
        a.t_hdr[t] = tot_hdr;
        a.t_tabs_rec[t] = tot_tabs;
        a.t_tabs_head[t] = tot_head;
        a.t_last_kind[t] = (uint8_t)(kall & 3u);
        a.t_last_ls[t] = ls_all ? ls_all - 1 : ~0ull;
    }
}

// The tile prefixes in three passes over groups of 1024 tiles (one tile per
// thread): each group's totals, then one workgroup over the groups (the line
// kind open at each group's start, and group bases), then each tile's kind
// carried in and its record, header and record-tab bases.
constexpr uint32_t GROUP = 1024;

// A group's view of its tiles: the kind carried in from earlier tiles of the
// group (TX_NONE before its first line start) and each tile's record tabs
// when that kind is known.
struct GroupTile {
    uint32_t carry;           // in-group carry key (0: none)
    uint64_t rec, hdr, known, pending, lsp;
};

__device__ inline GroupTile group_tile(const TextArgs &a, uint64_t i, uint32_t *l32, uint32_t *kall)
{
    GroupTile g{};
    const bool in = i < a.n_tiles;
    const uint32_t lk = in ? a.t_last_kind[i] : 0u;
    const uint32_t key = lk ? ((threadIdx.x + 1) << 2 | lk) : 0u;
    g.carry = block_excl_max<GROUP / 64>(key, l32, kall);
    if (in) {
        g.rec = a.t_rec[i];
        g.hdr = a.t_hdr[i];
        const uint32_t k = g.carry & 3u;
        g.known = a.t_tabs_rec[i] + (k == TX_RECORD ? a.t_tabs_head[i] : 0u);
        g.pending = k == TX_NONE ? a.t_tabs_head[i] : 0u;
        g.lsp = a.t_last_ls[i] == ~0ull ? 0ull : a.t_last_ls[i] + 1;
    }
    return g;
}

__global__ __launch_bounds__(GROUP) void text_groups_reduce_kernel(TextArgs a)
{
    __shared__ uint64_t l64[GROUP / 64];
    __shared__ uint32_t l32[GROUP / 64];
    const uint64_t ng = (a.n_tiles + GROUP - 1) / GROUP;
    for (uint64_t g = blockIdx.x; g < ng; g += gridDim.x) {
        uint32_t kall = 0;
        const GroupTile t = group_tile(a, g * GROUP + threadIdx.x, l32, &kall);
        uint64_t s_rec, s_hdr, s_known, s_pending, s_ls;
        block_excl_sum<GROUP / 64>(t.rec, l64, &s_rec);
        block_excl_sum<GROUP / 64>(t.hdr, l64, &s_hdr);
        block_excl_sum<GROUP / 64>(t.known, l64, &s_known);
        block_excl_sum<GROUP / 64>(t.pending, l64, &s_pending);
        block_excl_max<GROUP / 64>(t.lsp, l64, &s_ls);
        if (threadIdx.x == 0) {
            a.g_rec[g] = s_rec;
            a.g_hdr[g] = s_hdr;
            a.g_known[g] = s_known;
            a.g_pending[g] = s_pending;
            a.g_last_kind[g] = kall & 3u;
            a.g_last_ls[g] = s_ls;
        }
    }
}

// One workgroup over the groups: the kind carried into each, its bases, the
// batch totals and the last line start.
__global__ __launch_bounds__(1024) void text_groups_scan_kernel(TextArgs a)
{
    __shared__ uint64_t l64[16];
    __shared__ uint32_t l32[16];
    const uint64_t ng = (a.n_tiles + GROUP - 1) / GROUP;
    uint64_t run_rec = 0, run_hdr = 0, run_tabs = 0, run_ls = 0;
    uint32_t run_kind = TX_NONE;
    for (uint64_t c0 = 0; c0 < ng; c0 += 1024) {
        const uint64_t g = c0 + threadIdx.x;
        const bool in = g < ng;
        const uint32_t lk = in ? a.g_last_kind[g] : 0u;
        const uint32_t key = lk ? ((threadIdx.x + 1) << 2 | lk) : 0u;
        uint32_t kall = 0;
        const uint32_t kex = block_excl_max<16>(key, l32, &kall);
        const uint32_t carry = kex ? (kex & 3u) : run_kind;
        uint64_t rec = 0, hdr = 0, tabs = 0, lsp = 0;
        if (in) {
            rec = a.g_rec[g];
            hdr = a.g_hdr[g];
            tabs = a.g_known[g] + (carry == TX_RECORD ? a.g_pending[g] : 0ull);
            lsp = a.g_last_ls[g];
        }
        uint64_t s_rec, s_hdr, s_tabs, s_ls;
        const uint64_t e_rec = block_excl_sum<16>(rec, l64, &s_rec);
        const uint64_t e_hdr = block_excl_sum<16>(hdr, l64, &s_hdr);
        const uint64_t e_tabs = block_excl_sum<16>(tabs, l64, &s_tabs);
        block_excl_max<16>(lsp, l64, &s_ls);
        if (in) {
            a.g_carry[g] = carry;
            a.g_rec[g] = run_rec + e_rec;    // totals become bases in place
            a.g_hdr[g] = run_hdr + e_hdr;
            a.g_known[g] = run_tabs + e_tabs;
        }
        run_rec += s_rec;
        run_hdr += s_hdr;
        run_tabs += s_tabs;
        if (s_ls > run_ls) run_ls = s_ls;
        if (kall) run_kind = kall & 3u;
    }
    if (threadIdx.x == 0) {
        a.rec_base[a.n_tiles] = run_rec;
        a.hdr_base[a.n_tiles] = run_hdr;
        a.tab_base[a.n_tiles] = run_tabs;
        a.totals[TX_T_RECORDS] = run_rec;
        a.totals[TX_T_HEADERS] = run_hdr;
        a.totals[TX_T_TABS] = run_tabs;
        a.totals[TX_T_LAST_LS] = run_ls ? run_ls - 1 : ~0ull;
        a.totals[TX_T_LAST_KIND] = run_kind;
    }
}

__global__ __launch_bounds__(GROUP) void text_groups_down_kernel(TextArgs a)
{
    __shared__ uint64_t l64[GROUP / 64];
    __shared__ uint32_t l32[GROUP / 64];
    const uint64_t ng = (a.n_tiles + GROUP - 1) / GROUP;
    for (uint64_t g = blockIdx.x; g < ng; g += gridDim.x) {
        const uint64_t i = g * GROUP + threadIdx.x;
        uint32_t kall = 0;
        const GroupTile t = group_tile(a, i, l32, &kall);
        const uint32_t carry = t.carry ? (t.carry & 3u) : a.g_carry[g];
        const uint64_t tabs = t.known + (carry == TX_RECORD ? t.pending : 0ull);
        uint64_t tot;
        const uint64_t e_rec = block_excl_sum<GROUP / 64>(t.rec, l64, &tot);
        const uint64_t e_hdr = block_excl_sum<GROUP / 64>(t.hdr, l64, &tot);
        const uint64_t e_tabs = block_excl_sum<GROUP / 64>(tabs, l64, &tot);
        if (i < a.n_tiles) {
            a.t_carry[i] = (uint8_t)carry;
            a.rec_base[i] = a.g_rec[g] + e_rec;
            a.hdr_base[i] = a.g_hdr[g] + e_hdr;
            a.tab_base[i] = a.g_known[g] + e_tabs;
        }
    }
}

__global__ __launch_bounds__(THREADS) void text_write_kernel(TextArgs a)
{
    __shared__ uint32_t lds32[THREADS / 64];
    const uint64_t t = blockIdx.x;
    const uint64_t base = (a.first_tile + t) * TX_TILE + (uint64_t)threadIdx.x * PER;
    Chunk c;
    load_chunk(a.buf, a.len, base, c);
    const Summary s = summarize(a, c);
    const uint32_t key = s.last_kind ? ((threadIdx.x + 1) << 2 | s.last_kind) : 0u;
    const uint32_t kin_b = block_excl_max<THREADS / 64, uint32_t>(key, lds32, nullptr);
    const uint32_t kin = kin_b ? (kin_b & 3u) : a.t_carry[t];
    const uint32_t rec_tabs = s.tabs_rec + (kin == TX_RECORD ? s.tabs_head : 0u);
    uint64_t rec = a.rec_base[t] + block_excl_sum<THREADS / 64, uint32_t>(s.rec, lds32, nullptr);
    uint64_t hdr = a.hdr_base[t] + block_excl_sum<THREADS / 64, uint32_t>(s.hdr, lds32, nullptr);
    uint64_t tab = a.tab_base[t] + block_excl_sum<THREADS / 64, uint32_t>(rec_tabs, lds32, nullptr);
    uint32_t cur = kin;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint64_t p = c.base + j;
        if (p < a.start || p >= a.len) continue;
        const uint8_t ch = c.b[j];
        const uint8_t pc = j ? c.b[j - 1] : c.prev;
        if ((p == a.start || pc == '\n') && ch != '\n') {
            cur = kind_of(ch, a.comment, a.secondary);
            if (cur == TX_RECORD) {
                a.record_offsets[rec] = p;
                a.field_starts[rec] = tab;
                ++rec;
            } else {
                a.header_offsets[hdr++] = p;
            }
        } else if (ch == '\n' && p > a.start && pc != '\n' && cur == TX_RECORD) {
            a.record_ends[rec - 1] = p; // the newline of the record line open here
        }
        if (ch == '\t' && cur == TX_RECORD) a.tabs[tab++] = p;
    }
}

} // namespace

hipError_t launch_text_count(const TextArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(text_count_kernel, dim3((uint32_t)a.n_tiles), dim3(THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_text_tiles(const TextArgs &a, hipStream_t s)
{
    uint64_t ng = text_groups(a.n_tiles);
    const uint32_t grid = (uint32_t)(ng < 1024 ? (ng ? ng : 1) : 1024);
    hipLaunchKernelGGL(text_groups_reduce_kernel, dim3(grid), dim3(GROUP), 0, s, a);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(text_groups_scan_kernel, dim3(1), dim3(1024), 0, s, a);
    r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(text_groups_down_kernel, dim3(grid), dim3(GROUP), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_text_write(const TextArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(text_write_kernel, dim3((uint32_t)a.n_tiles), dim3(THREADS), 0, s, a);
    return hipGetLastError();
}

} // namespace frhip
