// bbx_tile.h -- LDS tiles for the neighbour gathers.
//
// A tile is the kBlock consecutive particles (cell-sorted storage) of one
// workgroup.  Storage is ordered by linear cell id (x fastest), so the
// 27-cell neighbourhoods of a tile's cells [cmin, cmax] are covered by nine
// contiguous storage ranges, one per (dy, dz) row offset:
//     cells [cmin + dy*nx + dz*nx*ny - 1, cmax + dy*nx + dz*nx*ny + 1].
// The bucket kernel merges them into at most nine disjoint intervals,
// records them per tile, stages the candidates in LDS and writes each
// bucket as 16-bit LDS-local indices.  Every later gather of the substep
// re-stages the same intervals (coalesced loads) and reads neighbours from
// LDS.  A tile whose intervals hold more than kTileCap particles keeps
// 32-bit storage indices and reads neighbours from global memory.
#pragma once

#include <cstddef>
#include <type_traits>

#include "bbx_device.h"

namespace bbx {

constexpr int kTileCap = kTileCapHost;  // staged particles per tile
// Kernels that stage two float4 arrays hold at most kTileCap2 particles so
// that two workgroups fit the 160 KB LDS of a CU; a staged tile above that
// reads its neighbours from global memory through tab_global.
constexpr int kTileCap2 = kTileCap2Host;
constexpr int kMaxIvl = 9;

// XCD-aware tile order.  Workgroups go to the chip's eight XCDs in
// round-robin order of blockIdx, and each XCD has its own L2.  Workgroup b
// runs tile start(x) + b / 8 of XCD x = b % 8, so every XCD works through a
// contiguous run of tiles and neighbouring tiles, which stage the same
// neighbour rows, meet in the same L2.  A bijection on [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int tile_of_block() {
    const int b = (int)blockIdx.x, nb = (int)gridDim.x;
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = b % kXcds, k = b / kXcds;
    return x * q + min(x, r) + k;
}
constexpr int kTabInts = kTabIntsHost;  // per-tile table stride in ints

// per-tile table, stored in global memory with the same 39-int layout
// (padded to kTabInts): intervals, staged flag, staged total, the storage
// start, length and LDS offset of each interval, then for each of the nine
// rows (see tab_row) LDS index - storage index inside that row's interval
struct TileTab {
    int n_ivl, staged, total;
    int start[kMaxIvl], len[kMaxIvl], off[kMaxIvl];
    int rdelta[9];
};
constexpr int kTabWords = 3 + 3 * kMaxIvl + 9;
static_assert(kTabWords <= kTabInts, "tile table stride");
static_assert(sizeof(TileTab) == kTabWords * sizeof(int), "TileTab layout");
static_assert(offsetof(TileTab, staged) == 1 * sizeof(int) && offsetof(TileTab, start) == 3 * sizeof(int) &&
                  offsetof(TileTab, len) == (3 + kMaxIvl) * sizeof(int) &&
                  offsetof(TileTab, off) == (3 + 2 * kMaxIvl) * sizeof(int),
              "TileTab word indices (tab_word)");

// the block's threads copy a __shared__ TileTab to global memory
__device__ __forceinline__ void tab_store(int *__restrict__ g, const TileTab &t) {
    const int *src = reinterpret_cast<const int *>(&t);
    if (threadIdx.x < kTabWords) g[threadIdx.x] = src[threadIdx.x];
}

// the block's threads copy the table into a __shared__ TileTab
__device__ __forceinline__ void tab_load_shared(const int *__restrict__ g, TileTab &t) {
    int *dst = reinterpret_cast<int *>(&t);
    if (threadIdx.x < kTabWords) dst[threadIdx.x] = g[threadIdx.x];
}

// LDS position of storage index j (j inside one of the intervals)
__device__ __forceinline__ int tab_local(const TileTab &t, int j) {
    for (int k = 0; k < t.n_ivl; k++)
        if (j >= t.start[k] && j < t.start[k] + t.len[k]) return t.off[k] + (j - t.start[k]);
    return -1;
}

// A staged tile's bucket entries are LDS byte offsets (slot << kSlotShift,
// slot < kTileCap, so they fit 16 bits): a gather adds no shift per entry.
constexpr int kSlotShift = 4;  // 16-byte slots
__device__ __forceinline__ float4 lds_at(const float4 *base, int byte_off) {
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + byte_off);
}

// storage index of LDS-local index jl (jl < t.total)
__device__ __forceinline__ int tab_global(const TileTab &t, int jl) {
    int k = 0;
    while (k + 1 < t.n_ivl && jl >= t.off[k] + t.len[k]) k++;
    return t.start[k] + (jl - t.off[k]);
}

// Storage range [a, b) of row r = (dz+1)*3 + (dy+1) of a tile whose home
// cells span [cmin, cmax]: cells [cmin + dy*nx + dz*nx*ny - 1, cmax + ... + 1]
// clipped to the grid; a = -1 for an empty row.
__device__ __forceinline__ void tab_row(const GridGeom &g, int cmin, int cmax,
                                        const int *__restrict__ start, int r, int &a, int &b) {
    a = -1;
    b = -1;
    if (cmax < 0) return;
    const int nx = g.n[0], nxy = g.n[0] * g.n[1];
    const int dz = r / 3 - 1, dy = r % 3 - 1;
    const int last = (int)g.total - 1;
    long long lo = (long long)cmin + dy * nx + (long long)dz * nxy - 1;
    long long hi = (long long)cmax + dy * nx + (long long)dz * nxy + 1;
    if (lo < 0) lo = 0;
    if (hi > last) hi = last;
    if (lo > hi) return;
    const int ra = start[lo], rb = start[hi + 1];  // start[] has total + 1 entries
    if (rb <= ra) return;
    a = ra;
    b = rb;
}

// Merge the nine row ranges of a tile into its interval table (single
// thread, no memory traffic).  t.rdelta[r] receives, for each row, LDS index
// - storage index of that row's interval.
__device__ __forceinline__ void tab_merge(const int *row_a, const int *row_b, int cap, TileTab &t) {
    int *const rdelta = t.rdelta;
    int s0[kMaxIvl], s1[kMaxIvl], m = 0;
    for (int r = 0; r < 9; r++) {
        rdelta[r] = 0;
        const int a = row_a[r];
        if (a < 0) continue;
        // insertion by start
        int p = m++;
        while (p > 0 && s0[p - 1] > a) {
            s0[p] = s0[p - 1];
            s1[p] = s1[p - 1];
            p--;
        }
        s0[p] = a;
        s1[p] = row_b[r];
    }
    // merge overlapping / touching ranges
    int k = -1;
    for (int q = 0; q < m; q++) {
        if (k >= 0 && s0[q] <= t.start[k] + t.len[k]) {
            int e = max(t.start[k] + t.len[k], s1[q]);
            t.len[k] = e - t.start[k];
        } else {
            k++;
            t.start[k] = s0[q];
            t.len[k] = s1[q] - s0[q];
        }
    }
    t.n_ivl = k + 1;
    int off = 0;
    for (int q = 0; q < t.n_ivl; q++) {
        t.off[q] = off;
        off += t.len[q];
    }
    t.total = off;
    t.staged = off <= cap ? 1 : 0;
    for (int r = 0; r < 9; r++)
        if (row_a[r] >= 0)
            for (int q = 0; q < t.n_ivl; q++)
                if (row_a[r] >= t.start[q] && row_a[r] < t.start[q] + t.len[q]) rdelta[r] = t.off[q] - t.start[q];
}

// tab_merge for the common case where the non-empty rows' ranges come in
// order of their starts (row r's cells begin dy*nx + dz*nx*ny - 1 past the
// tile's first cell, increasing with r when the grid has three or more cells
// along y, and start[] is non-decreasing): one pass in row order merges them
// into the same table as tab_merge.  Returns false (table untouched) when
// the rows are out of order; tab_merge then sorts them.
__device__ __forceinline__ bool tab_merge_ordered(const int *row_a, const int *row_b, int cap, TileTab &t) {
    int prev = -1;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        if (row_a[r] < 0) continue;
        if (row_a[r] < prev) return false;
        prev = row_a[r];
    }
    int k = -1, end = 0, off = 0;
    int ivl[9];
#pragma unroll
    for (int r = 0; r < 9; r++) {
        const int a = row_a[r], b = row_b[r];
        ivl[r] = -1;
        if (a < 0) continue;
        if (k >= 0 && a <= end) {
            end = max(end, b);
        } else {
            if (k >= 0) {
                t.len[k] = end - t.start[k];
                off += t.len[k];
            }
            k++;
            t.start[k] = a;
            t.off[k] = off;
            end = b;
        }
        ivl[r] = k;
    }
    if (k >= 0) {
        t.len[k] = end - t.start[k];
        off += t.len[k];
    }
    t.n_ivl = k + 1;
    t.total = off;
    t.staged = off <= cap ? 1 : 0;
#pragma unroll
    for (int r = 0; r < 9; r++) t.rdelta[r] = ivl[r] >= 0 ? t.off[ivl[r]] - t.start[ivl[r]] : 0;
    return true;
}

// Run f(std::true_type) for an LDS-staged tile and f(std::false_type)
// otherwise, so code inside f that reads neighbours from LDS or from global
// memory is compiled separately and never selects between the two address
// spaces at run time (a select compiles to a slower generic load).
template <typename F>
__device__ __forceinline__ void with_staging(bool staged, F f) {
    if (staged)
        f(std::true_type{});
    else
        f(std::false_type{});
}

// Bucket entries of particle i come a quad (4 entries, one 8 B or 16 B
// load) at a time.  The first D quads can be requested early -- before the
// tile is staged, which they do not depend on -- so their latency overlaps
// the staging; entries past the bucket count are loaded but never used
// (every particle has 25 quads of table).
constexpr int kNbrDepth = 4;  // quads (16 entries) in flight per lane in the hot gathers

// A quad as loaded: a staged tile's four 16-bit entries stay packed in .x and
// .y until they are used (quad_unpack), so nothing reads a loaded register
// right after its load and the loads in flight are never waited for early;
// otherwise four 32-bit entries.
template <int D>
struct NbrQuads {
    int4 q[D];
};

__device__ __forceinline__ int4 nbr_quad(const unsigned short *__restrict__ n16,
                                         const int *__restrict__ n32, bool staged, int i,
                                         int stride, int q) {
    const size_t at = (size_t)q * stride + i;
    if (staged) {
        const uint2 v = reinterpret_cast<const uint2 *>(n16)[at];
        return make_int4((int)v.x, (int)v.y, 0, 0);
    }
    return reinterpret_cast<const int4 *>(n32)[at];
}

__device__ __forceinline__ int4 quad_unpack(const int4 &r, bool staged) {
    if (staged)
        return make_int4(r.x & 0xffff, (int)((unsigned)r.x >> 16), r.y & 0xffff, (int)((unsigned)r.y >> 16));
    return r;
}

// quads q0, q0 + step, ..., q0 + (D - 1) step of particle i (q0 + (D-1) step < 25)
template <int D, int kStep = 1>
__device__ __forceinline__ NbrQuads<D> nbr_prefetch(const unsigned short *__restrict__ n16,
                                                    const int *__restrict__ n32, bool staged,
                                                    int i, int stride, int q0 = 0) {
    NbrQuads<D> p;
#pragma unroll
    for (int d = 0; d < D; d++) p.q[d] = nbr_quad(n16, n32, staged, i, stride, q0 + d * kStep);
    return p;
}

// The four entries of one (unpacked) quad go through body(j, live) without
// branches; live = false for an entry at or past the bucket count.  Every
// slot of a quad the walk reaches holds a valid index: the bucket build
// zero-fills the unused slots of a bucket's last quad, and index 0 is valid
// in either index space.  So by default a dead entry keeps its slot's index
// (no select per entry); with kSelect it gets `dead` instead (the particle's
// own index: a sum over x_j - x_i then gets exact zeros from dead entries
// unmasked).  Straight-line code lets the four entries' neighbour loads
// issue together instead of one load-wait per entry.  The body must make a
// dead entry contribute nothing (or an exact zero).
template <bool kSelect, typename Body>
__device__ __forceinline__ void quad_body(const int4 &v, int k, int cnt, Body &body, int dead) {
    const int rem = cnt - k;  // entries of the bucket from this quad on
    if constexpr (kSelect) {
        body(rem > 0 ? v.x : dead, rem > 0);
        body(rem > 1 ? v.y : dead, rem > 1);
        body(rem > 2 ? v.z : dead, rem > 2);
        body(rem > 3 ? v.w : dead, rem > 3);
    } else {
        body(v.x, rem > 0);
        body(v.y, rem > 1);
        body(v.z, rem > 2);
        body(v.w, rem > 3);
    }
}

// Walk bucket slots 0..cnt-1 of particle i in bucket order, starting from
// the prefetched first D quads and keeping the next D quads in flight while
// the current D are processed (quads 0, step, 2 step, ... from quad `first0`;
// step > 1 splits a bucket between lanes).  body(j, live) gets an LDS-local
// index (staged tile) or a storage index.  The loads of the next quads are
// unconditional: a quad past the bucket's last one reloads the last (a cache
// hit), so no load sits in a branch that would have to wait for it.
template <int D, int kStep = 1, bool kSelect = false, typename Body>
__device__ __forceinline__ void for_each_neighbour_from(const NbrQuads<D> &first,
                                                        const unsigned short *__restrict__ n16,
                                                        const int *__restrict__ n32, bool staged,
                                                        int i, int cnt, int stride, Body body,
                                                        int first0 = 0, int dead = 0) {
    const int nq = (cnt + 3) >> 2;
    const int last = nq > 0 ? nq - 1 : 0;
    int4 cur[D];
#pragma unroll
    for (int d = 0; d < D; d++) cur[d] = first.q[d];
    for (int q = first0; q < nq; q += D * kStep) {
        int4 nxt[D];
#pragma unroll
        for (int d = 0; d < D; d++) nxt[d] = nbr_quad(n16, n32, staged, i, stride, min(q + (D + d) * kStep, last));
#pragma unroll
        for (int d = 0; d < D; d++) {
            const int qc = q + d * kStep;
            if (qc < nq) quad_body<kSelect>(quad_unpack(cur[d], staged), qc * 4, cnt, body, dead);
        }
#pragma unroll
        for (int d = 0; d < D; d++) cur[d] = nxt[d];
    }
}

template <int D, typename Body>
__device__ __forceinline__ void for_each_neighbour(const unsigned short *__restrict__ n16,
                                                   const int *__restrict__ n32, bool staged,
                                                   int i, int cnt, int stride, Body body) {
    const NbrQuads<D> first = nbr_prefetch<D>(n16, n32, staged, i, stride);
    for_each_neighbour_from<D>(first, n16, n32, staged, i, cnt, stride, body);
}

// The same with dead entries reading index `dead`.
template <int D, typename Body>
__device__ __forceinline__ void for_each_neighbour_dead(const unsigned short *__restrict__ n16,
                                                        const int *__restrict__ n32, bool staged,
                                                        int i, int cnt, int stride, Body body, int dead) {
    const NbrQuads<D> first = nbr_prefetch<D>(n16, n32, staged, i, stride);
    for_each_neighbour_from<D, 1, true>(first, n16, n32, staged, i, cnt, stride, body, 0, dead);
}

// LDS DMA (gfx950 global_load_lds): each active lane of the wave copies
// 16 bytes from its own global address to LDS at wave_dst + lane * 16,
// wave_dst being the same for the whole wave.  No VGPR round trip and no
// LDS store instruction; every load of the tile can be in flight at once.
typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ void lds_dma16(const void *src, float4 *wave_dst) {
    __builtin_amdgcn_global_load_lds(src, (lds_void *)wave_dst, 16, 0, 0);
}

// Stage the tile's intervals with LDS DMA: copy(j, e) is called for storage
// index j by lanes whose LDS slot is e + lane, e being the slot of the wave's
// lane 0 (wave-uniform).  Each interval is cut into 64-slot runs, one per
// wave per pass.  The caller then waits (lds_dma_wait) and synchronises.
template <typename Copy>
__device__ __forceinline__ void tab_stage_dma(const TileTab &t, Copy copy) {
    const int lane = (int)threadIdx.x & 63;
    const int w0 = (int)threadIdx.x - lane;
    const int nivl = t.n_ivl;
    for (int k = 0; k < nivl; k++) {
        const int len = t.len[k], s0 = t.start[k], o0 = t.off[k];
        for (int b = w0; b < len; b += (int)blockDim.x)
            if (b + lane < len) copy(s0 + b + lane, o0 + b);
    }
}

// The tile table held in registers, one word per lane (lane w < kTabWords
// holds word w of the TileTab layout), loaded by every wave for itself so
// that no workgroup barrier is needed before the staging.  tab_word reads a
// word by a wave-uniform index.
__device__ __forceinline__ int tab_load_wave(const int *__restrict__ g) {
    const int lane = (int)threadIdx.x & 63;
    return lane < kTabWords ? g[lane] : 0;
}
__device__ __forceinline__ int tab_word(int tw, int w) { return __builtin_amdgcn_readlane(tw, w); }

// tab_stage_dma from a table held in registers (tab_load_wave)
template <typename Copy>
__device__ __forceinline__ void tab_stage_dma_wave(int tw, Copy copy) {
    const int lane = (int)threadIdx.x & 63;
    const int w0 = (int)threadIdx.x - lane;
    const int nivl = tab_word(tw, 0);
    for (int k = 0; k < nivl; k++) {
        const int s0 = tab_word(tw, 3 + k), len = tab_word(tw, 3 + kMaxIvl + k),
                  o0 = tab_word(tw, 3 + 2 * kMaxIvl + k);
        for (int b = w0; b < len; b += (int)blockDim.x)
            if (b + lane < len) copy(s0 + b + lane, o0 + b);
    }
}

// The wave's outstanding loads (LDS DMA included) have landed; a barrier
// must follow before other waves read them.
__device__ __forceinline__ void lds_dma_wait() { __builtin_amdgcn_s_waitcnt(0); }

// Cooperative copy of the tile's intervals into LDS: dst[off + e] = fn(start + e).
// The LDS image is the concatenation of the intervals, so thread t copies
// LDS slots t, t + blockDim, ...; each thread issues kBatch global loads
// before its first LDS store, keeping that many loads in flight instead of
// waiting out one memory latency per staged particle.
// The interval of a slot is found with a fixed chain of compare-selects over
// the (block-uniform) interval offsets held in registers: no search loop and
// no table reads per slot.
// store(e, v) writes the value loaded for LDS slot e (tab_stage below stores
// it whole into one array).
template <typename Fn, typename Store>
__device__ __forceinline__ void tab_stage_with(const TileTab &t, Fn fn, Store store) {
    using T = decltype(fn(0));
    constexpr int kBatch = sizeof(T) > 16 ? 4 : 8;  // registers held per batch
    const int total = t.total, nivl = t.n_ivl;
    int off[kMaxIvl], shift[kMaxIvl];  // LDS offset and storage - LDS shift of each interval
#pragma unroll
    for (int k = 0; k < kMaxIvl; k++) {
        off[k] = k < nivl ? t.off[k] : 0x7fffffff;
        shift[k] = k < nivl ? t.start[k] - t.off[k] : 0;
    }
    for (int base = threadIdx.x; base < total; base += kBatch * blockDim.x) {
        int src[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            const int e = base + b * (int)blockDim.x;
            int d = shift[0];
#pragma unroll
            for (int k = 1; k < kMaxIvl; k++) d = e >= off[k] ? shift[k] : d;
            src[b] = e < total ? e + d : -1;
        }
        T v[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; b++)
            if (src[b] >= 0) v[b] = fn(src[b]);
#pragma unroll
        for (int b = 0; b < kBatch; b++)
            if (src[b] >= 0) store(base + b * (int)blockDim.x, v[b]);
    }
}

template <typename T, typename Fn>
__device__ __forceinline__ void tab_stage(const TileTab &t, T *__restrict__ dst, Fn fn) {
    tab_stage_with(t, fn, [&](int e, const T &v) { dst[e] = v; });
}

}  // namespace bbx
