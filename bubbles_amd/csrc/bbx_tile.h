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

#include "bbx_device.h"

namespace bbx {

constexpr int kTileCap = kTileCapHost;  // staged particles per tile
constexpr int kMaxIvl = 9;
constexpr int kTabInts = kTabIntsHost;  // per-tile table stride in ints

// per-tile table, stored in global memory with the same 30-int layout
// (padded to kTabInts): intervals, staged flag, staged total, then the
// storage start, length and LDS offset of each interval
struct TileTab {
    int n_ivl, staged, total;
    int start[kMaxIvl], len[kMaxIvl], off[kMaxIvl];
};
constexpr int kTabWords = 3 + 3 * kMaxIvl;
static_assert(sizeof(TileTab) == kTabWords * sizeof(int), "TileTab layout");

__device__ __forceinline__ void tab_store(int *__restrict__ g, const TileTab &t) {
    const int *src = reinterpret_cast<const int *>(&t);
    for (int k = 0; k < kTabWords; k++) g[k] = src[k];
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

// Build the interval table of a tile from the min/max home cell of its
// particles (single thread).  rdelta[(dz+1)*3 + (dy+1)] receives, for each
// row offset, LDS index - storage index of that row's interval.
__device__ __forceinline__ void tab_build(const GridGeom &g, int cmin, int cmax,
                                          const int *__restrict__ start,
                                          const int *__restrict__ count, int cap, TileTab &t,
                                          int *__restrict__ rdelta) {
    t.n_ivl = 0;
    t.total = 0;
    t.staged = 1;
    for (int r = 0; r < 9; r++) rdelta[r] = 0;
    if (cmax < 0) return;
    const int nx = g.n[0], nxy = g.n[0] * g.n[1];
    const int last = (int)g.total - 1;
    int s0[kMaxIvl], s1[kMaxIvl], m = 0;
    int row_a[9];
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++) {
            row_a[(dz + 1) * 3 + (dy + 1)] = -1;
            long long lo = (long long)cmin + dy * nx + (long long)dz * nxy - 1;
            long long hi = (long long)cmax + dy * nx + (long long)dz * nxy + 1;
            if (lo < 0) lo = 0;
            if (hi > last) hi = last;
            if (lo > hi) continue;
            int a = start[lo], b = start[hi] + count[hi];
            if (b <= a) continue;
            row_a[(dz + 1) * 3 + (dy + 1)] = a;
            // insertion by start
            int p = m++;
            while (p > 0 && s0[p - 1] > a) {
                s0[p] = s0[p - 1];
                s1[p] = s1[p - 1];
                p--;
            }
            s0[p] = a;
            s1[p] = b;
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

// Walk bucket slots 0..cnt-1 of particle i, software-pipelined: the next U
// bucket entries are loaded while the current U neighbours are processed, so
// each lane keeps several loads in flight instead of one dependent chain per
// neighbour.  body(j) gets an LDS-local index (staged tile) or a storage
// index, in bucket order.
template <int U, typename Body>
__device__ __forceinline__ void for_each_neighbour(const unsigned short *__restrict__ n16,
                                                   const int *__restrict__ n32, bool staged,
                                                   int i, int cnt, int stride, Body body) {
    auto entry = [&](int k) -> int {
        const size_t at = (size_t)k * stride + i;
        return staged ? (int)n16[at] : n32[at];
    };
    int cur[U];
#pragma unroll
    for (int u = 0; u < U; u++) cur[u] = u < cnt ? entry(u) : 0;
    for (int k = 0; k < cnt; k += U) {
        int nxt[U];
#pragma unroll
        for (int u = 0; u < U; u++) nxt[u] = (k + U + u) < cnt ? entry(k + U + u) : 0;
#pragma unroll
        for (int u = 0; u < U; u++)
            if (k + u < cnt) body(cur[u]);
#pragma unroll
        for (int u = 0; u < U; u++) cur[u] = nxt[u];
    }
}

// Cooperative copy of the tile's intervals into LDS: dst[off + e] = fn(start + e).
// The LDS image is the concatenation of the intervals, so thread t copies
// LDS slots t, t + blockDim, ...; each thread issues kBatch global loads
// before its first LDS store, keeping that many loads in flight instead of
// waiting out one memory latency per staged particle.
template <typename T, typename Fn>
__device__ __forceinline__ void tab_stage(const TileTab &t, T *__restrict__ dst, Fn fn) {
    constexpr int kBatch = 8;
    const int total = t.total;
    int k = 0;  // interval holding the current slot (slots only increase)
    for (int base = threadIdx.x; base < total; base += kBatch * blockDim.x) {
        int src[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; b++) {
            const int e = base + b * (int)blockDim.x;
            src[b] = -1;
            if (e < total) {
                while (e >= t.off[k] + t.len[k]) k++;
                src[b] = t.start[k] + (e - t.off[k]);
            }
        }
        T v[kBatch];
#pragma unroll
        for (int b = 0; b < kBatch; b++)
            if (src[b] >= 0) v[b] = fn(src[b]);
#pragma unroll
        for (int b = 0; b < kBatch; b++)
            if (src[b] >= 0) dst[base + b * (int)blockDim.x] = v[b];
    }
}

}  // namespace bbx
