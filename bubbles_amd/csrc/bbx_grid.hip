// bbx_grid.hip -- neighbour-grid build on gfx950.
//
// Replaces the reference's per-cell linked-list redistribution
// (UpdateGridDistributionKernel / SwapGridStatesKernel / UpdateParticlesBuckets,
// src/equations/sph_equations3.cpp:487-539 over src/core/grid.h:389-500,422-447)
// with a counting sort into cell-contiguous storage:
//
//   1. hash      one thread per particle: FP64 cell hash, sort key, atomic count
//   2. scan      exclusive prefix sum of the cell counts (wave64 shuffles + LDS)
//   3. scatter   place each particle in its cell's segment (arrival order)
//   4. rank+move rank every particle inside its cell by key and move x, v
//   5. buckets   per particle, walk the 27 cells in (y,x,z) order, apply the
//                reference's FP64 cutoff, keep the first 100 -- fused with the
//                density sum over the same members
//
// Reference chain order is reproduced by the key: a full distribution orders
// each cell by particle id (DistributeByParticle); an incremental one
// (DistributeToCellOpt) concatenates the 27 neighbour cells' old chains in
// (y,x,z) order, i.e. key = (direction of the old cell seen from the new cell,
// old storage slot), since storage is the old chain order.
#include "bbx_mlp_device.h"
#include "bbx_tile.h"

namespace bbx {

// ---------------------------------------------------------------------------
// fills and scan
// ---------------------------------------------------------------------------
__global__ void k_fill_int(int *__restrict__ p, int value, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

int launch_fill_int(bbx_solver *s, int *p, int value, size_t n) {
    if (n == 0) return BBX_OK;
    size_t blocks = (n + kBlock - 1) / kBlock;
    k_fill_int<<<dim3((unsigned)blocks), dim3(kBlock), 0, s->stream>>>(p, value, n);
    BBX_HIP_CHECK(hipGetLastError());
    return BBX_OK;
}

// two int ranges set to zero by one launch: blocks [0, nb0) cover a, the
// rest cover b
__global__ void k_zero_int2(int *__restrict__ a, size_t na, size_t nb0, int *__restrict__ b, size_t nb) {
    const size_t blk = blockIdx.x;
    if (blk < nb0) {
        const size_t i = blk * blockDim.x + threadIdx.x;
        if (i < na) a[i] = 0;
    } else {
        const size_t i = (blk - nb0) * blockDim.x + threadIdx.x;
        if (i < nb) b[i] = 0;
    }
}

static int launch_zero_int2(bbx_solver *s, int *a, size_t na, int *b, size_t nb) {
    if (!b || nb == 0) return launch_fill_int(s, a, 0, na);
    const size_t ba = (na + kBlock - 1) / kBlock, bb = (nb + kBlock - 1) / kBlock;
    k_zero_int2<<<dim3((unsigned)(ba + bb)), dim3(kBlock), 0, s->stream>>>(a, na, ba, b, nb);
    BBX_HIP_CHECK(hipGetLastError());
    return BBX_OK;
}

// Block-wide exclusive scan of kScanItems items per thread.  Each workgroup
// scans kBlock*kScanItems elements and writes its total to block_sums.
__global__ void __launch_bounds__(kBlock) k_scan_block(const int *__restrict__ in,
                                                       int *__restrict__ out,
                                                       int *__restrict__ block_sums, size_t n) {
    __shared__ int wave_tot[kBlock / 64];
    const size_t base = (size_t)blockIdx.x * kBlock * kScanItems + (size_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + k;
        v[k] = i < n ? in[i] : 0;
        sum += v[k];
    }
    // inclusive scan of thread sums inside the wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
    }
    if (lane == 63) wave_tot[wave] = incl;
    __syncthreads();
    int wave_off = 0;
    for (int w = 0; w < wave; w++) wave_off += wave_tot[w];
    int run = wave_off + incl - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        size_t i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (threadIdx.x == kBlock - 1) {
        int tot = 0;
        for (int w = 0; w < kBlock / 64; w++) tot += wave_tot[w];
        block_sums[blockIdx.x] = tot;
    }
}

__global__ void k_scan_add(int *__restrict__ out, const int *__restrict__ block_off, size_t n) {
    const size_t per = (size_t)kBlock * kScanItems;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += block_off[i / per];
}

// Second pass of a scan over at most kScanFuseBlocks first-pass blocks:
// workgroup w adds to its kBlock * kScanItems elements (those of first-pass
// block w) the sum of the first-pass totals before it, which it reduces
// itself, so the totals need no scan of their own.
constexpr int kScanFuseBlocks = 4 * kBlock;
__global__ void __launch_bounds__(kBlock) k_scan_add_reduce(int *__restrict__ out,
                                                            const int *__restrict__ block_sums, size_t n) {
    __shared__ int wave_tot[kBlock / 64];
    __shared__ int offset;
    const int w = (int)blockIdx.x;
    int part = 0;
    for (int t = (int)threadIdx.x; t < w; t += kBlock) part += block_sums[t];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
    if (lane == 0) wave_tot[wave] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int k = 0; k < kBlock / 64; k++) tot += wave_tot[k];
        offset = tot;
    }
    __syncthreads();
    const int add = offset;
    const size_t base = (size_t)w * kBlock * kScanItems;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        const size_t i = base + (size_t)k * kBlock + threadIdx.x;
        if (i < n) out[i] += add;
    }
}

// Exclusive scan, recursive over block totals.  Needs scan_tmp sized by
// scan_scratch_elems(n).
static size_t blocks_for(size_t n) { return (n + (size_t)kBlock * kScanItems - 1) / ((size_t)kBlock * kScanItems); }

size_t scan_scratch_elems(size_t n) {
    size_t total = 0;
    while (n > 1) {
        size_t b = blocks_for(n);
        total += 2 * b;
        n = b;
    }
    return total + 2;
}

static int scan_rec(bbx_solver *s, const int *in, int *out, size_t n, int *tmp) {
    size_t b = blocks_for(n);
    int *sums = tmp;
    int *sums_scanned = tmp + b;
    k_scan_block<<<dim3((unsigned)b), dim3(kBlock), 0, s->stream>>>(in, out, sums, n);
    BBX_HIP_CHECK(hipGetLastError());
    if (b > 1 && b <= (size_t)kScanFuseBlocks) {
        k_scan_add_reduce<<<dim3((unsigned)b), dim3(kBlock), 0, s->stream>>>(out, sums, n);
        BBX_HIP_CHECK(hipGetLastError());
    } else if (b > 1) {
        BBX_CHECK(scan_rec(s, sums, sums_scanned, b, tmp + 2 * b));
        size_t blocks = (n + kBlock - 1) / kBlock;
        k_scan_add<<<dim3((unsigned)blocks), dim3(kBlock), 0, s->stream>>>(out, sums_scanned, n);
        BBX_HIP_CHECK(hipGetLastError());
    }
    return BBX_OK;
}

int launch_scan(bbx_solver *s, const int *in, int *out, size_t n) {
    if (n == 0) return BBX_OK;
    if (scan_scratch_elems(n) > s->scan_tmp_elems) return fail(BBX_E_RANGE, "scan scratch too small");
    return scan_rec(s, in, out, n, s->scan_tmp);
}

// ---------------------------------------------------------------------------
// 1. hash + key + count
// ---------------------------------------------------------------------------
// Full mode (flag set, or first distribution): key = particle id.
// Incremental mode: key = (27-neighbour direction index of the old cell as
// seen from the new cell, old slot); a particle whose cell moved by more than
// one in any axis, or that was in no chain, stays in no chain (the reference
// only looks at the 27 neighbours' chains).
// Append mode (append_from >= 0, after new particles were placed in slots
// append_from..n-1): every old particle keeps its chain and place (key = old
// slot), every new one joins the chain of the cell it hashes into, after the
// old members in the order added (DistributeByParticleList + AddToChain,
// src/core/grid.h:358-387,128-140).
__global__ void __launch_bounds__(kBlock) k_hash(SimConst c, const float4 *__restrict__ pos,
                                                 const int *__restrict__ cell_old,
                                                 const int *__restrict__ flag, int force_full,
                                                 int append_from,
                                                 int *__restrict__ newc, int *__restrict__ hcell,
                                                 unsigned long long *__restrict__ key,
                                                 int *__restrict__ count,
                                                 int *__restrict__ rank_tmp,
                                                 int *__restrict__ counters) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int none = (int)c.g.total;
    const bool valid = s < c.n;
    // the wave stays whole for the aggregated count below
    const float4 p = valid ? pos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    int u[3], bad = 0;
    int cid = hash_cell(c.g, p, u, &bad);
    bool full = append_from < 0 && (force_full || (*flag != 0));
    int nc;
    unsigned long long k;
    if (append_from >= 0) {
        k = (unsigned long long)s;
        if (s < append_from)
            nc = valid ? cell_old[s] : none;
        else
            nc = bad ? none : cid;
    } else if (full) {
        nc = bad ? none : cid;
        k = (unsigned long long)(unsigned)pid_of(p);
    } else {
        const int co = valid ? cell_old[s] : none;
        nc = none;
        k = (unsigned long long)s;
        if (!bad && co != none) {
            int uo[3];
            unlinear(c.g, co, uo);
            int dx = uo[0] - u[0], dy = uo[1] - u[1], dz = uo[2] - u[2];
            if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) {
                nc = cid;
                unsigned d = (unsigned)((dy + 1) * 9 + (dx + 1) * 3 + (dz + 1));
                k = ((unsigned long long)d << 32) | (unsigned long long)(unsigned)s;
            }
        }
    }
    // Storage is sorted by cell, so lanes with the same new cell mostly form
    // runs: one atomic per run reserves its slots (the order inside a cell
    // is set later by the key, not by arrival).
    const int lane = threadIdx.x & 63;
    const int cellv = valid ? nc : -1;
    const int prev = __shfl_up(cellv, 1, 64);
    const bool head = lane == 0 || prev != cellv;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int head_lane = 63 - __builtin_clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const int next_head = after ? __builtin_ctzll(after) : 64;
    int base = 0;
    if (head && valid) base = atomicAdd(&count[nc], next_head - lane);
    base = __shfl(base, head_lane, 64);
    if (!valid) return;
    if (bad) atomicAdd(&counters[CNT_BAD_HASH], 1);
    newc[s] = nc;
    // a point outside the grid has no home cell: its bucket stays empty
    hcell[s] = bad ? -1 : cid;
    key[s] = k;
    rank_tmp[s] = base + (lane - head_lane);
}

// ---------------------------------------------------------------------------
// 3. scatter into cell segments (arrival order), record full/lost stats
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_scatter(SimConst c, const int *__restrict__ newc,
                                                    const int *__restrict__ rank_tmp,
                                                    const int *__restrict__ start,
                                                    const int *__restrict__ count,
                                                    int *__restrict__ tmp_idx,
                                                    int *__restrict__ flag, int force_full,
                                                    int append,
                                                    int *__restrict__ counters) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) {
        counters[CNT_LOST] = count[c.g.total];
        if (!append) {  // an append leaves the higher-level flag to the next substep
            counters[CNT_FULL] = (force_full || *flag) ? 1 : 0;
            *flag = 0;  // ResetHigherLevel
        }
    }
    if (s >= c.n) return;
    tmp_idx[start[newc[s]] + rank_tmp[s]] = s;
}

// ---------------------------------------------------------------------------
// 4. rank inside the cell by key and move the particle state
// ---------------------------------------------------------------------------
// Keys are unique inside a cell (particle id, or (direction, old slot), or
// old slot), so rank = number of smaller keys in the cell is a permutation
// of the cell's segment at every occupancy.  A workgroup whose particles all
// lie in one crowded cell stages the cell's keys through LDS a workgroup-
// width chunk at a time (every lane then reads the same LDS word: a
// broadcast); any other workgroup compares against global memory.
constexpr int kRankLds = 512;  // cells above this use the LDS chunks when uniform
__global__ void __launch_bounds__(kBlock) k_rank_move(SimConst c, const int *__restrict__ tmp_idx,
                                                      const int *__restrict__ newc,
                                                      const int *__restrict__ hcell,
                                                      const unsigned long long *__restrict__ key,
                                                      const int *__restrict__ start,
                                                      const int *__restrict__ count,
                                                      const float4 *__restrict__ pos_in,
                                                      const float4 *__restrict__ vel_in,
                                                      float4 *__restrict__ pos_out,
                                                      float4 *__restrict__ vel_out,
                                                      int *__restrict__ cell_out,
                                                      int *__restrict__ home_out) {
    __shared__ unsigned long long skey[kBlock];
    __shared__ int cell0;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < c.n;
    const int s = valid ? tmp_idx[t] : 0;
    const int cc = valid ? newc[s] : -1;
    const unsigned long long k = valid ? key[s] : 0ull;
    const int b = valid ? start[cc] : 0;
    const int e = valid ? b + count[cc] : 0;
    // thread 0 is valid in every launched workgroup (the grid is ceil(n / kBlock))
    if (threadIdx.x == 0) cell0 = cc;
    // Thread t holds the key of storage slot t, so the keys of the cell
    // members inside this workgroup's slots come from LDS; only members in
    // neighbouring workgroups' slots are read from global memory.
    const int blk0 = (int)(blockIdx.x * blockDim.x);
    skey[threadIdx.x] = k;
    __syncthreads();
    const bool uniform = __syncthreads_and(!valid || cc == cell0) != 0;
    const int c0 = cell0;
    const int bb = uniform ? start[c0] : 0;
    const int ee = uniform ? bb + count[c0] : 0;
    int rank = 0;
    if (uniform && ee - bb > kRankLds) {
        // every valid thread has [b, e) == [bb, ee); the trip count and the
        // barriers are the same for the whole workgroup
        for (int q0 = bb; q0 < ee; q0 += kBlock) {
            const int q = q0 + (int)threadIdx.x;
            if (q < ee) skey[threadIdx.x] = key[tmp_idx[q]];
            __syncthreads();
            if (valid) {
                const int m = min(kBlock, ee - q0);
                for (int u = 0; u < m; u++) rank += skey[u] < k ? 1 : 0;
            }
            __syncthreads();
        }
    } else if (valid) {
        for (int q = b; q < e; q++) {
            const int u = q - blk0;
            const unsigned long long kq = (u >= 0 && u < kBlock) ? skey[u] : key[tmp_idx[q]];
            rank += kq < k ? 1 : 0;
        }
    }
    if (!valid) return;
    const int dest = b + rank;
    pos_out[dest] = pos_in[s];
    vel_out[dest] = vel_in[s];
    cell_out[dest] = cc;
    home_out[dest] = hcell[s];
}

// ---------------------------------------------------------------------------
// 5. tile tables: one wave per tile of kBlock particles
// ---------------------------------------------------------------------------
// The tile's home cells span [cmin, cmax]; its nine neighbour rows become at
// most nine merged storage intervals (tab_row, tab_merge).  Every gather of
// the substep reads this table.  Storage is sorted by cell except for the
// particles in no chain (home -1) at its end, so unless the tile's last
// particle is one of those, cmin and cmax are the homes of its first and
// last particles (otherwise a wave reduction over the tile finds them).  The
// row ranges of lanes 0-8 are gathered by shuffles and merged by lane 0.
constexpr int kTabWaves = kBlock / 64;
__global__ void __launch_bounds__(kBlock) k_tile_tables(SimConst c, const int *__restrict__ home,
                                                        const int *__restrict__ start, int ntiles,
                                                        int *__restrict__ tile_tab) {
    __shared__ TileTab tt[kTabWaves];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tile = blockIdx.x * kTabWaves + w;  // wave-uniform
    if (tile >= ntiles) return;
    const int t0 = tile * kBlock, t1 = min(t0 + kBlock, c.n) - 1;
    int lo = 0x7fffffff, hi = home[t1];
    if (hi >= 0) {
        lo = home[t0];
    } else {
#pragma unroll
        for (int k = 0; k < kBlock / 64; k++) {
            const int i = t0 + k * 64 + lane;
            const int hc = i < c.n ? home[i] : -1;
            if (hc >= 0) {
                lo = min(lo, hc);
                hi = max(hi, hc);
            }
        }
        lo = wave_min_i32(lo);
        hi = wave_max_i32(hi);
    }
    int a = -1, b = -1;
    if (lane < 9) tab_row(c.g, hi >= 0 ? lo : 0, hi, start, lane, a, b);
    int ra[9], rb[9];
#pragma unroll
    for (int r = 0; r < 9; r++) {
        ra[r] = __shfl(a, r, 64);
        rb[r] = __shfl(b, r, 64);
    }
    if (lane == 0 && !tab_merge_ordered(ra, rb, c.tile_cap, tt[w])) tab_merge(ra, rb, c.tile_cap, tt[w]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < kTabWords) tile_tab[(size_t)tile * kTabInts + lane] = reinterpret_cast<const int *>(&tt[w])[lane];
}

// ---------------------------------------------------------------------------
// 5. buckets + density (+ Tait pressure for SPH)
// ---------------------------------------------------------------------------
// ComputeDensityFor (src/equations/sph_equations3.cpp:25-58) sums poly6 over
// the bucket in bucket order; the sum runs as the bucket is built.
// The tile's candidate intervals are staged in LDS; the bucket of a staged
// tile is written as LDS-local indices (nbr16), otherwise as storage indices
// (nbr32).  Candidates are visited in the reference's order: neighbour cells
// y outer, x, z inner (src/core/grid.h:568-591), each cell in chain order.
// The learned warm start fused into the bucket kernel (kFeat > 0): the
// features of include/bbx_features.h come from the finished bucket, the
// density and the particle's own state, so they need nothing another
// particle computes in this launch.  kFeat == 2 then infers the network on
// the workgroup's 256 particles and starts the substep from its pressures.
struct WarmIn {
    StepConst st;
    const float4 *__restrict__ vel;       // own velocity, w = previous converged pressure
    int store_feat;                       // write the features to feat
    float *__restrict__ feat;             // [16][stride]
    const float *__restrict__ P;          // network parameters (kFeat == 2)
    float *__restrict__ pwarm;            // network pressure
    float *__restrict__ rho_star;         // rho* = rho at k = 0
    float4 *__restrict__ xq;              // (x, p0 / rho^2)
};

template <bool kPressure, int kFeat>
__global__ void __launch_bounds__(kBlock) k_buckets_density(SimConst c,
                                                            const float4 *__restrict__ pos,
                                                            const int *__restrict__ home,
                                                            const int *__restrict__ start,
                                                            const int *__restrict__ count,
                                                            unsigned short *__restrict__ nbr16,
                                                            int *__restrict__ nbr32,
                                                            int *__restrict__ nbr_cnt,
                                                            float *__restrict__ rho,
                                                            float *__restrict__ pres,
                                                            const int *__restrict__ tile_tab,
                                                            int *__restrict__ counters,
                                                            WarmIn wi) {
    // staged candidate positions (16-byte LDS DMA); eight slack slots let
    // the candidate test read whole groups of four, and the group after
    // them, past the end of a row
    __shared__ float4 sx[kTileCap + 8];
    __shared__ TileTab tt;
    const int tile = tile_of_block();
    // The tile table is read first, into each wave's lanes: the staging
    // below then waits for nothing but it, not for the row bounds requested
    // in between, and needs no workgroup barrier (the LDS copy that the rest
    // of the kernel reads is visible after the barrier behind the staging).
    const int tw = tab_load_wave(tile_tab + (size_t)tile * kTabInts);
    const int i = tile * kBlock + threadIdx.x;
    const bool valid = i < c.n;
    const int hc = valid ? home[i] : -1;
    // The particle's position and its nine row bounds do not depend on the
    // tile table or the staging: they are requested first.  Row r =
    // (dz+1)*3 + (dy+1): the cells x-1, x, x+1 of a row are contiguous in
    // storage; A[r] = storage start of cell x-1, and O[r] packs the offsets
    // of cells x-1 (zero), x, x+1 and of the row end as bytes 0-3 (rows
    // outside the grid are empty).  A row of at most 64 candidates whose cells hold at most 32
    // each is "narrow": it is tested into a bit mask; otherwise the particle
    // walks cell by cell.
    const int nx = c.g.n[0], ny = c.g.n[1], nz = c.g.n[2], nxy = nx * ny;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    int u[3] = {0, 0, 0};
    int A[9];
    unsigned O[9];
    bool narrow = true;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        A[r] = 0;
        O[r] = 0u;
    }
    if (valid) pi = pos[i];
    // All 36 row-bound loads are issued before any is used (indices clamped
    // into start[0 .. total]; rows outside the grid and cells past the x
    // bounds read a valid word and ignore it), and the staging is issued
    // behind them, so the latencies overlap instead of following one another
    // row by row.
    int b[9][4];
    if (hc >= 0) {
        unlinear_fast(c.g, c.inv_nx, c.inv_nxy, hc, u);
        const int last = (int)c.g.total;
#pragma unroll
        for (int r = 0; r < 9; r++) {
            const int yy = u[1] + (r % 3) - 1, zz = u[2] + (r / 3) - 1;
            const bool in = yy >= 0 && yy < ny && zz >= 0 && zz < nz;
            const int cx = in ? u[0] + yy * nx + zz * nxy : 0;
#pragma unroll
            for (int q = 0; q < 4; q++) b[r][q] = start[min(max(cx - 1 + q, 0), last)];
        }
    }
    if (threadIdx.x < kTabWords) reinterpret_cast<int *>(&tt)[threadIdx.x] = tw;
    const bool staged = tab_word(tw, 1) != 0;
    if (staged) tab_stage_dma_wave(tw, [&](int j, int e) { lds_dma16(pos + j, sx + e); });
    if (hc >= 0) {
#pragma unroll
        for (int r = 0; r < 9; r++) {
            const int yy = u[1] + (r % 3) - 1, zz = u[2] + (r / 3) - 1;
            if (yy >= 0 && yy < ny && zz >= 0 && zz < nz) {
                const int b1 = b[r][1], b2 = b[r][2];
                const int b0 = u[0] > 0 ? b[r][0] : b1;
                const int b3 = u[0] + 1 < nx ? b[r][3] : b2;
                A[r] = b0;
                O[r] = ((unsigned)(b1 - b0) << 8) | ((unsigned)(b2 - b0) << 16) | ((unsigned)(b3 - b0) << 24);
                narrow = narrow && (b3 - b0 <= c.row_max) && (b1 - b0 <= 32) && (b2 - b1 <= 32) &&
                         (b3 - b2 <= 32);
            }
        }
    }
    lds_dma_wait();
    __syncthreads();
    if (kFeat != 2 && !valid) return;  // the inference below needs the whole block at its barriers
    float fv[kFeat ? BBX_NFEAT : 1];
    float di = 0.f;
    // The particle's work is instantiated once for LDS-staged and once for
    // global-memory tiles, so no load selects between the two address spaces
    // (that would compile to a slower generic load).
    auto per_particle = [&](auto staged_tag) {
        constexpr bool kStaged = decltype(staged_tag)::value;
        int cnt = 0, dropped = 0;
        float sum = 0.f, comp = 0.f;
        float m1 = 0.f, m2 = 0.f, m3 = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;  // features
        // The current quad of bucket entries.  Staged: a 64-bit shift
        // register of 16-bit entries held as two words (entry k of the quad
        // ends in bits 16k..16k+15); global: four 32-bit entries.
        unsigned qlo = 0u, qhi = 0u;
        int e0 = 0, e1 = 0, e2 = 0, e3 = 0;
        auto store_quad = [&](int k, int nfill) {  // k = entries so far, nfill = entries in the quad
            const size_t at = (size_t)((k - 1) >> 2) * c.stride + i;
            if (kStaged) {
                unsigned long long q = ((unsigned long long)qhi << 32) | qlo;
                q >>= 16 * (4 - nfill);  // a partial quad moves down to slot 0
                reinterpret_cast<uint2 *>(nbr16)[at] = make_uint2((unsigned)q, (unsigned)(q >> 32));
            } else {
                int4 v = make_int4(e0, e1, e2, e3);
                if (nfill == 3) v = make_int4(e1, e2, e3, 0);
                else if (nfill == 2) v = make_int4(e2, e3, 0, 0);
                else if (nfill == 1) v = make_int4(e3, 0, 0, 0);
                reinterpret_cast<int4 *>(nbr32)[at] = v;
            }
        };
        // append accepted candidate j (storage index; LDS index j + sh) to the
        // bucket (emit: up to the cap).  Only the index is handled here: the
        // emission loops run as long as the wave's longest cell, so the
        // density sum runs afterwards over the finished bucket, one uniform
        // pass.
        auto append = [&](int j, int sh) {
            if (kStaged) {  // the entry is the candidate's LDS byte offset
                qlo = __builtin_amdgcn_alignbit(qhi, qlo, 16);
                qhi = __builtin_amdgcn_alignbit((unsigned)(j + sh) << kSlotShift, qhi, 16);
            } else {
                e0 = e1;
                e1 = e2;
                e2 = e3;
                e3 = j;
            }
            cnt++;
            if ((cnt & 3) == 0) store_quad(cnt, 4);
        };
        auto emit = [&](int j, int sh) {
            if (cnt >= kBucket)
                dropped++;
            else
                append(j, sh);
        };
        if (hc >= 0) {
            int sh[9];
#pragma unroll
            for (int r = 0; r < 9; r++) sh[r] = kStaged ? tt.rdelta[r] : 0;
            bool emitted = false;
            if (narrow) {
                // Pass 1: each row's candidates, four at a time with no
                // branches, into an acceptance mask.  The FP32 test uses
                // t = |pi - pj|^2 - band_c (the offset folded into the first
                // FMA): t < 0 accepts; a candidate with |t| <= band_w (rare)
                // makes the lane re-test its row with the reference's FP64
                // rule.  The sign bits of a group form a nibble that enters a
                // 64-bit shift register {hi, lo} at its top.
                unsigned long long M[9];
                int total = 0;
#pragma unroll
                for (int r = 0; r < 9; r++) {
                    const int len = (int)(O[r] >> 24);
                    unsigned lo = 0u, hi = 0u;
                    float bdist = 3.4e38f;
                    int t0 = 0;
                    // Two groups of four per step: the second group's loads
                    // are in flight while the first group's arithmetic runs
                    // (one step of a row reads up to seven candidates past
                    // its end: slack slots when staged, clamped otherwise;
                    // their bits are cleared below).
                    for (; t0 < len; t0 += 8) {
                        float4 pc[8];
#pragma unroll
                        for (int q = 0; q < 8; q++) {
                            if (kStaged)
                                pc[q] = sx[A[r] + sh[r] + t0 + q];
                            else
                                pc[q] = pos[min(A[r] + t0 + q, c.n - 1)];
                        }
                        float tq[8];
#pragma unroll
                        for (int q = 0; q < 8; q++) tq[q] = cut_offset(pi, pc[q], c.band_c);
#pragma unroll
                        for (int q = 0; q < 8; q += 2) bdist = fminf(fminf(bdist, fabsf(tq[q])), fabsf(tq[q + 1]));
                        unsigned byte = __float_as_uint(tq[7]) >> 31;  // bit q: candidate t0 + q
#pragma unroll
                        for (int q = 6; q >= 0; q--) byte = __builtin_amdgcn_alignbit(byte, __float_as_uint(tq[q]), 31);
                        lo = __builtin_amdgcn_alignbit(hi, lo, 8);
                        hi = __builtin_amdgcn_alignbit(byte, hi, 8);
                    }
                    // t0 = 8 * steps; candidate t sits at bit 64 - t0 + t
                    unsigned long long m = ((unsigned long long)hi << 32) | lo;
                    m = t0 > 0 ? m >> (64 - t0) : 0ull;
                    if (len < 64) m &= (1ull << len) - 1ull;
                    if (bdist <= c.band_w) {
                        // candidates inside the band: the exact FP64 rule
                        for (int t = 0; t < len; t++) {
                            const float4 pj = kStaged ? sx[A[r] + sh[r] + t] : pos[A[r] + t];
                            if (fabsf(cut_offset(pi, pj, c.band_c)) <= c.band_w) {
                                const unsigned long long bit = 1ull << t;
                                m = within_std_fp64(pi, pj, c.h2d) ? (m | bit) : (m & ~bit);
                            }
                        }
                    }
                    M[r] = m;
                    total += __popcll(m);
                }
                // Pass 2: emit in the reference's order, cells (y, x, z), chain
                // order inside each cell.  x runs as a loop with its cell bounds
                // selected, which keeps the kernel's code (nine emission sites
                // instead of 27) small.  A narrow row's cells hold <= 32
                // candidates, so each cell's bits fit one 32-bit word.  The
                // bucket cap cannot bind here (a lane over it takes the
                // cell-by-cell path below, which truncates).
                if (total <= kBucket) {
                    emitted = true;
#pragma unroll
                    for (int dy = 0; dy < 3; dy++)
#pragma unroll 1
                        for (int dx = 0; dx < 3; dx++)
#pragma unroll
                            for (int dz = 0; dz < 3; dz++) {
                                const int r = dz * 3 + dy;
                                const int lo = (int)__builtin_amdgcn_ubfe(O[r], 8 * dx, 8);
                                const int hi = (int)__builtin_amdgcn_ubfe(O[r], 8 * dx + 8, 8);
                                const int w = hi - lo;
                                if (w <= 0) continue;
                                unsigned m = (unsigned)(M[r] >> lo);
                                if (w < 32) m = __builtin_amdgcn_ubfe(m, 0, w);
                                while (m) {
                                    const int b = __builtin_ctz(m);
                                    m &= m - 1u;
                                    append(A[r] + lo + b, sh[r]);
                                }
                            }
                }
            }
            if (!emitted) {
                // a wide row, a crowded cell or a bucket over the cap: cell by cell
                for (int dy = 0; dy < 3; dy++)
                    for (int dx = 0; dx < 3; dx++) {
                        const int x = u[0] + dx - 1;
                        if (x < 0 || x >= nx) continue;
                        for (int dz = 0; dz < 3; dz++) {
                            const int yy = u[1] + dy - 1, zz = u[2] + dz - 1;
                            if (yy < 0 || yy >= ny || zz < 0 || zz >= nz) continue;
                            const int cell = x + yy * nx + zz * nxy;
                            const int sr = sh[dz * 3 + dy];
                            const int b = start[cell], e = b + count[cell];
                            for (int j = b; j < e; j++) {
                                const float4 pj = kStaged ? sx[j + sr] : pos[j];
                                if (in_bucket(c, pi, pj)) emit(j, sr);
                            }
                        }
                    }
            }
        }
        if (cnt & 3) store_quad(cnt, cnt & 3);
        // ComputeDensityFor over the bucket in bucket order, read back from
        // the table this lane has just written (its own quads: no other lane
        // or wave writes them, and no earlier load in this kernel cached them)
        __builtin_amdgcn_s_waitcnt(0);
        // The particle's own index (its LDS slot in the centre row).  Dead
        // entries of the last quad read the particle itself: its staged copy
        // is its position bit for bit, so they add exact zeros to the
        // feature sums below, as the particle's own entry does, and only
        // the density masks them.
        const int self = kStaged ? (i + tt.rdelta[4]) << kSlotShift : i;
        for_each_neighbour_dead<kNbrDepth>(nbr16, nbr32, kStaged, i, cnt, c.stride, [&](int j, bool live) {
            const float4 pj = kStaged ? lds_at(sx, j) : pos[j];
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            const float w = poly6_w_member(c, d2);
            kahan_add(sum, comp, live ? w : 0.f);  // a dead entry adds an exact zero
            if constexpr (kFeat > 0) {
                // distance moments r, r^2, r^3 and the centroid offset over
                // the members (powers of 1/h applied once afterwards)
                const float r = fast_sqrt(d2);
                m1 += r;
                m2 += d2;
                m3 = fmaf(r, d2, m3);
                cx -= dx;
                cy -= dy;
                cz -= dz;
            }
        }, self);
        nbr_cnt[i] = cnt;
        di = sum * c.mass;
        rho[i] = di;
        if constexpr (kFeat > 0) {
            const float4 vi = wi.vel[i];
            const float inv = 1.f / (cnt > 1 ? (float)(cnt - 1) : 1.f);
            float wall = 3.4e38f;
            const float p3[3] = {pi.x, pi.y, pi.z};
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float lo = p3[a] - (c.tank_c[a] - c.tank_h[a]);
                const float hi = (c.tank_c[a] + c.tank_h[a]) - p3[a];
                wall = fminf(wall, fminf(lo, hi));
            }
            const float vs = wi.st.dt * c.inv_h;
            const float s1 = c.inv_h * inv, s2 = s1 * c.inv_h, s3 = s2 * c.inv_h;
            cx *= c.inv_h;
            cy *= c.inv_h;
            cz *= c.inv_h;
            fv[BBX_F_DENSITY_ERR] = di / c.rho0 - 1.f;
            fv[BBX_F_RESERVED1] = 0.f;
            fv[BBX_F_COUNT] = (float)cnt;
            fv[BBX_F_M1] = m1 * s1;
            fv[BBX_F_M2] = m2 * s2;
            fv[BBX_F_M3] = m3 * s3;
            fv[BBX_F_CX] = cx * inv;
            fv[BBX_F_CY] = cy * inv;
            fv[BBX_F_CZ] = cz * inv;
            fv[BBX_F_CNORM] = sqrtf(cx * cx + cy * cy + cz * cz) * inv;
            fv[BBX_F_PPREV] = wi.st.delta != 0.f ? vi.w / (wi.st.delta * c.rho0) : 0.f;
            fv[BBX_F_VY] = vi.y * vs;
            fv[BBX_F_VNORM] = sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z) * vs;
            fv[BBX_F_RESERVED2] = 0.f;
            fv[BBX_F_RESERVED3] = 0.f;
            fv[BBX_F_WALL] = fminf(fmaxf(wall * c.inv_h, 0.f), 4.f);
            if (kFeat == 1 || wi.store_feat) {
                const size_t S = (size_t)c.stride;
#pragma unroll
                for (int f = 0; f < BBX_NFEAT; f++) wi.feat[(size_t)f * S + i] = fv[f];
            }
        }
        if (kPressure) {
            // ComputePressureValue (src/equations/sph_equations3.cpp:7-18)
            float p = c.eos_scale / c.eos_exp * (powf(di / c.rho0, c.eos_exp) - 1.f);
            if (p < 0.f) p *= c.nps;
            pres[i] = p;
        }
        if (dropped) atomicAdd(&counters[CNT_OVERFLOW], dropped);
    };
    if (valid) with_staging(staged, per_particle);
    if constexpr (kFeat == 2) {
        // The block's 256 feature rows go through LDS (over the staged
        // candidates, which every thread has finished with) into the MFMA
        // layout; each wave infers its 64 particles as four 16-sample tiles.
        constexpr int kRow = BBX_NFEAT + 1;  // padded feature row
        float *lf = reinterpret_cast<float *>(sx);  // 256 rows
        float *lp = lf + kBlock * kRow;             // 256 warm-start pressures
        float *lb = lp + kBlock;                    // b1 | b2 | w3
        unsigned short *lw = reinterpret_cast<unsigned short *>(lb + 3 * kH);  // BF16 weights of both layers
        static_assert((kBlock * (kRow + 1) + 3 * kH) * 4 + kLdsWeights * 2 <= (int)sizeof(float4) * kTileCap &&
                          (kBlock * (kRow + 1) + 3 * kH) % 4 == 0,
                      "warm-start LDS layout over the staged candidates");
        __syncthreads();
#pragma unroll
        for (int f = 0; f < BBX_NFEAT; f++) lf[threadIdx.x * kRow + f] = valid ? fv[f] : 0.f;
        stage_infer_weights(wi.P, lw, lb, (int)threadIdx.x, kBlock);
        __syncthreads();
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const int qq = lane >> 4, jj = lane & 15;
        bf16x4 w1[4], w2[4][4];
        load_weight_operands_lds(lw, lane, w1, w2);
        float mean[4], istd[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            mean[v] = wi.P[OFF_MEAN + 4 * qq + v];
            istd[v] = wi.P[OFF_ISTD + 4 * qq + v];
        }
        const float b3 = wi.P[OFF_B3], oscale = wi.P[OFF_OSCALE], alpha = wi.P[OFF_ALPHA];
        const float pcap = wi.P[OFF_PCAP], resid = wi.P[OFF_RESID];
#pragma unroll
        for (int half = 0; half < 2; half++) {  // two pairs of tiles keep the registers below 128
            float xv[2][4];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int smp = wv * 64 + (2 * half + t) * 16 + jj;
#pragma unroll
                for (int v = 0; v < 4; v++) xv[t][v] = lf[smp * kRow + 4 * qq + v];
            }
            float yn[2];
            forward_infer_tiles<2>(w1, w2, lb, lb + kH, lb + 2 * kH, b3, xv, mean, istd, qq, yn);
            if (lane < 16) {
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int smp = wv * 64 + (2 * half + t) * 16 + jj;
                    // the same expression as the standalone inference kernel
                    float y = yn[t] * oscale;
                    if (resid != 0.f) y += resid * lf[smp * kRow + BBX_F_PPREV];
                    y = alpha * fminf(fmaxf(y, 0.f), pcap);
                    lp[smp] = y * wi.st.warm_scale;
                }
            }
        }
        __syncthreads();
        if (valid) {
            // PredictVelocityAndPositionFor at k = 0 (src/equations/pcisph_equations3.cpp:3-28)
            // with the network's pressure instead of 0: rho* = rho, p = p0
            const float p = lp[threadIdx.x];
            const float d2 = di * di;
            pres[i] = p;
            wi.pwarm[i] = p;
            wi.rho_star[i] = di;
            wi.xq[i] = make_float4(pi.x, pi.y, pi.z, (fabsf(d2) < 1e-8f) ? 0.f : p / d2);
        }
    }
}

// density over the existing buckets (after an upload of positions)
__global__ void __launch_bounds__(kBlock) k_density(SimConst c, const float4 *__restrict__ pos,
                                                    const unsigned short *__restrict__ nbr16,
                                                    const int *__restrict__ nbr32,
                                                    const int *__restrict__ nbr_cnt,
                                                    const int *__restrict__ tile_tab,
                                                    float *__restrict__ rho) {
    __shared__ float4 sx[kTileCap];
    __shared__ TileTab tt;
    const int tile = tile_of_block();
    tab_load_shared(tile_tab + (size_t)tile * kTabInts, tt);
    __syncthreads();
    const bool staged = tt.staged != 0;
    if (staged) tab_stage(tt, sx, [&](int j) { return pos[j]; });
    __syncthreads();
    const int i = tile * kBlock + threadIdx.x;
    if (i >= c.n) return;
    const float4 pi = pos[i];
    float sum = 0.f, comp = 0.f;
    const int cnt = nbr_cnt[i];
    with_staging(staged, [&](auto tag) {
        constexpr bool kS = decltype(tag)::value;
        for_each_neighbour<2>(nbr16, nbr32, kS, i, cnt, c.stride, [&](int j, bool live) {
            const float4 pj = kS ? lds_at(sx, j) : pos[j];
            float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float w = poly6_w(c, dx * dx + dy * dy + dz * dz);
            kahan_add(sum, comp, live ? w : 0.f);  // a dead entry adds an exact zero
        });
    });
    rho[i] = sum * c.mass;
}

static unsigned nblocks(int n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// UpdateGridDistribution: full or incremental chain update, then buckets.
int distribute_grid(bbx_solver *s, bool force_full, int append_from, int *clear2, size_t clear2_n) {
    SimConst &c = s->sc;
    const int n = s->n;
    const size_t cells = (size_t)c.g.total + 1;
    BBX_CHECK(launch_zero_int2(s, s->cell_count, cells, clear2, clear2_n));
    if (n > 0) {
        {
            BBX_PROF_BEGIN("grid_hash");
            k_hash<<<nblocks(n), kBlock, 0, s->stream>>>(c, s->posA, s->cellA, s->flag,
                                                         force_full ? 1 : 0, append_from, s->newc,
                                                         s->hash_tmp, s->key, s->cell_count,
                                                         s->rank_tmp, s->counters);
            BBX_HIP_CHECK(hipGetLastError());
            BBX_PROF_END();
        }
        {
            BBX_PROF_BEGIN("grid_scan");
            BBX_CHECK(launch_scan(s, s->cell_count, s->cell_start, cells));
            BBX_PROF_END();
        }
        {
            BBX_PROF_BEGIN("grid_scatter");
            k_scatter<<<nblocks(n), kBlock, 0, s->stream>>>(c, s->newc, s->rank_tmp,
                                                            s->cell_start, s->cell_count,
                                                            s->tmp_idx, s->flag,
                                                            force_full ? 1 : 0, append_from >= 0 ? 1 : 0,
                                                            s->counters);
            BBX_HIP_CHECK(hipGetLastError());
            BBX_PROF_END();
        }
        {
            BBX_PROF_BEGIN("grid_rank_move");
            k_rank_move<<<nblocks(n), kBlock, 0, s->stream>>>(c, s->tmp_idx, s->newc,
                                                              s->hash_tmp, s->key,
                                                              s->cell_start, s->cell_count,
                                                              s->posA, s->velA, s->posB,
                                                              s->velB, s->cellB, s->home);
            BBX_HIP_CHECK(hipGetLastError());
            BBX_PROF_END();
        }
        std::swap(s->posA, s->posB);
        std::swap(s->velA, s->velB);
        std::swap(s->cellA, s->cellB);
    }
    return BBX_OK;
}

int build_buckets(bbx_solver *s, bool with_pressure, int feat_mode, const StepConst *st) {
    if (feat_mode && (with_pressure || !st))
        return fail(BBX_E_INVALID, "warm-start features need a PCISPH bucket build and step constants");
    if (feat_mode == 2 && !s->mlp) return fail(BBX_E_INVALID, "no MLP attached");
    SimConst &c = s->sc;
    const int n = s->n;
    if (n > 0) {
        {
            BBX_PROF_BEGIN("tile_tables");
            k_tile_tables<<<(unsigned)((s->ntiles + kTabWaves - 1) / kTabWaves), kBlock, 0, s->stream>>>(
                c, s->home, s->cell_start, s->ntiles, s->tile_tab);
            BBX_HIP_CHECK(hipGetLastError());
            BBX_PROF_END();
        }
        WarmIn wi{};
        if (st) wi.st = *st;
        wi.vel = s->velA;
        wi.store_feat = s->keep_features;
        wi.feat = s->feat;
        wi.P = feat_mode == 2 ? mlp_params_device(s->mlp) : nullptr;
        wi.pwarm = s->pwarm;
        wi.rho_star = s->rho_star;
        wi.xq = s->xq;
        BBX_PROF_BEGIN(feat_mode == 2 ? "buckets_density_warm"
                                      : (feat_mode == 1 ? "buckets_density_features" : "buckets_density"));
#define BBX_BUCKETS(P, F)                                                                                   \
    k_buckets_density<P, F><<<nblocks(n), kBlock, 0, s->stream>>>(c, s->posA, s->home, s->cell_start,       \
                                                                   s->cell_count, s->nbr16, s->nbr32,      \
                                                                   s->nbr_cnt, s->rho, s->pres,            \
                                                                   s->tile_tab, s->counters, wi)
        if (with_pressure)
            BBX_BUCKETS(true, 0);
        else if (feat_mode == 2)
            BBX_BUCKETS(false, 2);
        else if (feat_mode == 1)
            BBX_BUCKETS(false, 1);
        else
            BBX_BUCKETS(false, 0);
#undef BBX_BUCKETS
        BBX_HIP_CHECK(hipGetLastError());
        BBX_PROF_END();
    }
    return BBX_OK;
}

int distribute(bbx_solver *s, bool force_full, bool buckets, bool with_pressure, int append_from,
               int feat_mode, const StepConst *st) {
    if (feat_mode && !buckets) return fail(BBX_E_INVALID, "warm-start features need a bucket build");
    BBX_CHECK(distribute_grid(s, force_full, append_from));
    return buckets ? build_buckets(s, with_pressure, feat_mode, st) : BBX_OK;
}

int run_density_only(bbx_solver *s) {
    if (s->n == 0) return BBX_OK;
    BBX_PROF_BEGIN("density");
    k_density<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, s->posA, s->nbr16, s->nbr32, s->nbr_cnt,
                                                        s->tile_tab, s->rho);
    BBX_HIP_CHECK(hipGetLastError());
    BBX_PROF_END();
    return BBX_OK;
}

}  // namespace bbx
