// bbx_sph.hip -- SPH / PCISPH gathers, prediction and integration on gfx950.
//
// Every kernel is one thread per particle in cell-sorted storage order.  A
// workgroup's 256 particles form a tile: the tile's candidate intervals
// (recorded by the bucket kernel) are first copied into LDS with coalesced
// loads, then each thread walks its bucket -- LDS-local 16-bit indices in
// the quad-major table (four entries per load) -- reading neighbours from LDS.
// Tiles too large to stage walk 32-bit storage indices and read global
// memory instead.  Fusions relative to the reference's launch sequence
// (src/equations/pcisph_equations3.cpp:218-256):
//
//   nonpressure + predict(k=0)       ComputeNonPressureForceFor + PredictVelocityAndPositionFor
//   predicted density + pressure     PredictPressureFor, FP64 max-error reduction
//   pressure force + predict(k+1)    PredictPressureForceFor + PredictVelocityAndPositionFor
//   pressure force + integrate       PredictPressureForceFor + AccumulateForcesFor + TimeIntegrationFor
//
// Arithmetic is FP32; density and pressure-force sums use Kahan compensation.
#include "bbx_collider.h"
#include "bbx_mlp_device.h"
#include "bbx_tile.h"

namespace bbx {

static unsigned nblocks(int n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// block-wide max of v, one atomic per workgroup into its slot (whole block calls)
__device__ __forceinline__ void block_max_to_slot(unsigned v, unsigned *__restrict__ slots) {
    __shared__ unsigned wmax[kBlock / 64];
    v = wave_max_u32(v);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned m = wmax[0];
        for (int w = 1; w < kBlock / 64; w++) m = wmax[w] > m ? wmax[w] : m;
        if (m) atomicMax(&slots[blockIdx.x & (kSlots - 1)], m);
    }
}

// neighbour tables of the substep
struct NbrTab {
    const unsigned short *__restrict__ n16;
    const int *__restrict__ n32;
    const int *__restrict__ cnt;
    const int *__restrict__ tab;
};

// The tank alone takes the inlined reversed-box path; a collider list goes
// through the collider-set rule (same result for the tank as its only entry).
__device__ __forceinline__ void collide_any(const SimConst &c, float e, float3 &p, float3 &v) {
    if (c.ncoll == 0)
        collide(c, e, p, v);
    else
        resolve_colliders(c.coll, c.ncoll, c.radius, e, p, v);
}

// PredictVelocityAndPositionFor (src/equations/pcisph_equations3.cpp:3-28):
// v* = v + dt/m (f + f_p), x* = x + dt v*, collider with restitution 0.
__device__ __forceinline__ float4 predict_position(const SimConst &c, const StepConst &st,
                                                   float4 x, float4 v, float fx, float fy,
                                                   float fz) {
    float3 tv = make_float3(v.x + st.dt_over_m * fx, v.y + st.dt_over_m * fy,
                            v.z + st.dt_over_m * fz);
    float3 tp = make_float3(x.x + st.dt * tv.x, x.y + st.dt * tv.y, x.z + st.dt * tv.z);
    collide_any(c, 0.f, tp, tv);
    return make_float4(tp.x, tp.y, tp.z, 0.f);
}

// TimeIntegrationFor (src/equations/sph_equations3.cpp:283-339): symplectic
// Euler, collider with restitution 0.6, clamp into the domain, 0.9-cell flag.
__device__ __forceinline__ void integrate_one(const SimConst &c, float dt, float4 xi, float4 vi,
                                              float Fx, float Fy, float Fz, float3 &p, float3 &v,
                                              int *__restrict__ flag) {
    const float inv_m = 1.f / c.mass;
    v = make_float3(vi.x + dt * (Fx * inv_m), vi.y + dt * (Fy * inv_m), vi.z + dt * (Fz * inv_m));
    p = make_float3(xi.x + dt * v.x, xi.y + dt * v.y, xi.z + dt * v.z);
    collide_any(c, 0.6f, p, v);
    if (!inside_box(p.x, p.y, p.z, c.bmin, c.bmax)) {
        p.x = fminf(fmaxf(p.x, c.bmin[0] + c.radius), c.bmax[0] - c.radius);
        p.y = fminf(fmaxf(p.y, c.bmin[1] + c.radius), c.bmax[1] - c.radius);
        p.z = fminf(fmaxf(p.z, c.bmin[2] + c.radius), c.bmax[2] - c.radius);
    }
    float ddx = p.x - xi.x, ddy = p.y - xi.y, ddz = p.z - xi.z;
    if (sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) >= c.min_len * 0.9f) *flag = 1;
}

// ---------------------------------------------------------------------------
// non-pressure force (src/equations/sph_equations3.cpp:80-110)
// mode: 0 = store force only, 1 = + predict(k=0) with f_p = 0,
//       2 = + the warm start's pressure force f_p(p0) at rho* = rho, whose
//           (x, q0 = p0 / rho^2) the bucket kernel wrote to xq, and
//           predict(k=0) with it
// LDS: neighbour position (mode 2: with q0), (velocity, density).
//
// Two lanes per particle: the two LDS arrays allow only two workgroups per
// CU, so each particle's bucket is split between lanes 2p and 2p + 1 (even
// and odd quads) to double the waves in flight; the two partial sums are
// added with one lane exchange.
// ---------------------------------------------------------------------------
constexpr int kNpThreads = 2 * kBlock;

template <int kMode>
__global__ void __launch_bounds__(kNpThreads) __attribute__((amdgpu_waves_per_eu(4))) k_nonpressure(SimConst c, StepConst st, NbrTab nb,
                                                            const float4 *__restrict__ pos,
                                                            const float4 *__restrict__ vel,
                                                            const float *__restrict__ rho,
                                                            const float4 *__restrict__ xq,
                                                            float4 *__restrict__ force,
                                                            float4 *__restrict__ xstar,
                                                            float *__restrict__ rho_star) {
    constexpr bool kWarm = kMode == 2;
    __shared__ float4 sx[kTileCap2];
    __shared__ float4 sv[kTileCap2];
    __shared__ TileTab tt;
    const int tile = tile_of_block();
    tab_load_shared(nb.tab + (size_t)tile * kTabInts, tt);
    __syncthreads();
    const bool staged = tt.staged != 0;                  // buckets hold LDS-local indices
    const bool in_lds = staged && tt.total <= c.tile_cap2;  // and they fit this kernel's LDS
    const int pl = threadIdx.x >> 1, half = threadIdx.x & 1;
    const int i = tile * kBlock + pl;
    const bool valid = i < c.n;
    int cnt = 0;
    float4 xi = make_float4(0.f, 0.f, 0.f, 0.f), vi = xi;
    float rhoi = 0.f, qi = 0.f;
    // (x, q0) in the warm mode, else the positions
    const float4 *__restrict__ sxsrc = kWarm ? xq : pos;
    // this lane's first two quads: half, half + 2
    NbrQuads<2> first{};
    if (valid) {  // independent of the staging: requested before it
        cnt = nb.cnt[i];
        xi = pos[i];
        vi = vel[i];
        rhoi = rho[i];
        if (kWarm) qi = xq[i].w;
        first = nbr_prefetch<2, 2>(nb.n16, nb.n32, staged, i, c.stride, half);
    }
    // staged through registers: the (velocity, 1 / density) pair is assembled on
    // the way, and seven 4-byte LDS DMA streams measured slower here; both
    // arrays in one pass, so each slot's storage index is found once
    if (in_lds) {
        struct Slot {
            float4 a, b;
        };
        tab_stage_with(
            tt,
            [&](int j) {
                const float4 v = vel[j];
                return Slot{sxsrc[j], make_float4(v.x, v.y, v.z, fast_rcp(rho[j]))};  // 1 / rho_j once
            },
            [&](int e, const Slot &v) {
                sx[e] = v.a;
                sv[e] = v.b;
            });
    }
    __syncthreads();
    if (!valid) return;  // the lane pair shares validity
    const float coef = c.visc * c.mass2;
    const float kg = c.mass2 * c.spiky_grad;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    float px = 0.f, py = 0.f, pz = 0.f, epx = 0.f, epy = 0.f, epz = 0.f;  // warm: Kahan f_p
    with_staging(in_lds, [&](auto tag) {
        constexpr bool kL = decltype(tag)::value;
        auto body = [&](int j, bool live) {
            float4 a, b;
            if (kL) {
                a = lds_at(sx, j);
                b = lds_at(sv, j);
            } else {
                const int g = staged ? tab_global(tt, j >> kSlotShift) : j;
                a = sxsrc[g];
                const float4 v = vel[g];
                b = make_float4(v.x, v.y, v.z, fast_rcp(rho[g]));
            }
            const float dx = xi.x - a.x, dy = xi.y - a.y, dz = xi.z - a.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            const float r = fast_sqrt(d2);
            // masks instead of branches: every pair's arithmetic runs, a
            // dead pair contributes exact zeros, and the neighbour's LDS
            // reads stay unconditional (so they are issued early)
            const bool in = live & spiky_in(c, r);
            const float x = keep_if(in, 1.f - r * c.inv_h);
            // viscosity term mu m^2 (v_j - v_i) d2W(r) / rho_j
            const float s = coef * (c.spiky_lap * x) * b.w;
            fx += (b.x - vi.x) * s;
            fy += (b.y - vi.y) * s;
            fz += (b.z - vi.z) * s;
            if constexpr (kWarm) {
                // PredictPressureForceFor (src/equations/pcisph_equations3.cpp:114-155)
                // at the current positions with rho* = rho: pairs j != i with
                // 1e-8 <= r inside the spiky cutoff, the pressure-force
                // kernel's pair rule (r = d2 rsqrt(d2) is NaN for d2 = 0);
                // with d = x_i - x_j here, f_p = +sum g d
                const float ir = fast_rsq(d2);
                const float rp = d2 * ir;
                const bool inp = live & spiky_in(c, rp) & (rp >= 1e-8f);
                const float xp = 1.f - rp * c.inv_h;
                const float g = keep_if(inp, kg * (qi + a.w) * xp * xp * ir);
                kahan_add(px, epx, g * dx);
                kahan_add(py, epy, g * dy);
                kahan_add(pz, epz, g * dz);
            }
        };
        // quads half, half + 2, half + 4, ... of the bucket, two in flight
        for_each_neighbour_from<2, 2>(first, nb.n16, nb.n32, staged, i, cnt, c.stride, body, half);
    });
    // the pair's partial sums
    fx += __shfl_xor(fx, 1, 64);
    fy += __shfl_xor(fy, 1, 64);
    fz += __shfl_xor(fz, 1, 64);
    fx += c.mass * c.gx - c.drag * vi.x;
    fy += c.mass * c.gy - c.drag * vi.y;
    fz += c.mass * c.gz - c.drag * vi.z;
    float fpx = 0.f, fpy = 0.f, fpz = 0.f;
    if (kWarm) {
        // compensated partial sums (sum - compensation), then the pair's
        fpx = px - epx;
        fpy = py - epy;
        fpz = pz - epz;
        fpx += __shfl_xor(fpx, 1, 64);
        fpy += __shfl_xor(fpy, 1, 64);
        fpz += __shfl_xor(fpz, 1, 64);
    }
    if (half == 0) {
        force[i] = make_float4(fx, fy, fz, 0.f);
        if (kMode >= 1) {
            xstar[i] = predict_position(c, st, xi, vi, fx + fpx, fy + fpy, fz + fpz);
            rho_star[i] = rhoi;
        }
    }
}

// ---------------------------------------------------------------------------
// predicted density and pressure update (PredictPressureFor,
// src/equations/pcisph_equations3.cpp:60-90) + FP64 max-error reduction
// into the error slot of iteration k.  Does nothing once an earlier
// iteration integrated.  LDS: predicted positions.
// ---------------------------------------------------------------------------
// Two lanes per particle (quads half, half + 2, ...), 512 threads per tile:
// twice the waves of one lane per particle at the same LDS, to hide the
// gather's latency.
constexpr int kPpThreads = 2 * kBlock;
// Quads in flight per lane in the predicted-density gather: 1 measured 2%
// faster than 2 over the whole 1M run, and 3 slower still.
constexpr int kPpDepth = 1;
__global__ void __launch_bounds__(kPpThreads) k_predict_pressure(SimConst c, StepConst st, double delta,
                                                                 NbrTab nb,
                                                                 const float4 *__restrict__ xstar,
                                                                 const float4 *__restrict__ pos,
                                                                 float *__restrict__ pres,
                                                                 float *__restrict__ rho_star,
                                                                 float4 *__restrict__ xq,
                                                                 LoopState *__restrict__ loop, int k) {
    // predicted positions, copied by 16-byte LDS DMA (the w lane is the
    // unused fourth word of x*)
    __shared__ float4 sx[kTileCap];
    __shared__ unsigned long long wmax[kPpThreads / 64];
    if (k > 0 && loop->done[k - 1]) return;  // uniform early exit
    const int tile = tile_of_block();
    // Each wave reads the tile table into its own lanes, so no workgroup
    // barrier stands between the table and the staging: the particle's own
    // data and its first quads (LDS-index table; reloaded from the storage-
    // index table for an unstaged tile) are requested with it and stay in
    // flight while the staging is issued.
    // The loads are unconditional (a lane past the last particle reads the
    // last one's data and ignores it): with no branch between them and the
    // table word's first use, waiting for the table waits for it alone.
    const int tw = tab_load_wave(nb.tab + (size_t)tile * kTabInts);
    const int pl = threadIdx.x >> 1, half = threadIdx.x & 1;
    const int i = tile * kBlock + pl;
    const bool valid = i < c.n;
    const int ic = valid ? i : c.n - 1;  // the kernel runs only for n > 0
    NbrQuads<kPpDepth> first = nbr_prefetch<kPpDepth, 2>(nb.n16, nb.n32, true, ic, c.stride, half);
    const int cnt_ic = nb.cnt[ic];
    const float4 xi = xstar[ic], pi = pos[ic];
    const int cnt = valid ? cnt_ic : 0;
    const float pown = pres[ic];
    const float pold = (!st.first && half == 0) ? pown : 0.f;
    const bool staged = tab_word(tw, 1) != 0;
    if (staged) {
        tab_stage_dma_wave(tw, [&](int j, int e) { lds_dma16(xstar + j, sx + e); });
        lds_dma_wait();
    } else if (valid) {
        first = nbr_prefetch<kPpDepth, 2>(nb.n16, nb.n32, false, i, c.stride, half);
    }
    __syncthreads();
    unsigned long long ebits = 0ull;
    if (valid) {  // both lanes of a particle share validity
        float sum = 0.f, comp = 0.f;
        with_staging(staged, [&](auto tag) {
            constexpr bool kS = decltype(tag)::value;
            for_each_neighbour_from<kPpDepth, 2>(first, nb.n16, nb.n32, kS, i, cnt, c.stride, [&](int j, bool live) {
                float3 xj;
                if (kS) {
                    const float4 a = lds_at(sx, j);
                    xj = make_float3(a.x, a.y, a.z);
                } else {
                    const float4 g = xstar[j];
                    xj = make_float3(g.x, g.y, g.z);
                }
                float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                const float w = poly6_w(c, dx * dx + dy * dy + dz * dz);
                kahan_add(sum, comp, live ? w : 0.f);  // a dead entry adds an exact zero
            }, half);
        });
        // the pair's compensated partial sums; a + b == b + a, so both lanes
        // hold the same total
        const float mine = sum - comp;
        const float total = mine + __shfl_xor(mine, 1, 64);
        float density = total * c.mass;
        // FP64 error and pressure increment (delta is FP64 on the host)
        double err = (double)density - (double)c.rho0;
        double dp = delta * err;
        if (dp < 0.0) {
            dp *= (double)c.nps;
            err *= (double)c.nps;
        }
        if (half == 0) {
            float p = pold + (float)dp;
            pres[i] = p;
            rho_star[i] = density;
            float d2 = density * density;
            xq[i] = make_float4(pi.x, pi.y, pi.z, (fabsf(d2) < 1e-8f) ? 0.f : p / d2);
            // |err| with a cleared sign bit: -0.0 must not win the unsigned max
            // (the reference's AbsfMax keeps the first of equal squares)
            const double ae = fabs(err);
            ebits = (unsigned long long)__double_as_longlong(ae);
        }
    }
    ebits = wave_max_u64(ebits);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = ebits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = wmax[0];
        for (int w = 1; w < kPpThreads / 64; w++) m = wmax[w] > m ? wmax[w] : m;
        if (m) atomicMax(&loop->slot[k][blockIdx.x & (kSlots - 1)], m);
    }
}

// ---------------------------------------------------------------------------
// pressure force (PredictPressureForceFor, src/equations/pcisph_equations3.cpp:114-155)
// at the current positions with predicted densities (q = p / rho*^2); then
// either predict the next iteration's x* or accumulate and integrate
// (AccumulateForcesFor + TimeIntegrationFor).
// LDS: (position, q).
// ---------------------------------------------------------------------------
// Loop control (ComputePressureForceAndIntegrate, :218-256) is decided here
// on the device: for iteration k >= 0 the kernel integrates when the FP64
// max error of iteration k is below tol (ratio |max err| / rho0 < tol, as the
// reference computes it), when k is the last iteration, or in REFERENCE_EXACT
// mode; otherwise it predicts x* for iteration k + 1.  It does nothing when
// an earlier iteration already integrated.  k < 0 is the warm-start predict.
// Quads in flight per lane in the pressure-force gather: 2 measured faster
// than 3, 4 and 6 over the whole 1M run (0.4%, 1% and 3%).
constexpr int kPfDepth = 2;
__global__ void __launch_bounds__(kBlock) k_pressure_force(SimConst c, StepConst st, NbrTab nb,
                                                           const float4 *__restrict__ pos,
                                                           const float4 *__restrict__ vel,
                                                           float4 *force,
                                                           const float4 *__restrict__ xq,
                                                           const float *__restrict__ pres,
                                                           float4 *__restrict__ xstar,
                                                           float4 *__restrict__ fp_out,
                                                           float4 *__restrict__ pos_out,
                                                           float4 *__restrict__ vel_out,
                                                           int *__restrict__ flag,
                                                           unsigned *__restrict__ max_force,
                                                           LoopState *__restrict__ loop, int k,
                                                           int max_it, int exact, double tol,
                                                           LoopSignal *sig, unsigned seq,
                                                           const int *__restrict__ counters) {
    // (position, q = p / rho*^2) of the neighbours, copied by 16-byte LDS DMA
    __shared__ float4 sx[kTileCap];
    if (k > 0 && loop->done[k - 1]) {
        // an earlier iteration integrated: pass the flag on and exit (uniform)
        if (blockIdx.x == 0 && threadIdx.x == 0) loop->done[k] = 1;
        return;
    }
    const int tile = tile_of_block();
    const int lane = (int)threadIdx.x & 63;
    // Each wave reads the tile table and the iteration's error slots into its
    // own lanes and reduces the slots itself, so no workgroup barrier stands
    // before the staging.  The particle's own data and its first quads (LDS-
    // index table; reloaded from the storage-index table for an unstaged
    // tile) are requested with them, unconditionally (a lane past the last
    // particle reads the last one's data and ignores it), so that waiting for
    // the table or the slots waits for them alone.
    const int tw = tab_load_wave(nb.tab + (size_t)tile * kTabInts);
    const unsigned long long slot_err = (k >= 0 && lane < kSlots) ? loop->slot[k][lane] : 0ull;
    const int i = tile * kBlock + threadIdx.x;
    const bool valid = i < c.n;
    const int ic = valid ? i : c.n - 1;  // the kernel runs only for n > 0
    NbrQuads<kPfDepth> first = nbr_prefetch<kPfDepth>(nb.n16, nb.n32, true, ic, c.stride);
    const int cnt_ic = nb.cnt[ic];
    const float4 xi = pos[ic];
    const float qi = xq[ic].w;
    const float4 fi = force[ic], vi = vel[ic];
    const int cnt = valid ? cnt_ic : 0;
    const bool staged = tab_word(tw, 1) != 0;
    bool integrate = false;
    if (k >= 0) {
        // the iteration's FP64 max error: max over the slots
        const unsigned long long emax = wave_max_u64(slot_err);
        const double e = __longlong_as_double((long long)emax);
        const double ratio = e / c.rho0d;
        integrate = exact || k == max_it - 1 || fabs(ratio) < tol;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            loop->err[k] = emax;
            if (integrate) loop->done[k] = 1;
            // tell the waiting host what this iteration decided, and with
            // the integrating decision the substep's statistics (vector
            // stores, system scope, into pinned host memory; the counters
            // were final before the loop began)
            if (sig) {
                if (integrate) {
                    const unsigned long long tag = (unsigned long long)seq << 32;
                    const unsigned long long v[kPubWords] = {
                        tag | (emax & 0xffffffffull), tag | (emax >> 32),
                        tag | (unsigned)counters[CNT_OVERFLOW], tag | (unsigned)counters[CNT_LOST],
                        tag | (unsigned)counters[CNT_FULL]};
#pragma unroll
                    for (int q = 0; q < kPubWords; q++)
                        __hip_atomic_store(&sig->stats[q], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                }
                __hip_atomic_store(&sig->word, (seq << 9) | ((unsigned)k << 1) | (integrate ? 1u : 0u),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
    if (staged) {
        tab_stage_dma_wave(tw, [&](int j, int e) { lds_dma16(xq + j, sx + e); });
        lds_dma_wait();
    } else {
        first = nbr_prefetch<kPfDepth>(nb.n16, nb.n32, false, ic, c.stride);
    }
    __syncthreads();
    unsigned fbits = 0u;
    if (valid) {
        // pairs j != i with 1e-8 <= r and r inside the spiky cutoff
        // (src/equations/pcisph_equations3.cpp:128-150); r = d2 * rsqrt(d2)
        // is NaN for d2 = 0 (the particle itself or a coincident one), which
        // fails both tests
        const float kg = c.mass2 * c.spiky_grad;
        float tx = 0.f, ty = 0.f, tz = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
        with_staging(staged, [&](auto tag) {
            constexpr bool kS = decltype(tag)::value;
            for_each_neighbour_from<kPfDepth>(first, nb.n16, nb.n32, kS, i, cnt, c.stride, [&](int j, bool live) {
                const float4 a = kS ? lds_at(sx, j) : xq[j];  // (x_j, q_j)
                const float qj = a.w;
                const float dx = a.x - xi.x, dy = a.y - xi.y, dz = a.z - xi.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                const float ir = fast_rsq(d2);
                const float r = d2 * ir;
                const bool in = live && spiky_in(c, r) && r >= 1e-8f;
                const float x = 1.f - r * c.inv_h;
                // gradW(r, dir) = -dW(r) dir = spiky_grad x^2 (x_j - x_i)/r;
                // a dead pair adds an exact zero.  Masking the bits (rather
                // than ?:) keeps the pair's arithmetic unconditional, so the
                // compiler issues the neighbour's LDS read as one 16-byte
                // load and does not branch per pair.
                const float g = keep_if(in, kg * (qi + qj) * x * x * ir);
                kahan_add(tx, ex, g * dx);
                kahan_add(ty, ey, g * dy);
                kahan_add(tz, ez, g * dz);
            });
        });
        const float fpx = -tx, fpy = -ty, fpz = -tz;
        if (!integrate) {
            xstar[i] = predict_position(c, st, xi, vi, fi.x + fpx, fi.y + fpy, fi.z + fpz);
        } else {
            const float Fx = fi.x + fpx, Fy = fi.y + fpy, Fz = fi.z + fpz;
            float3 p, v;
            integrate_one(c, st.dt, xi, vi, Fx, Fy, Fz, p, v, flag);
            pos_out[i] = make_float4(p.x, p.y, p.z, xi.w);
            vel_out[i] = make_float4(v.x, v.y, v.z, pres[i]);
            force[i] = make_float4(Fx, Fy, Fz, 0.f);  // the lane read force[i] above
            fp_out[i] = make_float4(fpx, fpy, fpz, 0.f);
            fbits = __float_as_uint(sqrtf(Fx * Fx + Fy * Fy + Fz * Fz));
        }
    }
    if (integrate) block_max_to_slot(fbits, max_force);
}

// ---------------------------------------------------------------------------
// SPH (WCSPH) forces: ComputeAllForcesFor without the in-loop integration
// (src/equations/sph_equations3.cpp:184-278); integration is a separate pass.
// LDS: (position, p / rho^2), (velocity, density).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_sph_force(SimConst c, NbrTab nb,
                                                      const float4 *__restrict__ pos,
                                                      const float4 *__restrict__ vel,
                                                      const float *__restrict__ rho,
                                                      const float *__restrict__ pres,
                                                      float4 *__restrict__ force) {
    __shared__ float4 sx[kTileCap2];
    __shared__ float4 sv[kTileCap2];
    __shared__ TileTab tt;
    const int tile = tile_of_block();
    tab_load_shared(nb.tab + (size_t)tile * kTabInts, tt);
    __syncthreads();
    const bool staged = tt.staged != 0;
    const bool in_lds = staged && tt.total <= c.tile_cap2;
    auto xq = [&](int j) {
        float4 p = pos[j];
        float d = rho[j];
        return make_float4(p.x, p.y, p.z, pres[j] / (d * d));
    };
    auto vr = [&](int j) {
        float4 v = vel[j];
        return make_float4(v.x, v.y, v.z, rho[j]);
    };
    if (in_lds) {
        tab_stage(tt, sx, xq);
        tab_stage(tt, sv, vr);
    }
    __syncthreads();
    const int i = tile * kBlock + threadIdx.x;
    if (i >= c.n) return;
    const int self = staged ? tab_local(tt, i) << kSlotShift : i;
    const float4 xi = pos[i], vi = vel[i];
    const float di = rho[i];
    const float qi = pres[i] / (di * di);
    const float coef = c.visc * c.mass2;
    float fx = c.mass * c.gx - c.drag * vi.x;
    float fy = c.mass * c.gy - c.drag * vi.y;
    float fz = c.mass * c.gz - c.drag * vi.z;
    float tx = 0.f, ty = 0.f, tz = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
    const int cnt = nb.cnt[i];
    for_each_neighbour<2>(nb.n16, nb.n32, staged, i, cnt, c.stride, [&](int j, bool live) {
        // masks, not early returns: the pair's loads and arithmetic are
        // unconditional and a skipped pair adds exact zeros
        float4 xj, vj;
        if (in_lds) {
            xj = lds_at(sx, j);
            vj = lds_at(sv, j);
        } else {
            const int g = staged ? tab_global(tt, j >> kSlotShift) : j;
            xj = xq(g);
            vj = vr(g);
        }
        const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        const bool in = live & (j != self) & spiky_in(c, r);
        const float x = 1.f - r * c.inv_h;
        const bool pg = in & (r > 0.f) & !(fabsf(r) < 1e-8f);
        const float g = keep_if(pg, c.mass2 * (qi + xj.w) * c.spiky_grad * x * x / r);
        kahan_add(tx, ex, g * dx);
        kahan_add(ty, ey, g * dy);
        kahan_add(tz, ez, g * dz);
        const float s = keep_if(in, coef * c.spiky_lap * x / vj.w);
        fx += (vj.x - vi.x) * s;
        fy += (vj.y - vi.y) * s;
        fz += (vj.z - vi.z) * s;
    });
    force[i] = make_float4(fx - tx, fy - ty, fz - tz, 0.f);
}

// TimeIntegrationFor as a separate pass (SPH two-phase form)
__global__ void __launch_bounds__(kBlock) k_integrate(SimConst c, StepConst st,
                                                      const float4 *__restrict__ pos,
                                                      const float4 *__restrict__ vel,
                                                      const float4 *__restrict__ force,
                                                      float4 *__restrict__ pos_out,
                                                      float4 *__restrict__ vel_out,
                                                      int *__restrict__ flag,
                                                      unsigned *__restrict__ max_force) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned fbits = 0u;
    if (i < c.n) {
        const float4 xi = pos[i], vi = vel[i], fi = force[i];
        float3 p, v;
        integrate_one(c, st.dt, xi, vi, fi.x, fi.y, fi.z, p, v, flag);
        pos_out[i] = make_float4(p.x, p.y, p.z, xi.w);
        vel_out[i] = make_float4(v.x, v.y, v.z, vi.w);
        fbits = __float_as_uint(sqrtf(fi.x * fi.x + fi.y * fi.y + fi.z * fi.z));
    }
    block_max_to_slot(fbits, max_force);
}

// XSPH velocity smoothing (src/equations/sph_equations3.cpp:341-382), active
// only when pseudoViscosity * dt > 0.1.  LDS: position, (velocity, density).
__global__ void __launch_bounds__(kBlock) k_xsph(SimConst c, StepConst st, NbrTab nb,
                                                 const float4 *__restrict__ pos,
                                                 const float4 *__restrict__ vel,
                                                 const float *__restrict__ rho,
                                                 float4 *__restrict__ vel_out) {
    __shared__ float4 sx[kTileCap2];
    __shared__ float4 sv[kTileCap2];
    __shared__ TileTab tt;
    const int tile = tile_of_block();
    tab_load_shared(nb.tab + (size_t)tile * kTabInts, tt);
    __syncthreads();
    const bool staged = tt.staged != 0;
    const bool in_lds = staged && tt.total <= c.tile_cap2;
    auto vr = [&](int j) {
        float4 v = vel[j];
        return make_float4(v.x, v.y, v.z, rho[j]);
    };
    if (in_lds) {
        tab_stage(tt, sx, [&](int j) { return pos[j]; });
        tab_stage(tt, sv, vr);
    }
    __syncthreads();
    const int i = tile * kBlock + threadIdx.x;
    if (i >= c.n) return;
    const float4 xi = pos[i], vi = vel[i];
    float sx_ = 0.f, sy_ = 0.f, sz_ = 0.f, ws = 0.f;
    const float spiky_w = c.spiky_lap * c.h * c.h / 6.f;  // 15 / (pi h^3)
    const int cnt = nb.cnt[i];
    for_each_neighbour<2>(nb.n16, nb.n32, staged, i, cnt, c.stride, [&](int j, bool live) {
        float4 xj, vj;
        if (in_lds) {
            xj = lds_at(sx, j);
            vj = lds_at(sv, j);
        } else {
            const int g = staged ? tab_global(tt, j >> kSlotShift) : j;
            xj = pos[g];
            vj = vr(g);
        }
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float r = sqrtf(dx * dx + dy * dy + dz * dz);
        const float x = 1.f - r * c.inv_h;
        const float w = keep_if(live & spiky_in(c, r), spiky_w * x * x * x);
        const float wj = keep_if(live, c.mass / vj.w * w);
        ws += wj;
        sx_ += wj * vj.x;
        sy_ += wj * vj.y;
        sz_ += wj * vj.z;
    });
    if (ws > 0.f) {
        float inv = 1.f / ws;
        sx_ *= inv;
        sy_ *= inv;
        sz_ *= inv;
    }
    float f = fminf(fmaxf(st.dt * c.pseudo_visc, 0.f), 1.f);
    vel_out[i] = make_float4((1.f - f) * vi.x + f * sx_, (1.f - f) * vi.y + f * sy_,
                             (1.f - f) * vi.z + f * sz_, vi.w);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static NbrTab nbr_tab(const bbx_solver *s) {
    return NbrTab{s->nbr16, s->nbr32, s->nbr_cnt, s->tile_tab};
}

int run_nonpressure(bbx_solver *s, const StepConst &st, int mode) {
    if (s->n == 0) return BBX_OK;
    const unsigned nb = nblocks(s->n);
    const NbrTab t = nbr_tab(s);
#define BBX_NP(M)                                                                                     \
    k_nonpressure<M><<<nb, kNpThreads, 0, s->stream>>>(s->sc, st, t, s->posA, s->velA, s->rho, s->xq, \
                                                     s->force, s->xstar, s->rho_star)
    if (mode == 2) {
        BBX_PROF_BEGIN("nonpressure_warm");
        BBX_NP(2);
        BBX_HIP_CHECK(hipGetLastError());
        BBX_PROF_END();
    } else if (mode == 1) {
        BBX_PROF_BEGIN("nonpressure_predict");
        BBX_NP(1);
        BBX_HIP_CHECK(hipGetLastError());
        BBX_PROF_END();
    } else {
        BBX_PROF_BEGIN("nonpressure");
        BBX_NP(0);
        BBX_HIP_CHECK(hipGetLastError());
        BBX_PROF_END();
    }
#undef BBX_NP
    return BBX_OK;
}

int run_predict_pressure(bbx_solver *s, const StepConst &st, int k) {
    if (s->n == 0) return BBX_OK;
    if (k < 0 || k >= kMaxIt) return fail(BBX_E_INVALID, "iteration index out of range");
    // the pressure increment uses the host's FP64 delta (s->delta64)
    BBX_PROF_BEGIN_IT("predict_pressure", k);
    k_predict_pressure<<<nblocks(s->n), kPpThreads, 0, s->stream>>>(s->sc, st, s->delta64, nbr_tab(s), s->xstar,
                                                                s->posA, s->pres, s->rho_star, s->xq,
                                                                s->loop, k);
    BBX_HIP_CHECK(hipGetLastError());
    BBX_PROF_END();
    return BBX_OK;
}

int run_pressure_force(bbx_solver *s, const StepConst &st, int k, int max_it, int exact, double tol) {
    if (s->n == 0) return BBX_OK;
    if (k >= kMaxIt || max_it > kMaxIt) return fail(BBX_E_INVALID, "iteration index out of range");
    BBX_PROF_BEGIN_IT("pressure_force", k < 0 ? -1 : k);
    k_pressure_force<<<nblocks(s->n), kBlock, 0, s->stream>>>(
        s->sc, st, nbr_tab(s), s->posA, s->velA, s->force, s->xq, s->pres, s->xstar, s->fp, s->posB,
        s->velB, s->flag, s->max_force, s->loop, k < 0 ? -1 : k, max_it, exact, tol, s->d_signal,
        s->loop_seq & 0x7fffffu, s->counters);
    BBX_HIP_CHECK(hipGetLastError());
    BBX_PROF_END();
    return BBX_OK;
}

int run_sph_force(bbx_solver *s) {
    if (s->n == 0) return BBX_OK;
    BBX_PROF_BEGIN("sph_force");
    k_sph_force<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, nbr_tab(s), s->posA, s->velA, s->rho,
                                                         s->pres, s->force);
    BBX_HIP_CHECK(hipGetLastError());
    BBX_PROF_END();
    return BBX_OK;
}

int run_integrate(bbx_solver *s, const StepConst &st) {
    if (s->n == 0) return BBX_OK;
    BBX_PROF_BEGIN("integrate");
    k_integrate<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, st, s->posA, s->velA, s->force,
                                                         s->posB, s->velB, s->flag, s->max_force);
    BBX_HIP_CHECK(hipGetLastError());
    BBX_PROF_END();
    return BBX_OK;
}

int swap_after_integrate(bbx_solver *s) {
    std::swap(s->posA, s->posB);
    std::swap(s->velA, s->velB);
    return BBX_OK;
}

int run_xsph(bbx_solver *s, const StepConst &st) {
    if (s->n == 0) return BBX_OK;
    if (!(s->prm.pseudo_viscosity * (double)st.dt > 0.1)) return BBX_OK;
    BBX_PROF_BEGIN("xsph");
    k_xsph<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, st, nbr_tab(s), s->posA, s->velA, s->rho,
                                                    s->velB);
    BBX_HIP_CHECK(hipGetLastError());
    BBX_PROF_END();
    std::swap(s->velA, s->velB);
    return BBX_OK;
}


}  // namespace bbx
