// bbx_device.h -- device-side helpers: FP64 hash and bucket rule with the
// reference's rounding, FP32 SPH kernels, the reversed-box collider and
// wave64 reductions.  Included by the .hip translation units only.
#pragma once

#include "bbx_common.h"

namespace bbx {

__device__ __forceinline__ int pid_of(float4 p) { return __float_as_int(p.w); }

// Grid::GetHashedPosition per axis (src/core/grid.h:258-298): ExtremeEpsilon
// adds +-Epsilon (the float literal 0.0001f) only within 1e-8 of the bounds,
// then floor((p - min) / len).  IEEE FP64, no contraction.
__device__ __forceinline__ int hash_axis(double p, double bmin, double bmax, double minp,
                                         double len, int n, int *bad) {
#pragma clang fp contract(off)
    double eps = 0.0;
    double d0 = p - bmin;
    double d1 = p - bmax;
    if ((d0 > 0 ? d0 : -d0) < 1e-8) eps = (double)0.0001f;
    else if ((d1 > 0 ? d1 : -d1) < 1e-8) eps = -(double)0.0001f;
    p += eps;
    double dp = (p - minp) / len;
    double f = floor(dp);
    if (!(f >= 0.0)) { *bad = 1; return 0; }
    if (f >= (double)n) { *bad = 1; return n - 1; }
    return (int)f;
}

// Returns the linear cell id (x + y*nx + z*nx*ny) and the 3-D index; *bad is
// set when the point hashes outside the grid (the index is clamped).
__device__ __forceinline__ int hash_cell(const GridGeom &g, float4 p, int u[3], int *bad) {
    u[0] = hash_axis((double)p.x, g.bmin[0], g.bmax[0], g.minp[0], g.len[0], g.n[0], bad);
    u[1] = hash_axis((double)p.y, g.bmin[1], g.bmax[1], g.minp[1], g.len[1], g.n[1], bad);
    u[2] = hash_axis((double)p.z, g.bmin[2], g.bmax[2], g.minp[2], g.len[2], g.n[2], bad);
    return u[0] + u[1] * g.n[0] + u[2] * g.n[0] * g.n[1];
}

__device__ __forceinline__ void unlinear(const GridGeom &g, int id, int u[3]) {
    int plane = g.n[0] * g.n[1];
    u[2] = id / plane;
    int rem = id - u[2] * plane;
    u[1] = rem / g.n[0];
    u[0] = rem - u[1] * g.n[0];
}

// unlinear with the two divisions replaced by FP32 reciprocal products and
// one correction each: the FP32 quotient of a cell id (< 2^31) by n is off
// by less than one for every grid whose quotient fits 2^22, so one +-1 step
// makes it exact.
__device__ __forceinline__ void unlinear_fast(const GridGeom &g, float inv_nx, float inv_nxy, int id,
                                              int u[3]) {
    const int nx = g.n[0], nxy = g.n[0] * g.n[1];
    int z = (int)((float)id * inv_nxy);
    int rem = id - z * nxy;
    if (rem < 0) {
        z--;
        rem += nxy;
    } else if (rem >= nxy) {
        z++;
        rem -= nxy;
    }
    int y = (int)((float)rem * inv_nx);
    int x = rem - y * nx;
    if (x < 0) {
        y--;
        x += nx;
    } else if (x >= nx) {
        y++;
        x -= nx;
    }
    u[0] = x;
    u[1] = y;
    u[2] = z;
}

// IsWithinStd(Distance(pi, pj), h) in FP64 (src/core/kernel.cpp:229-234,
// src/core/geometry.h:632-633): d = sqrt(dx^2+dy^2+dz^2), d2 = d*d,
// inside iff !(|d2 - h2| < 1e-8 || d2 - h2 > 0).
// Not inlined: only pairs inside the FP32 band around the cutoff reach it.
__device__ __attribute__((noinline)) bool within_std_fp64(float4 a, float4 b, double h2) {
#pragma clang fp contract(off)
    double dx = (double)a.x - (double)b.x;
    double dy = (double)a.y - (double)b.y;
    double dz = (double)a.z - (double)b.z;
    double s = dx * dx + dy * dy + dz * dz;
    double d = __builtin_sqrt(s);
    double d2 = d * d;
    double of = d2 - h2;
    double aof = of > 0 ? of : -of;
    return !(aof < 1e-8 || of > 0);
}

// Bucket membership: FP32 squared distance decides outside a narrow band
// around the FP64 cutoff; inside the band the exact FP64 rule decides.
__device__ __forceinline__ bool in_bucket(const SimConst &c, float4 a, float4 b) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < c.lo2) return true;
    if (d2 > c.hi2) return false;
    return within_std_fp64(a, b, c.h2d);
}

// |a - b|^2 - off in FP32 with the offset folded into the first FMA.  Every
// rounding acts on a value of magnitude <= max(|a - b|^2, off), so the
// result is within 5 * 2^-24 * max(|a - b|^2, off) of the exact value: with
// off = band_c, a result outside +-band_w has the sign of the exact
// |a - b|^2 - (h^2 - 1e-8), i.e. decides the cutoff as the FP64 rule does.
__device__ __forceinline__ float cut_offset(float4 a, float4 b, float off) {
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    return fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, -off)));
}

// The reference's cutoff rule "zero when |of| < 1e-8 or of > 0"
// (src/core/kernel.cpp:123-132,181-226) is one comparison: for any
// non-NaN of it is "inside iff of <= -1e-8"; a NaN is outside.
__device__ __forceinline__ bool cutoff_in(float of) { return of <= -1e-8f; }

// poly6 W with its own cutoff (src/core/kernel.cpp:123-132), FP32.
__device__ __forceinline__ float poly6_w(const SimConst &c, float d2) {
    const float x = 1.f - d2 * c.inv_h2;
    return cutoff_in(d2 - c.h2) ? c.poly6 * x * x * x : 0.f;
}

// poly6 W of a bucket member at the positions its bucket was built from:
// the FP64 rule put it at d^2 <= h^2 - 1e-8, where the FP32 d2 is below h^2
// by far more than its rounding, so the kernel's own cutoff can only drop a
// term of relative size (1e-8 / h^2)^3 and is not evaluated.
__device__ __forceinline__ float poly6_w_member(const SimConst &c, float d2) {
    const float x = 1.f - d2 * c.inv_h2;
    return c.poly6 * x * x * x;
}

// spiky cutoff (src/core/kernel.cpp:181-226): zero when |r-h| < 1e-8 or r > h
__device__ __forceinline__ bool spiky_in(const SimConst &c, float r) { return cutoff_in(r - c.h); }

// v if keep, else +0 (also for a NaN or infinite v), without a branch
__device__ __forceinline__ float keep_if(bool keep, float v) {
    return __uint_as_float(__float_as_uint(v) & (0u - (unsigned)keep));
}

// Hardware reciprocal, square root and reciprocal square root (about 1 ulp,
// one instruction each) for the gather sums; the FP64 bucket rule and the
// hash never use them.
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Kahan compensated add (FP32)
__device__ __forceinline__ void kahan_add(float &s, float &comp, float v) {
    float y = v - comp;
    float t = s + y;
    comp = (t - s) - y;
    s = t;
}

// Reversed-box collider (ColliderSet3 with one container; CollisionHandle,
// BoxClosestPoint, IsInside: src/core/collider.cpp:3-44,238-264,
// src/shapes/box.cpp:74-172, src/core/shape.cpp:275-282), FP32.
__device__ __forceinline__ bool inside_box(float px, float py, float pz, const float lo[3],
                                           const float hi[3]) {
    bool rv = px >= lo[0] && px <= hi[0] && py >= lo[1] && py <= hi[1] && pz >= lo[2] &&
              pz <= hi[2];
    if (!rv) {
        float x = fminf(fabsf(lo[0] - px), fabsf(hi[0] - px));
        float y = fminf(fabsf(lo[1] - py), fabsf(hi[1] - py));
        float z = fminf(fabsf(lo[2] - pz), fabsf(hi[2] - pz));
        rv = x < 1e-6f || y < 1e-6f || z < 1e-6f;
    }
    return rv;
}

// Written with scalars and selects only: a run-time index into a local
// array (the chosen face) would put the array in scratch memory.
__device__ __forceinline__ void collide(const SimConst &c, float e, float3 &p, float3 &v) {
    const float hx = c.tank_h[0], hy = c.tank_h[1], hz = c.tank_h[2];
    const float lo[3] = {-hx, -hy, -hz};
    const float hi[3] = {hx, hy, hz};
    const float lx = p.x - c.tank_c[0], ly = p.y - c.tank_c[1], lz = p.z - c.tank_c[2];
    float clx, cly, clz;
    int axis = 0;
    float sgn = 1.f;  // outward normal of the chosen face: sgn * e_axis
    float dist;
    if (inside_box(lx, ly, lz, lo, hi)) {
        // nearest of the six planes, order +x,+y,+z,-x,-y,-z, strict <
        const float d[6] = {hx - lx, hy - ly, hz - lz, -hx - lx, -hy - ly, -hz - lz};
        float best = d[0] * d[0];
#pragma unroll
        for (int f = 1; f < 6; f++) {
            const float d2 = d[f] * d[f];
            if (d2 < best) {
                best = d2;
                axis = f % 3;
                sgn = (f < 3) ? 1.f : -1.f;
            }
        }
        const float face = sgn > 0.f ? (axis == 0 ? hx : (axis == 1 ? hy : hz))
                                     : (axis == 0 ? -hx : (axis == 1 ? -hy : -hz));
        clx = axis == 0 ? face : lx;
        cly = axis == 1 ? face : ly;
        clz = axis == 2 ? face : lz;
        dist = -sqrtf(best);
    } else {
        clx = fminf(fmaxf(lx, -hx), hx);
        cly = fminf(fmaxf(ly, -hy), hy);
        clz = fminf(fmaxf(lz, -hz), hz);
        const float cov[3] = {lx - clx, ly - cly, lz - clz};
        float maxc = cov[0];
#pragma unroll
        for (int f = 1; f < 6; f++) {
            const float cs = (f < 3) ? cov[f % 3] : -cov[f % 3];
            if (cs > maxc) {
                maxc = cs;
                axis = f % 3;
                sgn = (f < 3) ? 1.f : -1.f;
            }
        }
        const float dx = clx - lx, dy = cly - ly, dz = clz - lz;
        dist = sqrtf(dx * dx + dy * dy + dz * dz);
    }
    // reversed orientation: the normal points into the tank
    const float nx = axis == 0 ? -sgn : 0.f;
    const float ny = axis == 1 ? -sgn : 0.f;
    const float nz = axis == 2 ? -sgn : 0.f;
    // IsInside(reversed) = !(ClosestDistance < 0); penetrating if that or |d| < r
    const bool is_inside = !(dist < 0.f);
    if (!(is_inside || fabsf(dist) < c.radius)) return;
    const float tx = clx + c.tank_c[0] + c.radius * nx;
    const float ty = cly + c.tank_c[1] + c.radius * ny;
    const float tz = clz + c.tank_c[2] + c.radius * nz;
    const float ndr = nx * v.x + ny * v.y + nz * v.z;
    if (ndr < 0.f) {
        // relN = ndr * n -> -e * relN; relT unchanged (friction 0)
        const float rn = ndr * (-e) - ndr;  // change along n
        v.x += rn * nx;
        v.y += rn * ny;
        v.z += rn * nz;
    }
    p.x = tx;
    p.y = ty;
    p.z = tz;
}

// wave64 reductions
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
    for (int off = 32; off > 0; off >>= 1) {
        unsigned o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ int wave_min_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_sum_i32(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace bbx
