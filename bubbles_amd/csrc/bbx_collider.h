// bbx_collider.h -- device colliders: boxes and spheres under an affine
// transform, SDF grids (meshes collide through theirs), and the collider-set
// rule that picks one collider per particle.  FP32, one thread per particle.
// Included by the .hip translation units only.
#pragma once

#include <cfloat>

#include "bbx_common.h"

namespace bbx {

// ClosestPointQuery (src/core/collider.h:21-34)
struct Cpq {
    float3 p, n;
    float d;
    float3 v;
};

__device__ __forceinline__ float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float len3(float3 a) { return sqrtf(dot3(a, a)); }

// Transform::Point and ::Normal (src/core/transform.h:433-468): points by
// the matrix, normals by the transpose of the inverse
__device__ __forceinline__ float3 xf_point(const float *m, float3 p) {
    return f3(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
__device__ __forceinline__ float3 xf_normal(const float *mi, float3 n) {
    return f3(mi[0] * n.x + mi[4] * n.y + mi[8] * n.z, mi[1] * n.x + mi[5] * n.y + mi[9] * n.z,
              mi[2] * n.x + mi[6] * n.y + mi[10] * n.z);
}

// Inside(p, bounds) with the IsUnsafeZero margin (src/core/geometry.h:1774-1784)
__device__ __forceinline__ bool inside_bounds(float3 p, const float lo[3], const float hi[3]) {
    bool rv = p.x >= lo[0] && p.x <= hi[0] && p.y >= lo[1] && p.y <= hi[1] && p.z >= lo[2] && p.z <= hi[2];
    if (!rv) {
        const float x = fminf(fabsf(lo[0] - p.x), fabsf(hi[0] - p.x));
        const float y = fminf(fabsf(lo[1] - p.y), fabsf(hi[1] - p.y));
        const float z = fminf(fabsf(lo[2] - p.z), fabsf(hi[2] - p.z));
        rv = x < 1e-6f || y < 1e-6f || z < 1e-6f;
    }
    return rv;
}

// ---------------------------------------------------------------------------
// box (src/shapes/box.cpp:74-172): closest point on the six face planes
// inside (order +x,+y,+z,-x,-y,-z, strict <), the clamped point outside with
// the face whose normal has the largest cosine to (p - closest)
// ---------------------------------------------------------------------------
__device__ __forceinline__ Cpq box_query(const DevCollider &c, float3 pw) {
    const float3 pl = xf_point(c.mi, pw);
    const float hx = c.hx, hy = c.hy, hz = c.hz;
    const float lo[3] = {-hx, -hy, -hz}, hi[3] = {hx, hy, hz};
    float3 cl;
    int axis = 0;
    float sgn = 1.f;
    float dist;
    if (inside_bounds(pl, lo, hi)) {
        const float d0 = hx - pl.x, d1 = hy - pl.y, d2 = hz - pl.z;
        const float d3 = -hx - pl.x, d4 = -hy - pl.y, d5 = -hz - pl.z;
        float best = d0 * d0;
        if (d1 * d1 < best) { best = d1 * d1; axis = 1; }
        if (d2 * d2 < best) { best = d2 * d2; axis = 2; }
        if (d3 * d3 < best) { best = d3 * d3; axis = 0; sgn = -1.f; }
        if (d4 * d4 < best) { best = d4 * d4; axis = 1; sgn = -1.f; }
        if (d5 * d5 < best) { best = d5 * d5; axis = 2; sgn = -1.f; }
        const float face = sgn > 0.f ? (axis == 0 ? hx : (axis == 1 ? hy : hz))
                                     : (axis == 0 ? -hx : (axis == 1 ? -hy : -hz));
        cl = f3(axis == 0 ? face : pl.x, axis == 1 ? face : pl.y, axis == 2 ? face : pl.z);
        dist = -len3(sub3(cl, pl));
    } else {
        cl = f3(fminf(fmaxf(pl.x, -hx), hx), fminf(fmaxf(pl.y, -hy), hy), fminf(fmaxf(pl.z, -hz), hz));
        const float3 cov = sub3(pl, cl);
        float maxc = cov.x;
        if (cov.y > maxc) { maxc = cov.y; axis = 1; }
        if (cov.z > maxc) { maxc = cov.z; axis = 2; }
        if (-cov.x > maxc) { maxc = -cov.x; axis = 0; sgn = -1.f; }
        if (-cov.y > maxc) { maxc = -cov.y; axis = 1; sgn = -1.f; }
        if (-cov.z > maxc) { maxc = -cov.z; axis = 2; sgn = -1.f; }
        dist = len3(sub3(cl, pl));
    }
    const float ns = c.reverse ? -sgn : sgn;
    const float3 n = f3(axis == 0 ? ns : 0.f, axis == 1 ? ns : 0.f, axis == 2 ? ns : 0.f);
    Cpq q;
    q.p = xf_point(c.m, cl);
    q.n = xf_normal(c.mi, n);
    q.d = dist;
    q.v = f3(0.f, 0.f, 0.f);  // BoxClosestPoint passes no velocity
    return q;
}

// ---------------------------------------------------------------------------
// sphere (src/shapes/sphere.cpp:47-70, VelocityAt src/core/shape.cpp:300-305)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float sphere_distance(const DevCollider &c, float3 pw) {
    return len3(xf_point(c.mi, pw)) - c.radius;
}
__device__ __forceinline__ Cpq sphere_query(const DevCollider &c, float3 pw) {
    const float3 pl = xf_point(c.mi, pw);
    const float len = len3(pl);
    const float d = len - c.radius;
    float3 n = f3(0.f, 1.f, 0.f);
    if (!(pl.x == 0.f && pl.y == 0.f && pl.z == 0.f)) n = f3(pl.x / len, pl.y / len, pl.z / len);
    const float3 pn = mul3(n, c.radius);
    if (c.reverse) n = mul3(n, -1.f);
    Cpq q;
    q.p = xf_point(c.m, pn);
    q.n = xf_normal(c.mi, n);
    q.d = d;
    const float3 r = f3(q.p.x - c.m[3], q.p.y - c.m[7], q.p.z - c.m[11]);
    q.v = f3(c.lin[0] + (c.ang[1] * r.z - c.ang[2] * r.y), c.lin[1] + (c.ang[2] * r.x - c.ang[0] * r.z),
             c.lin[2] + (c.ang[0] * r.y - c.ang[1] * r.x));
    return q;
}

// ---------------------------------------------------------------------------
// SDF grid: FieldGrid::Sample / ::Gradient (src/core/grid.h:1093-1153),
// GetBarycentric (src/core/sampling.h:7-31), Trilerp (:38-44)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void barycentric(float x, int nodes, int &i, float &f) {
    const float s = floorf(x);
    const int high = nodes - 1;
    if (high == 0) { i = 0; f = 0.f; }
    else if (s < 0.f) { i = 0; f = 0.f; }
    else if (s > (float)(high - 1)) { i = high - 1; f = 1.f; }
    else { i = (int)s; f = x - s; }
}
__device__ __forceinline__ float lerpf(float a, float b, float t) { return (1.f - t) * a + t * b; }

__device__ __forceinline__ float sdf_sample(const DevCollider &c, float3 p) {
    int i, j, k;
    float fx, fy, fz;
    barycentric((p.x - c.ox) / c.sx, c.nx, i, fx);
    barycentric((p.y - c.oy) / c.sy, c.ny, j, fy);
    barycentric((p.z - c.oz) / c.sz, c.nz, k, fz);
    const int i1 = min(i + 1, c.nx - 1), j1 = min(j + 1, c.ny - 1), k1 = min(k + 1, c.nz - 1);
    const int sy = c.nx, sz = c.nx * c.ny;
    const float *f = c.field;
    const float f000 = f[i + j * sy + k * sz], f100 = f[i1 + j * sy + k * sz];
    const float f010 = f[i + j1 * sy + k * sz], f110 = f[i1 + j1 * sy + k * sz];
    const float f001 = f[i + j * sy + k1 * sz], f101 = f[i1 + j * sy + k1 * sz];
    const float f011 = f[i + j1 * sy + k1 * sz], f111 = f[i1 + j1 * sy + k1 * sz];
    const float b0 = lerpf(lerpf(f000, f100, fx), lerpf(f010, f110, fx), fy);
    const float b1 = lerpf(lerpf(f001, f101, fx), lerpf(f011, f111, fx), fy);
    return lerpf(b0, b1, fz);
}

__device__ __forceinline__ float3 sdf_gradient(const DevCollider &c, float3 p) {
    const float gx = (sdf_sample(c, f3(p.x + c.sx, p.y, p.z)) - sdf_sample(c, f3(p.x - c.sx, p.y, p.z))) *
                     (1.f / (2.f * c.sx));
    const float gy = (sdf_sample(c, f3(p.x, p.y + c.sy, p.z)) - sdf_sample(c, f3(p.x, p.y - c.sy, p.z))) *
                     (1.f / (2.f * c.sy));
    const float gz = (sdf_sample(c, f3(p.x, p.y, p.z + c.sz)) - sdf_sample(c, f3(p.x, p.y, p.z - c.sz))) *
                     (1.f / (2.f * c.sz));
    return f3(gx, gy, gz);
}

// ClosestPointBySDF (src/shapes/bvh.cpp:666-706): up to five steps of
// p <- p - sdf(p) grad/|grad|, stopping once |sdf| < 0.001; the distance is
// |p - start| (unsigned) and the query carries no velocity
__device__ __forceinline__ Cpq sdf_query(const DevCollider &c, float3 pw) {
    const float3 point = xf_point(c.mi, pw);
    float3 tp = point;
    float3 tn = f3(0.f, 1.f, 0.f);
    bool has_gradient = false;
    for (int it = 0; it < 5; it++) {
        const float sdf = sdf_sample(c, tp);
        if (fabsf(sdf) < 0.001f) break;
        tn = sdf_gradient(c, tp);
        const float len = len3(tn);
        if (len > 0.f) tn = f3(tn.x / len, tn.y / len, tn.z / len);
        tp = sub3(tp, mul3(tn, sdf));
        has_gradient = true;
    }
    if (!has_gradient) {
        tn = sdf_gradient(c, tp);
        const float len = len3(tn);
        if (len > 0.f) tn = f3(tn.x / len, tn.y / len, tn.z / len);
    }
    Cpq q;
    q.p = xf_point(c.m, tp);
    q.n = xf_normal(c.mi, tn);
    q.d = len3(sub3(point, tp));
    q.v = f3(0.f, 0.f, 0.f);
    return q;
}

// Shape::ClosestDistance (src/core/shape.cpp:206-228); a mesh takes its SDF
// sample (the reference walks the mesh BVH for this one value)
__device__ __forceinline__ float closest_distance(const DevCollider &c, float3 p) {
    if (c.kind == BBX_SHAPE_BOX) return box_query(c, p).d;
    if (c.kind == BBX_SHAPE_SPHERE) return sphere_distance(c, p);
    return sdf_sample(c, p);
}

// Shape::IsInside (src/core/shape.cpp:275-282): a filled SDF decides by its
// sign, other shapes by the closest distance and the orientation
__device__ __forceinline__ bool shape_inside(const DevCollider &c, float3 p) {
    if (c.kind == BBX_SHAPE_MESH || c.kind == BBX_SHAPE_SDF) return sdf_sample(c, p) < 0.f;
    return (c.reverse != 0) == !(closest_distance(c, p) < 0.f);
}

__device__ __forceinline__ Cpq closest_point(const DevCollider &c, float3 p) {
    if (c.kind == BBX_SHAPE_BOX) return box_query(c, p);
    if (c.kind == BBX_SHAPE_SPHERE) return sphere_query(c, p);
    return sdf_query(c, p);
}

// CollisionHandle + Collider3::IsPenetrating / OptmizedClosestPointCheck
// (src/core/collider.cpp:3-44,113-131)
__device__ __forceinline__ bool collision_handle(const DevCollider &c, float radius, float e, float3 &p,
                                                 float3 &v) {
    if (c.kind == BBX_SHAPE_MESH && !inside_bounds(p, c.bmin, c.bmax)) return false;
    const Cpq q = closest_point(c, p);
    if (!(shape_inside(c, p) || fabsf(q.d) < radius)) return false;
    const float3 n = q.n;
    const float3 target = add3(q.p, mul3(n, radius));
    const float3 rel = sub3(v, q.v);
    const float ndr = dot3(n, rel);
    float3 rel_n = mul3(n, ndr);
    float3 rel_t = sub3(rel, rel_n);
    if (ndr < 0.f) {
        const float3 d_rel_n = mul3(rel_n, -1.f - e);
        rel_n = mul3(rel_n, -e);
        if (dot3(rel_t, rel_t) > 0.f) {
            const float scale = len3(d_rel_n) / len3(rel_t);
            rel_t = mul3(rel_t, fmaxf(0.f, 1.f - c.friction * scale));
        }
        v = add3(add3(rel_n, rel_t), q.v);
    }
    p = target;
    return true;
}

// ColliderSet3::ResolveCollision (src/core/collider.cpp:238-264): the active
// collider with the smallest |ClosestDistance| (first on ties) handles p
__device__ __forceinline__ bool resolve_colliders(const DevCollider *__restrict__ cs, int nc, float radius,
                                                  float e, float3 &p, float3 &v) {
    int target = -1;
    float mind = FLT_MAX;
    for (int i = 0; i < nc; i++) {
        const DevCollider &c = cs[i];
        if (!c.active) continue;
        if (c.kind == BBX_SHAPE_MESH && !inside_bounds(p, c.bmin, c.bmax)) continue;
        const float d = fabsf(closest_distance(c, p));
        if (d < mind) {
            target = i;
            mind = d;
        }
    }
    if (target < 0) return false;
    return collision_handle(cs[target], radius, e, p, v);
}

}  // namespace bbx
