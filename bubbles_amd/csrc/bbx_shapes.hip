// bbx_shapes.hip -- collider shapes: SDF grids (uploaded or built from a
// triangle mesh on the device), device evaluation of the collider-set rule,
// and the solver's collider list.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bbx_collider.h"

using namespace bbx;

struct bbx_sdf {
    int nodes[3] = {0, 0, 0};
    double origin[3] = {0, 0, 0};
    double spacing[3] = {0, 0, 0};
    float *field = nullptr;  // device, x fastest
};

namespace {

constexpr int kTriTile = 128;  // triangles staged per LDS round (9 doubles each)
constexpr int kSdfBlock = 256;

__device__ __forceinline__ double sgn_d(double a) { return a > 0 ? 1.0 : (a < 0 ? -1.0 : 0.0); }
__device__ __forceinline__ double clamp_d(double v, double lo, double hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// DistanceTriangle (src/shapes/bvh.cpp:212-242), FP64 without contraction:
// squared distance to triangle abc (edge parameters clamped to [1e-4, 1])
__device__ double tri_dist2(const double *t, double px, double py, double pz) {
#pragma clang fp contract(off)
    const double ax = t[0], ay = t[1], az = t[2], bx = t[3], by = t[4], bz = t[5];
    const double cx = t[6], cy = t[7], cz = t[8];
    const double bax = bx - ax, bay = by - ay, baz = bz - az, pax = px - ax, pay = py - ay, paz = pz - az;
    const double cbx = cx - bx, cby = cy - by, cbz = cz - bz, pbx = px - bx, pby = py - by, pbz = pz - bz;
    const double acx = ax - cx, acy = ay - cy, acz = az - cz, pcx = px - cx, pcy = py - cy, pcz = pz - cz;
    // nor = ba x ac
    const double nx = bay * acz - baz * acy, ny = baz * acx - bax * acz, nz = bax * acy - bay * acx;
    // (e x nor) . p_e for the three edges
    const double s0 = (bay * nz - baz * ny) * pax + (baz * nx - bax * nz) * pay + (bax * ny - bay * nx) * paz;
    const double s1 = (cby * nz - cbz * ny) * pbx + (cbz * nx - cbx * nz) * pby + (cbx * ny - cby * nx) * pbz;
    const double s2 = (acy * nz - acz * ny) * pcx + (acz * nx - acx * nz) * pcy + (acx * ny - acy * nx) * pcz;
    if (sgn_d(s0) + sgn_d(s1) + sgn_d(s2) < 2.0) {
        const double t0 = clamp_d((bax * pax + bay * pay + baz * paz) / (bax * bax + bay * bay + baz * baz), 0.0001, 1.0);
        const double t1 = clamp_d((cbx * pbx + cby * pby + cbz * pbz) / (cbx * cbx + cby * cby + cbz * cbz), 0.0001, 1.0);
        const double t2 = clamp_d((acx * pcx + acy * pcy + acz * pcz) / (acx * acx + acy * acy + acz * acz), 0.0001, 1.0);
        const double ex = bax * t0 - pax, ey = bay * t0 - pay, ez = baz * t0 - paz;
        const double fx = cbx * t1 - pbx, fy = cby * t1 - pby, fz = cbz * t1 - pbz;
        const double gx = acx * t2 - pcx, gy = acy * t2 - pcy, gz = acz * t2 - pcz;
        const double d0 = ex * ex + ey * ey + ez * ez;
        const double d1 = fx * fx + fy * fy + fz * fz;
        const double d2 = gx * gx + gy * gy + gz * gz;
        return fmin(fmin(d0, d1), d2);
    }
    const double np = nx * pax + ny * pay + nz * paz;
    return np * np / (nx * nx + ny * ny + nz * nz);
}

// signed solid angle of triangle abc seen from p (Van Oosterom-Strackee);
// the sum over a closed mesh is 4 pi inside and 0 outside
__device__ double tri_solid_angle(const double *t, double px, double py, double pz) {
#pragma clang fp contract(off)
    const double ax = t[0] - px, ay = t[1] - py, az = t[2] - pz;
    const double bx = t[3] - px, by = t[4] - py, bz = t[5] - pz;
    const double cx = t[6] - px, cy = t[7] - py, cz = t[8] - pz;
    const double la = sqrt(ax * ax + ay * ay + az * az);
    const double lb = sqrt(bx * bx + by * by + bz * bz);
    const double lc = sqrt(cx * cx + cy * cy + cz * cz);
    const double det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
    const double den = la * lb * lc + (ax * bx + ay * by + az * bz) * lc + (ax * cx + ay * cy + az * cz) * lb +
                       (bx * cx + by * cy + bz * cz) * la;
    return 2.0 * atan2(det, den);
}

// SetNodeSDFKernel (src/core/shape.cpp:411-424) for a mesh: distance to the
// closest triangle, at least 1e-5, negative inside.  Inside is decided by
// the winding number of the closed mesh (the reference counts ray hits
// towards the mesh's bounds centre, src/core/shape.cpp:356-373).
__global__ void __launch_bounds__(kSdfBlock) k_sdf_from_mesh(const double *__restrict__ tri, int ntri, int nx,
                                                             int ny, int nz, double ox, double oy, double oz,
                                                             double dx, float *__restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double st[kTriTile * 9];
    const long long total = (long long)nx * ny * nz;
    const long long node = (long long)blockIdx.x * kSdfBlock + threadIdx.x;
    const bool live = node < total;
    double px = 0, py = 0, pz = 0;
    if (live) {
        const long long i = node % nx, j = (node / nx) % ny, k = node / ((long long)nx * ny);
        px = ox + dx * (double)i;
        py = oy + dx * (double)j;
        pz = oz + dx * (double)k;
    }
    double best = DBL_MAX;
    double wind = 0.0;
    for (int t0 = 0; t0 < ntri; t0 += kTriTile) {
        const int cnt = min(kTriTile, ntri - t0);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * 9; e += kSdfBlock) st[e] = tri[(size_t)t0 * 9 + e];
        __syncthreads();
        if (live) {
            for (int t = 0; t < cnt; t++) {
                best = fmin(best, tri_dist2(&st[t * 9], px, py, pz));
                wind += tri_solid_angle(&st[t * 9], px, py, pz);
            }
        }
    }
    if (live) {
        const bool interior = fabs(wind) > 2.0 * M_PI;  // |winding number| > 1/2
        const double psd = fmax(sqrt(best), 0.00001);
        out[node] = (float)(interior ? -psd : psd);
    }
}

__global__ void __launch_bounds__(256) k_sdf_sample(DevCollider c, const float4 *__restrict__ pts, int n,
                                                    float4 *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float3 p = make_float3(pts[i].x, pts[i].y, pts[i].z);
    const float3 g = sdf_gradient(c, p);
    out[i] = make_float4(sdf_sample(c, p), g.x, g.y, g.z);
}

__global__ void __launch_bounds__(256) k_resolve(const DevCollider *__restrict__ cs, int nc, float radius, float e,
                                                 float4 *__restrict__ pos, float4 *__restrict__ vel,
                                                 int *__restrict__ hit, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float3 p = make_float3(pos[i].x, pos[i].y, pos[i].z);
    float3 v = make_float3(vel[i].x, vel[i].y, vel[i].z);
    const bool h = resolve_colliders(cs, nc, radius, e, p, v);
    pos[i] = make_float4(p.x, p.y, p.z, 0.f);
    vel[i] = make_float4(v.x, v.y, v.z, 0.f);
    hit[i] = h ? 1 : 0;
}

// The same rule over the caller's FP64 arrays as staged in device memory:
// each coordinate rounds to FP32 exactly as the host conversion of
// k_resolve's caller does, and the result widens back to FP64.
__global__ void __launch_bounds__(256) k_resolve_f64(const DevCollider *__restrict__ cs, int nc, float radius,
                                                     float e, double *__restrict__ pos, double *__restrict__ vel,
                                                     int *__restrict__ hit, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *pp = pos + 3 * (size_t)i;
    double *vv = vel + 3 * (size_t)i;
    float3 p = make_float3((float)pp[0], (float)pp[1], (float)pp[2]);
    float3 v = make_float3((float)vv[0], (float)vv[1], (float)vv[2]);
    const bool h = resolve_colliders(cs, nc, radius, e, p, v);
    pp[0] = p.x;
    pp[1] = p.y;
    pp[2] = p.z;
    vv[0] = v.x;
    vv[1] = v.y;
    vv[2] = v.z;
    hit[i] = h ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Staging of bbx_colliders_resolve_noalloc: one file-scope device array per
// GPU, laid out as [collider table | positions | velocities | flags].  Its
// address is looked up once per device with hipGetSymbolAddress; host code
// never passes the array's name itself to a kernel or a copy.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocCollBytes = ((sizeof(DevCollider) * kMaxColliders + 255) / 256) * 256;
constexpr size_t kNoAllocVecBytes = (size_t)BBX_NOALLOC_MAX_POINTS * 3 * sizeof(double);
constexpr size_t kNoAllocBytes = kNoAllocCollBytes + 2 * kNoAllocVecBytes + (size_t)BBX_NOALLOC_MAX_POINTS * sizeof(int);
static_assert(kNoAllocVecBytes % 256 == 0, "scratch segments stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kNoAllocDevices] = {};

// Device address of the current device's scratch array (looked up on the
// first call made with that device current).
int scratch_for_current_device(unsigned char **out) {
    int dev = -1;
    BBX_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) return fail(BBX_E_RANGE, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BBX_E_HIP, std::string("scratch address: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return BBX_OK;
}

void sdf_fields(const bbx_sdf *s, DevCollider &d) {
    d.field = s->field;
    d.nx = s->nodes[0];
    d.ny = s->nodes[1];
    d.nz = s->nodes[2];
    d.ox = (float)s->origin[0];
    d.oy = (float)s->origin[1];
    d.oz = (float)s->origin[2];
    d.sx = (float)s->spacing[0];
    d.sy = (float)s->spacing[1];
    d.sz = (float)s->spacing[2];
}

int to_device(const bbx_collider &h, DevCollider &d) {
    std::memset(&d, 0, sizeof(d));
    if (h.kind < BBX_SHAPE_BOX || h.kind > BBX_SHAPE_SDF) return fail(BBX_E_INVALID, "unknown collider kind");
    d.kind = h.kind;
    d.reverse = h.reverse_orientation ? 1 : 0;
    d.active = h.active ? 1 : 0;
    d.friction = (float)h.friction;
    for (int k = 0; k < 12; k++) {
        d.m[k] = (float)h.to_world[k];
        d.mi[k] = (float)h.to_object[k];
    }
    d.hx = (float)(h.size[0] / 2.0);
    d.hy = (float)(h.size[1] / 2.0);
    d.hz = (float)(h.size[2] / 2.0);
    d.radius = (float)h.radius;
    for (int k = 0; k < 3; k++) {
        d.lin[k] = (float)h.linear_velocity[k];
        d.ang[k] = (float)h.angular_velocity[k];
        d.bmin[k] = (float)h.bounds_min[k];
        d.bmax[k] = (float)h.bounds_max[k];
    }
    if (h.kind == BBX_SHAPE_MESH || h.kind == BBX_SHAPE_SDF) {
        if (!h.sdf || !h.sdf->field) return fail(BBX_E_INVALID, "mesh and SDF colliders need an SDF");
        sdf_fields(h.sdf, d);
    }
    return BBX_OK;
}

int upload_colliders(const bbx_collider *list, int count, std::vector<DevCollider> &out) {
    if (count < 0 || count > kMaxColliders) return fail(BBX_E_RANGE, "collider count out of range (0..64)");
    if (count && !list) return fail(BBX_E_INVALID, "null collider list");
    out.resize((size_t)count);
    for (int i = 0; i < count; i++) BBX_CHECK(to_device(list[i], out[(size_t)i]));
    return BBX_OK;
}

}  // namespace

extern "C" {

int bbx_sdf_create(const int nodes[3], const double origin[3], const double spacing[3], const double *values,
                   bbx_sdf **out) {
    if (!out || !nodes || !origin || !spacing || !values) return fail(BBX_E_INVALID, "null argument");
    *out = nullptr;
    long long total = 1;
    for (int k = 0; k < 3; k++) {
        if (nodes[k] < 1 || nodes[k] > (1 << 15)) return fail(BBX_E_RANGE, "SDF node count out of range");
        if (!(spacing[k] > 0)) return fail(BBX_E_INVALID, "SDF spacing must be positive");
        total *= nodes[k];
    }
    if (total > (1LL << 31) - 1) return fail(BBX_E_RANGE, "SDF too large");
    bbx_sdf *s = new bbx_sdf();
    for (int k = 0; k < 3; k++) {
        s->nodes[k] = nodes[k];
        s->origin[k] = origin[k];
        s->spacing[k] = spacing[k];
    }
    std::vector<float> h((size_t)total);
    for (long long i = 0; i < total; i++) h[(size_t)i] = (float)values[i];
    hipError_t e = hipMalloc((void **)&s->field, sizeof(float) * (size_t)total);
    if (e == hipSuccess) e = hipMemcpy(s->field, h.data(), sizeof(float) * (size_t)total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (s->field) (void)hipFree(s->field);
        delete s;
        return fail(BBX_E_HIP, std::string("SDF upload: ") + hipGetErrorString(e));
    }
    *out = s;
    return BBX_OK;
}

int bbx_sdf_from_mesh(const double *vertices, int nvertices, const int *triangles, int ntriangles, double dx,
                      double margin, bbx_sdf **out) {
    if (!out) return fail(BBX_E_INVALID, "null output");
    *out = nullptr;
    if (!vertices || !triangles || nvertices < 3 || ntriangles < 1) return fail(BBX_E_INVALID, "empty mesh");
    if (!(dx > 0) || !(margin >= 0)) return fail(BBX_E_INVALID, "bad SDF spacing or margin");
    // bounds of the triangles' vertices (the mesh BVH root bound)
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    std::vector<double> tri((size_t)ntriangles * 9);
    for (int t = 0; t < ntriangles; t++)
        for (int c = 0; c < 3; c++) {
            const int v = triangles[3 * (size_t)t + c];
            if (v < 0 || v >= nvertices) return fail(BBX_E_INVALID, "triangle index out of range");
            for (int k = 0; k < 3; k++) {
                const double x = vertices[3 * (size_t)v + k];
                tri[9 * (size_t)t + 3 * c + k] = x;
                lo[k] = std::min(lo[k], x);
                hi[k] = std::max(hi[k], x);
            }
        }
    // GenerateShapeSDF (src/core/shape.cpp:478-505)
    double pmin[3], pmax[3];
    for (int k = 0; k < 3; k++) {
        const double scale = hi[k] - lo[k];
        pmin[k] = lo[k] - margin * scale;
        pmax[k] = hi[k] + margin * scale;
    }
    const double width = pmax[0] - pmin[0], height = pmax[1] - pmin[1], depth = pmax[2] - pmin[2];
    if (!(width > 0)) return fail(BBX_E_INVALID, "mesh has no extent along x");
    const double resd = std::ceil(width / dx);
    if (resd > 4096) return fail(BBX_E_RANGE, "SDF resolution above 4096 per axis");
    const int res = (int)resd;
    const double sp = width / (double)res;
    const int resy = (int)std::ceil(res * height / width);
    const int resz = (int)std::ceil(res * depth / width);
    if (resy > 4096 || resz > 4096) return fail(BBX_E_RANGE, "SDF resolution above 4096 per axis");
    if ((long long)(res + 1) * (resy + 1) * (resz + 1) > (1LL << 31) - 1) return fail(BBX_E_RANGE, "SDF too large");
    bbx_sdf *s = new bbx_sdf();
    s->nodes[0] = res + 1;
    s->nodes[1] = resy + 1;
    s->nodes[2] = resz + 1;
    for (int k = 0; k < 3; k++) {
        s->origin[k] = pmin[k];
        s->spacing[k] = sp;
    }
    const long long total = (long long)s->nodes[0] * s->nodes[1] * s->nodes[2];
    double *dtri = nullptr;
    hipError_t e = hipMalloc((void **)&s->field, sizeof(float) * (size_t)total);
    if (e == hipSuccess) e = hipMalloc((void **)&dtri, sizeof(double) * tri.size());
    if (e == hipSuccess) e = hipMemcpy(dtri, tri.data(), sizeof(double) * tri.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const unsigned blocks = (unsigned)((total + kSdfBlock - 1) / kSdfBlock);
        k_sdf_from_mesh<<<blocks, kSdfBlock>>>(dtri, ntriangles, s->nodes[0], s->nodes[1], s->nodes[2], pmin[0],
                                               pmin[1], pmin[2], sp, s->field);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (dtri) (void)hipFree(dtri);
    if (e != hipSuccess) {
        if (s->field) (void)hipFree(s->field);
        delete s;
        return fail(BBX_E_HIP, std::string("SDF from mesh: ") + hipGetErrorString(e));
    }
    *out = s;
    return BBX_OK;
}

int bbx_sdf_describe(const bbx_sdf *sdf, int nodes[3], double origin[3], double spacing[3]) {
    if (!sdf) return fail(BBX_E_INVALID, "null SDF");
    for (int k = 0; k < 3; k++) {
        if (nodes) nodes[k] = sdf->nodes[k];
        if (origin) origin[k] = sdf->origin[k];
        if (spacing) spacing[k] = sdf->spacing[k];
    }
    return BBX_OK;
}

int bbx_sdf_download(const bbx_sdf *sdf, double *values) {
    if (!sdf || !values) return fail(BBX_E_INVALID, "null argument");
    const size_t total = (size_t)sdf->nodes[0] * sdf->nodes[1] * sdf->nodes[2];
    std::vector<float> h(total);
    BBX_HIP_CHECK(hipMemcpy(h.data(), sdf->field, sizeof(float) * total, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < total; i++) values[i] = h[i];
    return BBX_OK;
}

int bbx_sdf_sample(const bbx_sdf *sdf, const double *points, int n, double *values, double *gradients) {
    if (!sdf || (n > 0 && (!points || !values))) return fail(BBX_E_INVALID, "null argument");
    if (n <= 0) return BBX_OK;
    DevCollider c;
    std::memset(&c, 0, sizeof(c));
    c.kind = BBX_SHAPE_SDF;
    sdf_fields(sdf, c);
    std::vector<float4> hp((size_t)n), ho((size_t)n);
    for (int i = 0; i < n; i++)
        hp[(size_t)i] = make_float4((float)points[3 * (size_t)i], (float)points[3 * (size_t)i + 1],
                                    (float)points[3 * (size_t)i + 2], 0.f);
    float4 *dp = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc((void **)&dp, sizeof(float4) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void **)&dout, sizeof(float4) * (size_t)n);
    if (e == hipSuccess) e = hipMemcpy(dp, hp.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_sdf_sample<<<(unsigned)((n + 255) / 256), 256>>>(c, dp, n, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(ho.data(), dout, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost);
    if (dp) (void)hipFree(dp);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return fail(BBX_E_HIP, std::string("SDF sample: ") + hipGetErrorString(e));
    for (int i = 0; i < n; i++) {
        values[i] = ho[(size_t)i].x;
        if (gradients) {
            gradients[3 * (size_t)i] = ho[(size_t)i].y;
            gradients[3 * (size_t)i + 1] = ho[(size_t)i].z;
            gradients[3 * (size_t)i + 2] = ho[(size_t)i].w;
        }
    }
    return BBX_OK;
}

void bbx_sdf_destroy(bbx_sdf *sdf) {
    if (!sdf) return;
    if (sdf->field) (void)hipFree(sdf->field);
    delete sdf;
}

int bbx_solver_set_colliders(bbx_solver *s, const bbx_collider *colliders, int count) {
    if (!s) return fail(BBX_E_INVALID, "null solver");
    std::vector<DevCollider> h;
    BBX_CHECK(upload_colliders(colliders, count, h));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));  // kernels may still read the old list
    if (count == 0) {
        if (s->coll) (void)hipFree(s->coll);
        s->coll = nullptr;
        s->ncoll = 0;
        fill_sim_const(s);
        return BBX_OK;
    }
    if (!s->coll) {
        hipError_t e = hipMalloc((void **)&s->coll, sizeof(DevCollider) * kMaxColliders);
        if (e != hipSuccess) {
            s->coll = nullptr;
            return fail(BBX_E_NOMEM, std::string("collider list: ") + hipGetErrorString(e));
        }
    }
    BBX_HIP_CHECK(hipMemcpy(s->coll, h.data(), sizeof(DevCollider) * (size_t)count, hipMemcpyHostToDevice));
    s->ncoll = count;
    fill_sim_const(s);
    return BBX_OK;
}

int bbx_colliders_resolve(const bbx_collider *colliders, int count, double radius, double restitution,
                          double *pos, double *vel, int n, int *collided) {
    if (n < 0 || (n > 0 && (!pos || !vel))) return fail(BBX_E_INVALID, "bad point arrays");
    if (count < 1) return fail(BBX_E_INVALID, "no colliders");
    std::vector<DevCollider> h;
    BBX_CHECK(upload_colliders(colliders, count, h));
    if (n == 0) return BBX_OK;
    std::vector<float4> hp((size_t)n), hv((size_t)n);
    std::vector<int> hh((size_t)n);
    for (int i = 0; i < n; i++) {
        hp[(size_t)i] = make_float4((float)pos[3 * (size_t)i], (float)pos[3 * (size_t)i + 1],
                                    (float)pos[3 * (size_t)i + 2], 0.f);
        hv[(size_t)i] = make_float4((float)vel[3 * (size_t)i], (float)vel[3 * (size_t)i + 1],
                                    (float)vel[3 * (size_t)i + 2], 0.f);
    }
    DevCollider *dc = nullptr;
    float4 *dp = nullptr, *dv = nullptr;
    int *dh = nullptr;
    hipError_t e = hipMalloc((void **)&dc, sizeof(DevCollider) * (size_t)count);
    if (e == hipSuccess) e = hipMalloc((void **)&dp, sizeof(float4) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void **)&dv, sizeof(float4) * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void **)&dh, sizeof(int) * (size_t)n);
    if (e == hipSuccess) e = hipMemcpy(dc, h.data(), sizeof(DevCollider) * (size_t)count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, hp.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dv, hv.data(), sizeof(float4) * (size_t)n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_resolve<<<(unsigned)((n + 255) / 256), 256>>>(dc, count, (float)radius, (float)restitution, dp, dv, dh,
                                                         n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(hp.data(), dp, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hv.data(), dv, sizeof(float4) * (size_t)n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(hh.data(), dh, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost);
    if (dc) (void)hipFree(dc);
    if (dp) (void)hipFree(dp);
    if (dv) (void)hipFree(dv);
    if (dh) (void)hipFree(dh);
    if (e != hipSuccess) return fail(BBX_E_HIP, std::string("collider resolve: ") + hipGetErrorString(e));
    for (int i = 0; i < n; i++) {
        pos[3 * (size_t)i] = hp[(size_t)i].x;
        pos[3 * (size_t)i + 1] = hp[(size_t)i].y;
        pos[3 * (size_t)i + 2] = hp[(size_t)i].z;
        vel[3 * (size_t)i] = hv[(size_t)i].x;
        vel[3 * (size_t)i + 1] = hv[(size_t)i].y;
        vel[3 * (size_t)i + 2] = hv[(size_t)i].z;
        if (collided) collided[i] = hh[(size_t)i];
    }
    return BBX_OK;
}

int bbx_colliders_resolve_noalloc(const bbx_collider *colliders, int count, double radius, double restitution,
                                  double *pos, double *vel, int n, int *collided) {
    if (n < 0 || (n > 0 && (!pos || !vel))) return fail(BBX_E_INVALID, "bad point arrays");
    if (n > BBX_NOALLOC_MAX_POINTS)
        return fail(BBX_E_RANGE, "bbx_colliders_resolve_noalloc takes at most " +
                                     std::to_string(BBX_NOALLOC_MAX_POINTS) + " points per call (got " +
                                     std::to_string(n) + ")");
    if (count < 1) return fail(BBX_E_INVALID, "no colliders");
    if (count > kMaxColliders) return fail(BBX_E_RANGE, "collider count out of range (0..64)");
    if (!colliders) return fail(BBX_E_INVALID, "null collider list");
    DevCollider table[kMaxColliders];
    for (int i = 0; i < count; i++) BBX_CHECK(to_device(colliders[i], table[i]));
    if (n == 0) return BBX_OK;
    unsigned char *base = nullptr;
    BBX_CHECK(scratch_for_current_device(&base));
    DevCollider *dc = (DevCollider *)base;
    double *dp = (double *)(base + kNoAllocCollBytes);
    double *dv = (double *)(base + kNoAllocCollBytes + kNoAllocVecBytes);
    int *dh = (int *)(base + kNoAllocCollBytes + 2 * kNoAllocVecBytes);
    const size_t vb = sizeof(double) * 3 * (size_t)n;
    hipError_t e = hipMemcpy(dc, table, sizeof(DevCollider) * (size_t)count, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, pos, vb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dv, vel, vb, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_resolve_f64<<<(unsigned)((n + 255) / 256), 256>>>(dc, count, (float)radius, (float)restitution, dp, dv,
                                                             dh, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(pos, dp, vb, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(vel, dv, vb, hipMemcpyDeviceToHost);
    if (e == hipSuccess && collided) e = hipMemcpy(collided, dh, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(BBX_E_HIP, std::string("collider resolve (no-alloc): ") + hipGetErrorString(e));
    return BBX_OK;
}

}  // extern "C"
