// bbx_sph.hip -- SPH / PCISPH gathers, prediction and integration on gfx950.
//
// Every kernel is one thread per particle in cell-sorted storage order and
// walks the particle's bucket (the reference neighbour list) from the
// slot-major table nbr[k * stride + i], so a wavefront reads bucket slot k of
// 64 consecutive particles with one coalesced load.  Fusions relative to the
// reference's launch sequence (src/equations/pcisph_equations3.cpp:218-256):
//
//   nonpressure + predict(k=0)       ComputeNonPressureForceFor + PredictVelocityAndPositionFor
//   predicted density + pressure     PredictPressureFor, FP64 max-error reduction
//   pressure force + predict(k+1)    PredictPressureForceFor + PredictVelocityAndPositionFor
//   pressure force + integrate       PredictPressureForceFor + AccumulateForcesFor + TimeIntegrationFor
//
// Arithmetic is FP32; density and pressure-force sums use Kahan compensation.
#include "bbx_device.h"

namespace bbx {

#define PROF_BEGIN(name)                                                \
    hipEvent_t ev_a_ = nullptr;                                         \
    int ev_id_ = -1;                                                    \
    if (s->prof.enabled) {                                              \
        ev_id_ = s->prof.name_id(name);                                 \
        s->prof.begin(s->stream, name, &ev_a_);                         \
    }
#define PROF_END()                                                      \
    if (s->prof.enabled) s->prof.end(s->stream, ev_id_, ev_a_);

static unsigned nblocks(int n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// PredictVelocityAndPositionFor (src/equations/pcisph_equations3.cpp:3-28):
// v* = v + dt/m (f + f_p), x* = x + dt v*, collider with restitution 0.
__device__ __forceinline__ float4 predict_position(const SimConst &c, const StepConst &st,
                                                   float4 x, float4 v, float fx, float fy,
                                                   float fz) {
    float3 tv = make_float3(v.x + st.dt_over_m * fx, v.y + st.dt_over_m * fy,
                            v.z + st.dt_over_m * fz);
    float3 tp = make_float3(x.x + st.dt * tv.x, x.y + st.dt * tv.y, x.z + st.dt * tv.z);
    collide(c, 0.f, tp, tv);
    return make_float4(tp.x, tp.y, tp.z, 0.f);
}

// ---------------------------------------------------------------------------
// non-pressure force (src/equations/sph_equations3.cpp:80-110)
// mode: 0 = store force only, 1 = + predict(k=0) with f_p = 0,
//       2 = + warm-start features
// ---------------------------------------------------------------------------
template <int kMode>
__global__ void __launch_bounds__(kBlock) k_nonpressure(SimConst c, StepConst st,
                                                        const float4 *__restrict__ pos,
                                                        const float4 *__restrict__ vel,
                                                        const float *__restrict__ rho,
                                                        const int *__restrict__ nbr,
                                                        const int *__restrict__ nbr_cnt,
                                                        float4 *__restrict__ force,
                                                        float4 *__restrict__ xstar,
                                                        float *__restrict__ rho_star,
                                                        float *__restrict__ feat,
                                                        unsigned long long *__restrict__ max_err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && max_err) *max_err = 0ull;
    if (i >= c.n) return;
    const float4 xi = pos[i];
    const float4 vi = vel[i];
    const float coef = c.visc * c.mass2;
    float fx = c.mass * c.gx - c.drag * vi.x;
    float fy = c.mass * c.gy - c.drag * vi.y;
    float fz = c.mass * c.gz - c.drag * vi.z;
    const int cnt = nbr_cnt[i];
    // feature accumulators
    float m1 = 0.f, m2 = 0.f, m3 = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, divv = 0.f;
    for (int k = 0; k < cnt; k++) {
        int j = nbr[(size_t)k * c.stride + i];
        float4 xj = pos[j];
        float4 vj = vel[j];
        float rj = rho[j];
        float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        float r = sqrtf(dx * dx + dy * dy + dz * dz);
        float w = 0.f, x = 0.f;
        if (spiky_in(c, r)) {
            x = 1.f - r * c.inv_h;
            w = c.spiky_lap * x;
        }
        float s = coef * w / rj;
        fx += (vj.x - vi.x) * s;
        fy += (vj.y - vi.y) * s;
        fz += (vj.z - vi.z) * s;
        if (kMode == 2 && j != i) {
            float u = r * c.inv_h;
            m1 += u;
            m2 += u * u;
            m3 += u * u * u;
            cx -= dx * c.inv_h;
            cy -= dy * c.inv_h;
            cz -= dz * c.inv_h;
            if (r > 0.f && x > 0.f) {
                // grad_i W = dW/dr (x_i - x_j)/r, dW/dr = -spiky_grad x^2
                float g = -c.spiky_grad * x * x / r;
                float dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
                divv += c.mass / rj * g * (dvx * dx + dvy * dy + dvz * dz);
            }
        }
    }
    force[i] = make_float4(fx, fy, fz, 0.f);
    if (kMode == 1) {
        xstar[i] = predict_position(c, st, xi, vi, fx, fy, fz);
        rho_star[i] = rho[i];
    }
    if (kMode == 2) {
        const size_t S = (size_t)c.stride;
        float nn = cnt > 1 ? (float)(cnt - 1) : 1.f;
        float inv = 1.f / nn;
        float wall = 3.4e38f;
        const float p3[3] = {xi.x, xi.y, xi.z};
        for (int a = 0; a < 3; a++) {
            float lo = p3[a] - (c.tank_c[a] - c.tank_h[a]);
            float hi = (c.tank_c[a] + c.tank_h[a]) - p3[a];
            wall = fminf(wall, fminf(lo, hi));
        }
        float vn = sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z);
        float fn = sqrtf(fx * fx + fy * fy + fz * fz);
        float a_scale = st.dt * st.dt / (c.mass * c.h);
        feat[BBX_F_DENSITY_ERR * S + i] = rho[i] / c.rho0 - 1.f;
        feat[BBX_F_DIVV * S + i] = st.dt * divv;
        feat[BBX_F_COUNT * S + i] = (float)cnt;
        feat[BBX_F_M1 * S + i] = m1 * inv;
        feat[BBX_F_M2 * S + i] = m2 * inv;
        feat[BBX_F_M3 * S + i] = m3 * inv;
        feat[BBX_F_CX * S + i] = cx * inv;
        feat[BBX_F_CY * S + i] = cy * inv;
        feat[BBX_F_CZ * S + i] = cz * inv;
        feat[BBX_F_CNORM * S + i] = sqrtf(cx * cx + cy * cy + cz * cz) * inv;
        feat[BBX_F_PPREV * S + i] = st.delta != 0.f ? vi.w / (st.delta * c.rho0) : 0.f;
        feat[BBX_F_VY * S + i] = st.dt * vi.y * c.inv_h;
        feat[BBX_F_VNORM * S + i] = st.dt * vn * c.inv_h;
        feat[BBX_F_AY * S + i] = fy * a_scale;
        feat[BBX_F_ANORM * S + i] = fn * a_scale;
        feat[BBX_F_WALL * S + i] = fminf(fmaxf(wall * c.inv_h, 0.f), 4.f);
        rho_star[i] = rho[i];
    }
}

// ---------------------------------------------------------------------------
// predicted density and pressure update (PredictPressureFor,
// src/equations/pcisph_equations3.cpp:60-90) + FP64 max-error reduction
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_predict_pressure(SimConst c, StepConst st, double delta,
                                                             const float4 *__restrict__ xstar,
                                                             const int *__restrict__ nbr,
                                                             const int *__restrict__ nbr_cnt,
                                                             float *__restrict__ pres,
                                                             float *__restrict__ rho_star,
                                                             float *__restrict__ q,
                                                             unsigned long long *__restrict__ max_err) {
    __shared__ unsigned long long wmax[kBlock / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long ebits = 0ull;
    if (i < c.n) {
        const float4 xi = xstar[i];
        float sum = 0.f, comp = 0.f;
        const int cnt = nbr_cnt[i];
        for (int k = 0; k < cnt; k++) {
            float4 xj = xstar[nbr[(size_t)k * c.stride + i]];
            float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            kahan_add(sum, comp, poly6_w(c, dx * dx + dy * dy + dz * dz));
        }
        float density = sum * c.mass;
        // FP64 error and pressure increment (delta is FP64 on the host)
        double err = (double)density - (double)c.rho0;
        double dp = delta * err;
        if (dp < 0.0) {
            dp *= (double)c.nps;
            err *= (double)c.nps;
        }
        float p = (st.first ? 0.f : pres[i]) + (float)dp;
        pres[i] = p;
        rho_star[i] = density;
        float d2 = density * density;
        q[i] = (fabsf(d2) < 1e-8f) ? 0.f : p / d2;
        double ae = err > 0 ? err : -err;
        ebits = (unsigned long long)__double_as_longlong(ae);
    }
    ebits = wave_max_u64(ebits);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = ebits;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = wmax[0];
        for (int w = 1; w < kBlock / 64; w++) m = wmax[w] > m ? wmax[w] : m;
        if (m) atomicMax(max_err, m);
    }
}

// ---------------------------------------------------------------------------
// pressure force (PredictPressureForceFor, src/equations/pcisph_equations3.cpp:114-155)
// at the current positions with predicted densities; then either predict the
// next iteration's x* (kIntegrate = 0) or accumulate and integrate
// (AccumulateForcesFor + TimeIntegrationFor, src/equations/sph_equations3.cpp:283-339).
// ---------------------------------------------------------------------------
template <int kIntegrate>
__global__ void __launch_bounds__(kBlock) k_pressure_force(SimConst c, StepConst st,
                                                           const float4 *__restrict__ pos,
                                                           const float4 *__restrict__ vel,
                                                           const float4 *__restrict__ force,
                                                           const float *__restrict__ q,
                                                           const float *__restrict__ pres,
                                                           const int *__restrict__ nbr,
                                                           const int *__restrict__ nbr_cnt,
                                                           float4 *__restrict__ xstar,
                                                           float4 *__restrict__ fp_out,
                                                           float4 *__restrict__ pos_out,
                                                           float4 *__restrict__ vel_out,
                                                           float4 *__restrict__ force_out,
                                                           int *__restrict__ flag,
                                                           unsigned *__restrict__ max_force,
                                                           unsigned long long *__restrict__ max_err) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (!kIntegrate && i == 0) *max_err = 0ull;
    unsigned fbits = 0u;
    if (i < c.n) {
        const float4 xi = pos[i];
        const float qi = q[i];
        float tx = 0.f, ty = 0.f, tz = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
        const int cnt = nbr_cnt[i];
        for (int k = 0; k < cnt; k++) {
            int j = nbr[(size_t)k * c.stride + i];
            if (j == i) continue;
            float4 xj = pos[j];
            float qj = q[j];
            float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
            float r = sqrtf(dx * dx + dy * dy + dz * dz);
            if (!spiky_in(c, r) || !(r > 0.f) || fabsf(r) < 1e-8f) continue;
            float x = 1.f - r * c.inv_h;
            // gradW(r, dir) = -dW(r) dir = spiky_grad x^2 (x_j - x_i)/r
            float g = c.mass2 * (qi + qj) * c.spiky_grad * x * x / r;
            kahan_add(tx, ex, g * dx);
            kahan_add(ty, ey, g * dy);
            kahan_add(tz, ez, g * dz);
        }
        const float fpx = -tx, fpy = -ty, fpz = -tz;
        const float4 fi = force[i];
        const float4 vi = vel[i];
        if (!kIntegrate) {
            xstar[i] = predict_position(c, st, xi, vi, fi.x + fpx, fi.y + fpy, fi.z + fpz);
        } else {
            float Fx = fi.x + fpx, Fy = fi.y + fpy, Fz = fi.z + fpz;
            float inv_m = 1.f / c.mass;
            float3 v = make_float3(vi.x + st.dt * (Fx * inv_m), vi.y + st.dt * (Fy * inv_m),
                                   vi.z + st.dt * (Fz * inv_m));
            float3 p = make_float3(xi.x + st.dt * v.x, xi.y + st.dt * v.y, xi.z + st.dt * v.z);
            collide(c, 0.6f, p, v);
            if (!inside_box(p.x, p.y, p.z, c.bmin, c.bmax)) {
                p.x = fminf(fmaxf(p.x, c.bmin[0] + c.radius), c.bmax[0] - c.radius);
                p.y = fminf(fmaxf(p.y, c.bmin[1] + c.radius), c.bmax[1] - c.radius);
                p.z = fminf(fmaxf(p.z, c.bmin[2] + c.radius), c.bmax[2] - c.radius);
            }
            float ddx = p.x - xi.x, ddy = p.y - xi.y, ddz = p.z - xi.z;
            if (sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) >= c.min_len * 0.9f) *flag = 1;
            pos_out[i] = make_float4(p.x, p.y, p.z, xi.w);
            vel_out[i] = make_float4(v.x, v.y, v.z, pres[i]);
            force_out[i] = make_float4(Fx, Fy, Fz, 0.f);
            fp_out[i] = make_float4(fpx, fpy, fpz, 0.f);
            fbits = __float_as_uint(sqrtf(Fx * Fx + Fy * Fy + Fz * Fz));
        }
    }
    if (kIntegrate) {
        fbits = wave_max_u32(fbits);
        if ((threadIdx.x & 63) == 0 && fbits) atomicMax(max_force, fbits);
    }
}

// ---------------------------------------------------------------------------
// SPH (WCSPH) forces: ComputeAllForcesFor without the in-loop integration
// (src/equations/sph_equations3.cpp:184-278); integration is a separate pass.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_sph_force(SimConst c, const float4 *__restrict__ pos,
                                                      const float4 *__restrict__ vel,
                                                      const float *__restrict__ rho,
                                                      const float *__restrict__ pres,
                                                      const int *__restrict__ nbr,
                                                      const int *__restrict__ nbr_cnt,
                                                      float4 *__restrict__ force) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const float4 xi = pos[i], vi = vel[i];
    const float di = rho[i];
    const float qi = pres[i] / (di * di);
    const float coef = c.visc * c.mass2;
    float fx = c.mass * c.gx - c.drag * vi.x;
    float fy = c.mass * c.gy - c.drag * vi.y;
    float fz = c.mass * c.gz - c.drag * vi.z;
    float tx = 0.f, ty = 0.f, tz = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
    const int cnt = nbr_cnt[i];
    for (int k = 0; k < cnt; k++) {
        int j = nbr[(size_t)k * c.stride + i];
        if (j == i) continue;
        float4 xj = pos[j], vj = vel[j];
        float dj = rho[j];
        float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
        float r = sqrtf(dx * dx + dy * dy + dz * dz);
        if (!spiky_in(c, r)) continue;
        float x = 1.f - r * c.inv_h;
        if (r > 0.f && !(fabsf(r) < 1e-8f)) {
            float g = c.mass2 * (qi + pres[j] / (dj * dj)) * c.spiky_grad * x * x / r;
            kahan_add(tx, ex, g * dx);
            kahan_add(ty, ey, g * dy);
            kahan_add(tz, ez, g * dz);
        }
        float s = coef * c.spiky_lap * x / dj;
        fx += (vj.x - vi.x) * s;
        fy += (vj.y - vi.y) * s;
        fz += (vj.z - vi.z) * s;
    }
    force[i] = make_float4(fx - tx, fy - ty, fz - tz, 0.f);
}

// TimeIntegrationFor (src/equations/sph_equations3.cpp:283-339)
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
        float inv_m = 1.f / c.mass;
        float3 v = make_float3(vi.x + st.dt * (fi.x * inv_m), vi.y + st.dt * (fi.y * inv_m),
                               vi.z + st.dt * (fi.z * inv_m));
        float3 p = make_float3(xi.x + st.dt * v.x, xi.y + st.dt * v.y, xi.z + st.dt * v.z);
        collide(c, 0.6f, p, v);
        if (!inside_box(p.x, p.y, p.z, c.bmin, c.bmax)) {
            p.x = fminf(fmaxf(p.x, c.bmin[0] + c.radius), c.bmax[0] - c.radius);
            p.y = fminf(fmaxf(p.y, c.bmin[1] + c.radius), c.bmax[1] - c.radius);
            p.z = fminf(fmaxf(p.z, c.bmin[2] + c.radius), c.bmax[2] - c.radius);
        }
        float ddx = p.x - xi.x, ddy = p.y - xi.y, ddz = p.z - xi.z;
        if (sqrtf(ddx * ddx + ddy * ddy + ddz * ddz) >= c.min_len * 0.9f) *flag = 1;
        pos_out[i] = make_float4(p.x, p.y, p.z, xi.w);
        vel_out[i] = make_float4(v.x, v.y, v.z, vi.w);
        fbits = __float_as_uint(sqrtf(fi.x * fi.x + fi.y * fi.y + fi.z * fi.z));
    }
    fbits = wave_max_u32(fbits);
    if ((threadIdx.x & 63) == 0 && fbits) atomicMax(max_force, fbits);
}

// XSPH velocity smoothing (src/equations/sph_equations3.cpp:341-382), active
// only when pseudoViscosity * dt > 0.1.
__global__ void __launch_bounds__(kBlock) k_xsph(SimConst c, StepConst st,
                                                 const float4 *__restrict__ pos,
                                                 const float4 *__restrict__ vel,
                                                 const float *__restrict__ rho,
                                                 const int *__restrict__ nbr,
                                                 const int *__restrict__ nbr_cnt,
                                                 float4 *__restrict__ vel_out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    const float4 xi = pos[i], vi = vel[i];
    float sx = 0.f, sy = 0.f, sz = 0.f, ws = 0.f;
    const float spiky_w = c.spiky_lap * c.h * c.h / 6.f;  // 15 / (pi h^3)
    const int cnt = nbr_cnt[i];
    for (int k = 0; k < cnt; k++) {
        int j = nbr[(size_t)k * c.stride + i];
        float4 xj = pos[j], vj = vel[j];
        float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        float r = sqrtf(dx * dx + dy * dy + dz * dz);
        float w = 0.f;
        if (spiky_in(c, r)) {
            float x = 1.f - r * c.inv_h;
            w = spiky_w * x * x * x;
        }
        float wj = c.mass / rho[j] * w;
        ws += wj;
        sx += wj * vj.x;
        sy += wj * vj.y;
        sz += wj * vj.z;
    }
    if (ws > 0.f) {
        float inv = 1.f / ws;
        sx *= inv; sy *= inv; sz *= inv;
    }
    float f = fminf(fmaxf(st.dt * c.pseudo_visc, 0.f), 1.f);
    vel_out[i] = make_float4((1.f - f) * vi.x + f * sx, (1.f - f) * vi.y + f * sy,
                             (1.f - f) * vi.z + f * sz, vi.w);
}

// warm start: p = MLP output, densityPredicted = density (k = 0 init)
__global__ void k_warm_init(SimConst c, const float *__restrict__ pwarm,
                            const float *__restrict__ rho, float *__restrict__ pres,
                            float *__restrict__ rho_star, float *__restrict__ q) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.n) return;
    float p = pwarm[i];
    float d = rho[i];
    pres[i] = p;
    rho_star[i] = d;
    float d2 = d * d;
    q[i] = (fabsf(d2) < 1e-8f) ? 0.f : p / d2;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int run_nonpressure(bbx_solver *s, const StepConst &st, int predict, int features) {
    if (s->n == 0) return BBX_OK;
    const unsigned nb = nblocks(s->n);
    if (features) {
        PROF_BEGIN("nonpressure_features");
        k_nonpressure<2><<<nb, kBlock, 0, s->stream>>>(s->sc, st, s->posA, s->velA, s->rho, s->nbr,
                                                       s->nbr_cnt, s->force, s->xstar, s->rho_star,
                                                       s->feat, s->max_err);
        BBX_HIP_CHECK(hipGetLastError());
        PROF_END();
    } else if (predict) {
        PROF_BEGIN("nonpressure_predict");
        k_nonpressure<1><<<nb, kBlock, 0, s->stream>>>(s->sc, st, s->posA, s->velA, s->rho, s->nbr,
                                                       s->nbr_cnt, s->force, s->xstar, s->rho_star,
                                                       s->feat, s->max_err);
        BBX_HIP_CHECK(hipGetLastError());
        PROF_END();
    } else {
        PROF_BEGIN("nonpressure");
        k_nonpressure<0><<<nb, kBlock, 0, s->stream>>>(s->sc, st, s->posA, s->velA, s->rho, s->nbr,
                                                       s->nbr_cnt, s->force, s->xstar, s->rho_star,
                                                       s->feat, nullptr);
        BBX_HIP_CHECK(hipGetLastError());
        PROF_END();
    }
    return BBX_OK;
}

int run_predict_pressure(bbx_solver *s, const StepConst &st) {
    if (s->n == 0) return BBX_OK;
    // the pressure increment uses the host's FP64 delta (s->delta64)
    PROF_BEGIN("predict_pressure");
    k_predict_pressure<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, st, s->delta64, s->xstar, s->nbr,
                                                                s->nbr_cnt, s->pres, s->rho_star,
                                                                s->q, s->max_err);
    BBX_HIP_CHECK(hipGetLastError());
    PROF_END();
    return BBX_OK;
}

int run_pressure_force(bbx_solver *s, const StepConst &st, int integrate) {
    if (s->n == 0) return BBX_OK;
    const unsigned nb = nblocks(s->n);
    if (integrate) {
        PROF_BEGIN("pressure_force_integrate");
        k_pressure_force<1><<<nb, kBlock, 0, s->stream>>>(s->sc, st, s->posA, s->velA, s->force, s->q,
                                                          s->pres, s->nbr, s->nbr_cnt, s->xstar, s->fp,
                                                          s->posB, s->velB, s->force, s->flag,
                                                          s->max_force, s->max_err);
        BBX_HIP_CHECK(hipGetLastError());
        PROF_END();
    } else {
        PROF_BEGIN("pressure_force_predict");
        k_pressure_force<0><<<nb, kBlock, 0, s->stream>>>(s->sc, st, s->posA, s->velA, s->force, s->q,
                                                          s->pres, s->nbr, s->nbr_cnt, s->xstar, s->fp,
                                                          s->posB, s->velB, s->force, s->flag,
                                                          s->max_force, s->max_err);
        BBX_HIP_CHECK(hipGetLastError());
        PROF_END();
    }
    return BBX_OK;
}

int run_sph_force(bbx_solver *s) {
    if (s->n == 0) return BBX_OK;
    PROF_BEGIN("sph_force");
    k_sph_force<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, s->posA, s->velA, s->rho, s->pres,
                                                         s->nbr, s->nbr_cnt, s->force);
    BBX_HIP_CHECK(hipGetLastError());
    PROF_END();
    return BBX_OK;
}

int run_integrate(bbx_solver *s, const StepConst &st) {
    if (s->n == 0) return BBX_OK;
    PROF_BEGIN("integrate");
    k_integrate<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, st, s->posA, s->velA, s->force,
                                                         s->posB, s->velB, s->flag, s->max_force);
    BBX_HIP_CHECK(hipGetLastError());
    PROF_END();
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
    PROF_BEGIN("xsph");
    k_xsph<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, st, s->posA, s->velA, s->rho, s->nbr,
                                                    s->nbr_cnt, s->velB);
    BBX_HIP_CHECK(hipGetLastError());
    PROF_END();
    std::swap(s->velA, s->velB);
    return BBX_OK;
}

int run_warm_init(bbx_solver *s) {
    if (s->n == 0) return BBX_OK;
    k_warm_init<<<nblocks(s->n), kBlock, 0, s->stream>>>(s->sc, s->pwarm, s->rho, s->pres,
                                                         s->rho_star, s->q);
    BBX_HIP_CHECK(hipGetLastError());
    return BBX_OK;
}

}  // namespace bbx
