// bbx_abi.hip -- C ABI of the MI355X SPH / PCISPH hot path: solver lifetime,
// Setup scalars (FP64 on the host), id-ordered transfers, substep control.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bbx_common.h"

namespace bbx {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

size_t scan_scratch_elems(size_t n);

// ---------------------------------------------------------------------------
// profiler
// ---------------------------------------------------------------------------
int Profiler::name_id(const std::string &name) {
    for (size_t i = 0; i < names.size(); i++)
        if (names[i] == name) return (int)i;
    names.emplace_back(name);
    total_ms.push_back(0.0);
    launches.push_back(0);
    return (int)names.size() - 1;
}

bool Profiler::wants(const char *name, int iter) const {
    if (!only.empty() && only != name) return false;
    if (only_iter != kNoIter && iter != only_iter) return false;
    return true;
}

static hipEvent_t pool_get(std::vector<hipEvent_t> &pool) {
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    // timing only: no system-scope fence (an L2 write-back and invalidate)
    // when the event is recorded
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

void Profiler::begin(hipStream_t s, hipEvent_t *a) {
    *a = pool_get(pool);
    if (*a) (void)hipEventRecord(*a, s);
}

void Profiler::end(hipStream_t s, int id, int iter, hipEvent_t a) {
    hipEvent_t b = pool_get(pool);
    if (!a || !b) return;
    (void)hipEventRecord(b, s);
    pending.push_back({id, iter, a, b});
}

static std::string role_name(const std::string &base, int iter, int iterations) {
    if (iter == kNoIter || iterations < 0) return base;
    if (base == "predict_pressure") return iter < iterations ? base : base + "_skipped";
    if (base == "pressure_force") {
        if (iter < 0) return base + "_warm";
        if (iter < iterations - 1) return base + "_predict";
        if (iter == iterations - 1) return base + "_integrate";
        return base + "_skipped";
    }
    return base;
}

int Profiler::resolve(int iterations, size_t upto) {
    const size_t m = upto < pending.size() ? upto : pending.size();
    for (size_t k = 0; k < m; k++) {
        const Pending &p = pending[k];
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            const int id = name_id(role_name(names[p.id], p.iter, iterations));
            total_ms[id] += ms;
            launches[id] += 1;
        }
        pool.push_back(p.a);
        pool.push_back(p.b);
    }
    pending.erase(pending.begin(), pending.begin() + (std::ptrdiff_t)m);
    return BBX_OK;
}

void Profiler::reset() {
    resolve();
    for (auto &t : total_ms) t = 0.0;
    for (auto &l : launches) l = 0;
}

void Profiler::destroy() {
    resolve();
    for (auto e : pool) (void)hipEventDestroy(e);
    pool.clear();
}

// ---------------------------------------------------------------------------
// Setup scalars in FP64 (host)
// ---------------------------------------------------------------------------
static inline double absd(double v) { return v > 0 ? v : -v; }
static inline double maxd(double a, double b) { return a < b ? b : a; }
static inline double mind(double a, double b) { return a < b ? a : b; }

// Grid::Build after UtilBuildGridForDomain (src/core/util.cpp:269-287,
// src/core/grid.h:626-666)
int build_grid_geom(const bbx_params &p, GridGeom *g) {
    double dmin[3], dmax[3];
    for (int k = 0; k < 3; k++) {
        double hx = p.tank_size[k] / 2.0;
        dmin[k] = -hx + p.tank_center[k];
        dmax[k] = hx + p.tank_center[k];
    }
    const double length = p.grid_spacing * p.grid_scale;
    if (!(length > 0)) return fail(BBX_E_INVALID, "grid spacing * scale must be > 0");
    const double hlen = 0.5 * length;
    const double inv_len = 1.0f / length;
    int m[3];
    for (int k = 0; k < 3; k++) {
        m[k] = (int)std::ceil(absd(dmax[k] - dmin[k]) * inv_len);
        m[k] = (int)(m[k] + (m[k] % 2) * length);
        if (m[k] <= 0) return fail(BBX_E_INVALID, "empty tank");
    }
    double pmin[3], pmax[3];
    for (int k = 0; k < 3; k++) {
        double center = (dmin[k] + dmax[k]) * 0.5;
        pmin[k] = center - m[k] * hlen;
        pmax[k] = center + m[k] * hlen;
    }
    // Build(resolution, pMin, pMax)
    double lower[3], high[3];
    double total = 1;
    for (int k = 0; k < 3; k++) {
        double dl = maxd(pmin[k], pmax[k]), du = mind(pmin[k], pmax[k]);
        lower[k] = (double)3.402823466e+38f;
        high[k] = -(double)3.402823466e+38f;
        if (dl < lower[k]) lower[k] = dl;
        if (dl > high[k]) high[k] = dl;
        if (du > high[k]) high[k] = du;
        if (du < lower[k]) lower[k] = du;
    }
    for (int k = 0; k < 3; k++) {
        double s = high[k] - lower[k];
        double len = s / (double)(unsigned)m[k];
        g->len[k] = len;
        g->n[k] = (int)std::ceil(s / len);
        double maxp = lower[k] + (double)(unsigned)g->n[k] * len;
        g->minp[k] = lower[k];
        g->bmin[k] = mind(lower[k], maxp);
        g->bmax[k] = maxd(lower[k], maxp);
        total *= g->n[k];
    }
    if (total > 2.0e9) return fail(BBX_E_RANGE, "grid has more than 2^31 cells");
    for (int k = 0; k < 3; k++)  // the device's FP32 cell-id decode is exact below 2^22 per axis
        if (g->n[k] >= (1 << 22)) return fail(BBX_E_RANGE, "grid has 2^22 or more cells along an axis");
    g->total = (unsigned)total;
    return BBX_OK;
}

static double poly6_w64(double h, double distance) {
    double h2 = h * h, h3 = h2 * h;
    double d2 = distance * distance;
    double of = d2 - h2;
    if (absd(of) < 1e-8 || of > 0) return 0;
    double x = 1.0 - d2 / h2;
    return 315.0 / (64.0 * M_PI * h3) * x * x * x;
}

static double spiky_dw64(double h, double distance) {
    double h4 = (h * h) * (h * h);
    double of = distance - h;
    if (absd(of) < 1e-8 || of > 0) return 0;
    double x = 1.0 - distance / h;
    return -45.0 / (M_PI * h4) * x * x;
}

// BCC lattice over [-1.5h, 1.5h]^3 (src/generator/bcclattice.cpp:5-76)
static void bcc_points(double kr, double spacing, int maxn, std::vector<double> &pts) {
    const double lo = -1.5 * kr, hi = 1.5 * kr;
    const double half = spacing / 2;
    const double ext = absd(hi - lo);
    bool off = false, stop = false;
    pts.clear();
    for (int k = 0; k * half <= ext && !stop; k++) {
        double z = k * half + lo;
        double o = off ? half : 0;
        for (int j = 0; j * spacing + o <= ext && !stop; j++) {
            double y = j * spacing + o + lo;
            for (int i = 0; i * spacing + o <= ext && !stop; i++) {
                if (maxn >= 0 && (int)(pts.size() / 3) >= maxn) { stop = true; break; }
                pts.push_back(i * spacing + o + lo);
                pts.push_back(y);
                pts.push_back(z);
            }
        }
        off = !off;
    }
}

// SphParticleSet3::ComputeMass (src/core/particle.h:557-582) and
// PciSphSolver3::ComputeDeltaDenom (src/solvers/pcisph_solver3.cpp:148-180)
int compute_mass_and_delta(bbx_solver *s) {
    const double rho0 = s->prm.target_density;
    const double spacing = s->prm.target_spacing;
    s->radius = maxd(0, spacing);
    s->kr = s->prm.relative_radius * spacing;
    if (!(s->kr > 0) || !(spacing > 0)) return fail(BBX_E_INVALID, "kernel radius must be > 0");
    std::vector<double> pts;
    bcc_points(s->kr, spacing, 1024, pts);
    const int c = (int)(pts.size() / 3);
    double nden = 0;
    for (int i = 0; i < c; i++) {
        double sum = 0;
        for (int j = 0; j < c; j++) {
            double dx = pts[3 * i] - pts[3 * j], dy = pts[3 * i + 1] - pts[3 * j + 1],
                   dz = pts[3 * i + 2] - pts[3 * j + 2];
            sum += poly6_w64(s->kr, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
        nden = maxd(nden, sum);
    }
    if (!(nden > 0)) return fail(BBX_E_INVALID, "zero number density");
    s->mass = maxd(0, rho0 / nden);
    double m = s->mass / rho0;
    m *= m;
    s->mass_over_rho2 = m;
    bcc_points(s->kr, spacing, -1, pts);
    const int np = (int)(pts.size() / 3);
    const double kr2 = s->kr * s->kr;
    double d1x = 0, d1y = 0, d1z = 0, denom2 = 0;
    for (int i = 0; i < np; i++) {
        double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
        double d2 = px * px + py * py + pz * pz;
        if (d2 < kr2) {
            double dist = std::sqrt(d2);
            double dx = 0, dy = 0, dz = 0;
            if (dist > 0) {
                double inv = (double)1 / dist;
                dx = px * inv; dy = py * inv; dz = pz * inv;
            }
            double w = -spiky_dw64(s->kr, dist);
            double gx = dx * w, gy = dy * w, gz = dz * w;
            d1x += gx; d1y += gy; d1z += gz;
            denom2 += gx * gx + gy * gy + gz * gz;
        }
    }
    double denom = 0;
    denom += -(d1x * d1x + d1y * d1y + d1z * d1z) - denom2;
    s->delta_denom = denom;
    return BBX_OK;
}

// PciSphSolver3::ComputeDelta (src/solvers/pcisph_solver3.cpp:182-189)
static double compute_delta(const bbx_solver *s, double dt) {
    double tss = dt * dt;
    double beta = 2.0 * s->mass_over_rho2 * tss;
    return absd(s->delta_denom) > 0 ? -1 / (beta * s->delta_denom) : 0;
}

void fill_sim_const(bbx_solver *s) {
    SimConst &c = s->sc;
    c.n = s->n;
    c.stride = s->stride;
    c.g = s->grid;
    const double h = s->kr;
    c.h = (float)h;
    c.h2 = (float)(h * h);
    c.inv_h = (float)(1.0 / h);
    c.inv_h2 = (float)(1.0 / (h * h));
    c.poly6 = (float)(315.0 / (64.0 * M_PI * h * h * h));
    c.spiky_grad = (float)(45.0 / (M_PI * h * h * h * h));
    c.spiky_lap = (float)(90.0 / (M_PI * h * h * h * h * h));
    c.h2d = h * h;
    c.rho0d = s->prm.target_density;
    // FP32 prefilter band: outside [T(1-4e-6), T(1+4e-6)] with T = h^2 - 1e-8 the
    // FP32 squared distance (relative error < 5 * 2^-24) decides; inside it the
    // FP64 rule decides.
    const double T = h * h - 1e-8;
    if (T > 1e-6) {
        c.lo2 = std::nextafter((float)(T * (1.0 - 4e-6)), 0.f);
        c.hi2 = std::nextafter((float)(T * (1.0 + 4e-6)), 3.4e38f);
    } else {
        c.lo2 = -1.f;         // always take the FP64 path
        c.hi2 = 3.4e38f;
    }
    // one-comparison form of the band: every d2 in [lo2, hi2] has
    // |d2 - band_c| <= band_w (d2 - band_c is exact near the cutoff)
    c.band_c = (float)(((double)c.lo2 + (double)c.hi2) / 2.0);
    c.band_w = (float)(((double)c.hi2 - (double)c.lo2) / 2.0 * 1.001);
    for (int k = 0; k < 4; k++) c.band_w = std::nextafter(c.band_w, 3.4e38f);
    c.mass = (float)s->mass;
    c.mass2 = (float)(s->mass * s->mass);
    c.rho0 = (float)s->prm.target_density;
    c.visc = (float)s->prm.viscosity;
    c.drag = (float)s->prm.drag_coefficient;
    c.nps = (float)s->prm.negative_pressure_scale;
    c.gx = (float)s->prm.gravity[0];
    c.gy = (float)s->prm.gravity[1];
    c.gz = (float)s->prm.gravity[2];
    c.radius = (float)s->radius;
    for (int k = 0; k < 3; k++) {
        c.tank_c[k] = (float)s->prm.tank_center[k];
        c.tank_h[k] = (float)(s->prm.tank_size[k] / 2.0);
        c.bmin[k] = (float)s->grid.bmin[k];
        c.bmax[k] = (float)s->grid.bmax[k];
    }
    c.min_len = (float)mind(s->grid.len[0], mind(s->grid.len[1], s->grid.len[2]));
    c.eos_scale = (float)(s->prm.target_density * s->prm.sound_speed * s->prm.sound_speed);
    c.eos_exp = (float)s->prm.eos_exponent;
    c.pseudo_visc = (float)s->prm.pseudo_viscosity;
    // LDS tile capacities; BBX_TILE_CAP (0 .. kTileCap) and BBX_TILE_CAP2
    // (0 .. kTileCap2) lower them to time and test the global-memory paths
    c.tile_cap = kTileCapHost;
    c.tile_cap2 = kTileCap2Host;
    c.ncoll = s->ncoll;
    c.coll = s->coll;
    if (const char *e = std::getenv("BBX_TILE_CAP")) {
        int v = std::atoi(e);
        c.tile_cap = v < 0 ? 0 : (v > kTileCapHost ? kTileCapHost : v);
    }
    // rows of up to 64 candidates take the two-pass mask path of the bucket
    // build; BBX_BUCKET_ROW_MAX (0 .. 64) lowers the limit to test the other path
    c.inv_nx = (float)(1.0 / (double)c.g.n[0]);
    c.inv_nxy = (float)(1.0 / ((double)c.g.n[0] * (double)c.g.n[1]));
    c.row_max = 64;
    if (const char *e = std::getenv("BBX_BUCKET_ROW_MAX")) {
        int v = std::atoi(e);
        c.row_max = v < 0 ? 0 : (v > 64 ? 64 : v);
    }
    if (const char *e = std::getenv("BBX_TILE_CAP2")) {
        int v = std::atoi(e);
        c.tile_cap2 = v < 0 ? 0 : (v > kTileCap2Host ? kTileCap2Host : v);
    }
}

// host-side hash: Grid::GetLinearHashedPosition (src/core/grid.h:258-321)
static long long host_hash(const GridGeom &g, const double p[3]) {
    int u[3];
    for (int i = 0; i < 3; i++) {
        double pi = p[i];
        double eps = 0;
        if (absd(pi - g.bmin[i]) < 1e-8) eps = (double)0.0001f;
        else if (absd(pi - g.bmax[i]) < 1e-8) eps = -(double)0.0001f;
        pi += eps;
        double f = std::floor((pi - g.minp[i]) / g.len[i]);
        if (!(f >= 0) || f >= (double)g.n[i]) return -1;
        u[i] = (int)f;
    }
    return (long long)u[0] + (long long)u[1] * g.n[0] + (long long)u[2] * g.n[0] * g.n[1];
}

}  // namespace bbx

using namespace bbx;

// ---------------------------------------------------------------------------
// allocation helpers
// ---------------------------------------------------------------------------
template <typename T>
static int dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(BBX_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    return BBX_OK;
}

template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

static void free_all(bbx_solver *s) {
    dfree(s->posA); dfree(s->posB); dfree(s->velA); dfree(s->velB);
    dfree(s->cellA); dfree(s->cellB); dfree(s->home);
    dfree(s->cell_count); dfree(s->cell_start); dfree(s->scan_tmp);
    dfree(s->key); dfree(s->newc); dfree(s->rank_tmp); dfree(s->tmp_idx); dfree(s->hash_tmp);
    dfree(s->nbr16); dfree(s->nbr32); dfree(s->nbr_cnt); dfree(s->tile_tab);
    dfree(s->rho); dfree(s->pres); dfree(s->rho_star); dfree(s->xq);
    dfree(s->force); dfree(s->fp); dfree(s->xstar); dfree(s->feat); dfree(s->pwarm);
    dfree(s->step); dfree(s->flag);
    s->loop = nullptr; s->max_force = nullptr; s->counters = nullptr;
    dfree(s->coll);
    s->ncoll = 0;
    if (s->h_step) (void)hipHostFree(s->h_step);
    if (s->h_flag) (void)hipHostFree(s->h_flag);
    if (s->h_signal) (void)hipHostFree(s->h_signal);
    s->h_step = nullptr; s->h_flag = nullptr;
    s->h_signal = nullptr; s->d_signal = nullptr;
    s->h_loop = nullptr; s->h_counters = nullptr; s->h_max_force = nullptr;
    for (int k = 0; k < 2; k++) {
        if (s->ev0[k]) (void)hipEventDestroy(s->ev0[k]);
        if (s->ev1[k]) (void)hipEventDestroy(s->ev1[k]);
        if (s->ev_read[k]) (void)hipEventDestroy(s->ev_read[k]);
        s->ev0[k] = s->ev1[k] = s->ev_read[k] = nullptr;
    }
    for (int k = 0; k < bbx_solver::kMarks; k++) {
        if (s->mark[k]) (void)hipEventDestroy(s->mark[k]);
        s->mark[k] = nullptr;
    }
    s->prof.destroy();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->stream = nullptr;
}

static int check_solver(bbx_solver *s) {
    if (!s) return fail(BBX_E_INVALID, "null solver");
    return BBX_OK;
}

static int check_setup(bbx_solver *s) {
    BBX_CHECK(check_solver(s));
    if (!s->setup_done) return fail(BBX_E_INVALID, "call bbx_solver_setup first");
    return BBX_OK;
}

// download the slot -> id map (pos.w of the current storage)
static int slot_ids(bbx_solver *s, std::vector<int> &ids, std::vector<float> *pos4 = nullptr) {
    std::vector<float> tmp((size_t)s->n * 4);
    BBX_HIP_CHECK(hipMemcpyAsync(tmp.data(), s->posA, sizeof(float4) * s->n, hipMemcpyDeviceToHost,
                                 s->stream));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    ids.resize(s->n);
    for (int i = 0; i < s->n; i++) {
        int id;
        std::memcpy(&id, &tmp[4 * (size_t)i + 3], sizeof(int));
        if (id < 0 || id >= s->n) return fail(BBX_E_INVALID, "corrupt particle id table");
        ids[i] = id;
    }
    if (pos4) pos4->swap(tmp);
    return BBX_OK;
}

extern "C" {

const char *bbx_last_error(void) { return g_last_error.c_str(); }

int bbx_version(int *major, int *minor) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    return BBX_OK;
}

int bbx_device_count(int *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    if (count) *count = n;
    return BBX_OK;
}

int bbx_default_params(bbx_params *p) {
    if (!p) return fail(BBX_E_INVALID, "null params");
    std::memset(p, 0, sizeof(*p));
    p->target_density = 1000.0;
    p->target_spacing = 0.1;
    p->relative_radius = 2.0;
    p->tank_size[0] = p->tank_size[1] = p->tank_size[2] = 1.0;
    p->grid_spacing = 0.1;
    p->grid_scale = 2.0;
    p->eos_exponent = 7.0;
    p->negative_pressure_scale = 0.0;
    p->viscosity = 0.04;
    p->pseudo_viscosity = 10.0;
    p->sound_speed = 100.0;
    p->drag_coefficient = 0.0001;
    p->gravity[1] = (double)(-9.8f);
    return BBX_OK;
}

int bbx_solver_create(int kind, const bbx_params *params, const double *pos, const double *vel,
                      int n, bbx_solver **out) {
    if (!out || !params) return fail(BBX_E_INVALID, "null argument");
    *out = nullptr;
    if (kind != BBX_PCISPH && kind != BBX_SPH) return fail(BBX_E_INVALID, "unknown solver kind");
    if (n < 0 || (n > 0 && !pos)) return fail(BBX_E_INVALID, "bad particle array");
    if (n >= (1 << 28)) return fail(BBX_E_RANGE, "more than 2^28 particles");
    int ndev = 0;
    bbx_device_count(&ndev);
    if (ndev <= 0) return fail(BBX_E_NODEVICE, "no HIP device visible");
    bbx_solver *s = new bbx_solver();
    s->kind = kind;
    s->prm = *params;
    s->n = n;
    s->stride = ((n + 63) / 64) * 64;
    if (s->stride == 0) s->stride = 64;
    int rc = build_grid_geom(s->prm, &s->grid);
    if (rc != BBX_OK) { delete s; return rc; }
    hipError_t he = hipGetDevice(&s->device);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        delete s;
        return fail(BBX_E_HIP, std::string("stream create: ") + hipGetErrorString(he));
    }
    const size_t N = (size_t)s->stride;
    const size_t cells = (size_t)s->grid.total + 1;
    s->scan_tmp_elems = scan_scratch_elems(cells);
    rc = BBX_OK;
    if (!rc) rc = dalloc(&s->posA, N);
    if (!rc) rc = dalloc(&s->posB, N);
    if (!rc) rc = dalloc(&s->velA, N);
    if (!rc) rc = dalloc(&s->velB, N);
    if (!rc) rc = dalloc(&s->cellA, N);
    if (!rc) rc = dalloc(&s->cellB, N);
    if (!rc) rc = dalloc(&s->home, N);
    if (!rc) rc = dalloc(&s->cell_count, cells);
    if (!rc) rc = dalloc(&s->cell_start, cells);
    if (!rc) rc = dalloc(&s->scan_tmp, s->scan_tmp_elems);
    if (!rc) rc = dalloc(&s->key, N);
    if (!rc) rc = dalloc(&s->newc, N);
    if (!rc) rc = dalloc(&s->rank_tmp, N);
    if (!rc) rc = dalloc(&s->tmp_idx, N);
    if (!rc) rc = dalloc(&s->hash_tmp, N);
    s->ntiles = (int)((N + kBlock - 1) / kBlock);
    if (!rc) rc = dalloc(&s->nbr16, N * kBucket);
    if (!rc) rc = dalloc(&s->nbr32, N * kBucket);
    if (!rc) rc = dalloc(&s->nbr_cnt, N);
    if (!rc) rc = dalloc(&s->tile_tab, (size_t)s->ntiles * kTabIntsHost);
    if (!rc) rc = dalloc(&s->rho, N);
    if (!rc) rc = dalloc(&s->pres, N);
    if (!rc) rc = dalloc(&s->rho_star, N);
    if (!rc) rc = dalloc(&s->xq, N);
    if (!rc) rc = dalloc(&s->force, N);
    if (!rc) rc = dalloc(&s->fp, N);
    if (!rc) rc = dalloc(&s->xstar, N);
    if (!rc) rc = dalloc(&s->feat, N * BBX_NFEAT);
    if (!rc) rc = dalloc(&s->pwarm, N);
    if (!rc) rc = dalloc(&s->step, 1);
    if (!rc) {
        s->loop = &s->step->loop;
        s->max_force = s->step->max_force;
        s->counters = s->step->counters;
    }
    if (!rc) rc = dalloc(&s->flag, 1);
    if (!rc && hipHostMalloc((void **)&s->h_step, sizeof(StepBlock)) != hipSuccess)
        rc = fail(BBX_E_NOMEM, "pinned alloc");
    if (!rc && hipHostMalloc((void **)&s->h_flag, sizeof(int)) != hipSuccess)
        rc = fail(BBX_E_NOMEM, "pinned alloc");
    if (!rc && hipHostMalloc((void **)&s->h_signal, sizeof(LoopSignal), hipHostMallocMapped | hipHostMallocCoherent) !=
                   hipSuccess)
        rc = fail(BBX_E_NOMEM, "pinned alloc");
    if (!rc) {
        std::memset(s->h_signal, 0, sizeof(LoopSignal));
        if (hipHostGetDevicePointer((void **)&s->d_signal, s->h_signal, 0) != hipSuccess)
            rc = fail(BBX_E_HIP, "device address of the loop signal");
        // BBX_LOOP_SIGNAL=0: no published decisions; the host reads every
        // decision back through a copy (the fallback path, for tests)
        if (const char *e = std::getenv("BBX_LOOP_SIGNAL"))
            if (std::atoi(e) == 0) s->d_signal = nullptr;
    }
    if (!rc) {
        s->h_loop = &s->h_step->loop;
        s->h_max_force = s->h_step->max_force;
        s->h_counters = s->h_step->counters;
    }
    // ev0 / ev1 bracket a substep's device time and synchronise nothing: they
    // are recorded without the system-scope fence (an L2 write-back and
    // invalidate, about 5 us at 1M particles, twice per substep).  ev_read
    // follows the statistics copy to pinned memory, which the host reads
    // once it has completed, so it keeps the fence.
    for (int k = 0; k < 2 && !rc; k++)
        if (hipEventCreateWithFlags(&s->ev0[k], hipEventDisableSystemFence) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev1[k], hipEventDisableSystemFence) != hipSuccess ||
            hipEventCreateWithFlags(&s->ev_read[k], hipEventDisableTiming) != hipSuccess)
            rc = fail(BBX_E_HIP, "event create");
    for (int k = 0; k < bbx_solver::kMarks && !rc; k++)
        if (hipEventCreateWithFlags(&s->mark[k], hipEventDisableSystemFence) != hipSuccess)
            rc = fail(BBX_E_HIP, "event create");
    if (rc) {
        free_all(s);
        delete s;
        return rc;
    }
    // upload: FP64 -> FP32, id in pos.w, velocity w = 0 (no previous pressure)
    std::vector<float4> hp(N), hv(N);
    for (int i = 0; i < n; i++) {
        float idf;
        std::memcpy(&idf, &i, sizeof(int));
        hp[i] = make_float4((float)pos[3 * i], (float)pos[3 * i + 1], (float)pos[3 * i + 2], idf);
        hv[i] = vel ? make_float4((float)vel[3 * i], (float)vel[3 * i + 1], (float)vel[3 * i + 2], 0.f)
                    : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    hipError_t e = hipMemcpy(s->posA, hp.data(), sizeof(float4) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->velA, hv.data(), sizeof(float4) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->force, 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(s->fp, 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(s->xstar, 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(s->rho, 0, sizeof(float) * N);
    if (e == hipSuccess) e = hipMemset(s->pres, 0, sizeof(float) * N);
    if (e == hipSuccess) e = hipMemset(s->rho_star, 0, sizeof(float) * N);
    if (e == hipSuccess) e = hipMemset(s->xq, 0, sizeof(float4) * N);
    if (e == hipSuccess) e = hipMemset(s->nbr_cnt, 0, sizeof(int) * N);
    if (e == hipSuccess) e = hipMemset(s->tile_tab, 0, sizeof(int) * (size_t)s->ntiles * kTabIntsHost);
    if (e == hipSuccess) e = hipMemset(s->feat, 0, sizeof(float) * N * BBX_NFEAT);
    if (e == hipSuccess) e = hipMemset(s->max_force, 0, sizeof(unsigned) * kSlots);
    if (e == hipSuccess) e = hipMemset(s->loop, 0, sizeof(LoopState));
    if (e == hipSuccess) e = hipMemset(s->flag, 0, sizeof(int));
    if (e == hipSuccess) e = hipMemset(s->counters, 0, sizeof(int) * CNT_N);
    if (e != hipSuccess) {
        free_all(s);
        delete s;
        return fail(BBX_E_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    // every slot starts in no chain until Setup distributes
    std::vector<int> none(N, (int)s->grid.total);
    e = hipMemcpy(s->cellA, none.data(), sizeof(int) * N, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->home, none.data(), sizeof(int) * N, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        free_all(s);
        delete s;
        return fail(BBX_E_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
    fill_sim_const(s);
    *out = s;
    return BBX_OK;
}

int bbx_solver_destroy(bbx_solver *s) {
    if (!s) return BBX_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    mlp_release(s->mlp);
    s->mlp = nullptr;
    free_all(s);
    delete s;
    return BBX_OK;
}

int bbx_solver_setup(bbx_solver *s) {
    BBX_CHECK(check_solver(s));
    BBX_CHECK(compute_mass_and_delta(s));
    fill_sim_const(s);
    BBX_CHECK(launch_fill_int(s, s->counters, 0, CNT_N));
    // full distribution into ascending-id chains; buckets come with the first substep
    BBX_CHECK(distribute(s, true, false, false));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    s->setup_done = true;
    return BBX_OK;
}

int bbx_solver_info(bbx_solver *s, bbx_info *out) {
    BBX_CHECK(check_solver(s));
    if (!out) return fail(BBX_E_INVALID, "null info");
    out->count = s->n;
    out->mass = s->mass;
    out->kernel_radius = s->kr;
    out->radius = s->radius;
    out->delta_denom = s->delta_denom;
    for (int k = 0; k < 3; k++) {
        out->grid.usizes[k] = (unsigned)s->grid.n[k];
        out->grid.cells_len[k] = s->grid.len[k];
        out->grid.min_point[k] = s->grid.minp[k];
        out->grid.bounds_min[k] = s->grid.bmin[k];
        out->grid.bounds_max[k] = s->grid.bmax[k];
    }
    out->grid.total = s->grid.total;
    return BBX_OK;
}

int bbx_solver_download(bbx_solver *s, int field, double *out) {
    BBX_CHECK(check_solver(s));
    if (!out) return fail(BBX_E_INVALID, "null output");
    if (s->n == 0) return BBX_OK;
    std::vector<int> ids;
    BBX_CHECK(slot_ids(s, ids));
    const size_t n = (size_t)s->n;
    auto vec4 = [&](const float4 *src) -> int {
        std::vector<float4> h(n);
        BBX_HIP_CHECK(hipMemcpy(h.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            double *o = &out[3 * (size_t)ids[i]];
            o[0] = h[i].x; o[1] = h[i].y; o[2] = h[i].z;
        }
        return BBX_OK;
    };
    auto scal = [&](const float *src) -> int {
        std::vector<float> h(n);
        BBX_HIP_CHECK(hipMemcpy(h.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) out[ids[i]] = h[i];
        return BBX_OK;
    };
    switch (field) {
        case BBX_FIELD_POSITION: return vec4(s->posA);
        case BBX_FIELD_VELOCITY: return vec4(s->velA);
        case BBX_FIELD_FORCE: return vec4(s->force);
        case BBX_FIELD_PRESSURE_FORCE: return vec4(s->fp);
        case BBX_FIELD_TEMP_POSITION: return vec4(s->xstar);
        case BBX_FIELD_DENSITY: return scal(s->rho);
        case BBX_FIELD_PRESSURE: return scal(s->pres);
        case BBX_FIELD_DENSITY_PREDICTED: return scal(s->rho_star);
        case BBX_FIELD_WARM_PRESSURE: return scal(s->pwarm);
        case BBX_FIELD_FEATURES: {
            std::vector<float> h(n * BBX_NFEAT);
            for (int f = 0; f < BBX_NFEAT; f++)
                BBX_HIP_CHECK(hipMemcpy(&h[f * n], s->feat + (size_t)f * s->stride, sizeof(float) * n,
                                        hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; i++)
                for (int f = 0; f < BBX_NFEAT; f++) out[(size_t)ids[i] * BBX_NFEAT + f] = h[f * n + i];
            return BBX_OK;
        }
        default: return fail(BBX_E_INVALID, "unknown field");
    }
}

// ---------------------------------------------------------------------------
// growing particle sets (ContinuousParticleSetBuilder, src/core/grid.h:1246-1450)
// ---------------------------------------------------------------------------
}  // extern "C"
namespace {
// replace *p by a buffer of new_n elements holding the first keep elements
template <typename T>
int regrow(bbx_solver *s, T *&p, size_t new_n, size_t keep) {
    T *q = nullptr;
    BBX_CHECK(dalloc(&q, new_n));
    if (keep && p) {
        hipError_t e = hipMemcpyAsync(q, p, sizeof(T) * keep, hipMemcpyDeviceToDevice, s->stream);
        if (e != hipSuccess) {
            dfree(q);
            return fail(BBX_E_HIP, std::string("regrow copy: ") + hipGetErrorString(e));
        }
    }
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    dfree(p);
    p = q;
    return BBX_OK;
}
}  // namespace
extern "C" {

int bbx_solver_reserve(bbx_solver *s, int capacity) {
    BBX_CHECK(check_solver(s));
    if (capacity < 0) return fail(BBX_E_INVALID, "negative capacity");
    const size_t S0 = (size_t)s->stride;
    const size_t S1 = (size_t)(((capacity + 63) / 64) * 64);
    if (S1 <= S0) return BBX_OK;
    const size_t n = (size_t)s->n;
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    // particle state kept (first n slots), scratch and per-substep tables not
    BBX_CHECK(regrow(s, s->posA, S1, n));
    BBX_CHECK(regrow(s, s->velA, S1, n));
    BBX_CHECK(regrow(s, s->cellA, S1, n));
    BBX_CHECK(regrow(s, s->home, S1, n));
    BBX_CHECK(regrow(s, s->force, S1, n));
    BBX_CHECK(regrow(s, s->fp, S1, n));
    BBX_CHECK(regrow(s, s->xstar, S1, n));
    BBX_CHECK(regrow(s, s->xq, S1, n));
    BBX_CHECK(regrow(s, s->rho, S1, n));
    BBX_CHECK(regrow(s, s->pres, S1, n));
    BBX_CHECK(regrow(s, s->rho_star, S1, n));
    BBX_CHECK(regrow(s, s->pwarm, S1, n));
    BBX_CHECK(regrow(s, s->posB, S1, 0));
    BBX_CHECK(regrow(s, s->velB, S1, 0));
    BBX_CHECK(regrow(s, s->cellB, S1, 0));
    BBX_CHECK(regrow(s, s->key, S1, 0));
    BBX_CHECK(regrow(s, s->newc, S1, 0));
    BBX_CHECK(regrow(s, s->rank_tmp, S1, 0));
    BBX_CHECK(regrow(s, s->tmp_idx, S1, 0));
    BBX_CHECK(regrow(s, s->hash_tmp, S1, 0));
    BBX_CHECK(regrow(s, s->nbr16, S1 * kBucket, 0));
    BBX_CHECK(regrow(s, s->nbr32, S1 * kBucket, 0));
    BBX_CHECK(regrow(s, s->nbr_cnt, S1, 0));
    // features are feature-major with the row stride: copy row by row
    {
        float *q = nullptr;
        BBX_CHECK(dalloc(&q, S1 * BBX_NFEAT));
        for (int f = 0; f < BBX_NFEAT && n; f++)
            BBX_HIP_CHECK(hipMemcpyAsync(q + (size_t)f * S1, s->feat + (size_t)f * S0, sizeof(float) * n,
                                         hipMemcpyDeviceToDevice, s->stream));
        BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
        dfree(s->feat);
        s->feat = q;
    }
    s->ntiles = (int)((S1 + kBlock - 1) / kBlock);
    BBX_CHECK(regrow(s, s->tile_tab, (size_t)s->ntiles * kTabIntsHost, 0));
    BBX_HIP_CHECK(hipMemset(s->nbr_cnt, 0, sizeof(int) * S1));
    BBX_HIP_CHECK(hipMemset(s->tile_tab, 0, sizeof(int) * (size_t)s->ntiles * kTabIntsHost));
    // slots beyond n are in no chain
    std::vector<int> none(S1 - n, (int)s->grid.total);
    BBX_HIP_CHECK(hipMemcpy(s->cellA + n, none.data(), sizeof(int) * none.size(), hipMemcpyHostToDevice));
    BBX_HIP_CHECK(hipMemcpy(s->home + n, none.data(), sizeof(int) * none.size(), hipMemcpyHostToDevice));
    s->stride = (int)S1;
    fill_sim_const(s);
    return BBX_OK;
}

int bbx_solver_capacity(bbx_solver *s, int *capacity) {
    BBX_CHECK(check_solver(s));
    if (!capacity) return fail(BBX_E_INVALID, "null output");
    *capacity = s->stride;
    return BBX_OK;
}

int bbx_solver_append(bbx_solver *s, const double *pos, const double *vel, int count) {
    BBX_CHECK(check_setup(s));
    if (count < 0) return fail(BBX_E_INVALID, "negative count");
    if (count == 0) return BBX_OK;
    if (!pos) return fail(BBX_E_INVALID, "null positions");
    if ((long long)s->n + count > 0x7fffffffLL / kBucket) return fail(BBX_E_RANGE, "too many particles");
    const int n0 = s->n;
    if (n0 + count > s->stride) BBX_CHECK(bbx_solver_reserve(s, std::max(2 * s->stride, n0 + count)));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    // AppendData (src/core/particle.h:232-240): ids n0.., zero force, and
    // no previous pressure for the warm start
    std::vector<float4> hp((size_t)count), hv((size_t)count);
    for (int k = 0; k < count; k++) {
        const int id = n0 + k;
        float idf;
        std::memcpy(&idf, &id, sizeof(int));
        hp[k] = make_float4((float)pos[3 * (size_t)k], (float)pos[3 * (size_t)k + 1], (float)pos[3 * (size_t)k + 2], idf);
        hv[k] = vel ? make_float4((float)vel[3 * (size_t)k], (float)vel[3 * (size_t)k + 1], (float)vel[3 * (size_t)k + 2], 0.f)
                    : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    BBX_HIP_CHECK(hipMemcpy(s->posA + n0, hp.data(), sizeof(float4) * count, hipMemcpyHostToDevice));
    BBX_HIP_CHECK(hipMemcpy(s->velA + n0, hv.data(), sizeof(float4) * count, hipMemcpyHostToDevice));
    BBX_HIP_CHECK(hipMemset(s->force + n0, 0, sizeof(float4) * count));
    BBX_HIP_CHECK(hipMemset(s->fp + n0, 0, sizeof(float4) * count));
    BBX_HIP_CHECK(hipMemset(s->rho + n0, 0, sizeof(float) * count));
    BBX_HIP_CHECK(hipMemset(s->pres + n0, 0, sizeof(float) * count));
    s->n = n0 + count;
    fill_sim_const(s);
    // DistributeByParticleList (src/core/grid.h:358-387): each new particle
    // goes to the tail of its cell's chain; buckets come with the next substep
    BBX_CHECK(launch_fill_int(s, s->counters, 0, CNT_N));
    BBX_CHECK(distribute(s, false, false, false, n0));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    return BBX_OK;
}

int bbx_solver_save(bbx_solver *s, const char *path, int flags) {
    BBX_CHECK(check_solver(s));
    const size_t n = (size_t)s->n;
    std::vector<double> pos(3 * n), vel(3 * n), dens(n);
    if (flags & BBX_SER_POSITION) BBX_CHECK(bbx_solver_download(s, BBX_FIELD_POSITION, pos.data()));
    if (flags & BBX_SER_VELOCITY) BBX_CHECK(bbx_solver_download(s, BBX_FIELD_VELOCITY, vel.data()));
    if (flags & BBX_SER_DENSITY) BBX_CHECK(bbx_solver_download(s, BBX_FIELD_DENSITY, dens.data()));
    return bbx_serializer_write(path, flags, s->prm.target_spacing, s->n, pos.data(), vel.data(),
                                dens.data(), s->mass, nullptr);
}

int bbx_solver_upload(bbx_solver *s, int field, const double *in) {
    BBX_CHECK(check_solver(s));
    if (!in) return fail(BBX_E_INVALID, "null input");
    if (s->n == 0) return BBX_OK;
    std::vector<int> ids;
    const size_t n = (size_t)s->n;
    if (field == BBX_FIELD_POSITION || field == BBX_FIELD_VELOCITY) {
        float4 *dst = field == BBX_FIELD_POSITION ? s->posA : s->velA;
        std::vector<float4> h(n);
        BBX_CHECK(slot_ids(s, ids));
        BBX_HIP_CHECK(hipMemcpy(h.data(), dst, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            const double *v = &in[3 * (size_t)ids[i]];
            h[i].x = (float)v[0]; h[i].y = (float)v[1]; h[i].z = (float)v[2];
        }
        BBX_HIP_CHECK(hipMemcpy(dst, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        return BBX_OK;
    }
    if (field == BBX_FIELD_PRESSURE) {
        BBX_CHECK(slot_ids(s, ids));
        std::vector<float4> h(n);
        std::vector<float> p(n);
        BBX_HIP_CHECK(hipMemcpy(h.data(), s->velA, sizeof(float4) * n, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) {
            p[i] = (float)in[ids[i]];
            h[i].w = p[i];
        }
        BBX_HIP_CHECK(hipMemcpy(s->velA, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        BBX_HIP_CHECK(hipMemcpy(s->pres, p.data(), sizeof(float) * n, hipMemcpyHostToDevice));
        return BBX_OK;
    }
    return fail(BBX_E_INVALID, "field cannot be uploaded");
}

int bbx_solver_get_flag(bbx_solver *s, int *flag) {
    BBX_CHECK(check_solver(s));
    if (!flag) return fail(BBX_E_INVALID, "null flag");
    BBX_HIP_CHECK(hipMemcpyAsync(s->h_flag, s->flag, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    *flag = *s->h_flag;
    return BBX_OK;
}

int bbx_solver_set_flag(bbx_solver *s, int flag) {
    BBX_CHECK(check_solver(s));
    BBX_CHECK(launch_fill_int(s, s->flag, flag ? 1 : 0, 1));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    return BBX_OK;
}

// One readback per call: counters, max |f| and (PCISPH) the loop state.
// iterations = index of the first iteration whose done flag is set, + 1;
// the reported error is that iteration's FP64 max error / rho0.
// Queue the copy of the step's counters, max-force slots and loop record
// into pinned host memory (one copy: they are one block).
static int queue_stats(bbx_solver *s) {
    BBX_HIP_CHECK(hipMemcpyAsync(s->h_step, s->step, kStepReadback, hipMemcpyDeviceToHost, s->stream));
    return BBX_OK;
}

// Statistics from the copies queue_stats made (they must have landed).
static void parse_stats(bbx_solver *s, bbx_stats *st, int loop_len) {
    st->full_redistribution = s->h_counters[CNT_FULL];
    st->bucket_overflows = s->h_counters[CNT_OVERFLOW];
    st->lost_particles = s->h_counters[CNT_LOST];
    st->iterations = 0;
    st->max_density_error = 0;
    if (loop_len > 0) {
        int it = loop_len;
        for (int k = 0; k < loop_len; k++)
            if (s->h_loop->done[k]) {
                it = k + 1;
                break;
            }
        double e;
        std::memcpy(&e, &s->h_loop->err[it - 1], sizeof(double));
        st->iterations = it;
        st->max_density_error = e / s->prm.target_density;
    }
}

static int read_stats(bbx_solver *s, bbx_stats *st, int loop_len = 0) {
    BBX_CHECK(queue_stats(s));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    if (st) parse_stats(s, st, loop_len);
    return BBX_OK;
}

int bbx_grid_distribute(bbx_solver *s, bbx_stats *stats) {
    BBX_CHECK(check_setup(s));
    BBX_CHECK(launch_fill_int(s, s->counters, 0, CNT_N));
    BBX_CHECK(distribute(s, false, true, s->kind == BBX_SPH));
    if (stats) std::memset(stats, 0, sizeof(*stats));
    BBX_CHECK(read_stats(s, stats));
    if (stats) stats->max_density_error = 0;
    if (s->prof.enabled) s->prof.resolve();
    return BBX_OK;
}

int bbx_grid_get_chains(bbx_solver *s, int *cell_start, int *members) {
    BBX_CHECK(check_solver(s));
    if (!cell_start || !members) return fail(BBX_E_INVALID, "null output");
    const size_t cells = (size_t)s->grid.total + 1;
    std::vector<int> start(cells);
    BBX_HIP_CHECK(hipMemcpy(start.data(), s->cell_start, sizeof(int) * cells, hipMemcpyDeviceToHost));
    std::vector<int> ids;
    if (s->n > 0) BBX_CHECK(slot_ids(s, ids));
    std::memcpy(cell_start, start.data(), sizeof(int) * cells);
    const int in_chains = start[cells - 1];
    for (int i = 0; i < in_chains; i++) members[i] = ids[i];
    return BBX_OK;
}

int bbx_grid_set_chains(bbx_solver *s, const int *cell_start, const int *members) {
    BBX_CHECK(check_solver(s));
    if (!cell_start || (!members && s->n > 0)) return fail(BBX_E_INVALID, "null input");
    const unsigned total = s->grid.total;
    const int in_chains = cell_start[total];
    if (in_chains < 0 || in_chains > s->n) return fail(BBX_E_INVALID, "bad chain CSR");
    const size_t n = (size_t)s->n;
    std::vector<int> ids;
    std::vector<float> pos4;
    if (n) BBX_CHECK(slot_ids(s, ids, &pos4));
    std::vector<float4> vel(n);
    if (n) BBX_HIP_CHECK(hipMemcpy(vel.data(), s->velA, sizeof(float4) * n, hipMemcpyDeviceToHost));
    std::vector<int> slot_of(n, -1);
    for (size_t i = 0; i < n; i++) slot_of[ids[i]] = (int)i;
    std::vector<float4> np(n), nv(n);
    std::vector<int> ncell(n, (int)total);
    std::vector<char> used(n, 0);
    size_t at = 0;
    for (unsigned c = 0; c < total; c++) {
        if (cell_start[c + 1] < cell_start[c]) return fail(BBX_E_INVALID, "bad chain CSR");
        for (int k = cell_start[c]; k < cell_start[c + 1]; k++) {
            int id = members[k];
            if (id < 0 || id >= s->n || used[id]) return fail(BBX_E_INVALID, "bad chain member");
            used[id] = 1;
            int sl = slot_of[id];
            std::memcpy(&np[at], &pos4[4 * (size_t)sl], sizeof(float4));
            nv[at] = vel[sl];
            ncell[at] = (int)c;
            at++;
        }
    }
    for (int id = 0; id < s->n; id++) {
        if (used[id]) continue;
        int sl = slot_of[id];
        std::memcpy(&np[at], &pos4[4 * (size_t)sl], sizeof(float4));
        nv[at] = vel[sl];
        at++;
    }
    if (n) {
        BBX_HIP_CHECK(hipMemcpy(s->posA, np.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        BBX_HIP_CHECK(hipMemcpy(s->velA, nv.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        BBX_HIP_CHECK(hipMemcpy(s->cellA, ncell.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    }
    // cell table for the new order (buckets are rebuilt by the next distribute)
    std::vector<int> cnt(total + 1, 0), st(total + 1, 0);
    for (unsigned c = 0; c < total; c++) cnt[c] = cell_start[c + 1] - cell_start[c];
    cnt[total] = s->n - in_chains;
    for (unsigned c = 0; c <= total; c++) st[c] = c < total ? cell_start[c] : in_chains;
    BBX_HIP_CHECK(hipMemcpy(s->cell_count, cnt.data(), sizeof(int) * (total + 1), hipMemcpyHostToDevice));
    BBX_HIP_CHECK(hipMemcpy(s->cell_start, st.data(), sizeof(int) * (total + 1), hipMemcpyHostToDevice));
    s->setup_done = true;
    return BBX_OK;
}

int bbx_grid_hash(bbx_solver *s, const double *points, int n, int *cell_ids) {
    BBX_CHECK(check_solver(s));
    if (n < 0 || (n > 0 && (!points || !cell_ids))) return fail(BBX_E_INVALID, "bad arrays");
    for (int i = 0; i < n; i++) cell_ids[i] = (int)host_hash(s->grid, &points[3 * (size_t)i]);
    return BBX_OK;
}

int bbx_grid_cell_neighbors(bbx_solver *s, int cell, int *out27, int *count) {
    BBX_CHECK(check_solver(s));
    if (!out27 || !count) return fail(BBX_E_INVALID, "null output");
    if (cell < 0 || (unsigned)cell >= s->grid.total) return fail(BBX_E_INVALID, "cell out of range");
    const int nx = s->grid.n[0], ny = s->grid.n[1], nz = s->grid.n[2];
    int u2 = cell / (nx * ny), rem = cell - u2 * nx * ny, u1 = rem / nx, u0 = rem - u1 * nx;
    int c = 0;
    for (int j = u1 - 1; j <= u1 + 1; j++) {
        if (j < 0 || j >= ny) continue;
        for (int i = u0 - 1; i <= u0 + 1; i++) {
            if (i < 0 || i >= nx) continue;
            for (int k = u2 - 1; k <= u2 + 1; k++) {
                if (k < 0 || k >= nz) continue;
                out27[c++] = i + j * nx + k * nx * ny;
            }
        }
    }
    *count = c;
    return BBX_OK;
}

static void desc_from_geom(const GridGeom &g, bbx_grid_desc *out) {
    for (int k = 0; k < 3; k++) {
        out->usizes[k] = (unsigned)g.n[k];
        out->cells_len[k] = g.len[k];
        out->min_point[k] = g.minp[k];
        out->bounds_min[k] = g.bmin[k];
        out->bounds_max[k] = g.bmax[k];
    }
    out->total = g.total;
}

int bbx_grid_describe(const bbx_params *params, bbx_grid_desc *out) {
    if (!params || !out) return fail(BBX_E_INVALID, "null argument");
    GridGeom g{};
    BBX_CHECK(build_grid_geom(*params, &g));
    desc_from_geom(g, out);
    return BBX_OK;
}

int bbx_grid_hash_desc(const bbx_grid_desc *grid, const double *points, int n, int *cell_ids, int *u) {
    if (!grid || n < 0 || (n > 0 && (!points || !cell_ids))) return fail(BBX_E_INVALID, "bad arguments");
    GridGeom g{};
    for (int k = 0; k < 3; k++) {
        g.n[k] = (int)grid->usizes[k];
        g.len[k] = grid->cells_len[k];
        g.minp[k] = grid->min_point[k];
        g.bmin[k] = grid->bounds_min[k];
        g.bmax[k] = grid->bounds_max[k];
    }
    g.total = grid->total;
    for (int i = 0; i < n; i++) {
        const double *p = &points[3 * (size_t)i];
        cell_ids[i] = (int)host_hash(g, p);
        if (u) {
            long long id = cell_ids[i];
            if (id < 0) {
                u[3 * i] = u[3 * i + 1] = u[3 * i + 2] = -1;
            } else {
                int plane = g.n[0] * g.n[1];
                u[3 * i + 2] = (int)(id / plane);
                int rem = (int)(id - (long long)u[3 * i + 2] * plane);
                u[3 * i + 1] = rem / g.n[0];
                u[3 * i] = rem - u[3 * i + 1] * g.n[0];
            }
        }
    }
    return BBX_OK;
}

int bbx_neighbors_get(bbx_solver *s, int *counts, int *out_ids) {
    BBX_CHECK(check_solver(s));
    if (!counts) return fail(BBX_E_INVALID, "null output");
    const size_t n = (size_t)s->n;
    if (!n) return BBX_OK;
    std::vector<int> ids;
    BBX_CHECK(slot_ids(s, ids));
    std::vector<int> cnt(n);
    BBX_HIP_CHECK(hipMemcpy(cnt.data(), s->nbr_cnt, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++) counts[ids[i]] = cnt[i];
    if (out_ids) {
        // per-tile tables translate LDS-local entries back to storage slots
        const int ntiles = (int)((n + kBlock - 1) / kBlock);
        std::vector<int> tab((size_t)ntiles * kTabIntsHost);
        BBX_HIP_CHECK(hipMemcpy(tab.data(), s->tile_tab, sizeof(int) * tab.size(), hipMemcpyDeviceToHost));
        // one quad row (entries 4q..4q+3 of every particle) at a time
        std::vector<int> quad32(4 * n);
        std::vector<unsigned short> quad16(4 * n);
        for (int k = 0; k < kBucket; k++) {
            if ((k & 3) == 0) {
                const size_t base = nbr_index(k, 0, s->stride);
                BBX_HIP_CHECK(hipMemcpy(quad32.data(), s->nbr32 + base, sizeof(int) * 4 * n,
                                        hipMemcpyDeviceToHost));
                BBX_HIP_CHECK(hipMemcpy(quad16.data(), s->nbr16 + base,
                                        sizeof(unsigned short) * 4 * n, hipMemcpyDeviceToHost));
            }
            for (size_t i = 0; i < n; i++) {
                int v = -1;
                if (k < cnt[i]) {
                    const int *t = &tab[(i / kBlock) * kTabIntsHost];
                    int j = -1;
                    if (t[1]) {  // staged tile: local index -> interval -> slot
                        const int jl = quad16[4 * i + (k & 3)] >> 4;  // LDS byte offset -> slot
                        for (int q = 0; q < t[0]; q++) {
                            const int st = t[3 + q], len = t[3 + 9 + q], off = t[3 + 18 + q];
                            if (jl >= off && jl < off + len) j = st + (jl - off);
                        }
                    } else {
                        j = quad32[4 * i + (k & 3)];
                    }
                    v = (j >= 0 && (size_t)j < n) ? ids[j] : -2;
                }
                out_ids[(size_t)ids[i] * kBucket + k] = v;
            }
        }
    }
    return BBX_OK;
}

int bbx_stage_density(bbx_solver *s) {
    BBX_CHECK(check_setup(s));
    BBX_CHECK(run_density_only(s));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    return BBX_OK;
}

int bbx_stage_nonpressure(bbx_solver *s) {
    BBX_CHECK(check_setup(s));
    StepConst st{};
    BBX_CHECK(run_nonpressure(s, st, 0));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    return BBX_OK;
}

int bbx_pcisph_delta(bbx_solver *s, double dt, double *delta) {
    BBX_CHECK(check_setup(s));
    if (!delta) return fail(BBX_E_INVALID, "null output");
    *delta = compute_delta(s, dt);
    return BBX_OK;
}

static int max_force_host(bbx_solver *s, double *mf) {
    // a substep's statistics readback already brought the slots (Advance
    // asks right after one); otherwise read them now
    if (!s->maxf_cached) {
        BBX_HIP_CHECK(hipMemcpyAsync(s->h_max_force, s->max_force, sizeof(unsigned) * kSlots,
                                     hipMemcpyDeviceToHost, s->stream));
        BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    }
    unsigned m = 0;  // non-negative floats order like their bit patterns
    for (int k = 0; k < kSlots; k++) m = s->h_max_force[k] > m ? s->h_max_force[k] : m;
    float f;
    std::memcpy(&f, &m, sizeof(float));
    *mf = (double)f;
    return BBX_OK;
}

// SphParticleSet3::ComputeNumberOfTimeSteps (src/core/particle.h:584-608)
static int time_steps(bbx_solver *s, double interval, double scale, unsigned *out) {
    double maxf = 0;
    BBX_CHECK(max_force_host(s, &maxf));
    double ref = s->kr;
    double limit = 0.40 * ref / s->prm.sound_speed;
    if (!(absd(maxf) < 1e-8)) {
        double by_force = 0.25 * std::sqrt(ref * s->mass / maxf);
        limit = mind(by_force, limit);
    }
    double target = scale * limit;
    *out = (unsigned)std::ceil(interval / target);
    return BBX_OK;
}

int bbx_number_of_time_steps(bbx_solver *s, double interval, double scale, unsigned *n) {
    BBX_CHECK(check_setup(s));
    if (!n) return fail(BBX_E_INVALID, "null output");
    return time_steps(s, interval, scale, n);
}

static int check_substep(bbx_solver *s, double dt, int max_it, int warm) {
    BBX_CHECK(check_setup(s));
    if (s->kind != BBX_PCISPH) return fail(BBX_E_INVALID, "not a PCISPH solver");
    if (max_it < 1 || max_it > kMaxIt)
        return fail(BBX_E_INVALID, "max_it must be in [1, " + std::to_string(kMaxIt) + "]");
    if (!(dt > 0)) return fail(BBX_E_INVALID, "dt must be > 0");
    if (warm != BBX_WARM_NONE && warm != BBX_WARM_MLP && warm != BBX_WARM_AUTO)
        return fail(BBX_E_INVALID, "unknown warm-start mode");
    if (warm != BBX_WARM_NONE && !s->mlp) return fail(BBX_E_INVALID, "no MLP attached");
    return BBX_OK;
}

// A PCISPH substep is queued in two parts.  The front does not depend on the
// previous substep's outcome: counters cleared and the particles
// redistributed over the grid (UpdateGridDistribution up to the bucket
// build).  The back needs that outcome (the warm-start decision and the
// loop's launch count): bucket build + density, non-pressure forces, the
// predict-correct loop, XSPH, and the statistics copied to pinned memory.
static int substep_front(bbx_solver *s, hipEvent_t start) {
    s->maxf_cached = false;
    if (start) BBX_HIP_CHECK(hipEventRecord(start, s->stream));
    // the step block is cleared by the launch that clears the cell counts
    return distribute_grid(s, false, -1, (int *)s->step, sizeof(StepBlock) / sizeof(int));
}

struct SubstepPlan {
    int warm = BBX_WARM_NONE;  // the start actually used (AUTO resolved)
    int launches = 0;          // loop launches queued
    bool by_copy = true;       // statistics through the copy to pinned memory
    int slot = 0;              // the copy's event
};

// Wait for the pressure-force kernel of iteration k (queued, of the current
// substep) to decide; *done = whether it integrates.  The host spins on the
// pinned word the kernel stores its decision in.  Should the stream drain
// without it (a decision the host cannot see), the done flag is read back
// through a copy instead.
static int wait_loop_decision(bbx_solver *s, int k, int *done) {
    const unsigned want = s->loop_seq & 0x7fffffu;
    auto decided = [&](unsigned w) { return (w >> 9) == want && ((w & 1u) || (int)((w >> 1) & 0xffu) >= k); };
    for (unsigned spin = 1; s->d_signal; spin++) {
        unsigned w = __atomic_load_n(&s->h_signal->word, __ATOMIC_ACQUIRE);
        if (decided(w)) {
            *done = (int)(w & 1u) && (int)((w >> 1) & 0xffu) <= k;
            return BBX_OK;
        }
        if ((spin & 4095u) == 0u) {
            const hipError_t q = hipStreamQuery(s->stream);
            if (q == hipErrorNotReady) continue;
            BBX_HIP_CHECK(q);
            w = __atomic_load_n(&s->h_signal->word, __ATOMIC_ACQUIRE);
            if (decided(w)) {
                *done = (int)(w & 1u) && (int)((w >> 1) & 0xffu) <= k;
                return BBX_OK;
            }
            break;
        }
    }
    BBX_HIP_CHECK(hipMemcpyAsync(s->h_flag, &s->loop->done[k], sizeof(int), hipMemcpyDeviceToHost, s->stream));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    *done = *s->h_flag;
    return BBX_OK;
}

// end: recorded behind the substep's work (nullptr: none); slot picks the
// statistics copy's event
static int substep_back(bbx_solver *s, int slot, hipEvent_t end, double dt, int max_it, double tol, int mode,
                        int warm, bool copy_stats, SubstepPlan *plan) {
    s->loop_seq++;  // numbers this substep's loop decisions
    if (warm == BBX_WARM_AUTO) {
        // The warm start costs about two thirds of a correction, so it only
        // pays where the plain loop needs several.  While the flow is calm
        // it is switched on after a plain substep that needed
        // BBX_WARM_AUTO_MIN_ITERATIONS or more corrections and off again
        // after BBX_WARM_AUTO_PROBE warm substeps in a row converged in one.
        // Once any substep needed BBX_WARM_AUTO_ACTIVE or more the flow is
        // active and it stays on: a warm-started loop that follows a plain
        // substep starts from the plain loop's lower converged pressures and
        // needs more corrections than one that follows a warm substep.
        bool on;
        if (s->warm_active)
            on = true;
        else
            on = s->last_warm ? s->warm_ones < BBX_WARM_AUTO_PROBE
                              : s->last_iterations >= BBX_WARM_AUTO_MIN_ITERATIONS;
        warm = on ? BBX_WARM_MLP : BBX_WARM_NONE;
    }
    const double delta = compute_delta(s, dt);
    s->delta64 = delta;
    StepConst st{};
    st.dt = (float)dt;
    st.dt_over_m = (float)(dt / s->mass);
    st.delta = (float)delta;
    st.warm_scale = (float)(delta * s->prm.target_density);
    if (warm == BBX_WARM_MLP) {
        // interaction + density, with the features and the network's
        // pressures p0 in the bucket kernel; then the viscosity gather also
        // sums f_p(p0, rho* = rho) and predicts k = 0
        BBX_CHECK(build_buckets(s, false, 2, &st));
        BBX_CHECK(run_nonpressure(s, st, 2));
    } else {
        // ResetHigherLevel + interaction + density
        BBX_CHECK(build_buckets(s, false, 0, &st));
        BBX_CHECK(run_nonpressure(s, st, 1));
    }
    // The predict-correct loop: the pressure-force kernel of iteration k
    // reads that iteration's FP64 error on the device and either integrates
    // or predicts iteration k + 1, and publishes that decision to the host
    // as soon as it has taken it (the first workgroup, at kernel start).
    // The host queues one iteration fewer than recent substeps needed, then
    // one at a time: it waits for the decision of the last queued one and,
    // if the loop goes on, queues the next while that pressure-force kernel
    // is still running, so the device does not wait for the host and a
    // launch that exits at entry is queued only when a substep needs two or
    // more iterations fewer than recent ones.  Once the loop has integrated,
    // the rest of the substep and the next substep's grid work are queued
    // behind the integrating kernel in the same way.  REFERENCE_EXACT stops after the first iteration (its error
    // is never recorded, so it never converges earlier and never runs
    // longer).
    const int exact = mode == BBX_MODE_REFERENCE_EXACT ? 1 : 0;
    const int cap = exact ? 1 : max_it;
    // An empty set converges in its first iteration with zero error, as the
    // reference's loop does over no particles; no kernel runs, so none would
    // publish a decision to wait for.
    int launches = s->n > 0 ? 0 : 1;
    // queued at once: one fewer than the smaller of the last two substeps'
    // counts (a stalled substep at the cap is not followed by a cap's worth
    // of launches that exit at entry)
    const int recent = s->prev_iterations > 0 && s->prev_iterations < s->last_iterations ? s->prev_iterations
                                                                                          : s->last_iterations;
    int goal = recent > 1 ? recent - 1 : 1;
    for (; s->n > 0;) {
        goal = goal < cap ? goal : cap;
        for (; launches < goal; launches++) {
            st.first = (launches == 0 && warm != BBX_WARM_MLP) ? 1 : 0;
            BBX_CHECK(run_predict_pressure(s, st, launches));
            BBX_CHECK(run_pressure_force(s, st, launches, max_it, exact, tol));
        }
        if (launches >= cap) break;
        int done = 0;
        BBX_CHECK(wait_loop_decision(s, launches - 1, &done));
        if (done) break;
        goal = launches + 1;
    }
    BBX_CHECK(swap_after_integrate(s));
    BBX_CHECK(run_xsph(s, st));
    if (end) BBX_HIP_CHECK(hipEventRecord(end, s->stream));
    // The statistics come with the integrating decision (LoopSignal), so no
    // copy is queued: a device-to-host copy and the system-scope fence of the
    // event behind it leave the device idle for about 20 us per substep.  The
    // copy remains where the max |f| slots are wanted too (the time-step rule
    // of Advance), without published decisions, and for an empty set.
    plan->by_copy = copy_stats || !s->d_signal || s->n == 0;
    plan->slot = slot;
    if (plan->by_copy) {
        BBX_CHECK(queue_stats(s));
        BBX_HIP_CHECK(hipEventRecord(s->ev_read[slot], s->stream));
    }
    plan->warm = warm;
    plan->launches = launches;
    return BBX_OK;
}

// The statistics the integrating decision of the current substep published
// (LoopSignal).  Should the stream drain without them all carrying the
// substep's number, that is an error: a later substep's work may already have
// cleared the device-side counters a copy would read.
static int published_stats(bbx_solver *s, bbx_stats *st) {
    const unsigned want = s->loop_seq & 0x7fffffu;
    unsigned long long v[kPubWords];
    unsigned w = 0u;
    auto current = [&]() {
        w = __atomic_load_n(&s->h_signal->word, __ATOMIC_ACQUIRE);
        if ((w >> 9) != want || !(w & 1u)) return false;
        for (int q = 0; q < kPubWords; q++) {
            v[q] = __atomic_load_n(&s->h_signal->stats[q], __ATOMIC_ACQUIRE);
            if ((unsigned)(v[q] >> 32) != want) return false;
        }
        return true;
    };
    for (unsigned spin = 1;; spin++) {
        if (current()) break;
        if ((spin & 4095u) == 0u) {
            const hipError_t q = hipStreamQuery(s->stream);
            if (q == hipErrorNotReady) continue;
            BBX_HIP_CHECK(q);
            if (current()) break;
            return fail(BBX_E_HIP, "the loop's statistics were not published");
        }
    }
    const unsigned long long bits = (v[PUB_ERR_LO] & 0xffffffffull) | ((v[PUB_ERR_HI] & 0xffffffffull) << 32);
    double e;
    std::memcpy(&e, &bits, sizeof(double));
    st->iterations = (int)((w >> 1) & 0xffu) + 1;
    st->max_density_error = e / s->prm.target_density;
    st->bucket_overflows = (int)(unsigned)v[PUB_OVERFLOW];
    st->lost_particles = (int)(unsigned)v[PUB_LOST];
    st->full_redistribution = (int)(unsigned)v[PUB_FULL];
    return BBX_OK;
}

// The back's statistics (everything but gpu_ms), and what the next
// substep's decisions read updated from them.  Published statistics need no
// wait for the substep's end; copied ones wait for the copy.
static int substep_stats(bbx_solver *s, const SubstepPlan &plan, bbx_stats *out) {
    bbx_stats local{};
    if (!plan.by_copy) {
        BBX_CHECK(published_stats(s, &local));
        s->maxf_cached = false;
    } else {
        BBX_HIP_CHECK(hipEventSynchronize(s->ev_read[plan.slot]));
        parse_stats(s, &local, plan.launches);
        s->maxf_cached = true;
    }
    s->prev_iterations = s->last_iterations;
    s->last_iterations = local.iterations;
    s->last_warm = plan.warm == BBX_WARM_MLP ? 1 : 0;
    s->warm_ones = s->last_warm && local.iterations == 1 ? s->warm_ones + 1 : 0;
    if (local.iterations >= BBX_WARM_AUTO_ACTIVE) s->warm_active = 1;
    local.warm_started = s->last_warm;
    *out = local;
    return BBX_OK;
}

// device time between two recorded events, once the later one has completed
static int elapsed_ms(hipEvent_t a, hipEvent_t b, double *ms) {
    BBX_HIP_CHECK(hipEventSynchronize(b));
    float t = 0.f;
    BBX_HIP_CHECK(hipEventElapsedTime(&t, a, b));
    *ms = t;
    return BBX_OK;
}

// copy_stats: also bring the max |f| slots to the host (for the time-step
// rule of the next Advance substep)
static int pcisph_substep(bbx_solver *s, double dt, int max_it, double tol, int mode, int warm,
                          bbx_stats *stats, bool copy_stats) {
    BBX_CHECK(check_substep(s, dt, max_it, warm));
    SubstepPlan plan;
    BBX_CHECK(substep_front(s, s->ev0[0]));
    BBX_CHECK(substep_back(s, 0, s->ev1[0], dt, max_it, tol, mode, warm, copy_stats, &plan));
    bbx_stats local{};
    BBX_CHECK(substep_stats(s, plan, &local));
    BBX_CHECK(elapsed_ms(s->ev0[0], s->ev1[0], &local.gpu_ms));
    if (stats) *stats = local;
    if (s->prof.enabled) s->prof.resolve(local.iterations, s->prof.pending.size());
    return BBX_OK;
}

int bbx_pcisph_substep(bbx_solver *s, double dt, int max_it, double tol, int mode, int warm,
                       bbx_stats *stats) {
    return pcisph_substep(s, dt, max_it, tol, mode, warm, stats, false);
}

int bbx_pcisph_run(bbx_solver *s, double dt, int count, int max_it, double tol, int mode, int warm,
                   bbx_stats *stats) {
    BBX_CHECK(check_substep(s, dt, max_it, warm));
    if (count < 0) return fail(BBX_E_INVALID, "count must be >= 0");
    // Substep k + 1's front is queued as soon as substep k's loop has
    // integrated, and its back as soon as the host has substep k's
    // statistics (published with that decision), so the device goes from one
    // substep to the next without waiting for the host.  The boundaries
    // between substeps are marked once each in a ring of events (the end of
    // substep k is the start of k + 1); substep j's device time is read when
    // its ring slot comes round again, or at the end.  The results are those
    // of `count` bbx_pcisph_substep calls.
    constexpr int M = bbx_solver::kMarks;
    int timed = 0;  // substeps before this one have their gpu_ms
    auto time_upto = [&](int limit) {  // gpu_ms of substeps [timed, limit)
        for (; timed < limit; timed++)
            BBX_CHECK(elapsed_ms(s->mark[timed % M], s->mark[(timed + 1) % M], &stats[timed].gpu_ms));
        return BBX_OK;
    };
    auto mark = [&](int m) {  // the boundary before substep m (m = count: the end)
        if (!stats) return BBX_OK;
        if (m >= M) BBX_CHECK(time_upto(m - M + 1));  // the slot's last user has ended
        BBX_HIP_CHECK(hipEventRecord(s->mark[m % M], s->stream));
        return BBX_OK;
    };
    if (count > 0) {
        BBX_CHECK(mark(0));
        BBX_CHECK(substep_front(s, nullptr));
    }
    for (int k = 0; k < count; k++) {
        SubstepPlan plan;
        BBX_CHECK(substep_back(s, k & 1, nullptr, dt, max_it, tol, mode, warm, false, &plan));
        BBX_CHECK(mark(k + 1));
        const size_t upto = s->prof.pending.size();
        if (k + 1 < count) BBX_CHECK(substep_front(s, nullptr));
        bbx_stats local{};
        BBX_CHECK(substep_stats(s, plan, &local));
        if (stats) stats[k] = local;
        if (s->prof.enabled) s->prof.resolve(local.iterations, upto);
    }
    if (stats && count > 0) BBX_CHECK(time_upto(count));
    return BBX_OK;
}

int bbx_pcisph_advance(bbx_solver *s, double interval, int max_it, double tol, int mode, int warm,
                       int *substeps) {
    BBX_CHECK(check_setup(s));
    double remaining = interval;
    int steps = 0;
    while (remaining > (double)0.0001f) {
        unsigned n = 0;
        BBX_CHECK(time_steps(s, remaining, 5.0, &n));
        if (n == 0) return fail(BBX_E_INVALID, "time-step rule returned 0");
        double dt = remaining / (double)n;
        BBX_CHECK(pcisph_substep(s, dt, max_it, tol, mode, warm, nullptr, true));
        remaining -= dt;
        steps++;
    }
    if (substeps) *substeps = steps;
    return BBX_OK;
}

int bbx_sph_substep(bbx_solver *s, double dt, bbx_stats *stats) {
    BBX_CHECK(check_setup(s));
    if (s->kind != BBX_SPH) return fail(BBX_E_INVALID, "not an SPH solver");
    if (!(dt > 0)) return fail(BBX_E_INVALID, "dt must be > 0");
    s->maxf_cached = false;
    BBX_HIP_CHECK(hipEventRecord(s->ev0[0], s->stream));
    BBX_CHECK(launch_fill_int(s, (int *)s->step, 0, sizeof(StepBlock) / sizeof(int)));
    BBX_CHECK(distribute(s, false, true, true));
    StepConst st{};
    st.dt = (float)dt;
    st.dt_over_m = (float)(dt / s->mass);
    BBX_CHECK(run_sph_force(s));
    BBX_CHECK(run_integrate(s, st));
    BBX_CHECK(swap_after_integrate(s));
    BBX_CHECK(run_xsph(s, st));
    BBX_HIP_CHECK(hipEventRecord(s->ev1[0], s->stream));
    bbx_stats local{};
    BBX_CHECK(read_stats(s, &local));
    s->maxf_cached = true;
    local.max_density_error = 0;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, s->ev0[0], s->ev1[0]);
    local.gpu_ms = ms;
    if (stats) *stats = local;
    if (s->prof.enabled) s->prof.resolve();
    return BBX_OK;
}

int bbx_sph_advance(bbx_solver *s, double interval, int *substeps) {
    BBX_CHECK(check_setup(s));
    double remaining = interval;
    int steps = 0;
    while (remaining > (double)0.0001f) {
        unsigned n = 0;
        BBX_CHECK(time_steps(s, remaining, 1.0, &n));
        if (n == 0) return fail(BBX_E_INVALID, "time-step rule returned 0");
        double dt = remaining / (double)n;
        BBX_CHECK(bbx_sph_substep(s, dt, nullptr));
        remaining -= dt;
        steps++;
    }
    if (substeps) *substeps = steps;
    return BBX_OK;
}

int bbx_solver_attach_mlp(bbx_solver *s, bbx_mlp *m) {
    BBX_CHECK(check_solver(s));
    if (s->mlp == m) return BBX_OK;
    // the solver's queued work may still read the old network
    if (s->mlp) BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    mlp_release(s->mlp);
    s->mlp = m;
    mlp_retain(m);
    return BBX_OK;
}

int bbx_solver_keep_features(bbx_solver *s, int keep) {
    BBX_CHECK(check_solver(s));
    s->keep_features = keep ? 1 : 0;
    return BBX_OK;
}

int bbx_pcisph_harvest(bbx_solver *s, double dt, int max_it, double tol, float *features,
                       float *targets, bbx_stats *stats) {
    BBX_CHECK(check_setup(s));
    if (!features || !targets) return fail(BBX_E_INVALID, "null output");
    const size_t n = (size_t)s->n;
    // features at the start of the substep: distribute + density + non-pressure
    // with feature output, then the plain substep from the same state.
    // The features are computed on a throwaway pass that leaves the state
    // untouched (distribute is idempotent for unchanged positions).
    BBX_CHECK(launch_fill_int(s, s->counters, 0, CNT_N));
    const double delta = compute_delta(s, dt);
    StepConst st{};
    st.dt = (float)dt;
    st.dt_over_m = (float)(dt / s->mass);
    st.delta = (float)delta;
    BBX_CHECK(distribute(s, false, true, false, -1, 1, &st));
    std::vector<int> ids;
    BBX_CHECK(slot_ids(s, ids));
    std::vector<float> h(n * BBX_NFEAT);
    for (int f = 0; f < BBX_NFEAT; f++)
        BBX_HIP_CHECK(hipMemcpy(&h[f * n], s->feat + (size_t)f * s->stride, sizeof(float) * n,
                                hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; i++)
        for (int f = 0; f < BBX_NFEAT; f++) features[(size_t)ids[i] * BBX_NFEAT + f] = h[f * n + i];
    bbx_stats local{};
    BBX_CHECK(bbx_pcisph_substep(s, dt, max_it, tol, BBX_MODE_TEXTBOOK, BBX_WARM_NONE, &local));
    // converged pressure / (delta rho0), by id (slots moved only inside the substep's distribute)
    BBX_CHECK(slot_ids(s, ids));
    std::vector<float> p(n);
    BBX_HIP_CHECK(hipMemcpy(p.data(), s->pres, sizeof(float) * n, hipMemcpyDeviceToHost));
    const double pref = delta * s->prm.target_density;
    for (size_t i = 0; i < n; i++) targets[ids[i]] = pref != 0 ? (float)(p[i] / pref) : 0.f;
    if (stats) *stats = local;
    return BBX_OK;
}

int bbx_profile_enable(bbx_solver *s, int enable) {
    BBX_CHECK(check_solver(s));
    s->prof.enabled = enable != 0;
    return BBX_OK;
}

int bbx_profile_select(bbx_solver *s, const char *kernel, int iteration) {
    BBX_CHECK(check_solver(s));
    static_assert(BBX_PROFILE_ALL_LAUNCHES == kNoIter, "profile sentinel");
    s->prof.only = kernel ? kernel : "";
    s->prof.only_iter = iteration;
    return BBX_OK;
}

int bbx_profile_reset(bbx_solver *s) {
    BBX_CHECK(check_solver(s));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    s->prof.reset();
    return BBX_OK;
}

int bbx_profile_get(bbx_solver *s, char *names, int names_cap, double *total_ms, int *launches,
                    int cap, int *count) {
    BBX_CHECK(check_solver(s));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    s->prof.resolve();
    int k = (int)s->prof.names.size();
    if (count) *count = k;
    std::string joined;
    for (int i = 0; i < k; i++) {
        if (i < cap) {
            if (total_ms) total_ms[i] = s->prof.total_ms[i];
            if (launches) launches[i] = s->prof.launches[i];
        }
        joined += s->prof.names[i];
        joined += '\n';
    }
    if (names && names_cap > 0) {
        std::strncpy(names, joined.c_str(), (size_t)names_cap - 1);
        names[names_cap - 1] = '\0';
    }
    return BBX_OK;
}

int bbx_solver_synchronize(bbx_solver *s) {
    BBX_CHECK(check_solver(s));
    BBX_HIP_CHECK(hipStreamSynchronize(s->stream));
    return BBX_OK;
}

}  // extern "C"
