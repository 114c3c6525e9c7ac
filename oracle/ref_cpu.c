/*
 * ref_cpu.c -- FP64 CPU restatement of the Bubbles SPH / PCISPH hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the oracle (the checker), never the
 * product.  Only tests/, the smoke check and the benchmark's CPU-baseline leg
 * load it.  Build with -ffp-contract=off so every expression rounds the way
 * the reference's host build does (no FMA on x86-64 without -mfma).
 *
 * Each function names the reference file:line whose arithmetic it restates;
 * expression order is kept (left-to-right evaluation, vec/f = vec*(1/f),
 * s*vec = vec*s) because the FP64 results are compared bit-for-bit where the
 * algorithm is integer/ordering work (hash, chains, buckets).
 */
#define _GNU_SOURCE
#include "ref_cpu.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* src/core/geometry.h:44-64 */
#define R_EPSILON ((double)0.0001f)
#define R_PI 3.14159265358979323846
#define R_INFINITY ((double)FLT_MAX)
/* src/core/particle.h:8-10 */
#define R_SPEED_FACTOR 0.40
#define R_FORCE_FACTOR 0.25

typedef struct { double x, y, z; } v3;

static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vmul(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
/* src/core/geometry.h:568-575: division multiplies by the reciprocal */
static inline v3 vdiv(v3 a, double f) { double inv = (double)1 / f; return mk(a.x * inv, a.y * inv, a.z * inv); }
static inline double vlen2(v3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline double vlen(v3 a) { return sqrt(vlen2(a)); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double vdist(v3 a, v3 b) { return vlen(vsub(a, b)); }
/* src/core/geometry.h:112-121 */
static inline double absf(double v) { return v > 0 ? v : -v; }
static inline int is_zero(double a) { return absf(a) < 1e-8; }
static inline int is_unsafe_zero(double a) { return absf(a) < 1e-6; }
static inline double fmaxr(double a, double b) { return a < b ? b : a; }
static inline double fminr(double a, double b) { return a < b ? a : b; }
static inline double clampr(double v, double lo, double hi) { if (v < lo) return lo; if (v > hi) return hi; return v; }

/* ------------------------------------------------------------------------ */
/* kernels: src/core/kernel.cpp:112-234                                      */
/* ------------------------------------------------------------------------ */
double bbo_poly6_W(double h, double distance) {
    double h2 = h * h, h3 = h2 * h;
    double d2 = distance * distance;
    double of = d2 - h2;
    if (is_zero(of) || of > 0) return 0;
    double x = 1.0 - d2 / h2;
    return 315.0 / (64.0 * R_PI * h3) * x * x * x;
}
double bbo_poly6_dW(double h, double distance) {
    double h2 = h * h, h3 = h2 * h, h5 = h2 * h3;
    double d2 = distance * distance;
    double of = distance - h;
    if (is_zero(of) || of > 0) return 0;
    double x = 1.0 - d2 / h2;
    return -945.0 / (32.0 * R_PI * h5) * distance * x * x;
}
double bbo_poly6_d2W(double h, double distance) {
    double h2 = h * h, h3 = h2 * h, h5 = h2 * h3;
    double d2 = distance * distance;
    double of = d2 - h2;
    if (is_zero(of) || of > 0) return 0;
    double x = d2 / h2;
    return 945.0 / (32.0 * R_PI * h5) * (1 - x) * (5 * x - 1);
}
double bbo_spiky_W(double h, double distance) {
    double h2 = h * h, h3 = h2 * h;
    double of = distance - h;
    if (is_zero(of) || of > 0) return 0;
    double x = 1.0 - distance / h;
    return 15.0 / (R_PI * h3) * x * x * x;
}
double bbo_spiky_dW(double h, double distance) {
    double h2 = h * h, h4 = h2 * h2;
    double of = distance - h;
    if (is_zero(of) || of > 0) return 0;
    double x = 1.0 - distance / h;
    return -45.0 / (R_PI * h4) * x * x;
}
double bbo_spiky_d2W(double h, double distance) {
    double h2 = h * h, h3 = h2 * h, h5 = h2 * h3;
    double of = distance - h;
    if (is_zero(of) || of > 0) return 0;
    double x = 1.0 - distance / h;
    return 90.0 / (R_PI * h5) * x;
}
int bbo_within_spiky(double distance, double radius) {
    double of = distance - radius;
    return !(is_zero(of) || of > 0);
}
int bbo_within_std(double distance, double radius) {
    double h2 = radius * radius;
    double d2 = distance * distance;
    double of = d2 - h2;
    return !(is_zero(of) || of > 0);
}
/* gradW(distance, dir) = -dW(distance) * dir  (src/core/kernel.cpp:221) */
static inline v3 spiky_grad(double h, double distance, v3 dir) {
    return vmul(dir, -bbo_spiky_dW(h, distance));
}

/* ------------------------------------------------------------------------ */
/* grid: src/core/grid.h:626-666 (Build), :258-321 (hash), :554-624 (lists) */
/* ------------------------------------------------------------------------ */
void bbo_grid_build(const unsigned res[3], const double dp0[3], const double dp1[3],
                    bbo_grid_desc *g) {
    double lower[3], high[3], p0[3], p1[3], maxp[3];
    g->total = 1;
    for (int k = 0; k < 3; k++) {
        lower[k] = R_INFINITY;
        high[k] = -R_INFINITY;
        p0[k] = fmaxr(dp0[k], dp1[k]);
        p1[k] = fminr(dp0[k], dp1[k]);
    }
    for (int k = 0; k < 3; k++) {
        double dl = p0[k], du = p1[k];
        if (dl < lower[k]) lower[k] = dl;
        if (dl > high[k]) high[k] = dl;
        if (du > high[k]) high[k] = du;
        if (du < lower[k]) lower[k] = du;
    }
    for (int k = 0; k < 3; k++) {
        double s = high[k] - lower[k];
        double len = s / (double)res[k];
        g->cells_len[k] = len;
        g->usizes[k] = (unsigned)(int)ceil(s / len);
        maxp[k] = lower[k] + (double)g->usizes[k] * g->cells_len[k];
        g->total *= g->usizes[k];
    }
    for (int k = 0; k < 3; k++) {
        g->min_point[k] = lower[k];
        g->bounds_min[k] = fminr(lower[k], maxp[k]);
        g->bounds_max[k] = fmaxr(lower[k], maxp[k]);
    }
}

/* src/core/util.cpp:269-287 */
void bbo_grid_for_domain(const double dmin[3], const double dmax[3], double spacing,
                         double spacing_scale, bbo_grid_desc *g) {
    double length = spacing * spacing_scale;
    double hlen = 0.5 * length;
    double inv_len = 1.0f / length;
    int m[3];
    double center[3], pmin[3], pmax[3];
    for (int k = 0; k < 3; k++) {
        double ext = absf(dmax[k] - dmin[k]);
        m[k] = (int)ceil(ext * inv_len);
    }
    for (int k = 0; k < 3; k++) m[k] = (int)(m[k] + (m[k] % 2) * length);
    for (int k = 0; k < 3; k++) {
        center[k] = (dmin[k] + dmax[k]) * 0.5;
        double half = m[k] * hlen;
        pmin[k] = center[k] - half;
        pmax[k] = center[k] + half;
    }
    unsigned res[3] = {(unsigned)m[0], (unsigned)m[1], (unsigned)m[2]};
    bbo_grid_build(res, pmin, pmax, g);
}

/* src/core/grid.h:258-298 (ExtremeEpsilon + GetHashedPosition) and :182-193 */
long long bbo_grid_hash(const bbo_grid_desc *g, const double p[3], int u[3]) {
    int uu[3];
    int bad = 0;
    for (int i = 0; i < 3; i++) {
        double pi = p[i];
        double eps = 0;
        if (is_zero(pi - g->bounds_min[i])) eps = R_EPSILON;
        else if (is_zero(pi - g->bounds_max[i])) eps = -R_EPSILON;
        pi += eps;
        double dp = (pi - g->min_point[i]) / g->cells_len[i];
        double f = floor(dp);
        int id;
        if (!(f >= 0) || f >= (double)g->usizes[i]) {
            bad = 1;
            id = (f >= (double)g->usizes[i]) ? (int)g->usizes[i] - 1 : 0;
        } else {
            id = (int)f;
        }
        uu[i] = id;
    }
    if (u) { u[0] = uu[0]; u[1] = uu[1]; u[2] = uu[2]; }
    if (bad) return -1;
    return (long long)((unsigned)uu[0] + (unsigned)uu[1] * g->usizes[0] +
                       (unsigned)uu[2] * g->usizes[0] * g->usizes[1]);
}

/* src/core/grid.h:554-593 (GetNeighborListFor, depth 1): y outer, x, z inner */
int bbo_grid_neighbors(const bbo_grid_desc *g, unsigned id, int *out) {
    unsigned plane = g->usizes[0] * g->usizes[1];
    int u2 = (int)(id / plane);
    unsigned rem = id - (unsigned)u2 * plane;
    int u1 = (int)(rem / g->usizes[0]);
    int u0 = (int)(rem % g->usizes[0]);
    int count = 0;
    for (int j = u1 - 1; j <= u1 + 1; j++) {
        if (j < 0 || j >= (int)g->usizes[1]) continue;
        for (int i = u0 - 1; i <= u0 + 1; i++) {
            if (i < 0 || i >= (int)g->usizes[0]) continue;
            for (int k = u2 - 1; k <= u2 + 1; k++) {
                if (k < 0 || k >= (int)g->usizes[2]) continue;
                out[count++] = (int)((unsigned)i + (unsigned)j * g->usizes[0] +
                                     (unsigned)k * plane);
            }
        }
    }
    return count;
}

/* ------------------------------------------------------------------------ */
/* BCC lattice: src/generator/bcclattice.cpp:5-76                            */
/* ------------------------------------------------------------------------ */
int bbo_bcc_generate(const double pmin[3], const double pmax[3], double spacing,
                     double *points, int maxn) {
    double half = spacing / 2;
    double width = absf(pmax[0] - pmin[0]);
    double height = absf(pmax[1] - pmin[1]);
    double depth = absf(pmax[2] - pmin[2]);
    int has_off = 0, stop = 0, n = 0;
    for (int k = 0; k * half <= depth && !stop; k++) {
        double z = k * half + pmin[2];
        double offset = has_off ? half : 0;
        for (int j = 0; j * spacing + offset <= height && !stop; j++) {
            double y = j * spacing + offset + pmin[1];
            for (int i = 0; i * spacing + offset <= width && !stop; i++) {
                double x = i * spacing + offset + pmin[0];
                if (maxn >= 0 && n >= maxn) { stop = 1; break; }
                if (points) { points[3 * n] = x; points[3 * n + 1] = y; points[3 * n + 2] = z; }
                n++;
            }
        }
        has_off = !has_off;
    }
    return n;
}

/* src/core/particle.h:557-582 (SphParticleSet3::ComputeMass, 1024 points) */
double bbo_compute_mass(double target_density, double target_spacing, double kernel_radius) {
    double pts[3 * 1024];
    double bmin[3], bmax[3];
    for (int k = 0; k < 3; k++) { bmin[k] = -1.5 * kernel_radius; bmax[k] = 1.5 * kernel_radius; }
    int c = bbo_bcc_generate(bmin, bmax, target_spacing, pts, 1024);
    double nden = 0;
    for (int i = 0; i < c; i++) {
        v3 pi = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        double sum = 0;
        for (int j = 0; j < c; j++) {
            v3 pj = mk(pts[3 * j], pts[3 * j + 1], pts[3 * j + 2]);
            sum += bbo_poly6_W(kernel_radius, vlen(vsub(pi, pj)));
        }
        nden = fmaxr(nden, sum);
    }
    return target_density / nden;
}

/* src/solvers/pcisph_solver3.cpp:148-180 (ComputeDeltaDenom) */
double bbo_delta_denom(double spacing, double kernel_radius) {
    double bmin[3], bmax[3];
    for (int k = 0; k < 3; k++) { bmin[k] = -1.5 * kernel_radius; bmax[k] = 1.5 * kernel_radius; }
    int n = bbo_bcc_generate(bmin, bmax, spacing, NULL, -1);
    double *pts = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    bbo_bcc_generate(bmin, bmax, spacing, pts, -1);
    double kr2 = kernel_radius * kernel_radius;
    double denom = 0, denom2 = 0;
    v3 denom1 = mk(0, 0, 0);
    for (int i = 0; i < n; i++) {
        v3 point = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        double d2 = vlen2(point);
        if (d2 < kr2) {
            double distance = sqrt(d2);
            v3 dir = (distance > 0) ? vdiv(point, distance) : mk(0, 0, 0);
            v3 grad = spiky_grad(kernel_radius, distance, dir);
            denom1 = vadd(denom1, grad);
            denom2 += vdot(grad, grad);
        }
    }
    free(pts);
    denom += -vdot(denom1, denom1) - denom2;
    return denom;
}

/* ------------------------------------------------------------------------ */
/* parallel-for with the reference's static slicing (src/cuda/cutil.h:294-338) */
/* ------------------------------------------------------------------------ */
typedef void (*range_fn)(void *ctx, int j0, int j1);
typedef struct { range_fn fn; void *ctx; int j0, j1; } pf_task;
static void *pf_thread(void *a) { pf_task *t = (pf_task *)a; t->fn(t->ctx, t->j0, t->j1); return NULL; }

static int ncores(void) { long n = sysconf(_SC_NPROCESSORS_ONLN); return n > 0 ? (int)n : 1; }

static void parallel_for(int threads, int start, int end, range_fn fn, void *ctx) {
    if (start > end) return;
    if (threads == 1) { fn(ctx, start, end); return; }
    int num = ncores();
    num = threads < 0 ? num : (threads < num ? threads : num);
    if (threads == 0) num = ncores();
    int n = end - start + 1;
    int slice = (int)lround(n / (double)num);
    if (slice < 1) slice = 1;
    pthread_t tid[1024];
    pf_task task[1024];
    int nt = 0;
    int i0 = start, i1 = start + slice < end ? start + slice : end;
    for (int i = 0; i + 1 < num && i0 < end && nt < 1023; i++) {
        task[nt].fn = fn; task[nt].ctx = ctx; task[nt].j0 = i0; task[nt].j1 = i1;
        pthread_create(&tid[nt], NULL, pf_thread, &task[nt]);
        nt++;
        i0 = i1;
        i1 = i1 + slice < end ? i1 + slice : end;
    }
    if (i0 < end) {
        task[nt].fn = fn; task[nt].ctx = ctx; task[nt].j0 = i0; task[nt].j1 = end;
        pthread_create(&tid[nt], NULL, pf_thread, &task[nt]);
        nt++;
    }
    for (int i = 0; i < nt; i++) pthread_join(tid[i], NULL);
}

/* ------------------------------------------------------------------------ */
/* simulation state                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int head[2], tail[2], len[2]; /* [0]=head (main nodes), [1]=headAux (aux nodes) */
    int active;
} cell_t;

struct bbo_sim {
    bbo_params prm;
    int n;
    bbo_grid_desc grid;
    int *nbr;      /* total*27 */
    int *nbr_cnt;  /* total */
    cell_t *cells;
    int *next[2];  /* next node index per particle, main / aux node arrays */
    int *bucket_ids; /* n*100 */
    int *bucket_cnt;
    v3 *pos, *vel, *force, *inter, *tpos, *tvel, *fp, *smooth;
    double *density, *pressure, *derr, *dpred;
    double mass, kernel_radius, radius, mass_over_rho2, delta_denom;
    int frame_index;
    int flag;   /* requiresHigherLevelUpdate */
    int threads;
    /* tank (reversed box) */
    v3 tank_c, tank_h;
    /* collider list replacing the tank (bbo_sim_set_colliders) */
    bbo_collider *coll;
    int ncoll;
    /* per-call parameters passed to range workers */
    double arg_dt, arg_delta;
    int arg_first, arg_record;
    int bucket_overflows;
};

void bbo_default_params(bbo_params *p) {
    memset(p, 0, sizeof(*p));
    /* DefaultSphSolverData3 (src/solvers/sph_solver3.cpp:124-151) */
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
    p->gravity[0] = 0.0;
    p->gravity[1] = (double)(-9.8f); /* vec3f(0.f, -9.8f, 0.f) */
    p->gravity[2] = 0.0;
    p->threads = 1;
}

bbo_sim *bbo_sim_create(const bbo_params *p, int n, const double *pos, const double *vel) {
    bbo_sim *s = (bbo_sim *)calloc(1, sizeof(bbo_sim));
    s->prm = *p;
    s->n = n;
    s->threads = p->threads;
    double dmin[3], dmax[3];
    for (int k = 0; k < 3; k++) {
        double hx = p->tank_size[k] / 2.0;
        /* BoxGetBounds: ObjectToWorld(Bounds3f(-h, h)) with a translation */
        dmin[k] = -hx + p->tank_center[k];
        dmax[k] = hx + p->tank_center[k];
    }
    s->tank_c = mk(p->tank_center[0], p->tank_center[1], p->tank_center[2]);
    s->tank_h = mk(p->tank_size[0] / 2.0, p->tank_size[1] / 2.0, p->tank_size[2] / 2.0);
    bbo_grid_for_domain(dmin, dmax, p->grid_spacing, p->grid_scale, &s->grid);
    unsigned total = s->grid.total;
    s->nbr = (int *)malloc(sizeof(int) * 27 * (size_t)total);
    s->nbr_cnt = (int *)malloc(sizeof(int) * (size_t)total);
    for (unsigned c = 0; c < total; c++) s->nbr_cnt[c] = bbo_grid_neighbors(&s->grid, c, &s->nbr[27 * (size_t)c]);
    s->cells = (cell_t *)calloc(total, sizeof(cell_t));
    for (int b = 0; b < 2; b++) s->next[b] = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    s->bucket_ids = (int *)malloc(sizeof(int) * BBO_BUCKET_SIZE * (size_t)(n > 0 ? n : 1));
    s->bucket_cnt = (int *)calloc((size_t)(n > 0 ? n : 1), sizeof(int));
    size_t nv = (size_t)(n > 0 ? n : 1);
    s->pos = (v3 *)calloc(nv, sizeof(v3));
    s->vel = (v3 *)calloc(nv, sizeof(v3));
    s->force = (v3 *)calloc(nv, sizeof(v3));
    s->inter = (v3 *)calloc(nv, sizeof(v3));
    s->tpos = (v3 *)calloc(nv, sizeof(v3));
    s->tvel = (v3 *)calloc(nv, sizeof(v3));
    s->fp = (v3 *)calloc(nv, sizeof(v3));
    s->smooth = (v3 *)calloc(nv, sizeof(v3));
    s->density = (double *)calloc(nv, sizeof(double));
    s->pressure = (double *)calloc(nv, sizeof(double));
    s->derr = (double *)calloc(nv, sizeof(double));
    s->dpred = (double *)calloc(nv, sizeof(double));
    for (int i = 0; i < n; i++) {
        s->pos[i] = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        if (vel) s->vel[i] = mk(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    }
    s->mass = 1e-3;
    s->radius = 1e-3;
    return s;
}

void bbo_sim_destroy(bbo_sim *s) {
    if (!s) return;
    free(s->coll);
    free(s->nbr); free(s->nbr_cnt); free(s->cells);
    free(s->next[0]); free(s->next[1]);
    free(s->bucket_ids); free(s->bucket_cnt);
    free(s->pos); free(s->vel); free(s->force); free(s->inter); free(s->tpos); free(s->tvel);
    free(s->fp); free(s->smooth); free(s->density); free(s->pressure); free(s->derr); free(s->dpred);
    free(s);
}

void bbo_sim_set_threads(bbo_sim *s, int threads) { s->threads = threads; }
int bbo_sim_count(const bbo_sim *s) { return s->n; }
int bbo_sim_get_flag(const bbo_sim *s) { return s->flag; }
void bbo_sim_set_flag(bbo_sim *s, int flag) { s->flag = flag; }

void bbo_sim_info(const bbo_sim *s, double *mass, double *kr, double *radius, double *dd,
                  bbo_grid_desc *grid) {
    if (mass) *mass = s->mass;
    if (kr) *kr = s->kernel_radius;
    if (radius) *radius = s->radius;
    if (dd) *dd = s->delta_denom;
    if (grid) *grid = s->grid;
}

static v3 *vfield(const bbo_sim *s, const char *f) {
    if (!strcmp(f, "position")) return s->pos;
    if (!strcmp(f, "velocity")) return s->vel;
    if (!strcmp(f, "force")) return s->force;
    if (!strcmp(f, "interaction")) return s->inter;
    if (!strcmp(f, "temp_position")) return s->tpos;
    if (!strcmp(f, "temp_velocity")) return s->tvel;
    if (!strcmp(f, "pressure_force")) return s->fp;
    return NULL;
}
static double *sfield(const bbo_sim *s, const char *f) {
    if (!strcmp(f, "density")) return s->density;
    if (!strcmp(f, "pressure")) return s->pressure;
    if (!strcmp(f, "density_error")) return s->derr;
    if (!strcmp(f, "density_predicted")) return s->dpred;
    return NULL;
}
void bbo_sim_get(const bbo_sim *s, const char *field, double *out) {
    v3 *v = vfield(s, field);
    if (v) { for (int i = 0; i < s->n; i++) { out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z; } return; }
    double *d = sfield(s, field);
    if (d) memcpy(out, d, sizeof(double) * (size_t)s->n);
}
void bbo_sim_set(bbo_sim *s, const char *field, const double *in) {
    v3 *v = vfield(s, field);
    if (v) { for (int i = 0; i < s->n; i++) v[i] = mk(in[3 * i], in[3 * i + 1], in[3 * i + 2]); return; }
    double *d = sfield(s, field);
    if (d) memcpy(d, in, sizeof(double) * (size_t)s->n);
}

/* ---- chains (src/core/grid.h:20-156) ---- */
static inline int chain_sel(const cell_t *c) { return c->active ? 1 : 0; }

static void cell_reset(cell_t *c) {
    c->head[0] = c->head[1] = -1;
    c->tail[0] = c->tail[1] = -1;
    c->len[0] = c->len[1] = 0;
    c->active = 0;
}
static void add_to_chain(bbo_sim *s, cell_t *c, int b, int node) {
    if (c->head[b] < 0) { c->head[b] = node; c->tail[b] = node; }
    else { s->next[b][c->tail[b]] = node; c->tail[b] = node; }
    s->next[b][node] = -1;
    c->len[b]++;
}

/* DistributeByParticle (src/core/grid.h:389-407), after DistributeResetCell on all */
static void distribute_full(bbo_sim *s) {
    for (unsigned c = 0; c < s->grid.total; c++) cell_reset(&s->cells[c]);
    for (int i = 0; i < s->n; i++) {
        double p[3] = {s->pos[i].x, s->pos[i].y, s->pos[i].z};
        long long h = bbo_grid_hash(&s->grid, p, NULL);
        if (h < 0) continue; /* outside the grid: the reference would fault */
        cell_t *c = &s->cells[h];
        add_to_chain(s, c, chain_sel(c), i);
    }
}

/* DistributeToCellOpt (src/core/grid.h:449-494) for a range of cells */
typedef struct { bbo_sim *s; long long *hash; } opt_ctx;
static void opt_range(void *vctx, int c0, int c1) {
    opt_ctx *o = (opt_ctx *)vctx;
    bbo_sim *s = o->s;
    for (int cid = c0; cid < c1; cid++) {
        cell_t *cell = &s->cells[cid];
        int nb = cell->active ? 0 : 1; /* next chain */
        cell->head[nb] = -1; cell->tail[nb] = -1; cell->len[nb] = 0;
        const int *nbr = &s->nbr[27 * (size_t)cid];
        for (int q = 0; q < s->nbr_cnt[cid]; q++) {
            cell_t *ncell = &s->cells[nbr[q]];
            int b = chain_sel(ncell);
            int node = ncell->head[b];
            int size = ncell->len[b];
            for (int j = 0; j < size; j++) {
                if (o->hash[node] == cid) add_to_chain(s, cell, nb, node);
                node = s->next[b][node];
            }
        }
    }
}
static void hash_range(void *vctx, int i0, int i1) {
    opt_ctx *o = (opt_ctx *)vctx;
    for (int i = i0; i < i1; i++) {
        double p[3] = {o->s->pos[i].x, o->s->pos[i].y, o->s->pos[i].z};
        o->hash[i] = bbo_grid_hash(&o->s->grid, p, NULL);
    }
}

/* DistributeParticleBucket (src/core/grid.h:422-447) */
static void bucket_range(void *vctx, int i0, int i1) {
    bbo_sim *s = (bbo_sim *)vctx;
    for (int pid = i0; pid < i1; pid++) {
        v3 pi = s->pos[pid];
        double p[3] = {pi.x, pi.y, pi.z};
        long long cid = bbo_grid_hash(&s->grid, p, NULL);
        int *ids = &s->bucket_ids[(size_t)pid * BBO_BUCKET_SIZE];
        int count = 0;
        if (cid >= 0) {
            const int *nbr = &s->nbr[27 * (size_t)cid];
            for (int q = 0; q < s->nbr_cnt[cid]; q++) {
                cell_t *cell = &s->cells[nbr[q]];
                int b = chain_sel(cell);
                int node = cell->head[b];
                int size = cell->len[b];
                for (int j = 0; j < size; j++) {
                    double distance = vdist(pi, s->pos[node]);
                    if (bbo_within_std(distance, s->kernel_radius)) {
                        if (count < BBO_BUCKET_SIZE) ids[count++] = node;
                        else __atomic_add_fetch(&s->bucket_overflows, 1, __ATOMIC_RELAXED); /* dropped silently by the reference */
                    }
                    node = s->next[b][node];
                }
            }
        }
        s->bucket_cnt[pid] = count;
    }
}

static int count_in_chains(const bbo_sim *s) {
    int c = 0;
    for (unsigned i = 0; i < s->grid.total; i++) c += s->cells[i].len[chain_sel(&s->cells[i])];
    return c;
}

/* UpdateGridDistributionCPU (src/equations/sph_equations3.cpp:388-423) then
 * ResetHigherLevel (src/solvers/pcisph_solver3.cpp:47) */
void bbo_stage_distribute(bbo_sim *s, bbo_stats *st) {
    int full = 0;
    if (s->flag) {
        distribute_full(s);
        full = 1;
    } else {
        opt_ctx o;
        o.s = s;
        o.hash = (long long *)malloc(sizeof(long long) * (size_t)(s->n > 0 ? s->n : 1));
        parallel_for(s->threads, 0, s->n, hash_range, &o);
        if (s->frame_index == 0) {
            /* DistributeToCell on every cell: ascending id order */
            for (unsigned c = 0; c < s->grid.total; c++) cell_reset(&s->cells[c]);
            for (int i = 0; i < s->n; i++) {
                if (o.hash[i] < 0) continue;
                cell_t *c = &s->cells[o.hash[i]];
                add_to_chain(s, c, chain_sel(c), i);
            }
            full = 1;
        } else {
            parallel_for(s->threads, 0, (int)s->grid.total, opt_range, &o);
            for (unsigned c = 0; c < s->grid.total; c++) s->cells[c].active = 1 - s->cells[c].active;
        }
        free(o.hash);
    }
    s->bucket_overflows = 0;
    parallel_for(s->threads, 0, s->n, bucket_range, s);
    s->frame_index = 1;
    s->flag = 0;
    if (st) {
        st->full_redistribution = full;
        st->bucket_overflows = s->bucket_overflows;
        st->lost_particles = s->n - count_in_chains(s);
    }
}

void bbo_sim_get_chains(const bbo_sim *s, int *cell_start, int *members) {
    int at = 0;
    for (unsigned c = 0; c < s->grid.total; c++) {
        const cell_t *cell = &s->cells[c];
        int b = chain_sel(cell);
        cell_start[c] = at;
        int node = cell->head[b];
        for (int j = 0; j < cell->len[b]; j++) { members[at++] = node; node = s->next[b][node]; }
    }
    cell_start[s->grid.total] = at;
}

void bbo_sim_set_chains(bbo_sim *s, const int *cell_start, const int *members) {
    for (unsigned c = 0; c < s->grid.total; c++) {
        cell_t *cell = &s->cells[c];
        cell_reset(cell);
        for (int j = cell_start[c]; j < cell_start[c + 1]; j++) add_to_chain(s, cell, 0, members[j]);
    }
    s->frame_index = 1;
}

/* ContinuousParticleSetBuilder::Commit with a mapped grid
 * (src/core/grid.h:1410-1425): AppendData (src/core/particle.h:232-240)
 * gives the new particles ids n.., then DistributeByParticleList
 * (src/core/grid.h:358-387) appends each, in order, to the tail of the
 * active chain of the cell it hashes into (AddToChain, :128-140) and builds
 * its bucket. */
static void *grow(void *p, size_t elem, size_t n) {
    void *q = realloc(p, elem * (n > 0 ? n : 1));
    return q;
}
void bbo_sim_append(bbo_sim *s, int count, const double *pos, const double *vel) {
    if (count <= 0) return;
    const int n0 = s->n, n1 = n0 + count;
    const size_t nv = (size_t)n1;
    for (int b = 0; b < 2; b++) s->next[b] = (int *)grow(s->next[b], sizeof(int), nv);
    s->bucket_ids = (int *)grow(s->bucket_ids, sizeof(int) * BBO_BUCKET_SIZE, nv);
    s->bucket_cnt = (int *)grow(s->bucket_cnt, sizeof(int), nv);
    v3 **vs[] = {&s->pos, &s->vel, &s->force, &s->inter, &s->tpos, &s->tvel, &s->fp, &s->smooth};
    for (size_t k = 0; k < sizeof vs / sizeof vs[0]; k++) {
        *vs[k] = (v3 *)grow(*vs[k], sizeof(v3), nv);
        memset(*vs[k] + n0, 0, sizeof(v3) * (size_t)count);
    }
    double **ds[] = {&s->density, &s->pressure, &s->derr, &s->dpred};
    for (size_t k = 0; k < sizeof ds / sizeof ds[0]; k++) {
        *ds[k] = (double *)grow(*ds[k], sizeof(double), nv);
        memset(*ds[k] + n0, 0, sizeof(double) * (size_t)count);
    }
    memset(s->bucket_cnt + n0, 0, sizeof(int) * (size_t)count);
    for (int k = 0; k < count; k++) {
        s->pos[n0 + k] = mk(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]);
        if (vel) s->vel[n0 + k] = mk(vel[3 * k], vel[3 * k + 1], vel[3 * k + 2]);
    }
    s->n = n1;
    for (int i = n0; i < n1; i++) {
        double p[3] = {s->pos[i].x, s->pos[i].y, s->pos[i].z};
        long long h = bbo_grid_hash(&s->grid, p, NULL);
        if (h < 0) continue; /* outside the grid: the reference asserts */
        cell_t *c = &s->cells[h];
        add_to_chain(s, c, chain_sel(c), i);
    }
    parallel_for(s->threads, n0, n1, bucket_range, s);
}

/* ContinuousParticleSetBuilder::MapGrid (src/core/grid.h:1286-1321): every
 * cell that some particle hashes into, in ascending id (the order a
 * std::set<unsigned> iterates), each with the positions of the particles of
 * its active chain, in chain order, at the time of the call.  Returns the
 * number of mapped cells; cells / offsets (ncells + 1) / positions (3 per
 * mapped position) may be NULL to count only, *npos gets the total number of
 * mapped positions. */
int bbo_map_grid(const bbo_sim *s, int *cells, int *offsets, double *positions, int *npos) {
    unsigned char *mark = (unsigned char *)calloc(s->grid.total > 0 ? s->grid.total : 1, 1);
    for (int i = 0; i < s->n; i++) {
        double p[3] = {s->pos[i].x, s->pos[i].y, s->pos[i].z};
        long long h = bbo_grid_hash(&s->grid, p, NULL);
        if (h >= 0) mark[h] = 1;
    }
    int nc = 0, np = 0;
    for (unsigned c = 0; c < s->grid.total; c++) {
        if (!mark[c]) continue;
        const cell_t *cell = &s->cells[c];
        const int b = chain_sel(cell);
        if (cells) {
            cells[nc] = (int)c;
            offsets[nc] = np;
            int node = cell->head[b];
            for (int j = 0; j < cell->len[b]; j++) {
                positions[3 * (np + j)] = s->pos[node].x;
                positions[3 * (np + j) + 1] = s->pos[node].y;
                positions[3 * (np + j) + 2] = s->pos[node].z;
                node = s->next[b][node];
            }
        }
        nc++;
        np += cell->len[b];
    }
    if (offsets) offsets[nc] = np;
    if (npos) *npos = np;
    free(mark);
    return nc;
}

/* The refill rule of MapGridEmit / MapGridEmitToOther (src/core/grid.h:
 * 1323-1395) with AddParticle's capacity test (:1397-1408).  Cells are
 * visited in the mapped order.  A cell whose chain (in this simulation's
 * grid) holds MaximumParticlesPerBucket (100) or more particles is passed
 * over.  Otherwise its first min(100 - chain length, mapped count) mapped
 * positions are tried in order; a position is kept unless a particle of the
 * cell's chain, as the chain stood when the call began, lies at a distance
 * below d (the reference's Distance, FP64).  Keeping a position that would
 * take the particle count past max_particles ends the call with
 * *capacity_hit = 1; the positions kept before it are still committed.
 * Returns how many positions were kept, written to out (3 doubles each, at
 * most out_cap of them). */
int bbo_map_grid_emit(const bbo_sim *s, int ncells, const int *cells, const int *offsets,
                      const double *positions, double d, int max_particles, double *out, int out_cap,
                      int *capacity_hit) {
    int kept = 0;
    *capacity_hit = 0;
    for (int k = 0; k < ncells; k++) {
        const cell_t *cell = &s->cells[cells[k]];
        const int b = chain_sel(cell);
        const int len = cell->len[b];
        if (len >= BBO_BUCKET_SIZE) continue;
        const int avail = offsets[k + 1] - offsets[k];
        const int tries = BBO_BUCKET_SIZE - len < avail ? BBO_BUCKET_SIZE - len : avail;
        for (int t = 0; t < tries; t++) {
            const double *q = &positions[3 * (offsets[k] + t)];
            const v3 pi = mk(q[0], q[1], q[2]);
            int clear = 1;
            int node = cell->head[b];
            for (int j = 0; j < len && clear; j++) {
                if (vdist(s->pos[node], pi) < d) clear = 0;
                node = s->next[b][node];
            }
            if (!clear) continue;
            if (s->n + kept + 1 > max_particles) {
                *capacity_hit = 1;
                return kept;
            }
            if (kept < out_cap) {
                out[3 * kept] = pi.x;
                out[3 * kept + 1] = pi.y;
                out[3 * kept + 2] = pi.z;
            }
            kept++;
        }
    }
    return kept;
}

void bbo_sim_get_buckets(const bbo_sim *s, int *counts, int *ids) {
    memcpy(counts, s->bucket_cnt, sizeof(int) * (size_t)s->n);
    if (ids) memcpy(ids, s->bucket_ids, sizeof(int) * BBO_BUCKET_SIZE * (size_t)s->n);
}

/* ---- setup: src/solvers/pcisph_solver3.cpp:103-146, src/core/particle.h:528-556 ---- */
void bbo_sim_setup(bbo_sim *s) {
    double rho0 = s->prm.target_density;
    double spacing = s->prm.target_spacing;
    /* SetTargetDensity / SetTargetSpacing / SetRelativeKernelRadius each call
     * ComputeMass; only the last call's inputs matter for the result. */
    s->radius = fmaxr(0, spacing);
    s->kernel_radius = s->prm.relative_radius * spacing;
    s->mass = fmaxr(0, bbo_compute_mass(rho0, spacing, s->kernel_radius));
    double m = s->mass / rho0;
    m *= m;
    s->mass_over_rho2 = m;
    s->delta_denom = bbo_delta_denom(spacing, s->kernel_radius);
    distribute_full(s);
    s->flag = 0;
    s->frame_index = 1;
}

/* src/solvers/pcisph_solver3.cpp:182-189 */
double bbo_compute_delta(const bbo_sim *s, double dt) {
    double tss = dt * dt;
    double beta = 2.0 * s->mass_over_rho2 * tss;
    return absf(s->delta_denom) > 0 ? -1 / (beta * s->delta_denom) : 0;
}

/* src/core/particle.h:584-608 */
unsigned bbo_number_of_time_steps(const bbo_sim *s, double interval, double scale) {
    double max_force = 0.0;
    double ref = s->kernel_radius;
    for (int i = 0; i < s->n; i++) max_force = fmaxr(max_force, vlen(s->force[i]));
    double limit = R_SPEED_FACTOR * ref / s->prm.sound_speed;
    if (!is_zero(max_force)) {
        double by_force = R_FORCE_FACTOR * sqrt(ref * s->mass / max_force);
        limit = fminr(by_force, limit);
    }
    double target = scale * limit;
    return (unsigned)ceil(interval / target);
}

/* ---- physics per particle ---- */
/* src/core/interaction.h:69 + src/equations/sph_equations3.cpp:63-78 */
static void inter_range(void *vctx, int i0, int i1) {
    bbo_sim *s = (bbo_sim *)vctx;
    v3 g = mk(s->prm.gravity[0], s->prm.gravity[1], s->prm.gravity[2]);
    for (int i = i0; i < i1; i++) s->inter[i] = vadd(mk(0, 0, 0), g);
}
void bbo_stage_interaction(bbo_sim *s) { parallel_for(s->threads, 0, s->n, inter_range, s); }

/* src/equations/sph_equations3.cpp:7-18 */
static double pressure_value(const bbo_sim *s, double di) {
    double rho0 = s->prm.target_density;
    double eos_scale = rho0 * s->prm.sound_speed * s->prm.sound_speed;
    double p = eos_scale / s->prm.eos_exponent * (pow((di / rho0), s->prm.eos_exponent) - 1.0);
    if (p < 0) p *= s->prm.negative_pressure_scale;
    return p;
}

/* src/equations/sph_equations3.cpp:25-58 */
static void density_range(void *vctx, int i0, int i1) {
    bbo_sim *s = (bbo_sim *)vctx;
    for (int i = i0; i < i1; i++) {
        v3 pi = s->pos[i];
        const int *ids = &s->bucket_ids[(size_t)i * BBO_BUCKET_SIZE];
        double sum = 0;
        for (int q = 0; q < s->bucket_cnt[i]; q++) sum += bbo_poly6_W(s->kernel_radius, vdist(pi, s->pos[ids[q]]));
        double di = sum * s->mass;
        if (s->arg_first) s->pressure[i] = pressure_value(s, di);
        s->density[i] = di;
    }
}
void bbo_stage_density(bbo_sim *s, int compute_pressure) {
    s->arg_first = compute_pressure;
    parallel_for(s->threads, 0, s->n, density_range, s);
}

/* src/equations/sph_equations3.cpp:80-110 */
static void nonpressure_range(void *vctx, int i0, int i1) {
    bbo_sim *s = (bbo_sim *)vctx;
    double mass = s->mass, mass2 = mass * mass, h = s->kernel_radius;
    for (int i = i0; i < i1; i++) {
        v3 pi = s->pos[i], vi = s->vel[i];
        v3 fi = mk(0, 0, 0);
        fi = vadd(fi, vmul(s->inter[i], mass));
        fi = vadd(fi, vmul(vi, -s->prm.drag_coefficient));
        const int *ids = &s->bucket_ids[(size_t)i * BBO_BUCKET_SIZE];
        for (int q = 0; q < s->bucket_cnt[i]; q++) {
            int j = ids[q];
            double dist = vdist(pi, s->pos[j]);
            v3 t = vmul(vsub(s->vel[j], vi), s->prm.viscosity * mass2);
            t = vmul(t, bbo_spiky_d2W(h, dist));
            t = vdiv(t, s->density[j]);
            fi = vadd(fi, t);
        }
        s->force[i] = fi;
    }
}
void bbo_stage_nonpressure(bbo_sim *s) { parallel_for(s->threads, 0, s->n, nonpressure_range, s); }

/* ---- box collider: src/core/collider.cpp:3-44,123-144,238-264;
 *      src/shapes/box.cpp:74-172; src/core/shape.cpp:275-282 ---- */
static int inside_box(v3 p, v3 lo, v3 hi) { /* src/core/geometry.h:1774-1784 */
    int rv = (p.x >= lo.x && p.x <= hi.x && p.y >= lo.y && p.y <= hi.y && p.z >= lo.z && p.z <= hi.z);
    if (!rv) {
        double x0 = absf(lo.x - p.x), x1 = absf(hi.x - p.x);
        double y0 = absf(lo.y - p.y), y1 = absf(hi.y - p.y);
        double z0 = absf(lo.z - p.z), z1 = absf(hi.z - p.z);
        rv = is_unsafe_zero(fminr(x0, x1)) || is_unsafe_zero(fminr(y0, y1)) || is_unsafe_zero(fminr(z0, z1));
    }
    return rv;
}
/* Transform::Point for a pure translation matrix (src/core/transform.h:456-468) */
static inline v3 xf_point(v3 p, v3 t) {
    return mk(1.0 * p.x + 0.0 * p.y + 0.0 * p.z + t.x,
              0.0 * p.x + 1.0 * p.y + 0.0 * p.z + t.y,
              0.0 * p.x + 0.0 * p.y + 1.0 * p.z + t.z);
}
static inline v3 xf_normal(v3 n) { /* inverse-transpose of a translation is identity */
    return mk(1.0 * n.x + 0.0 * n.y + 0.0 * n.z,
              0.0 * n.x + 1.0 * n.y + 0.0 * n.z,
              0.0 * n.x + 0.0 * n.y + 1.0 * n.z);
}
static const double PLANE_N[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

static v3 plane_closest(v3 p, v3 point, v3 n) { /* src/shapes/box.cpp:21-25 */
    v3 r = vsub(p, point);
    return vadd(vsub(r, vmul(n, vdot(r, n))), point);
}

static double box_closest_distance(const bbo_sim *s, v3 point) {
    v3 neg_c = mk(-s->tank_c.x, -s->tank_c.y, -s->tank_c.z);
    v3 pl = xf_point(point, neg_c);
    v3 lo = mk(-s->tank_h.x, -s->tank_h.y, -s->tank_h.z), hi = s->tank_h;
    v3 closest;
    int neg = 0;
    if (inside_box(pl, lo, hi)) {
        closest = plane_closest(pl, hi, mk(1, 0, 0));
        double d2 = vlen2(vsub(closest, pl));
        for (int i = 1; i < 6; i++) {
            v3 n = mk(PLANE_N[i][0], PLANE_N[i][1], PLANE_N[i][2]);
            v3 local = plane_closest(pl, i < 3 ? hi : lo, n);
            double ld2 = vlen2(vsub(local, pl));
            if (ld2 < d2) { closest = local; d2 = ld2; }
        }
        neg = 1;
    } else {
        closest = mk(clampr(pl.x, lo.x, hi.x), clampr(pl.y, lo.y, hi.y), clampr(pl.z, lo.z, hi.z));
    }
    return neg ? -vdist(pl, closest) : vdist(pl, closest);
}

static void box_closest_point(const bbo_sim *s, v3 point, v3 *qpoint, v3 *qnormal, double *qdist) {
    v3 neg_c = mk(-s->tank_c.x, -s->tank_c.y, -s->tank_c.z);
    v3 pl = xf_point(point, neg_c);
    v3 lo = mk(-s->tank_h.x, -s->tank_h.y, -s->tank_h.z), hi = s->tank_h;
    v3 closest, normal = mk(1, 0, 0);
    int neg = 0;
    if (inside_box(pl, lo, hi)) {
        closest = plane_closest(pl, hi, mk(1, 0, 0));
        double d2 = vlen2(vsub(closest, pl));
        for (int i = 1; i < 6; i++) {
            v3 n = mk(PLANE_N[i][0], PLANE_N[i][1], PLANE_N[i][2]);
            v3 local = plane_closest(pl, i < 3 ? hi : lo, n);
            double ld2 = vlen2(vsub(local, pl));
            if (ld2 < d2) { closest = local; normal = n; d2 = ld2; }
        }
        neg = 1;
    } else {
        closest = mk(clampr(pl.x, lo.x, hi.x), clampr(pl.y, lo.y, hi.y), clampr(pl.z, lo.z, hi.z));
        v3 cov = vsub(pl, closest);
        double maxc = vdot(cov, normal);
        for (int i = 1; i < 6; i++) {
            v3 n = mk(PLANE_N[i][0], PLANE_N[i][1], PLANE_N[i][2]);
            double cosine = vdot(n, cov);
            if (cosine > maxc) { normal = n; maxc = cosine; }
        }
    }
    double distance = vdist(closest, pl);
    if (neg) distance *= -1;
    normal = vmul(normal, -1); /* reverseOrientation */
    *qpoint = xf_point(closest, s->tank_c);
    *qnormal = xf_normal(normal);
    *qdist = distance;
}

/* ColliderSet3::ResolveCollision with the single container collider */
static void tank_collision(const bbo_sim *s, double radius, double e, v3 *pos, v3 *vel) {
    /* nearest collider by |ClosestDistance|: only one, so it always wins
     * unless the distance is not < Infinity (FLT_MAX) */
    double d = absf(box_closest_distance(s, *pos));
    if (!(d < R_INFINITY)) return;
    v3 cp, cn;
    double sd;
    box_closest_point(s, *pos, &cp, &cn, &sd);
    /* IsInside for a reversed box: reverseOrientation == !(ClosestDistance < 0) */
    int is_inside = (box_closest_distance(s, *pos) < 0) ? 0 : 1;
    if (!(is_inside || absf(sd) < radius)) return;
    v3 target = vadd(cp, vmul(cn, radius));
    v3 colv = mk(0, 0, 0);
    v3 rel = vsub(*vel, colv);
    double ndr = vdot(cn, rel);
    v3 reln = vmul(cn, ndr);
    v3 relt = vsub(rel, reln);
    if (ndr < 0) {
        v3 drn = vmul(reln, (-1.0 - e));
        reln = vmul(reln, -e);
        if (vlen2(relt) > 0) {
            double scale = vlen(drn) / vlen(relt);
            double fscale = 1.0 - 0.0 * scale; /* frictionCoefficient 0 */
            double frs = fmaxr(0, fscale);
            relt = vmul(relt, frs);
        }
        *vel = vadd(vadd(reln, relt), colv);
    }
    *pos = target;
}

/* ---- general colliders ---- */
/* Transform::Point / ::Normal (src/core/transform.h:433-468), affine rows */
static inline v3 cx_point(const double *m, v3 p) {
    return mk(m[0] * p.x + m[1] * p.y + m[2] * p.z + m[3], m[4] * p.x + m[5] * p.y + m[6] * p.z + m[7],
              m[8] * p.x + m[9] * p.y + m[10] * p.z + m[11]);
}
static inline v3 cx_normal(const double *mi, v3 n) {
    return mk(mi[0] * n.x + mi[4] * n.y + mi[8] * n.z, mi[1] * n.x + mi[5] * n.y + mi[9] * n.z,
              mi[2] * n.x + mi[6] * n.y + mi[10] * n.z);
}
static inline v3 cx_v3(const double *a) { return mk(a[0], a[1], a[2]); }
static inline v3 vcross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

/* BoxClosestPoint (src/shapes/box.cpp:111-172); the distance is
 * BoxClosestDistance's (:74-109), which picks the same closest point */
static void cbox_query(const bbo_collider *c, v3 point, v3 *qp, v3 *qn, double *qd) {
    v3 pl = cx_point(c->to_object, point);
    double hx = c->size[0] / 2.0, hy = c->size[1] / 2.0, hz = c->size[2] / 2.0;
    v3 lo = mk(-hx, -hy, -hz), hi = mk(hx, hy, hz);
    v3 closest, normal = mk(1, 0, 0);
    int neg = 0;
    if (inside_box(pl, lo, hi)) {
        closest = plane_closest(pl, hi, mk(1, 0, 0));
        double d2 = vlen2(vsub(closest, pl));
        for (int i = 1; i < 6; i++) {
            v3 n = mk(PLANE_N[i][0], PLANE_N[i][1], PLANE_N[i][2]);
            v3 local = plane_closest(pl, i < 3 ? hi : lo, n);
            double ld2 = vlen2(vsub(local, pl));
            if (ld2 < d2) { closest = local; normal = n; d2 = ld2; }
        }
        neg = 1;
    } else {
        closest = mk(clampr(pl.x, lo.x, hi.x), clampr(pl.y, lo.y, hi.y), clampr(pl.z, lo.z, hi.z));
        v3 cov = vsub(pl, closest);
        double maxc = vdot(cov, normal);
        for (int i = 1; i < 6; i++) {
            v3 n = mk(PLANE_N[i][0], PLANE_N[i][1], PLANE_N[i][2]);
            double cosine = vdot(n, cov);
            if (cosine > maxc) { normal = n; maxc = cosine; }
        }
    }
    double distance = vdist(closest, pl);
    if (neg) distance *= -1;
    if (c->reverse_orientation) normal = vmul(normal, -1);
    *qp = cx_point(c->to_world, closest);
    *qn = cx_normal(c->to_object, normal);
    *qd = distance;
}

/* SphereClosestDistance / SphereClosestPoint (src/shapes/sphere.cpp:47-70),
 * VelocityAt (src/core/shape.cpp:300-305) */
static double csphere_distance(const bbo_collider *c, v3 point) {
    v3 pl = cx_point(c->to_object, point);
    return vdist(pl, mk(0, 0, 0)) - c->radius;
}
static void csphere_query(const bbo_collider *c, v3 point, v3 *qp, v3 *qn, double *qd, v3 *qv) {
    double d = csphere_distance(c, point);
    v3 pl = cx_point(c->to_object, point);
    v3 n = mk(0, 1, 0);
    if (!(pl.x == 0 && pl.y == 0 && pl.z == 0)) n = vdiv(pl, vlen(pl));
    v3 pn = vmul(n, c->radius);
    if (c->reverse_orientation) n = vmul(n, -1);
    v3 p = cx_point(c->to_world, pn);
    *qp = p;
    *qn = cx_normal(c->to_object, n);
    *qd = d;
    v3 r = vsub(p, mk(c->to_world[3], c->to_world[7], c->to_world[11]));
    *qv = vadd(cx_v3(c->linear_velocity), vcross(cx_v3(c->angular_velocity), r));
}

/* GetBarycentric (src/core/sampling.h:7-31) */
static void barycentric(double x, int nodes, int *i, double *f) {
    double s = floor(x);
    int high = nodes - 1;
    if (high == 0) { *i = 0; *f = 0; }
    else if (s < 0) { *i = 0; *f = 0; }
    else if (s > (double)(high - 1)) { *i = high - 1; *f = 1; }
    else { *i = (int)s; *f = x - s; }
}
static inline double lerpd(double a, double b, double t) { return (1 - t) * a + t * b; }

/* FieldGrid::Sample (src/core/grid.h:1093-1131) with Trilerp (src/core/sampling.h:33-44) */
double bbo_sdf_sample(const bbo_sdf *g, const double p[3]) {
    int ii[3], jj[3];
    double w[3];
    for (int k = 0; k < 3; k++) {
        double nrm = (p[k] - g->origin[k]) / g->spacing[k];
        barycentric(nrm, g->nodes[k], &ii[k], &w[k]);
        jj[k] = ii[k] + 1 < g->nodes[k] - 1 ? ii[k] + 1 : g->nodes[k] - 1;
    }
    const int sy = g->nodes[0], sz = g->nodes[0] * g->nodes[1];
    const double *f = g->values;
#define SDFV(a, b, c) f[(a) + (b) * sy + (c) * sz]
    double f000 = SDFV(ii[0], ii[1], ii[2]), f100 = SDFV(jj[0], ii[1], ii[2]);
    double f010 = SDFV(ii[0], jj[1], ii[2]), f110 = SDFV(jj[0], jj[1], ii[2]);
    double f001 = SDFV(ii[0], ii[1], jj[2]), f101 = SDFV(jj[0], ii[1], jj[2]);
    double f011 = SDFV(ii[0], jj[1], jj[2]), f111 = SDFV(jj[0], jj[1], jj[2]);
#undef SDFV
    double b0 = lerpd(lerpd(f000, f100, w[0]), lerpd(f010, f110, w[0]), w[1]);
    double b1 = lerpd(lerpd(f001, f101, w[0]), lerpd(f011, f111, w[0]), w[1]);
    return lerpd(b0, b1, w[2]);
}

/* FieldGrid::Gradient (src/core/grid.h:1133-1153): central differences */
void bbo_sdf_gradient(const bbo_sdf *g, const double p[3], double out[3]) {
    for (int i = 0; i < 3; i++) {
        double a[3] = {p[0], p[1], p[2]}, b[3] = {p[0], p[1], p[2]};
        a[i] += g->spacing[i];
        b[i] -= g->spacing[i];
        double inv = 1.0 / (2.0 * g->spacing[i]);
        out[i] = (bbo_sdf_sample(g, a) - bbo_sdf_sample(g, b)) * inv;
    }
}
static double sdf_at(const bbo_sdf *g, v3 p) { double q[3] = {p.x, p.y, p.z}; return bbo_sdf_sample(g, q); }
static v3 sdf_grad(const bbo_sdf *g, v3 p) {
    double q[3] = {p.x, p.y, p.z}, o[3];
    bbo_sdf_gradient(g, q, o);
    return mk(o[0], o[1], o[2]);
}

/* ClosestPointBySDF (src/shapes/bvh.cpp:666-706) */
static void csdf_query(const bbo_collider *c, v3 pworld, v3 *qp, v3 *qn, double *qd) {
    v3 tn = mk(0, 1, 0);
    v3 point = cx_point(c->to_object, pworld);
    v3 tp = point;
    int has_gradient = 0;
    for (int i = 0; i < 5; i++) {
        double sdf = sdf_at(&c->sdf, tp);
        if (absf(sdf) < 0.001) break;
        tn = sdf_grad(&c->sdf, tp);
        double len = vlen(tn);
        if (len > 0) tn = vdiv(tn, len);
        tp = vsub(tp, vmul(tn, sdf));
        has_gradient = 1;
    }
    if (!has_gradient) {
        tn = sdf_grad(&c->sdf, tp);
        double len = vlen(tn);
        if (len > 0) tn = vdiv(tn, len);
    }
    *qp = cx_point(c->to_world, tp);
    *qn = cx_normal(c->to_object, tn);
    *qd = vdist(point, tp);
}

/* Shape::ClosestDistance (src/core/shape.cpp:206-228); a mesh takes its SDF
 * sample where the reference walks its BVH */
double bbo_collider_closest_distance(const bbo_collider *c, const double p[3]) {
    v3 q = mk(p[0], p[1], p[2]), cp, cn;
    double d;
    switch (c->kind) {
        case BBO_SHAPE_BOX: cbox_query(c, q, &cp, &cn, &d); return d;
        case BBO_SHAPE_SPHERE: return csphere_distance(c, q);
        default: return bbo_sdf_sample(&c->sdf, p);
    }
}
/* Shape::IsInside (src/core/shape.cpp:275-282) */
int bbo_collider_is_inside(const bbo_collider *c, const double p[3]) {
    if (c->kind == BBO_SHAPE_MESH || c->kind == BBO_SHAPE_SDF) return bbo_sdf_sample(&c->sdf, p) < 0;
    return (c->reverse_orientation != 0) == !(bbo_collider_closest_distance(c, p) < 0);
}
void bbo_collider_closest_point(const bbo_collider *c, const double p[3], double point[3], double normal[3],
                                double *signed_distance, double velocity[3]) {
    v3 q = mk(p[0], p[1], p[2]), cp, cn, cv = mk(0, 0, 0);
    double d;
    if (c->kind == BBO_SHAPE_BOX) cbox_query(c, q, &cp, &cn, &d);
    else if (c->kind == BBO_SHAPE_SPHERE) csphere_query(c, q, &cp, &cn, &d, &cv);
    else csdf_query(c, q, &cp, &cn, &d);
    point[0] = cp.x; point[1] = cp.y; point[2] = cp.z;
    normal[0] = cn.x; normal[1] = cn.y; normal[2] = cn.z;
    velocity[0] = cv.x; velocity[1] = cv.y; velocity[2] = cv.z;
    *signed_distance = d;
}

static int in_bounds(const bbo_collider *c, v3 p) {
    return inside_box(p, cx_v3(c->bounds_min), cx_v3(c->bounds_max));
}

/* CollisionHandle with Collider3::IsPenetrating / OptmizedClosestPointCheck
 * (src/core/collider.cpp:3-44,113-131) */
static int collision_handle(const bbo_collider *c, double radius, double e, v3 *pos, v3 *vel) {
    if (c->kind == BBO_SHAPE_MESH && !in_bounds(c, *pos)) return 0;
    double p[3] = {pos->x, pos->y, pos->z}, qp[3], qn[3], qv[3], qd;
    bbo_collider_closest_point(c, p, qp, qn, &qd, qv);
    if (!(bbo_collider_is_inside(c, p) || absf(qd) < radius)) return 0;
    v3 n = cx_v3(qn), colv = cx_v3(qv);
    v3 target = vadd(cx_v3(qp), vmul(n, radius));
    v3 rel = vsub(*vel, colv);
    double ndr = vdot(n, rel);
    v3 reln = vmul(n, ndr);
    v3 relt = vsub(rel, reln);
    if (ndr < 0) {
        v3 drn = vmul(reln, (-1.0 - e));
        reln = vmul(reln, -e);
        if (vlen2(relt) > 0) {
            double scale = vlen(drn) / vlen(relt);
            double fscale = 1.0 - c->friction * scale;
            relt = vmul(relt, fmaxr(0, fscale));
        }
        *vel = vadd(vadd(reln, relt), colv);
    }
    *pos = target;
    return 1;
}

/* ColliderSet3::ResolveCollision (src/core/collider.cpp:238-264) */
static int resolve_set(const bbo_collider *cs, int count, double radius, double e, v3 *pos, v3 *vel) {
    int target = -1;
    double mind = R_INFINITY;
    for (int i = 0; i < count; i++) {
        const bbo_collider *c = &cs[i];
        if (!c->active) continue;
        if (c->kind == BBO_SHAPE_MESH && !in_bounds(c, *pos)) continue;
        double p[3] = {pos->x, pos->y, pos->z};
        double d = absf(bbo_collider_closest_distance(c, p));
        if (d < mind) { target = i; mind = d; }
    }
    if (target < 0) return 0;
    return collision_handle(&cs[target], radius, e, pos, vel);
}

void bbo_colliders_resolve(const bbo_collider *cs, int count, double radius, double restitution, double *pos,
                           double *vel, int n, int *collided) {
    for (int i = 0; i < n; i++) {
        v3 p = mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        v3 v = mk(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        int h = resolve_set(cs, count, radius, restitution, &p, &v);
        pos[3 * i] = p.x; pos[3 * i + 1] = p.y; pos[3 * i + 2] = p.z;
        vel[3 * i] = v.x; vel[3 * i + 1] = v.y; vel[3 * i + 2] = v.z;
        if (collided) collided[i] = h;
    }
}

void bbo_sim_set_colliders(bbo_sim *s, const bbo_collider *cs, int count) {
    free(s->coll);
    s->coll = NULL;
    s->ncoll = 0;
    if (count <= 0) return;
    s->coll = (bbo_collider *)malloc(sizeof(bbo_collider) * (size_t)count);
    memcpy(s->coll, cs, sizeof(bbo_collider) * (size_t)count);
    s->ncoll = count;
}

static void resolve_collision(const bbo_sim *s, double radius, double e, v3 *pos, v3 *vel) {
    if (s->ncoll > 0) resolve_set(s->coll, s->ncoll, radius, e, pos, vel);
    else tank_collision(s, radius, e, pos, vel);
}

/* ---- mesh SDF: SetNodeSDFKernel / GenerateShapeSDF
 *      (src/core/shape.cpp:411-424,478-505) ---- */
static inline double sgnd(double a) { int t = a < 0 ? -1 : 0; return a > 0 ? 1 : t; }
/* DistanceTriangle (src/shapes/bvh.cpp:212-242), squared */
static double tri_dist2(const double *t, double px, double py, double pz) {
    const double ax = t[0], ay = t[1], az = t[2], bx = t[3], by = t[4], bz = t[5];
    const double cx = t[6], cy = t[7], cz = t[8];
    const double bax = bx - ax, bay = by - ay, baz = bz - az, pax = px - ax, pay = py - ay, paz = pz - az;
    const double cbx = cx - bx, cby = cy - by, cbz = cz - bz, pbx = px - bx, pby = py - by, pbz = pz - bz;
    const double acx = ax - cx, acy = ay - cy, acz = az - cz, pcx = px - cx, pcy = py - cy, pcz = pz - cz;
    const double nx = bay * acz - baz * acy, ny = baz * acx - bax * acz, nz = bax * acy - bay * acx;
    const double s0 = (bay * nz - baz * ny) * pax + (baz * nx - bax * nz) * pay + (bax * ny - bay * nx) * paz;
    const double s1 = (cby * nz - cbz * ny) * pbx + (cbz * nx - cbx * nz) * pby + (cbx * ny - cby * nx) * pbz;
    const double s2 = (acy * nz - acz * ny) * pcx + (acz * nx - acx * nz) * pcy + (acx * ny - acy * nx) * pcz;
    if (sgnd(s0) + sgnd(s1) + sgnd(s2) < 2.0) {
        double t0 = clampr((bax * pax + bay * pay + baz * paz) / (bax * bax + bay * bay + baz * baz), 0.0001, 1.0);
        double t1 = clampr((cbx * pbx + cby * pby + cbz * pbz) / (cbx * cbx + cby * cby + cbz * cbz), 0.0001, 1.0);
        double t2 = clampr((acx * pcx + acy * pcy + acz * pcz) / (acx * acx + acy * acy + acz * acz), 0.0001, 1.0);
        double ex = bax * t0 - pax, ey = bay * t0 - pay, ez = baz * t0 - paz;
        double fx = cbx * t1 - pbx, fy = cby * t1 - pby, fz = cbz * t1 - pbz;
        double gx = acx * t2 - pcx, gy = acy * t2 - pcy, gz = acz * t2 - pcz;
        double d0 = ex * ex + ey * ey + ez * ez, d1 = fx * fx + fy * fy + fz * fz, d2 = gx * gx + gy * gy + gz * gz;
        return fminr(fminr(d0, d1), d2);
    }
    double np = nx * pax + ny * pay + nz * paz;
    return np * np / (nx * nx + ny * ny + nz * nz);
}
double bbo_triangle_distance(const double a[3], const double b[3], const double c[3], const double p[3]) {
    double t[9] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
    return sqrt(tri_dist2(t, p[0], p[1], p[2]));
}
/* signed solid angle of a triangle seen from p (Van Oosterom-Strackee) */
static double tri_solid_angle(const double *t, double px, double py, double pz) {
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

typedef struct {
    const double *tri;
    int ntri, nx, ny;
    double o[3], dx;
    double *values;
} sdf_ctx;
static void sdf_range(void *vctx, int i0, int i1) {
    sdf_ctx *c = (sdf_ctx *)vctx;
    for (int node = i0; node < i1; node++) {
        int i = node % c->nx, j = (node / c->nx) % c->ny, k = node / (c->nx * c->ny);
        double px = c->o[0] + c->dx * (double)i, py = c->o[1] + c->dx * (double)j, pz = c->o[2] + c->dx * (double)k;
        double best = DBL_MAX, wind = 0;
        for (int t = 0; t < c->ntri; t++) {
            best = fminr(best, tri_dist2(&c->tri[9 * t], px, py, pz));
            wind += tri_solid_angle(&c->tri[9 * t], px, py, pz);
        }
        int interior = absf(wind) > 2.0 * R_PI;
        double psd = fmaxr(sqrt(best), 0.00001);
        c->values[node] = interior ? -psd : psd;
    }
}

int bbo_sdf_from_mesh(const double *vertices, int nvertices, const int *triangles, int ntriangles, double dx,
                      double margin, int threads, int nodes[3], double origin[3], double spacing[3],
                      double *values) {
    if (nvertices < 3 || ntriangles < 1 || !(dx > 0)) return -1;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    double *tri = (double *)malloc(sizeof(double) * 9 * (size_t)ntriangles);
    for (int t = 0; t < ntriangles; t++)
        for (int c = 0; c < 3; c++) {
            int v = triangles[3 * t + c];
            if (v < 0 || v >= nvertices) { free(tri); return -1; }
            for (int k = 0; k < 3; k++) {
                double x = vertices[3 * v + k];
                tri[9 * t + 3 * c + k] = x;
                lo[k] = fminr(lo[k], x);
                hi[k] = fmaxr(hi[k], x);
            }
        }
    double pmin[3], pmax[3];
    for (int k = 0; k < 3; k++) {
        double scale = hi[k] - lo[k];
        pmin[k] = lo[k] - margin * scale;
        pmax[k] = hi[k] + margin * scale;
    }
    double width = pmax[0] - pmin[0], height = pmax[1] - pmin[1], depth = pmax[2] - pmin[2];
    int res = (int)ceil(width / dx);
    double sp = width / (double)res;
    int resy = (int)ceil(res * height / width);
    int resz = (int)ceil(res * depth / width);
    nodes[0] = res + 1; nodes[1] = resy + 1; nodes[2] = resz + 1;
    for (int k = 0; k < 3; k++) { origin[k] = pmin[k]; spacing[k] = sp; }
    if (values) {
        sdf_ctx c = {tri, ntriangles, nodes[0], nodes[1], {pmin[0], pmin[1], pmin[2]}, sp, values};
        parallel_for(threads, 0, nodes[0] * nodes[1] * nodes[2], sdf_range, &c);
    }
    free(tri);
    return 0;
}

/* ---- PCISPH: src/equations/pcisph_equations3.cpp ---- */
static void predict_range(void *vctx, int i0, int i1) { /* :3-28 */
    bbo_sim *s = (bbo_sim *)vctx;
    double dt = s->arg_dt;
    for (int i = i0; i < i1; i++) {
        if (s->arg_first) {
            s->dpred[i] = s->density[i];
            s->fp[i] = mk(0, 0, 0);
            s->derr[i] = 0;
            s->pressure[i] = 0;
        }
        v3 tv = vadd(s->vel[i], vmul(vadd(s->force[i], s->fp[i]), dt / s->mass));
        v3 tp = vadd(s->pos[i], vmul(tv, dt));
        resolve_collision(s, s->radius, 0, &tp, &tv);
        s->tvel[i] = tv;
        s->tpos[i] = tp;
    }
}
void bbo_stage_predict(bbo_sim *s, double dt, int is_first) {
    s->arg_dt = dt;
    s->arg_first = is_first;
    parallel_for(s->threads, 0, s->n, predict_range, s);
}

static void predict_pressure_range(void *vctx, int i0, int i1) { /* :60-90 */
    bbo_sim *s = (bbo_sim *)vctx;
    double delta = s->arg_delta;
    for (int i = i0; i < i1; i++) {
        v3 pi = s->tpos[i];
        double ppi = s->pressure[i];
        const int *ids = &s->bucket_ids[(size_t)i * BBO_BUCKET_SIZE];
        double ws = 0;
        for (int q = 0; q < s->bucket_cnt[i]; q++) ws += bbo_poly6_W(s->kernel_radius, vdist(pi, s->tpos[ids[q]]));
        double density = ws * s->mass;
        double err = density - s->prm.target_density;
        double pressure = delta * err;
        if (pressure < 0) {
            pressure *= s->prm.negative_pressure_scale;
            err *= s->prm.negative_pressure_scale;
        }
        ppi += pressure;
        s->pressure[i] = ppi;
        s->dpred[i] = density;
        if (s->arg_record) s->derr[i] = err;
    }
}
void bbo_stage_predict_pressure(bbo_sim *s, double delta, int record_error) {
    s->arg_delta = delta;
    s->arg_record = record_error;
    parallel_for(s->threads, 0, s->n, predict_pressure_range, s);
}

static void pressure_force_range(void *vctx, int i0, int i1) { /* :114-155 */
    bbo_sim *s = (bbo_sim *)vctx;
    double mass = s->mass, mass2 = mass * mass, h = s->kernel_radius;
    for (int i = i0; i < i1; i++) {
        v3 pi = s->pos[i];
        double poi = s->pressure[i];
        double di = s->dpred[i];
        double di2 = di * di;
        v3 fi = mk(0, 0, 0), ti = mk(0, 0, 0);
        const int *ids = &s->bucket_ids[(size_t)i * BBO_BUCKET_SIZE];
        for (int q = 0; q < s->bucket_cnt[i]; q++) {
            int j = ids[q];
            if (j == i) continue;
            v3 pj = s->pos[j];
            double dj = s->dpred[j];
            double dj2 = dj * dj;
            double poj = s->pressure[j];
            double dist = vdist(pi, pj);
            int valid = bbo_within_spiky(dist, h);
            if (valid && !is_zero(dj2)) {
                if (dist > 0 && !is_zero(dist)) {
                    v3 dir = vdiv(vsub(pj, pi), dist);
                    v3 g = spiky_grad(h, dist, dir);
                    ti = vadd(ti, vmul(g, mass2 * (poi / di2 + poj / dj2)));
                }
            }
        }
        s->fp[i] = vsub(fi, ti);
    }
}
void bbo_stage_pressure_force(bbo_sim *s) { parallel_for(s->threads, 0, s->n, pressure_force_range, s); }

/* TimeIntegrationFor (src/equations/sph_equations3.cpp:283-339) */
static void integrate_one(bbo_sim *s, int i, double dt) {
    v3 vi = s->vel[i], fi = s->force[i], pi = s->pos[i];
    v3 oi = pi;
    v3 aT = mk(0, 0, 0);
    v3 aF = vdiv(fi, s->mass);
    v3 at = vadd(aT, aF);
    vi = vadd(vi, vmul(at, dt));
    pi = vadd(pi, vmul(vi, dt));
    resolve_collision(s, s->radius, 0.6, &pi, &vi);
    v3 lo = mk(s->grid.bounds_min[0], s->grid.bounds_min[1], s->grid.bounds_min[2]);
    v3 hi = mk(s->grid.bounds_max[0], s->grid.bounds_max[1], s->grid.bounds_max[2]);
    if (!inside_box(pi, lo, hi)) {
        double r = s->radius;
        pi = mk(clampr(pi.x, lo.x + r, hi.x - r), clampr(pi.y, lo.y + r, hi.y - r),
                clampr(pi.z, lo.z + r, hi.z - r));
    }
    double dist = vdist(pi, oi);
    double minlen = fminr(s->grid.cells_len[0], fminr(s->grid.cells_len[1], s->grid.cells_len[2]));
    if (dist >= minlen * 0.9) s->flag = 1;
    s->pos[i] = pi;
    s->vel[i] = vi;
}

static void acc_int_range(void *vctx, int i0, int i1) { /* :179-197 */
    bbo_sim *s = (bbo_sim *)vctx;
    for (int i = i0; i < i1; i++) {
        s->force[i] = vadd(s->force[i], s->fp[i]);
        integrate_one(s, i, s->arg_dt);
    }
}
void bbo_stage_accumulate_integrate(bbo_sim *s, double dt) {
    s->arg_dt = dt;
    parallel_for(s->threads, 0, s->n, acc_int_range, s);
}

/* XSPH: src/equations/sph_equations3.cpp:341-382,466-482 */
static void xsph_agg_range(void *vctx, int i0, int i1) {
    bbo_sim *s = (bbo_sim *)vctx;
    double h = s->kernel_radius;
    for (int i = i0; i < i1; i++) {
        v3 pi = s->pos[i];
        v3 sm = mk(0, 0, 0);
        double wsum = 0;
        const int *ids = &s->bucket_ids[(size_t)i * BBO_BUCKET_SIZE];
        for (int q = 0; q < s->bucket_cnt[i]; q++) {
            int j = ids[q];
            double dist = vdist(pi, s->pos[j]);
            double wj = s->mass / s->density[j] * bbo_spiky_W(h, dist);
            wsum += wj;
            sm = vadd(sm, vmul(s->vel[j], wj));
        }
        if (wsum > 0) sm = vmul(sm, 1.0 / wsum);
        s->smooth[i] = sm;
    }
}
static void xsph_int_range(void *vctx, int i0, int i1) {
    bbo_sim *s = (bbo_sim *)vctx;
    double f = clampr(s->arg_dt * s->prm.pseudo_viscosity, 0.0, 1.0);
    for (int i = i0; i < i1; i++) {
        v3 vi = s->vel[i], sv = s->smooth[i];
        /* Lerp(t, a, b) = (1 - t) * a + t * b */
        s->vel[i] = vadd(vmul(vi, 1 - f), vmul(sv, f));
    }
}
void bbo_stage_xsph(bbo_sim *s, double dt) {
    double scale = s->prm.pseudo_viscosity * dt;
    if (scale > 0.1) {
        s->arg_dt = dt;
        parallel_for(s->threads, 0, s->n, xsph_agg_range, s);
        parallel_for(s->threads, 0, s->n, xsph_int_range, s);
    }
}

static double absf_max(double x, double y) { return (x * x > y * y) ? x : y; }

static double max_error_ratio(const bbo_sim *s) {
    double m = 0;
    for (int i = 0; i < s->n; i++) m = absf_max(m, s->derr[i]);
    return m / s->prm.target_density;
}

/* AdvanceTimeStep (src/solvers/pcisph_solver3.cpp:42-65) + the loop of
 * ComputePressureForceAndIntegrate (src/equations/pcisph_equations3.cpp:218-256) */
void bbo_pcisph_substep(bbo_sim *s, double dt, int max_it, double tol, int mode,
                        const double *warm, bbo_stats *st) {
    bbo_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    bbo_stage_distribute(s, st);
    bbo_stage_interaction(s);
    bbo_stage_density(s, 1);
    bbo_stage_nonpressure(s);
    double delta = bbo_compute_delta(s, dt);
    int it = 0;
    double ratio = 0;
    for (int k = 0; k < max_it; k++) {
        if (k == 0 && warm) {
            for (int i = 0; i < s->n; i++) {
                s->dpred[i] = s->density[i];
                s->derr[i] = 0;
                s->pressure[i] = warm[i];
            }
            bbo_stage_pressure_force(s);
            bbo_stage_predict(s, dt, 0);
        } else {
            bbo_stage_predict(s, dt, k == 0);
        }
        /* the error is always recorded for reporting; REFERENCE_EXACT ignores it */
        bbo_stage_predict_pressure(s, delta, 1);
        bbo_stage_pressure_force(s);
        it = k + 1;
        ratio = max_error_ratio(s);
        if (mode == BBO_MODE_REFERENCE_EXACT) break; /* error array reads 0 (F4) */
        if (absf(ratio) < tol) break;
    }
    st->iterations = it;
    st->max_density_error = ratio;
    bbo_stage_accumulate_integrate(s, dt);
    bbo_stage_xsph(s, dt);
}

/* PciSphSolver3::Advance (src/solvers/pcisph_solver3.cpp:67-89), scale 5
 * (kDefaultTimeStepLimitScale, src/solvers/pcisph_solver2.cpp:7) */
int bbo_pcisph_advance(bbo_sim *s, double interval, int max_it, double tol, int mode) {
    double remaining = interval;
    int steps = 0;
    while (remaining > R_EPSILON) {
        unsigned n = bbo_number_of_time_steps(s, remaining, 5.0);
        double dt = remaining / (double)n;
        bbo_pcisph_substep(s, dt, max_it, tol, mode, NULL, NULL);
        remaining -= dt;
        steps++;
    }
    return steps;
}

/* ---- SPH: ComputeAllForcesFor (src/equations/sph_equations3.cpp:184-278) ---- */
static void sph_force_one(bbo_sim *s, int i) {
    double mass = s->mass, mass2 = mass * mass, h = s->kernel_radius;
    v3 pi = s->pos[i], vi = s->vel[i];
    double poi = s->pressure[i], di = s->density[i], di2 = di * di;
    v3 fi = mk(0, 0, 0), ti = mk(0, 0, 0);
    fi = vadd(fi, vmul(s->inter[i], mass));
    fi = vadd(fi, vmul(vi, -s->prm.drag_coefficient));
    const int *ids = &s->bucket_ids[(size_t)i * BBO_BUCKET_SIZE];
    for (int q = 0; q < s->bucket_cnt[i]; q++) {
        int j = ids[q];
        if (j == i) continue;
        v3 pj = s->pos[j];
        double dj = s->density[j], dj2 = dj * dj;
        v3 vj = s->vel[j];
        double poj = s->pressure[j];
        double dist = vdist(pi, pj);
        if (bbo_within_spiky(dist, h)) {
            if (dist > 0 && !is_zero(dist)) {
                v3 dir = vdiv(vsub(pj, pi), dist);
                v3 g = spiky_grad(h, dist, dir);
                ti = vadd(ti, vmul(g, mass2 * (poi / di2 + poj / dj2)));
            }
            v3 t = vmul(vsub(vj, vi), s->prm.viscosity * mass2);
            t = vmul(t, bbo_spiky_d2W(h, dist));
            t = vdiv(t, dj);
            fi = vadd(fi, t);
        }
    }
    s->force[i] = vsub(fi, ti);
}
static void sph_force_range(void *vctx, int i0, int i1) {
    bbo_sim *s = (bbo_sim *)vctx;
    for (int i = i0; i < i1; i++) sph_force_one(s, i);
}
static void sph_int_range(void *vctx, int i0, int i1) {
    bbo_sim *s = (bbo_sim *)vctx;
    for (int i = i0; i < i1; i++) integrate_one(s, i, s->arg_dt);
}

/* force phase of ComputeAllForcesFor alone (two-phase form), for tests */
void bbo_stage_sph_force(bbo_sim *s) { parallel_for(s->threads, 0, s->n, sph_force_range, s); }

/* AdvanceTimeStep(SphSolver3*) (src/solvers/sph_solver3.cpp:48-67) */
void bbo_sph_substep(bbo_sim *s, double dt, int sph_mode, bbo_stats *st) {
    bbo_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    bbo_stage_distribute(s, st);
    bbo_stage_interaction(s);
    bbo_stage_density(s, 1);
    s->arg_dt = dt;
    if (sph_mode == BBO_SPH_SEQUENTIAL) {
        for (int i = 0; i < s->n; i++) { sph_force_one(s, i); integrate_one(s, i, dt); }
    } else {
        parallel_for(s->threads, 0, s->n, sph_force_range, s);
        parallel_for(s->threads, 0, s->n, sph_int_range, s);
    }
    bbo_stage_xsph(s, dt);
}

int bbo_sph_advance(bbo_sim *s, double interval, int sph_mode) {
    double remaining = interval;
    int steps = 0;
    while (remaining > R_EPSILON) {
        unsigned n = bbo_number_of_time_steps(s, remaining, 1.0);
        double dt = remaining / (double)n;
        bbo_sph_substep(s, dt, sph_mode, NULL);
        remaining -= dt;
        steps++;
    }
    return steps;
}

/* ---- learned warm-start features (include/bbx_features.h) ----
 * The state at the start of a substep after the distribution and density
 * stages: bucket, density, own position and velocity and the previous
 * substep's converged pressure (s->pressure).  This library's own inputs;
 * the reference has no learned component. */
#include "../include/bbx_features.h"
void bbo_features(bbo_sim *s, double dt, double *out) {
    double h = s->kernel_radius, rho0 = s->prm.target_density;
    double delta = bbo_compute_delta(s, dt);
    double pref = delta * rho0;
    for (int i = 0; i < s->n; i++) {
        v3 pi = s->pos[i], vi = s->vel[i];
        const int *ids = &s->bucket_ids[(size_t)i * BBO_BUCKET_SIZE];
        int cnt = s->bucket_cnt[i];
        double m1 = 0, m2 = 0, m3 = 0;
        v3 c = mk(0, 0, 0);
        for (int q = 0; q < cnt; q++) {
            int j = ids[q];
            if (j == i) continue;
            v3 d = vsub(s->pos[j], pi);
            double r = vlen(d);
            double u = r / h;
            m1 += u; m2 += u * u; m3 += u * u * u;
            c = vadd(c, vmul(d, 1.0 / h));
        }
        double nn = cnt > 1 ? (double)(cnt - 1) : 1.0;
        double *f = &out[(size_t)i * BBX_NFEAT];
        double wall = 1e30;
        for (int k = 0; k < 3; k++) {
            double pk = k == 0 ? pi.x : (k == 1 ? pi.y : pi.z);
            double ck = k == 0 ? s->tank_c.x : (k == 1 ? s->tank_c.y : s->tank_c.z);
            double hk = k == 0 ? s->tank_h.x : (k == 1 ? s->tank_h.y : s->tank_h.z);
            double dlo = (pk - (ck - hk)), dhi = ((ck + hk) - pk);
            if (dlo < wall) wall = dlo;
            if (dhi < wall) wall = dhi;
        }
        f[BBX_F_DENSITY_ERR] = s->density[i] / rho0 - 1.0;
        f[BBX_F_RESERVED1] = 0.0;
        f[BBX_F_COUNT] = (double)cnt;
        f[BBX_F_M1] = m1 / nn;
        f[BBX_F_M2] = m2 / nn;
        f[BBX_F_M3] = m3 / nn;
        f[BBX_F_CX] = c.x / nn;
        f[BBX_F_CY] = c.y / nn;
        f[BBX_F_CZ] = c.z / nn;
        f[BBX_F_CNORM] = vlen(c) / nn;
        f[BBX_F_PPREV] = pref != 0 ? s->pressure[i] / pref : 0;
        f[BBX_F_VY] = dt * vi.y / h;
        f[BBX_F_VNORM] = dt * vlen(vi) / h;
        f[BBX_F_RESERVED2] = 0.0;
        f[BBX_F_RESERVED3] = 0.0;
        f[BBX_F_WALL] = fminr(fmaxr(wall / h, 0.0), 4.0);
    }
}
