// bbx_common.h -- internal declarations shared by the HIP translation units
// of the MI355X SPH / PCISPH hot path (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "bubbles_amd.h"
#include "bbx_features.h"

namespace bbx {

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define BBX_HIP_CHECK(expr)                                                         \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess)                                                       \
            return ::bbx::fail(BBX_E_HIP, std::string(#expr) + ": " +               \
                                              hipGetErrorString(e_));               \
    } while (0)

#define BBX_CHECK(expr)                                                             \
    do {                                                                            \
        int r_ = (expr);                                                            \
        if (r_ != BBX_OK) return r_;                                                \
    } while (0)

// ---------------------------------------------------------------------------
// constants shared by host and device
// ---------------------------------------------------------------------------
constexpr int kBucket = BBX_BUCKET_SIZE;   // reference bucket cap
constexpr int kBlock = 256;                // threads per workgroup (4 wave64)
constexpr int kScanItems = 8;              // items per thread in the cell scan
constexpr int kTabIntsHost = 48;           // ints per tile table (see bbx_tile.h)
// LDS tile capacity in particles (bbx_tile.h): 2528 float4 (+ the tile
// table) lets four single-array or two two-array workgroups share a CU's
// 160 KB of LDS
constexpr int kTileCapHost = 2528;
constexpr int kTileCap2Host = 2528;        // the same for two-array kernels

// Neighbour tables are quad-major: bucket entries 4q..4q+3 of particle i are
// adjacent at [(q * stride + i) * 4 + (k & 3)], so one 8-byte (u16) or
// 16-byte (int) load per lane fetches four entries and a wave's load of
// quad q is one contiguous 512 B / 1 KB run.
__host__ __device__ inline size_t nbr_index(int k, int i, int stride) {
    return ((size_t)(k >> 2) * (size_t)stride + (size_t)i) * 4 + (size_t)(k & 3);
}

// Grid geometry in FP64 (the hash is computed in the reference's precision).
struct GridGeom {
    int n[3];
    unsigned total;
    double len[3];
    double minp[3];
    double bmin[3];
    double bmax[3];
};

// Per-solver scalars passed to kernels by value.
// One collider on the device, FP32 (built on the host from bbx_collider).
struct DevCollider {
    int kind, reverse, active, pad;
    float friction;
    float m[12], mi[12];     // ObjectToWorld / WorldToObject, rows 0..2
    float hx, hy, hz;        // box half sizes
    float radius;            // sphere
    float lin[3], ang[3];    // sphere velocity (VelocityAt)
    float bmin[3], bmax[3];  // mesh bounds
    const float *field;      // SDF nodes, x fastest
    int nx, ny, nz, pad2;
    float ox, oy, oz;        // SDF origin
    float sx, sy, sz;        // SDF spacing
};
constexpr int kMaxColliders = 64;

struct SimConst {
    int n;             // particles
    int stride;        // neighbour-list row stride (>= n, multiple of 64)
    GridGeom g;
    // kernel radius and derived FP32 coefficients
    float h, h2, inv_h, inv_h2;
    float poly6;       // 315 / (64 pi h^3)
    float spiky_grad;  // 45 / (pi h^4)
    float spiky_lap;   // 90 / (pi h^5)
    double h2d;        // FP64 kernel radius squared for the bucket rule
    double rho0d;      // FP64 rest density (convergence ratio)
    float lo2, hi2;    // FP32 prefilter band around the FP64 cutoff
    float band_c, band_w;  // the same band as |d2 - band_c| <= band_w (slightly wider)
    float mass, mass2, rho0, visc, drag, nps;
    float gx, gy, gz;
    float radius;      // collision radius
    float tank_c[3], tank_h[3];
    float bmin[3], bmax[3];
    float min_len;     // min cell length (0.9-cell flag)
    float eos_scale, eos_exp;
    float pseudo_visc;
    int tile_cap;      // tiles with more staged candidates read global memory
    int tile_cap2;     // the same for kernels that stage two arrays
    int row_max;       // bucket build: rows up to this many candidates use the mask path (<= 64)
    float inv_nx, inv_nxy;  // 1 / n[0], 1 / (n[0] n[1]) (cell id decode)
    int ncoll;         // 0: the tank (reversed box above); else the collider list
    const DevCollider *coll;
};

// Per-call step constants.
struct StepConst {
    float dt, dt_over_m;
    float delta;       // PCISPH stiffness (FP64 on host, rounded for the update)
    int first;         // k == 0 without warm start: pressure starts at 0
    float warm_scale;  // delta * rho0: warm-start network output -> pressure
};

// predict-correct loop state on the device: per-iteration FP64 max error
// bits and done flags (iteration k's pressure-force kernel integrates and
// sets done[k] once the error of iteration k is below the tolerance)
// Global maxima are reduced per workgroup and then with atomics into one of
// kSlots slots (workgroup index mod kSlots), so no single address takes an
// atomic from every workgroup; readers take the max over the slots.
constexpr int kMaxIt = 256;
constexpr int kSlots = 32;
static_assert(kMaxIt <= 256, "the loop signal carries the iteration index in 8 bits");
struct LoopState {
    unsigned long long err[kMaxIt];              // max over the slots (host readback)
    int done[kMaxIt];
    unsigned long long slot[kMaxIt][kSlots];     // per-slot FP64 bits of max |err|
};

// The pinned block the predict-correct loop publishes into (mapped: the
// kernels store into it by its device address).  `word` takes every
// decision, (seq << 9) | (k << 1) | integrate, seq numbering the substep.
// The integrating decision first stores the substep's statistics into
// `stats`, each word (seq << 32) | value: a word of another substep carries
// another seq, so the host needs no ordering between the stores to tell
// that all of them are current.
enum { PUB_ERR_LO = 0, PUB_ERR_HI, PUB_OVERFLOW, PUB_LOST, PUB_FULL, kPubWords };
struct LoopSignal {
    unsigned word;
    unsigned pad;
    unsigned long long stats[kPubWords];
};

// device-side counters
enum {
    CNT_OVERFLOW = 0,   // bucket inserts dropped by the cap
    CNT_LOST = 1,       // particles in no chain
    CNT_FULL = 2,       // last distribution was a full pass
    CNT_BAD_HASH = 3,   // points hashing outside the grid
    CNT_N = 8
};

// The per-substep device state in one allocation: cleared by one fill at the
// start of a substep and read back by one copy (the first kStepReadback
// bytes: counters, max-force slots, the loop's per-iteration errors and done
// flags); the per-slot error maxima follow.
struct StepBlock {
    int counters[CNT_N];
    unsigned max_force[kSlots];
    LoopState loop;
};
constexpr size_t kStepReadback = offsetof(StepBlock, loop) + offsetof(LoopState, slot);
static_assert(sizeof(StepBlock) % sizeof(int) == 0, "StepBlock is cleared as ints");

// ---------------------------------------------------------------------------
// kernel profiler (HIP events on the solver stream)
// ---------------------------------------------------------------------------
constexpr int kNoIter = -1000;  // launch outside the predict-correct loop

struct Profiler {
    bool enabled = false;
    std::string only;        // time only this kernel (empty: every kernel)
    int only_iter = kNoIter; // and only its launch of this iteration (kNoIter: all)
    struct Pending { int id, iter; hipEvent_t a, b; };
    std::vector<std::string> names;
    std::vector<double> total_ms;
    std::vector<int> launches;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    int name_id(const std::string &name);
    bool wants(const char *name, int iter) const;
    void begin(hipStream_t s, hipEvent_t *a);
    void end(hipStream_t s, int id, int iter, hipEvent_t a);
    // iterations >= 0 names the loop kernels by role: predict_pressure or
    // predict_pressure_skipped; pressure_force_{warm,predict,integrate,skipped}.
    // Only the first `upto` pending records are resolved (those of one
    // substep when the next one's front is already queued).
    int resolve(int iterations = -1, size_t upto = SIZE_MAX);
    void reset();
    void destroy();
};

}  // namespace bbx

// Bracket one kernel launch with profiler events (needs `s`, a bbx_solver *).
#define BBX_PROF_BEGIN_IT(name, it)                                          \
    hipEvent_t ev_a_ = nullptr;                                              \
    int ev_id_ = -1;                                                         \
    const int ev_it_ = (it);                                                 \
    const bool ev_on_ = s->prof.enabled && s->prof.wants(name, ev_it_);      \
    if (ev_on_) {                                                            \
        ev_id_ = s->prof.name_id(name);                                      \
        s->prof.begin(s->stream, &ev_a_);                                    \
    }
#define BBX_PROF_BEGIN(name) BBX_PROF_BEGIN_IT(name, ::bbx::kNoIter)
#define BBX_PROF_END() \
    if (ev_on_) s->prof.end(s->stream, ev_id_, ev_it_, ev_a_);

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------
struct bbx_mlp;

struct bbx_solver {
    int kind = BBX_PCISPH;
    bbx_params prm{};
    int n = 0;
    int stride = 0;
    double mass = 1e-3, kr = 0, radius = 1e-3, delta_denom = 0, mass_over_rho2 = 0;
    double delta64 = 0;   // PCISPH delta of the current substep (FP64)
    // h_max_force holds the max-force slots of the last completed substep,
    // read back with its statistics, and no substep has been queued since
    bool maxf_cached = false;
    int last_iterations = 0;  // iterations of the previous PCISPH substep
    int prev_iterations = 0;  // and of the one before it
    int last_warm = 0;        // whether it started from the warm-start network
    int warm_ones = 0;        // warm substeps in a row that converged in one correction
    int warm_active = 0;      // a substep has needed BBX_WARM_AUTO_ACTIVE or more corrections
    bbx::GridGeom grid{};
    bbx::SimConst sc{};
    hipStream_t stream = nullptr;
    int device = 0;
    bool setup_done = false;

    // storage-order particle state.  A = sorted by cell (current),
    // B = same order after integration (input of the next redistribution).
    float4 *posA = nullptr, *posB = nullptr;   // xyz, w = particle id bits
    float4 *velA = nullptr, *velB = nullptr;   // xyz, w = last converged pressure
    int *cellA = nullptr, *cellB = nullptr;    // chain membership (total = none)
    int *home = nullptr;                       // hashed cell (neighbour lookups)
    // cell table (total + 1 entries, the last one collects particles in no chain)
    int *cell_count = nullptr, *cell_start = nullptr;
    int *scan_tmp = nullptr;
    size_t scan_tmp_elems = 0;
    // sort scratch (indexed by old slot / new position)
    unsigned long long *key = nullptr;
    int *newc = nullptr, *rank_tmp = nullptr, *tmp_idx = nullptr, *hash_tmp = nullptr;
    // neighbour lists (bucket slot k of particle i at nbr_index(k, i, stride),
    // k < 100): 16-bit LDS-local indices for staged tiles, 32-bit storage
    // indices otherwise
    unsigned short *nbr16 = nullptr;
    int *nbr32 = nullptr, *nbr_cnt = nullptr;
    int *tile_tab = nullptr;   // per-tile interval tables (kTabInts ints each)
    int ntiles = 0;
    // per-particle fields
    float *rho = nullptr, *pres = nullptr, *rho_star = nullptr;
    float4 *xq = nullptr;        // position xyz and q = p / rho*^2 (the pressure-force gather input)
    float4 *force = nullptr, *fp = nullptr, *xstar = nullptr;
    float *feat = nullptr;       // [16][stride]
    float *pwarm = nullptr;      // MLP output
    int keep_features = 0;       // warm substeps also store their features (bbx_solver_keep_features)
    // reductions / flags: loop, max_force and counters point into one
    // StepBlock allocation, h_loop, h_max_force and h_counters into its
    // pinned copy
    bbx::StepBlock *step = nullptr, *h_step = nullptr;
    bbx::LoopState *loop = nullptr;          // predict-correct loop state (device)
    bbx::LoopState *h_loop = nullptr;
    unsigned *max_force = nullptr;           // kSlots FP32 bits of max |f|
    int *counters = nullptr;                 // CNT_*
    int *h_counters = nullptr;
    unsigned *h_max_force = nullptr;
    int *flag = nullptr;                     // requiresHigherLevelUpdate
    int *h_flag = nullptr;                   // pinned: one int read back alone
    // Loop decisions (and, with the integrating one, the substep's
    // statistics) published to the host as they are taken by the pressure-
    // force kernel (bbx::LoopSignal; d_signal is its device address).
    bbx::LoopSignal *h_signal = nullptr, *d_signal = nullptr;
    unsigned loop_seq = 0;
    // substep timing: start and end of the queued work, and the end of the
    // statistics readback; two sets, so that a run of PCISPH substeps can
    // queue the next substep's grid work before it reads this one's
    hipEvent_t ev0[2] = {}, ev1[2] = {}, ev_read[2] = {};
    // a run of PCISPH substeps marks each boundary between substeps once (the
    // end of one is the start of the next) in a ring of kMarks events, read
    // kMarks - 1 substeps later, so the host never waits for a substep's end
    static constexpr int kMarks = 8;
    hipEvent_t mark[kMarks] = {};


    bbx_mlp *mlp = nullptr;
    bbx::Profiler prof;
    // collider list (bbx_solver_set_colliders); none: the params' tank
    bbx::DevCollider *coll = nullptr;
    int ncoll = 0;
};

// ---------------------------------------------------------------------------
// launchers (defined in the .hip files, called by the ABI layer)
// ---------------------------------------------------------------------------
namespace bbx {
int build_grid_geom(const bbx_params &p, GridGeom *g);
int compute_mass_and_delta(bbx_solver *s);
void fill_sim_const(bbx_solver *s);

int launch_fill_int(bbx_solver *s, int *p, int value, size_t n);
int launch_scan(bbx_solver *s, const int *in, int *out, size_t n);
// append_from >= 0: the append pass after new particles were placed in
// slots append_from..n-1 (old chains kept, new particles at their tails)
// feat_mode (with build_buckets): 0 none; 1 the bucket kernel also computes
// the warm-start features (stored in feat); 2 it also infers the attached
// network and starts the substep from its pressures (p0 into pres and pwarm,
// (x, p0 / rho^2) into xq, rho* = rho); st carries dt, delta and the scale
int distribute(bbx_solver *s, bool force_full, bool buckets, bool with_pressure, int append_from = -1,
               int feat_mode = 0, const StepConst *st = nullptr);
// the two halves of distribute: the grid (hash, scan, scatter, move) and the
// tile tables + bucket build (the latter is where the warm start is chosen)
// clear2 / clear2_n: a second int range set to zero by the same launch that
// clears the cell counts (the substep's step block)
int distribute_grid(bbx_solver *s, bool force_full, int append_from, int *clear2 = nullptr,
                    size_t clear2_n = 0);
int build_buckets(bbx_solver *s, bool with_pressure, int feat_mode, const StepConst *st);
int run_density_only(bbx_solver *s);
// mode: 0 force only, 1 force + predict(k = 0), 2 force + the warm start's
// pressure force + predict(k = 0) (the bucket kernel inferred p0)
int run_nonpressure(bbx_solver *s, const StepConst &st, int mode);
int run_predict_pressure(bbx_solver *s, const StepConst &st, int k);
// k < 0: predict only (warm start); otherwise integrate when iteration k is
// converged (or k == max_it - 1, or exact), else predict iteration k + 1
int run_pressure_force(bbx_solver *s, const StepConst &st, int k, int max_it, int exact, double tol);
int run_sph_force(bbx_solver *s);
int run_integrate(bbx_solver *s, const StepConst &st);
int run_xsph(bbx_solver *s, const StepConst &st);
int swap_after_integrate(bbx_solver *s);
const float *mlp_params_device(const bbx_mlp *m);
void mlp_retain(bbx_mlp *m);
void mlp_release(bbx_mlp *m);
}  // namespace bbx
