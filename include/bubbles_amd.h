/*
 * bubbles_amd.h -- C ABI of the MI355X (gfx950) SPH / PCISPH hot path.
 *
 * The reference exposes this path only as a C++ class API (PciSphSolver3,
 * SphSolver3, Grid3, ParticleSet3, Bucket; src/core/pcisph_solver.h:56-83,
 * src/core/sph_solver.h:76-96, src/core/grid.h:197-600,
 * src/core/particle.h:154-401).  Each entry point below is the flat form of
 * one of those calls; the citation names the reference interface it
 * replaces.  The C++ facade in include/bubbles/ (same class names) and any
 * FFI (ctypes, cgo, JNI) bind to these symbols.
 *
 * Conventions
 *   - every function returns BBX_OK (0) or a negative BBX_E* status; the
 *     message of the last failure on the calling thread is bbx_last_error();
 *   - host arrays are plain pointers in the reference's precision (double)
 *     and are indexed by particle id (the order particles were uploaded in),
 *     never by device storage order;
 *   - one solver owns one HIP stream and all of its device memory; calls on
 *     one handle are not thread-safe; nothing calls exit().
 */
#ifndef BUBBLES_AMD_H
#define BUBBLES_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define BBX_OK 0
#define BBX_E_INVALID (-1)   /* bad argument or call order            */
#define BBX_E_HIP (-2)       /* HIP runtime error                      */
#define BBX_E_NOMEM (-3)     /* device allocation failed               */
#define BBX_E_RANGE (-4)     /* size beyond a documented limit         */
#define BBX_E_NODEVICE (-5)  /* no gfx950 device visible               */

#define BBX_BUCKET_SIZE 100  /* MaximumParticlesPerBucket, src/core/particle.h:10 */

/* solver kinds */
#define BBX_PCISPH 0
#define BBX_SPH 1

/* PCISPH loop modes (see ComputePressureForceAndIntegrate,
 * src/equations/pcisph_equations3.cpp:218-256).  REFERENCE_EXACT runs one
 * predict-correct iteration per substep, which is what the shipped loop does
 * because its error array is never written; TEXTBOOK iterates until the max
 * compression error / rho0 < tol or max_it (1..256) is reached. */
#define BBX_MODE_REFERENCE_EXACT 0
#define BBX_MODE_TEXTBOOK 1

/* pressure warm start: NONE starts every substep from p = 0, MLP from the
 * attached network's prediction.  AUTO uses the network only where it can
 * pay for itself (it costs about two thirds of a correction): while the
 * flow is calm it switches the network on after a plain substep that needed
 * BBX_WARM_AUTO_MIN_ITERATIONS or more corrections and off again after
 * BBX_WARM_AUTO_PROBE warm substeps in a row converged in a single
 * correction; once any substep has needed BBX_WARM_AUTO_ACTIVE or more
 * corrections the network stays on for the rest of the solver's life.  The
 * correction loop and its tolerance are the same in every mode. */
#define BBX_WARM_NONE 0
#define BBX_WARM_MLP 1
#define BBX_WARM_AUTO 2
#define BBX_WARM_AUTO_MIN_ITERATIONS 3
#define BBX_WARM_AUTO_PROBE 8
#define BBX_WARM_AUTO_ACTIVE 4

/* downloadable / uploadable per-particle fields (by particle id) */
#define BBX_FIELD_POSITION 0          /* 3 doubles */
#define BBX_FIELD_VELOCITY 1          /* 3 doubles */
#define BBX_FIELD_FORCE 2             /* 3 doubles: last total force */
#define BBX_FIELD_DENSITY 3           /* 1 double */
#define BBX_FIELD_PRESSURE 4          /* 1 double */
#define BBX_FIELD_DENSITY_PREDICTED 5 /* 1 double */
#define BBX_FIELD_PRESSURE_FORCE 6    /* 3 doubles */
#define BBX_FIELD_TEMP_POSITION 7     /* 3 doubles: predicted x* */
#define BBX_FIELD_FEATURES 8          /* 16 doubles: warm-start features */
#define BBX_FIELD_WARM_PRESSURE 9     /* 1 double: MLP initial pressure of the last warm substep */

typedef struct bbx_solver bbx_solver;
typedef struct bbx_mlp bbx_mlp;

/* Setup inputs + SphSolverData3 fields (src/core/sph_solver.h:31-50,
 * DefaultSphSolverData3 src/solvers/sph_solver3.cpp:124-151). */
typedef struct bbx_params {
    double target_density;          /* Setup(targetDensity, ...)            */
    double target_spacing;          /* Setup(..., targetSpacing, ...)       */
    double relative_radius;         /* Setup(..., relativeRadius, ...)      */
    double tank_center[3];          /* reversed MakeBox container centre    */
    double tank_size[3];            /* container size                       */
    double grid_spacing;            /* UtilBuildGridForDomain(b, spacing, s) */
    double grid_scale;
    double eos_exponent;            /* 7                                    */
    double negative_pressure_scale; /* 0                                    */
    double viscosity;               /* 0.04                                 */
    double pseudo_viscosity;        /* 10                                   */
    double sound_speed;             /* 100                                  */
    double drag_coefficient;        /* 1e-4                                 */
    double gravity[3];              /* (0, -9.8f, 0)                        */
} bbx_params;

typedef struct bbx_grid_desc {      /* Grid3 after Build (src/core/grid.h:626-666) */
    unsigned usizes[3];
    unsigned total;
    double cells_len[3];
    double min_point[3];
    double bounds_min[3];
    double bounds_max[3];
} bbx_grid_desc;

typedef struct bbx_info {
    int count;
    double mass;                    /* ParticleSet3::GetMass                */
    double kernel_radius;           /* SphParticleSet3::GetKernelRadius     */
    double radius;                  /* ParticleSet3::GetRadius              */
    double delta_denom;             /* PciSphSolver3::deltaDenom            */
    bbx_grid_desc grid;
} bbx_info;

typedef struct bbx_stats {
    int iterations;                 /* predict-correct iterations this substep   */
    int full_redistribution;        /* 1 if the 0.9-cell flag forced a full pass */
    int bucket_overflows;           /* neighbours dropped by the 100 cap         */
    int lost_particles;             /* particles outside every cell chain        */
    double max_density_error;       /* max compression error / rho0 (FP64)       */
    double gpu_ms;                  /* device time of the substep                */
    int warm_started;               /* 1 if the warm-start network started it    */
    int pad;
} bbx_stats;

/* ---- library ---- */
const char *bbx_last_error(void);
int bbx_version(int *major, int *minor);
int bbx_device_count(int *count);

/* ---- defaults ----
 * DefaultSphSolverData3(true) (src/solvers/sph_solver3.cpp:124-151) */
int bbx_default_params(bbx_params *out);

/* ---- solver lifetime ----
 * replaces PciSphSolver3::Initialize / SphSolver3::Initialize plus
 * SphParticleSet3FromBuilder(ParticleSetBuilder3::AddParticle...) and
 * UtilBuildGridForDomain (src/solvers/pcisph_solver3.cpp:9-14,
 * src/core/particle.cpp:21-27, src/core/util.cpp:269-287).
 * pos/vel: n*3 doubles (vel may be NULL = zero). */
int bbx_solver_create(int kind, const bbx_params *params, const double *pos,
                      const double *vel, int n, bbx_solver **out);
int bbx_solver_destroy(bbx_solver *s);

/* PciSphSolver3::Setup / SphSolver3::Setup (src/solvers/pcisph_solver3.cpp:103-146,
 * src/solvers/sph_solver3.cpp:96-122): mass, delta denominator, full
 * distribution into ascending-id chains. */
int bbx_solver_setup(bbx_solver *s);
int bbx_solver_info(bbx_solver *s, bbx_info *out);

/* ParticleSet3::Get/SetParticle* (src/core/particle.h:316-386), by id. */
int bbx_solver_download(bbx_solver *s, int field, double *out);
int bbx_solver_upload(bbx_solver *s, int field, const double *in);

/* SphParticleSet3::requiresHigherLevelUpdate (src/core/particle.h:525-527) */
int bbx_solver_get_flag(bbx_solver *s, int *flag);
int bbx_solver_set_flag(bbx_solver *s, int flag);

/* ---- grid and neighbour query ----
 * UpdateGridDistributionGPU (src/equations/sph_equations3.cpp:511-539):
 * full or incremental chain update (flag-dependent), then every particle's
 * bucket; then ResetHigherLevel. */
int bbx_grid_distribute(bbx_solver *s, bbx_stats *stats);
/* Cell::GetChain walk for every cell (src/core/grid.h:56-69): CSR in chain
 * order; cell_start has total+1 entries, members up to count entries. */
int bbx_grid_get_chains(bbx_solver *s, int *cell_start, int *members);
/* Replace the chain state (cell membership and order) with the given CSR. */
int bbx_grid_set_chains(bbx_solver *s, const int *cell_start, const int *members);
/* Grid3::GetLinearHashedPosition (src/core/grid.h:307-321) for n points;
 * returns -1 for points outside the grid. */
int bbx_grid_hash(bbx_solver *s, const double *points, int n, int *cell_ids);
/* Grid3::GetNeighborsOf (src/core/grid.h:502-510): <=27 ids, (y,x,z) order */
int bbx_grid_cell_neighbors(bbx_solver *s, int cell, int *out27, int *count);
/* UtilBuildGridForDomain + Grid3::Build without a solver (src/core/util.cpp:269-287,
 * src/core/grid.h:626-666): the grid the params describe. */
int bbx_grid_describe(const bbx_params *params, bbx_grid_desc *out);
/* Grid3::GetHashedPosition / GetLinearHashedPosition on a described grid
 * (src/core/grid.h:258-321); u may be NULL (else 3 ints per point); ids are
 * -1 for points outside the grid. */
int bbx_grid_hash_desc(const bbx_grid_desc *grid, const double *points, int n, int *cell_ids,
                       int *u);
/* ParticleSet3::GetParticleBucket for all particles (src/core/particle.h:276-279,
 * Bucket::Get :52-59): counts[n], ids[n*100] (row i = particle i). */
int bbx_neighbors_get(bbx_solver *s, int *counts, int *ids);

/* ---- stages (one reference wrapper each; used by parity tests) ---- */
int bbx_stage_density(bbx_solver *s);          /* ComputeParticleInteraction + ComputeDensityGPU (src/equations/sph_equations3.cpp:541-553,585-597) */
int bbx_stage_nonpressure(bbx_solver *s);      /* ComputeNonPressureForceGPU (:608-620) */

/* ---- solver steps ----
 * One PCISPH substep of fixed dt: AdvanceTimeStep (src/solvers/pcisph_solver3.cpp:42-65). */
int bbx_pcisph_substep(bbx_solver *s, double dt, int max_it, double tol, int mode, int warm,
                       bbx_stats *stats);
/* `count` consecutive substeps of fixed dt, as `count` calls of
 * bbx_pcisph_substep (the reference's frame loop calls AdvanceTimeStep the
 * same way, src/solvers/pcisph_solver3.cpp:67-89) without returning to the
 * caller in between; stats (may be NULL) receives one entry per substep,
 * gpu_ms being the device time between the substep's boundaries with its
 * neighbours (the substeps follow one another on the device). */
int bbx_pcisph_run(bbx_solver *s, double dt, int count, int max_it, double tol, int mode, int warm,
                   bbx_stats *stats);
/* PciSphSolver3::Advance (src/solvers/pcisph_solver3.cpp:67-89): reference
 * substep rule with scale 5. */
int bbx_pcisph_advance(bbx_solver *s, double interval, int max_it, double tol, int mode,
                       int warm, int *substeps);
/* SphSolver3 AdvanceTimeStep / Advance (src/solvers/sph_solver3.cpp:48-93),
 * race-free two-phase force/integration. */
int bbx_sph_substep(bbx_solver *s, double dt, bbx_stats *stats);
int bbx_sph_advance(bbx_solver *s, double interval, int *substeps);
/* PciSphSolver3::ComputeDelta (src/solvers/pcisph_solver3.cpp:182-185) */
int bbx_pcisph_delta(bbx_solver *s, double dt, double *delta);
/* SphParticleSet3::ComputeNumberOfTimeSteps (src/core/particle.h:584-608) */
int bbx_number_of_time_steps(bbx_solver *s, double interval, double scale, unsigned *n);

/* ---- learned pressure warm start (no reference counterpart) ----
 * 16 -> 64 -> 64 -> 1 MLP, GELU, BF16 weights, FP32 accumulate, on MFMA. */
int bbx_mlp_create(int seed, bbx_mlp **out);
/* Destroying a network that is attached to a solver is deferred: it is freed
 * when the last solver it is attached to detaches it or is destroyed. */
int bbx_mlp_destroy(bbx_mlp *m);
/* params layout: W1[64][16] b1[64] W2[64][64] b2[64] W3[64] b3[1] in FP32
 * (master copy, trained), then feature mean[16], feature inv_std[16],
 * out_scale[1], alpha[1], pcap[1], resid[1] (not trained).  The prediction in
 * target units is net * out_scale + resid * feature[BBX_F_PPREV]; the solver
 * starts from p = delta*rho0 * alpha * clamp(prediction, 0, pcap). */
int bbx_mlp_param_count(int *count);
int bbx_mlp_get_params(bbx_mlp *m, float *out);
int bbx_mlp_set_params(bbx_mlp *m, const float *in);
/* y[n] = MLP(x[n*16]) on the device (host arrays). */
int bbx_mlp_infer(bbx_mlp *m, const float *x, int n, float *y);
/* Training loss: the expectile loss 0.5*mean(w (y - t)^2) with w = 2(1 - tau)
 * where y > t and w = 2 tau where y < t (tau = 0.5, the default, is the mean
 * squared error).  tau < 0.5 makes over-estimates costlier: the network then
 * learns a low expectile of the target.  tau must lie in (0, 1). */
int bbx_mlp_set_expectile(bbx_mlp *m, float tau);
/* gradients of the training loss w.r.t. the FP32 params (host arrays) */
int bbx_mlp_gradients(bbx_mlp *m, const float *x, const float *t, int n, float *grad,
                      float *loss);
/* one Adam step on a batch (host arrays); returns the batch loss */
int bbx_mlp_train_step(bbx_mlp *m, const float *x, const float *t, int n, float lr,
                       float *loss);
/* train for `steps` Adam steps on device-resident samples drawn from the
 * given host arrays (copied once); batch sampled by a counter-based RNG */
int bbx_mlp_train(bbx_mlp *m, const float *x, const float *t, int n, int steps, int batch,
                  float lr, int seed, float *final_loss);
/* Attach (or, with m == NULL, detach) the warm-start network; the solver
 * holds a reference until it detaches the network or is destroyed. */
int bbx_solver_attach_mlp(bbx_solver *s, bbx_mlp *m);
/* Keep the features of every warm-started substep for BBX_FIELD_FEATURES
 * downloads (default off: the fused bucket kernel then stores none, which
 * saves 64 bytes per particle of writes).  No counterpart in the reference. */
int bbx_solver_keep_features(bbx_solver *s, int keep);
/* Harvest (features, converged pressure / P) pairs of one TEXTBOOK substep:
 * runs the substep and writes n*16 features and n targets (by particle id). */
int bbx_pcisph_harvest(bbx_solver *s, double dt, int max_it, double tol, float *features,
                       float *targets, bbx_stats *stats);

/* ---- growing particle sets ----
 * ContinuousParticleSetBuilder3 (src/core/grid.h:1246-1450).  Capacity is
 * the number of particles the device buffers hold (a multiple of 64).
 * reserve keeps the state of the current particles; append grows the
 * capacity itself when needed. */
int bbx_solver_reserve(bbx_solver *s, int capacity);
int bbx_solver_capacity(bbx_solver *s, int *capacity);
/* Commit + DistributeByParticleList (src/core/grid.h:1410-1425,358-387):
 * the `count` particles (pos 3*count doubles, vel NULL = zero) get ids
 * n .. n+count-1 and join the tail of the chain of the cell they hash into,
 * in the order given, after the chain's current members; a particle outside
 * the grid is in no chain.  Buckets and densities are rebuilt by the next
 * substep.  Needs bbx_solver_setup first. */
int bbx_solver_append(bbx_solver *s, const double *pos, const double *vel, int count);

/* ---- colliders ----
 * ColliderSet3 / Collider3 / Shape (src/core/collider.cpp:3-44,109-264,
 * src/shapes/box.cpp:74-172, src/shapes/sphere.cpp:47-70,
 * src/shapes/bvh.cpp:212-242,666-706, src/core/shape.cpp:206-305,356-488,
 * src/core/grid.h:1093-1153).  A solver without a collider list collides
 * with the reversed box of its params (the tank); a list replaces it and is
 * resolved per particle as ColliderSet3::ResolveCollision does: the active
 * collider with the smallest |ClosestDistance| handles the particle. */
#define BBX_SHAPE_BOX 0     /* MakeBox: size = edge lengths */
#define BBX_SHAPE_SPHERE 1  /* MakeSphere: radius */
#define BBX_SHAPE_MESH 2    /* MakeMesh: collides through its SDF, bounds early-out */
#define BBX_SHAPE_SDF 3     /* a shape given only by an SDF grid */

typedef struct bbx_sdf bbx_sdf;  /* FieldGrid3f, vertex centred, on the device */

typedef struct bbx_collider {
    int kind;                    /* BBX_SHAPE_* */
    int reverse_orientation;     /* Shape::reverseOrientation */
    int active;                  /* Collider3::isActive */
    int pad;
    double friction;             /* Collider3::frictionCoefficient */
    double to_world[12];         /* ObjectToWorld, rows 0..2 of the 4x4 matrix */
    double to_object[12];        /* WorldToObject (its inverse) */
    double size[3];              /* box edge lengths */
    double radius;               /* sphere radius */
    double linear_velocity[3];   /* Shape::SetVelocities (sphere VelocityAt) */
    double angular_velocity[3];
    double bounds_min[3];        /* mesh: bounds of its vertices (OptmizedClosestPointCheck) */
    double bounds_max[3];
    const bbx_sdf *sdf;          /* MESH and SDF kinds */
} bbx_collider;

/* A FieldGrid3f from values: nodes[k] points per axis (x fastest), node
 * (i,j,k) at origin + spacing * (i,j,k). */
int bbx_sdf_create(const int nodes[3], const double origin[3], const double spacing[3],
                   const double *values, bbx_sdf **out);
/* GenerateShapeSDF for a mesh (src/core/shape.cpp:469-488,411-424): grid
 * over the mesh bounds grown by margin * extent, spacing dx adjusted to the
 * x extent; each node holds the distance to the closest triangle
 * (DistanceTriangle, at least 1e-5), negative inside the mesh.  Triangles
 * are 3 vertex indices each.  Built on the device. */
int bbx_sdf_from_mesh(const double *vertices, int nvertices, const int *triangles, int ntriangles,
                      double dx, double margin, bbx_sdf **out);
int bbx_sdf_describe(const bbx_sdf *sdf, int nodes[3], double origin[3], double spacing[3]);
int bbx_sdf_download(const bbx_sdf *sdf, double *values);
/* FieldGrid::Sample and ::Gradient on the device (gradients may be NULL) */
int bbx_sdf_sample(const bbx_sdf *sdf, const double *points, int n, double *values, double *gradients);
void bbx_sdf_destroy(bbx_sdf *sdf);
/* Replace the solver's colliders (count 0: back to the params' tank).  The
 * SDFs must outlive their use by the solver. */
int bbx_solver_set_colliders(bbx_solver *s, const bbx_collider *colliders, int count);
/* ColliderSet3::ResolveCollision(radius, restitution, &p, &v) on the device
 * for n points (in place); collided[i] = 1 where the particle was moved
 * (collided may be NULL). */
int bbx_colliders_resolve(const bbx_collider *colliders, int count, double radius, double restitution,
                          double *pos, double *vel, int n, int *collided);
/* bbx_colliders_resolve without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (tight serving loops, harnesses that forbid hipMalloc).  The staging of
 * the collider table, positions, velocities and flags lives in one static
 * device array per GPU (about 3.4 MB, held until the process ends), so a
 * call takes at most BBX_NOALLOC_MAX_POINTS points (BBX_E_RANGE beyond;
 * split larger sets).  Results are bit-identical to bbx_colliders_resolve.
 * Runs on the current device's default stream and returns when done.
 * Calls that overlap on one GPU (several host threads) share that array
 * and race: serialise them. */
#define BBX_NOALLOC_MAX_POINTS 65536
int bbx_colliders_resolve_noalloc(const bbx_collider *colliders, int count, double radius,
                                  double restitution, double *pos, double *vel, int n, int *collided);

/* ---- frame I/O in the reference's serializer format ----
 * Flag bits as SERIALIZER_* (src/third/serializer.h:7-17). */
#define BBX_SER_POSITION 0x01
#define BBX_SER_VELOCITY 0x02
#define BBX_SER_DENSITY 0x04
#define BBX_SER_BOUNDARY 0x08
#define BBX_SER_NORMAL 0x10
#define BBX_SER_MASS 0x20
#define BBX_SER_LAYERS 0x40
#define BBX_SER_XYZ 0x80
#define BBX_SER_BOUNDARY_EXCLUSIVE 0x100
/* SaveSphParticleSet (src/third/serializer.cpp:875-911): appends one
 * FluidBegin ... FluidEnd block with the requested fields of n particles
 * (host arrays: positions/velocities 3n doubles, densities n, one mass for
 * all, boundary n ints or NULL), numbers printed with %g.  Supported flags:
 * position, velocity, density, mass, boundary, boundary-exclusive (normals,
 * layers and the legacy xyz form are rejected). */
int bbx_serializer_write(const char *path, int flags, double spacing, int n, const double *positions,
                         const double *velocities, const double *densities, double mass,
                         const int *boundary);
/* The same for a solver's particles in id order: position, velocity and
 * density downloaded from the device (FP32 values widened to double), the
 * solver's mass and target spacing. */
int bbx_solver_save(bbx_solver *s, const char *path, int flags);
/* SerializerLoadParticles3 (src/third/serializer.cpp:543-560) over
 * SerializerFindFluidSection + _SerializerLoadParticles3 (:130-190,444-540):
 * reads the first FluidBegin block.  *flags is the caller's default and is
 * replaced by the block's "Format"; up to `capacity` particles are stored
 * in the non-NULL arrays (capacity < 0: count only).  Numbers are parsed as
 * the reference's ParseFloat does (src/third/obj_loader.cpp:55-170,270-277). */
int bbx_serializer_read(const char *path, int capacity, int *count, int *flags, double *positions,
                        double *velocities, double *densities, double *masses, int *boundary);
/* ParseFloat of one token (0 when it is not a number). */
int bbx_serializer_parse_number(const char *text, double *value);

/* ---- profiling ----
 * Average device time (ms) and launch count of every kernel of this solver
 * since the last reset, measured with HIP events on the solver's stream. */
int bbx_profile_enable(bbx_solver *s, int enable);
int bbx_profile_reset(bbx_solver *s);
/* Time only the launches of `kernel` (NULL or "": every kernel) and, when
 * iteration != BBX_PROFILE_ALL_LAUNCHES, only its launch for predict-correct
 * iteration `iteration` (-1: the warm-start launch).  Fewer events keep the
 * timed region closer to an unprofiled run.  Loop kernels are reported by
 * role: predict_pressure[_skipped], pressure_force_{warm,predict,integrate,
 * skipped} (a launch is skipped once an earlier iteration integrated). */
#define BBX_PROFILE_ALL_LAUNCHES (-1000)
int bbx_profile_select(bbx_solver *s, const char *kernel, int iteration);
/* names: '\n'-separated kernel names; ms/launches arrays of up to cap entries */
int bbx_profile_get(bbx_solver *s, char *names, int names_cap, double *total_ms,
                    int *launches, int cap, int *count);
int bbx_solver_synchronize(bbx_solver *s);
/* STREAM-style device copy of `bytes` (read + write), `iters` launches:
 * achievable HBM bandwidth (GB/s, 2*bytes per launch) and the average launch
 * time.  Used to re-measure the bandwidth roof and to calibrate the
 * FETCH_SIZE / WRITE_SIZE counters.  No counterpart in the reference. */
int bbx_stream_copy(double bytes, int iters, double *gbps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
