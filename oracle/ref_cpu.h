/*
 * ref_cpu.h -- FP64 CPU restatement of the Bubbles SPH / PCISPH hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the HIP product
 * path is checked against.  Only the test-suite, the smoke check and the CPU
 * baseline leg of the benchmark may load it; nothing in the product links it.
 *
 * It restates, in plain C with IEEE double arithmetic and no FP contraction,
 * the reference's CPU path (SetSystemUseCPU):
 *   grid build / hash / 27-cell lists ... src/core/grid.h:182-193,258-321,554-666
 *   chain distribution (full, incremental) src/core/grid.h:389-500
 *   per-particle buckets (cap 100) ........ src/core/grid.h:422-447, src/core/particle.h:22-60
 *   poly6 / spiky kernels ................. src/core/kernel.cpp:112-234
 *   density, interaction, non-pressure .... src/equations/sph_equations3.cpp:7-110
 *   SPH all-forces + integration .......... src/equations/sph_equations3.cpp:184-339
 *   PCISPH predict-correct loop ........... src/equations/pcisph_equations3.cpp:3-256
 *   mass, delta, time-step rule ........... src/core/particle.h:557-608,
 *                                           src/solvers/pcisph_solver3.cpp:103-189
 *   box collider (reversed tank) .......... src/core/collider.cpp:3-44,238-264,
 *                                           src/shapes/box.cpp:74-172, src/core/shape.cpp:275-282
 *   collider sets (boxes, spheres, SDFs) .. src/core/collider.cpp:109-264, src/shapes/sphere.cpp:47-70,
 *                                           src/shapes/bvh.cpp:212-242,666-706, src/core/grid.h:1093-1153
 * Data structures follow the reference (linked-list cell chains with two node
 * arrays, 100-slot buckets, AoS doubles) so the CPU timing is representative.
 */
#ifndef BBX_ORACLE_REF_CPU_H
#define BBX_ORACLE_REF_CPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define BBO_BUCKET_SIZE 100

/* PCISPH convergence modes.  REFERENCE_EXACT reproduces the shipped loop,
 * whose error array is never written so it always stops after one
 * iteration (src/equations/pcisph_equations3.cpp:17,79-89,238-246).
 * TEXTBOOK stores the (negative-scaled) density error and iterates until
 * max error / rho0 < tol or maxIt is reached. */
enum { BBO_MODE_REFERENCE_EXACT = 0, BBO_MODE_TEXTBOOK = 1 };

/* SPH force/integration semantics.  The reference integrates each particle
 * inside the force loop (src/equations/sph_equations3.cpp:277).  SEQUENTIAL
 * replays that loop on one thread in id order (the reference's
 * single-thread CPU result); TWO_PHASE computes all forces first and then
 * integrates (race-free; what the GPU path does). */
enum { BBO_SPH_TWO_PHASE = 0, BBO_SPH_SEQUENTIAL = 1 };

typedef struct bbo_params {
    /* Setup(targetDensity, targetSpacing, relativeRadius, ...) */
    double target_density;
    double target_spacing;
    double relative_radius;
    /* reversed box container: MakeBox(Translate(center), size, true) */
    double tank_center[3];
    double tank_size[3];
    /* UtilBuildGridForDomain(container bounds, grid_spacing, grid_scale) */
    double grid_spacing;
    double grid_scale;
    /* SphSolverData3 (DefaultSphSolverData3 values unless overridden) */
    double eos_exponent;
    double negative_pressure_scale;
    double viscosity;
    double pseudo_viscosity;
    double sound_speed;
    double drag_coefficient;
    double gravity[3];
    int threads;          /* ParallelFor thread count; <=0 means all cores */
} bbo_params;

typedef struct bbo_stats {
    int iterations;               /* predict-correct iterations run */
    double max_density_error;     /* max(err)/rho0 after the last iteration (textbook measure) */
    int full_redistribution;      /* 1 if this substep used a full distribution */
    int bucket_overflows;         /* particles whose bucket hit the 100 cap */
    int lost_particles;           /* particles in no chain after redistribution */
} bbo_stats;

typedef struct bbo_sim bbo_sim;

/* ---- kernels (src/core/kernel.cpp) ---- */
double bbo_poly6_W(double h, double r);
double bbo_poly6_dW(double h, double r);
double bbo_poly6_d2W(double h, double r);
double bbo_spiky_W(double h, double r);
double bbo_spiky_dW(double h, double r);
double bbo_spiky_d2W(double h, double r);
int bbo_within_std(double distance, double radius);
int bbo_within_spiky(double distance, double radius);

/* ---- grid (src/core/grid.h, src/core/util.cpp:269-287) ---- */
typedef struct bbo_grid_desc {
    unsigned usizes[3];
    unsigned total;
    double cells_len[3];
    double min_point[3];
    double bounds_min[3];
    double bounds_max[3];
} bbo_grid_desc;

void bbo_grid_build(const unsigned res[3], const double p0[3], const double p1[3],
                    bbo_grid_desc *out);
void bbo_grid_for_domain(const double dmin[3], const double dmax[3], double spacing,
                         double spacing_scale, bbo_grid_desc *out);
/* returns linear cell id; u (may be NULL) receives the 3-D index; returns -1
 * (and clamps u) when the point hashes outside the grid */
long long bbo_grid_hash(const bbo_grid_desc *g, const double p[3], int u[3]);
int bbo_grid_neighbors(const bbo_grid_desc *g, unsigned cell, int *out27);

/* ---- BCC helpers (src/generator/bcclattice.cpp, src/core/particle.h:557-582,
 *      src/solvers/pcisph_solver3.cpp:148-180) ---- */
int bbo_bcc_generate(const double pmin[3], const double pmax[3], double spacing,
                     double *points, int maxn);
double bbo_compute_mass(double target_density, double target_spacing, double kernel_radius);
double bbo_delta_denom(double target_spacing, double kernel_radius);

/* ---- simulation ---- */
void bbo_default_params(bbo_params *p);
bbo_sim *bbo_sim_create(const bbo_params *p, int n, const double *pos, const double *vel);
void bbo_sim_destroy(bbo_sim *s);
/* PciSphSolver3::Setup / SphSolver3::Setup: mass, delta denominator, full
 * distribution (ascending id chains), frame_index = 1 */
void bbo_sim_setup(bbo_sim *s);
/* ContinuousParticleSetBuilder::Commit + DistributeByParticleList */
void bbo_sim_append(bbo_sim *s, int count, const double *pos, const double *vel);
void bbo_sim_info(const bbo_sim *s, double *mass, double *kernel_radius, double *radius,
                  double *delta_denom, bbo_grid_desc *grid);
void bbo_sim_set_threads(bbo_sim *s, int threads);

/* state access (arrays indexed by particle id) */
void bbo_sim_get(const bbo_sim *s, const char *field, double *out);
void bbo_sim_set(bbo_sim *s, const char *field, const double *in);
int bbo_sim_get_flag(const bbo_sim *s);
void bbo_sim_set_flag(bbo_sim *s, int flag);
int bbo_sim_count(const bbo_sim *s);

/* chains: CSR view of every cell's active chain in chain order.
 * cell_start has total+1 entries; members has n entries (lost particles are
 * absent, so cell_start[total] may be < n). */
void bbo_sim_get_chains(const bbo_sim *s, int *cell_start, int *members);
void bbo_sim_set_chains(bbo_sim *s, const int *cell_start, const int *members);
/* buckets: counts[n], ids[n*100] (row i holds particle i's bucket) */
void bbo_sim_get_buckets(const bbo_sim *s, int *counts, int *ids);
/* ContinuousParticleSetBuilder::MapGrid and the MapGridEmit refill rule
 * (src/core/grid.h:1286-1408); see ref_cpu.c. */
int bbo_map_grid(const bbo_sim *s, int *cells, int *offsets, double *positions, int *npos);
int bbo_map_grid_emit(const bbo_sim *s, int ncells, const int *cells, const int *offsets,
                      const double *positions, double d, int max_particles, double *out, int out_cap,
                      int *capacity_hit);

/* ---- stages (each one is one reference wrapper) ---- */
void bbo_stage_distribute(bbo_sim *s, bbo_stats *st);      /* UpdateGridDistributionCPU + ResetHigherLevel */
void bbo_stage_interaction(bbo_sim *s);                     /* ComputeParticleInteractionCPU */
void bbo_stage_density(bbo_sim *s, int compute_pressure);   /* ComputeDensityCPU */
void bbo_stage_nonpressure(bbo_sim *s);                     /* ComputeNonPressureForceCPU */
void bbo_stage_predict(bbo_sim *s, double dt, int is_first);
void bbo_stage_predict_pressure(bbo_sim *s, double delta, int record_error);
void bbo_stage_pressure_force(bbo_sim *s);
void bbo_stage_accumulate_integrate(bbo_sim *s, double dt);
void bbo_stage_xsph(bbo_sim *s, double dt);
void bbo_stage_sph_force(bbo_sim *s);                       /* ComputeAllForcesFor, force phase */
double bbo_compute_delta(const bbo_sim *s, double dt);
unsigned bbo_number_of_time_steps(const bbo_sim *s, double interval, double scale);

/* ---- solver steps ---- */
/* one PCISPH substep of fixed dt (AdvanceTimeStep, src/solvers/pcisph_solver3.cpp:42-65).
 * warm_pressure (may be NULL): initial pressure per id; when given, the loop
 * starts from p = warm_pressure and f_p = pressure force of it at the current
 * density instead of p = 0, f_p = 0. */
void bbo_pcisph_substep(bbo_sim *s, double dt, int max_it, double tol, int mode,
                        const double *warm_pressure, bbo_stats *st);
/* PciSphSolver3::Advance: reference substep rule (scale 5) with the given loop
 * settings; returns number of substeps */
int bbo_pcisph_advance(bbo_sim *s, double interval, int max_it, double tol, int mode);
/* SPH (WCSPH) substep and Advance (scale 1) */
void bbo_sph_substep(bbo_sim *s, double dt, int sph_mode, bbo_stats *st);
int bbo_sph_advance(bbo_sim *s, double interval, int sph_mode);

/* ---- colliders: ColliderSet3 / Collider3 / Shape (src/core/collider.cpp:3-44,
 * 109-264; src/shapes/box.cpp:74-172; src/shapes/sphere.cpp:47-70;
 * src/shapes/bvh.cpp:212-242,666-706; src/core/shape.cpp:206-305,411-424,
 * 478-505; src/core/grid.h:1093-1153) ---- */
enum { BBO_SHAPE_BOX = 0, BBO_SHAPE_SPHERE = 1, BBO_SHAPE_MESH = 2, BBO_SHAPE_SDF = 3 };
typedef struct bbo_sdf {          /* FieldGrid3f, vertex centred */
    int nodes[3];
    double origin[3], spacing[3];
    const double *values;         /* x fastest; owned by the caller */
} bbo_sdf;
typedef struct bbo_collider {
    int kind, reverse_orientation, active, pad;
    double friction;
    double to_world[12], to_object[12];   /* rows 0..2 of ObjectToWorld / WorldToObject */
    double size[3];
    double radius;
    double linear_velocity[3], angular_velocity[3];
    double bounds_min[3], bounds_max[3];  /* mesh bounds */
    bbo_sdf sdf;
} bbo_collider;
/* DistanceTriangle (src/shapes/bvh.cpp:212-242) */
double bbo_triangle_distance(const double a[3], const double b[3], const double c[3], const double p[3]);
/* GenerateShapeSDF for a mesh: with values NULL only the grid is described;
 * else values gets every node (distance >= 1e-5, negative inside by the
 * winding number). */
int bbo_sdf_from_mesh(const double *vertices, int nvertices, const int *triangles, int ntriangles, double dx,
                      double margin, int threads, int nodes[3], double origin[3], double spacing[3],
                      double *values);
double bbo_sdf_sample(const bbo_sdf *g, const double p[3]);
void bbo_sdf_gradient(const bbo_sdf *g, const double p[3], double out[3]);
/* Shape::ClosestPoint / ::ClosestDistance / ::IsInside */
void bbo_collider_closest_point(const bbo_collider *c, const double p[3], double point[3], double normal[3],
                                double *signed_distance, double velocity[3]);
double bbo_collider_closest_distance(const bbo_collider *c, const double p[3]);
int bbo_collider_is_inside(const bbo_collider *c, const double p[3]);
/* ColliderSet3::ResolveCollision for n points (in place) */
void bbo_colliders_resolve(const bbo_collider *cs, int count, double radius, double restitution, double *pos,
                           double *vel, int n, int *collided);
/* the simulation's collider list (count 0: the params' tank); the list is
 * copied, SDF values are not */
void bbo_sim_set_colliders(bbo_sim *s, const bbo_collider *cs, int count);

/* learned warm-start features (FP64), 16 per particle, computed from the
 * state after density + non-pressure force of a substep; see features.h */
void bbo_features(bbo_sim *s, double dt, double *out);

#ifdef __cplusplus
}
#endif
#endif
