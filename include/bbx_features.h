/*
 * bbx_features.h -- per-particle input layout of the learned PCISPH pressure
 * warm-start.  The reference has no learned component; this layout is this
 * library's own and is shared by the HIP bucket kernel (which computes the
 * features, and for a warm-started substep infers the network, while it
 * builds the bucket and the density) and the CPU oracle.
 *
 * All features describe the state at the start of a substep: the particle's
 * bucket (reference neighbour set, src/core/grid.h:422-447) and density, its
 * own velocity and position, and the previous substep's converged pressure;
 * h is the kernel radius, rho0 the rest density, dt the substep and
 * P = delta * rho0 the pressure one PCISPH correction assigns to a 100%
 * density error (delta from src/solvers/pcisph_solver3.cpp:182-189).  Sums
 * run over bucket members j != i; n' = max(count - 1, 1).  No feature reads a
 * neighbour's velocity, so all of them are ready before the viscosity
 * gather, and the warm start's pressure force rides on that gather.
 *
 * Three slots are reserved and hold 0: a velocity divergence (as the SPH sum
 * over neighbour velocities it needs a gather the fused kernel avoids; as
 * the continuity estimate from each particle's density change over the last
 * substep it made the warm-started loop need more corrections) and two more
 * velocity components (no gain).
 *
 * The network target is the converged PCISPH pressure of the substep / P.
 */
#ifndef BBX_FEATURES_H
#define BBX_FEATURES_H

#define BBX_NFEAT 16

enum {
    BBX_F_DENSITY_ERR = 0, /* rho_i / rho0 - 1                                  */
    BBX_F_RESERVED1 = 1,   /* 0                                                   */
    BBX_F_COUNT = 2,       /* bucket count (self included)                        */
    BBX_F_M1 = 3,          /* sum (r/h)   / n'                                    */
    BBX_F_M2 = 4,          /* sum (r/h)^2 / n'                                    */
    BBX_F_M3 = 5,          /* sum (r/h)^3 / n'                                    */
    BBX_F_CX = 6,          /* sum (x_j - x_i)/h / n', x component                 */
    BBX_F_CY = 7,          /*   y component                                       */
    BBX_F_CZ = 8,          /*   z component                                       */
    BBX_F_CNORM = 9,       /* |sum (x_j - x_i)/h| / n'                            */
    BBX_F_PPREV = 10,      /* previous substep's converged pressure / P           */
    BBX_F_VY = 11,         /* dt * v_y / h                                        */
    BBX_F_VNORM = 12,      /* dt * |v| / h                                        */
    BBX_F_RESERVED2 = 13,  /* 0                                                   */
    BBX_F_RESERVED3 = 14,  /* 0                                                   */
    BBX_F_WALL = 15        /* distance to the nearest tank wall / h, in [0, 4]    */
};

/* network shape: BBX_NFEAT -> 64 -> 64 -> 1, GELU (tanh form) on hidden layers */
#define BBX_MLP_HIDDEN 64

#endif
