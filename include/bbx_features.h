/*
 * bbx_features.h -- per-particle input layout of the learned PCISPH pressure
 * warm-start.  The reference has no learned component; this layout is this
 * library's own and is shared by the HIP feature kernel and the CPU oracle.
 *
 * All features are computed from the state after the density and
 * non-pressure-force passes of a substep, over the particle's bucket
 * (reference neighbour set, src/core/grid.h:422-447), with h the kernel
 * radius, m the particle mass, rho0 the rest density, dt the substep and
 * P = delta * rho0 the pressure one PCISPH correction assigns to a 100%
 * density error (delta from src/solvers/pcisph_solver3.cpp:182-189).
 * Sums run over bucket members j != i; n' = max(count - 1, 1).
 *
 * The network target is the converged PCISPH pressure of the substep / P.
 */
#ifndef BBX_FEATURES_H
#define BBX_FEATURES_H

#define BBX_NFEAT 16

enum {
    BBX_F_DENSITY_ERR = 0, /* rho_i / rho0 - 1                                  */
    BBX_F_DIVV = 1,        /* dt * sum m/rho_j (v_j - v_i) . gradW_spiky(x_i-x_j) */
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
    BBX_F_AY = 13,         /* dt^2 * f_y / (m h)   (non-pressure force)           */
    BBX_F_ANORM = 14,      /* dt^2 * |f| / (m h)                                  */
    BBX_F_WALL = 15        /* distance to the nearest tank wall / h, in [0, 4]    */
};

/* network shape: BBX_NFEAT -> 64 -> 64 -> 1, GELU (tanh form) on hidden layers */
#define BBX_MLP_HIDDEN 64

#endif
