/*
 * dg_oracle.c -- CPU oracle for deep-green's compressible-Euler step.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference Rust solver
 * (src/rs/src/physics.rs, boundaries.rs, grid.rs, main.rs) used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * The product library never links, loads or calls it.
 *
 * Two instantiations are built from dg_oracle_impl.h: *_f32 restates the
 * reference's f32 arithmetic (same loop ranges, operand order and association,
 * compiled with -ffp-contract=off), *_f64 runs the same expressions in double.
 * The reference's compile-time constants (globals.rs: GAMMA, CFL, G) are
 * runtime parameters here.
 *
 * Pinned by tests/test_oracle_*.py against: golden vectors written by the
 * reference's own Fortran physics module (oracle/_ref, tests/golden/), closed
 * forms (uniform state, single-face Rusanov flux, round trips), and
 * invariants (axis permutation, mirror symmetry, conservation).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define DGO_PERIODIC 0 /* grid.rs:24-28 BoundaryType order */
#define DGO_OUTFLOW 1
#define DGO_WALL 2

#define DGO_LIM_NONE 0 /* reference: unlimited central slopes (physics.rs:198-215) */
#define DGO_LIM_MINMOD 1
#define DGO_LIM_VANLEER 2

typedef struct dgo_params {
    int64_t nx, ny, nz;   /* ghosted sizes: interior + 2*nghosts (grid.rs:64-66) */
    int32_t nghosts;      /* 2 */
    int32_t bcs[6];       /* x-, x+, y-, y+, z-, z+ (init.rs:45-50) */
    double ds;
    double gamma, cfl, g; /* globals.rs:3-7 as parameters */
    int32_t gravity;      /* Simulation::gravity (grid.rs:40) */
    int32_t gravity_k_shift; /* 0: k is 0-based as physics.rs:368; 1: 1-based */
    int32_t dt_nan_fallback; /* 1: min() error -> 1.0 (physics.rs:191-194) */
    int32_t nthreads;        /* OpenMP threads for the passes */
    int32_t ref_threading;   /* 1: xtr/flux/add/min serial, as in physics.rs */
    int32_t limiter;         /* DGO_LIM_*; 0 is the reference scheme */
    int32_t heat_mask;       /* conducting lower walls: bit 0 x-, 1 y-, 2 z- (0: none) */
    double t_out, v_out, rs; /* heat source constants (src/cu/include/constants.h:63-66) */
} dgo_params;

#define R float
#define SFX f32
#define RSQRT sqrtf
#define RFABS fabsf
#define RFMAX fmaxf
#include "dg_oracle_impl.h"
#undef R
#undef SFX
#undef RSQRT
#undef RFABS
#undef RFMAX

#define R double
#define SFX f64
#define RSQRT sqrt
#define RFABS fabs
#define RFMAX fmax
#include "dg_oracle_impl.h"
#undef R
#undef SFX
#undef RSQRT
#undef RFABS
#undef RFMAX

size_t dgo_params_size(void) { return sizeof(dgo_params); }
