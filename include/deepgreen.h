/*
 * deepgreen.h -- C ABI of the MI355X (gfx950) compressible-Euler step library.
 *
 * This is the drop-in boundary for the reference's hot path: the Rust host
 * (src/rs/src/main.rs, init.rs, grid.rs) keeps its state layout, its
 * set-up and its time loop, and replaces
 *     physics::compute_step(info, grid, &mut dt)      src/rs/src/physics.rs:595
 * and the loop around it
 *     compute(info, grid, total_time, max_steps, tout) src/rs/src/main.rs:38-68
 * by the calls below.  Plain C types only; every call returns 0 or a negative
 * dg_status and never aborts.  A dg_ctx is used by one host thread at a time.
 *
 * Host arrays crossing this boundary use the reference layout: one array per
 * field, C order over the ghosted grid (n+4 cells per axis, k fastest), the
 * layout of Array3<f32> in grid.rs:79-109, with conserved values extensive
 * (multiplied by the cell volume ds^3, physics.rs:143-175).  The element type
 * is double for DG_FP64 and float for DG_FP32 / DG_FP32_BULK.
 */
#ifndef DEEPGREEN_H
#define DEEPGREEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dg_status {
    DG_OK = 0,
    DG_ERR_INVALID = -1,     /* bad argument or configuration */
    DG_ERR_DEVICE = -2,      /* HIP runtime error */
    DG_ERR_ALLOC = -3,       /* device or host allocation failed */
    DG_ERR_NONFINITE = -4,   /* dt is not finite (only when dt_nan_fallback = 0) */
    DG_ERR_STATE = -5        /* no state set yet */
} dg_status;

/* BoundaryType, grid.rs:24-28 (same numbering).  DG_CONDUCTING_WALL is the
 * CUDA variant's BC::CONDUCTING_WALL (src/cu/include/constants.h:44): a wall
 * for the flow; on a lower face (x-, y-, z-) the first interior cell also
 * exchanges heat through the wall (src/cu/include/kernels.h:561-661; the
 * CUDA upper-face branches change only ghost cells).  Constants: dg_set_heat. */
typedef enum dg_boundary { DG_PERIODIC = 0, DG_OUTFLOW = 1, DG_WALL = 2, DG_CONDUCTING_WALL = 3 } dg_boundary;

typedef enum dg_precision {
    DG_FP64 = 0,      /* state, stencil, dt and sums in double */
    DG_FP32 = 1,      /* everything in float, dt as the reference f32 solver */
    DG_FP32_BULK = 2  /* float state and stencil; dt reduction and sums in double */
} dg_precision;

typedef enum dg_math {
    DG_MATH_EXACT = 0, /* the reference's operation order, IEEE div/sqrt, no FMA */
    DG_MATH_FAST = 1   /* same scheme, reassociated with FMA, hardware reciprocals and
                          reciprocal square roots (FP64: refined by Newton steps);
                          FP32 and FP32-bulk need 1e-5 <= ds <= 1e4 */
} dg_math;

/* Slope used by the MUSCL-Hancock reconstruction.  NONE is the reference's
 * unlimited central difference (physics.rs:198-215). */
typedef enum dg_limiter { DG_LIMITER_NONE = 0, DG_LIMITER_MINMOD = 1, DG_LIMITER_VANLEER = 2 } dg_limiter;

typedef enum dg_state_kind { DG_PRIMITIVE = 0, DG_CONSERVED = 1 } dg_state_kind;

/* Replaces Simulation (grid.rs:31-41) plus the compile-time constants of
 * globals.rs:3-7, which become parameters. */
typedef struct dg_config {
    int64_t n[3];          /* interior cells per axis (Simulation::new arguments) */
    int32_t nghosts;       /* must be 2 (the reference's only value) */
    int32_t bc[6];         /* x-, x+, y-, y+, z-, z+ ; periodic faces come in pairs */
    double ds;             /* cell size */
    double gamma;          /* GAMMA */
    double cfl;            /* CFL */
    double g;              /* G, used when gravity != 0 */
    int32_t gravity;       /* Simulation::gravity: z-flux gravity term, physics.rs:366-372 */
    int32_t precision;     /* dg_precision */
    int32_t math;          /* dg_math */
    int32_t limiter;       /* dg_limiter */
    int32_t dt_nan_fallback; /* 1: a NaN wave speed gives dt = CFL*1.0 (physics.rs:191-194) */
    int32_t device;        /* HIP device ordinal */
    int32_t tile_y;        /* 0 = automatic; else rows per workgroup (4, 8 or 16) */
    int32_t chunk_x;       /* 0 = automatic (long chunks, ~512 workgroups); else planes per workgroup along x */
} dg_config;

typedef struct dg_ctx dg_ctx;

/* Fill *cfg with the reference defaults: gamma 5/3, CFL 0.2, G -0.1, no
 * gravity, periodic everywhere, FP32 exact, no limiter (globals.rs:3-7). */
void dg_config_default(dg_config *cfg);

/* Replaces Simulation::new + SimulationGrid::new (grid.rs:44-76, 169-221):
 * allocates the device state (two conserved buffers) and a HIP stream. */
int dg_create(const dg_config *cfg, dg_ctx **out);
void dg_destroy(dg_ctx *ctx);
/* Message of the context's last failure; with ctx NULL, the reason of the
 * calling thread's last failed dg_create ("null context" if there was none). */
const char *dg_last_error(const dg_ctx *ctx);

/* Upload a state from host arrays f[0..4] in the reference layout.
 * kind = DG_PRIMITIVE: (rho, vx, vy, vz, p); the library then applies
 * calc_conserved (physics.rs:143-175).  kind = DG_CONSERVED: (mass,
 * momentum_x, momentum_y, momentum_z, energy).  Only interior cells are read. */
int dg_set_state(dg_ctx *ctx, const void *const f[5], int kind);

/* Download the current state into host arrays in the reference layout.
 * Interior cells are written; ghost cells are filled with the boundary maps
 * (boundaries.rs:5-99, handle_walls :197-407) for both kinds. */
int dg_get_state(dg_ctx *ctx, void *const f[5], int kind);

/* One call of physics::compute_step (physics.rs:595-681): computes dt from the
 * current state (returned in *dt_used), advances the state by one step. */
int dg_step(dg_ctx *ctx, double *dt_used);

/* dg_step that also returns what the reference's grid holds after
 * compute_step besides the conserved state, i.e. the groups its state dump
 * writes (io.rs:5-196, save_to_disk after compute_step, main.rs:56-62):
 * prim[5] the primitives of the step's start (rho, vx, vy, vz, p; ghosts
 * filled, grid.rs:121-135), grad[5] the gradients of rho, vx, vy, vz, p and
 * flux[5] the fluxes of mass, momentum x/y/z, energy (physics.rs:198-215,
 * 284-409, after their ghost updates, grid.rs:153-167).  Host arrays in the
 * reference layout: prim (n+4)^3, grad and flux (n+4)^3 x 3 with the axis
 * last (Array4, grid.rs:79-109).  Any pointer may be NULL.  The arrays are
 * computed by separate dump kernels in the reference's arithmetic (exact
 * math, whatever the context's setting) from the state the step starts from;
 * the first call allocates 40 ghosted arrays on the device, kept until
 * dg_destroy.  Fluxes do not include the conducting-wall heat source. */
int dg_step_dump(dg_ctx *ctx, double *dt_used, void *const prim[5], void *const grad[5], void *const flux[5]);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): one dg_step on caller-owned device buffers that calls no
 * allocating routine.  in[5] and out[5] are device arrays of the interior
 * conserved state (C order, n[0] x n[1] x n[2], the layout dg_device_state
 * exposes; out must not alias in).  dt is computed from `in` (as after
 * dg_set_state), the advanced state is written to out, and, if dt_used is not
 * NULL, the call synchronizes `stream` (a hipStream_t; NULL = the null stream)
 * and returns the dt (with dt_nan_fallback = 0 a non-finite dt returns
 * DG_ERR_NONFINITE before the step, as dg_step).  With dt_used NULL the call
 * only enqueues kernels (capturable).  The partial results and scalars live in
 * a static device array of the library, one per GPU, held until the process
 * ends: grids whose step needs more than DG_NOALLOC_MAX_SLOTS workgroups (or
 * conducting walls) are rejected with DG_ERR_INVALID, and calls that overlap
 * on one GPU race on that array. */
#define DG_NOALLOC_MAX_SLOTS 8192
int dg_step_noalloc(const dg_config *cfg, const void *const in[5], void *const out[5], void *stream,
                    double *dt_used);

/* The loop of main.rs:51-67 without output: steps while
 * !(t > t_end || steps >= max_steps), accumulating t += dt in the precision's
 * time type (float for DG_FP32, as the reference).  Runs from a captured HIP
 * graph with the dt kept on the device.  t and the step count continue from
 * previous calls; *t_out and *steps_out receive the totals. */
int dg_run(dg_ctx *ctx, int64_t max_steps, double t_end, double *t_out, int64_t *steps_out);

/* Constants of the conducting-wall heat source: outside temperature (K),
 * outside wind speed v_out (m/s; hc = 10 - v_out + 10 sqrt(v_out)) and the
 * specific gas constant (defaults 273, 0, 287: src/cu/include/constants.h:63-66).
 * Each step then adds dt ds (ds Q / V) to the energy of the first interior cell
 * along a conducting lower wall, Q = (T_out - p/(RS rho)) ds^2 / (0.008/0.2 +
 * 1/hc), V = ds^3 (p, rho of that cell at the start of the step). */
int dg_set_heat(dg_ctx *ctx, double t_out_kelvin, double v_out, double rs);

/* Reset the run clock (t = 0, steps = 0) kept by dg_run. */
int dg_reset_clock(dg_ctx *ctx);

/* FP64 sums over the interior of the five conserved fields of the current
 * state (total mass, momentum, energy).  No reference counterpart: the
 * reference has no conservation diagnostic. */
int dg_conservation_sums(dg_ctx *ctx, double out[5]);

/* dt the next step will use (CFL * min over cells, physics.rs:177-196). */
int dg_next_dt(dg_ctx *ctx, double *dt);

/* Device pointers of the current conserved state: five arrays over the
 * interior only, C order (n[0], n[1], n[2]).  Valid until the next step. */
int dg_device_state(dg_ctx *ctx, void *ptrs[5]);

/* Timing harness: runs nsteps plain (non-graph) steps on the context's stream
 * with a HIP event pair around every launch; reports the mean device time of
 * the fused step kernel and of the per-step dt kernels (finalize; with
 * conducting walls also the heat and reduction kernels), and the total wall
 * time of the nsteps steps (also measured with events). */
int dg_time_steps(dg_ctx *ctx, int64_t nsteps, double *ms_total, double *ms_step_kernel,
                  double *ms_reduce_kernel);

/* Library and kernel facts for reporting. */
int dg_kernel_info(dg_ctx *ctx, int *tile_y, int *chunk_x, int64_t *workgroups,
                   int *lds_bytes, int *threads_per_workgroup);

int dg_device_count(int *count);

/* Architecture name of a HIP device (e.g. "gfx950", NUL-terminated in arch)
 * and its number of compute units. */
int dg_device_info(int device, char *arch, int len, int *compute_units);

/* ---------------------------------------------------------------- ensemble --
 * Many independent grids advanced together: one persistent step kernel per
 * step strides over the (instance, tile) work items of every instance,
 * followed by one dt/clock kernel with a workgroup per instance; dg_ensemble_run
 * replays these from a captured HIP graph.  No reference counterpart: the
 * reference steps one grid per process (main.rs:51-67).  Every instance
 * computes exactly what dg_step / dg_run compute for the same dg_config and
 * state (same kernels, bit for bit), whatever the tiling. */
typedef struct dg_ensemble dg_ensemble;

/* cfgs[0..count-1]: one configuration per instance.  precision, math,
 * limiter, device and tile_y must agree across instances (tile_y 0 = 8);
 * n, bc, ds, gamma, cfl, g, gravity, dt_nan_fallback and chunk_x are per
 * instance.  grid: persistent workgroups per step launch (0 = as many as fit
 * on the device at once). */
int dg_ensemble_create(const dg_config *cfgs, int count, int grid, dg_ensemble **out);
void dg_ensemble_destroy(dg_ensemble *ens);
const char *dg_ensemble_last_error(const dg_ensemble *ens);

/* dg_set_state / dg_get_state of instance i. */
int dg_ensemble_set_state(dg_ensemble *ens, int i, const void *const f[5], int kind);
int dg_ensemble_get_state(dg_ensemble *ens, int i, void *const f[5], int kind);

/* Up to max_steps steps of every instance; an instance whose own clock
 * passes t_end (the test of dg_run) stops advancing while the others go on.
 * Without a stop time (t_end infinite) a run of up to 1024 steps is a single
 * launch of one captured graph (cached per length). */
int dg_ensemble_run(dg_ensemble *ens, int64_t max_steps, double t_end);

/* Run clock of instance i (as dg_run reports it) and the dt of its next step. */
int dg_ensemble_clock(dg_ensemble *ens, int i, double *t, int64_t *steps, double *next_dt);

/* dg_conservation_sums of instance i. */
int dg_ensemble_conservation_sums(dg_ensemble *ens, int i, double out[5]);

/* Timing harness: nsteps steps launched one by one with HIP events around
 * every step kernel; mean device time per step kernel and per dt kernel. */
int dg_ensemble_time_steps(dg_ensemble *ens, int64_t nsteps, double *ms_step_kernel, double *ms_dt_kernel);

/* New tiling without touching the state: tile_y for all instances (0 = keep),
 * chunk_x[i] per instance (chunk_x NULL = keep; 0 = automatic), grid
 * persistent workgroups (0 = as many as fit), waves = occupancy target of
 * the FP32 fast step kernel (3 or 4 waves per SIMD; 0 = keep; default 4).
 * Results do not change. */
int dg_ensemble_retile(dg_ensemble *ens, int tile_y, const int32_t *chunk_x, int grid, int waves);

/* Interior cells over all instances, work items per step, persistent
 * workgroups, rows per workgroup. */
int dg_ensemble_info(dg_ensemble *ens, int64_t *cells, int *items, int *grid, int *tile_y);

/* One step with a given dt instead of the CFL dt (the dt of the step after
 * is the CFL dt again).  No reference counterpart (compute_step ignores its
 * incoming dt, physics.rs:595-599); used to generate training data. */
int dg_step_dt(dg_ctx *ctx, double dt);

/* FP32 grids: the state of `coarse` becomes the 2x restriction of the state of
 * `fine` (n_fine = 2 n_coarse per axis; each coarse cell holds the sum of its
 * 8 children, the conserved values being extensive, physics.rs:143-175). */
int dg_restrict(dg_ctx *fine, dg_ctx *coarse);

/* ------------------------------------------------ learned flux correction --
 * A 3D CNN 5 -> 32 -> 32 -> 15 (3x3x3 kernels, GELU after layers 1 and 2;
 * bf16 MFMA with fp32 accumulation; fp32 master weights and Adam) maps the
 * normalized primitive state of a periodic FP32 grid to corrections of the
 * five flux components on the three upper faces of every cell.  A learned
 * step adds them to the plain step's face fluxes, so the totals are conserved
 * to rounding; every face of a cell whose corrected density or pressure would
 * not be positive keeps the plain flux.  No reference counterpart.
 *
 * Parameters (fp32, dg_cnn_param_count() values): W1 [27][5][32], b1 [32],
 * W2 [27][32][32], b2 [32], W3 [27][32][15], b3 [15]; tap = (dx*3 + dy)*3 + dz.
 * Normalization (15 floats): input shift[5] and scale[5] of (rho, vx, vy, vz,
 * p), output scale[5] of the five flux components. */
typedef struct dg_cnn dg_cnn;

int dg_cnn_create(int device, uint64_t seed, dg_cnn **out);
void dg_cnn_destroy(dg_cnn *m);
const char *dg_cnn_last_error(const dg_cnn *m);
int64_t dg_cnn_param_count(void);
int dg_cnn_get_params(dg_cnn *m, float *params);
int dg_cnn_set_params(dg_cnn *m, const float *params);
int dg_cnn_get_normalization(dg_cnn *m, float norm[15]);
int dg_cnn_set_normalization(dg_cnn *m, const float norm[15]);

/* Network output y [batch][dims-6 cells][15] for inputs x [batch][dims cells][5]
 * (host fp32, channels last, already normalized; valid convolutions). */
int dg_cnn_forward(dg_cnn *m, int batch, const int32_t dims[3], const float *x, float *y);
/* Parameter gradients of sum(y * dy) (dy [batch][dims-6 cells][15]); grads may
 * be NULL (kept on the device for dg_cnn_adam_step). */
int dg_cnn_backward(dg_cnn *m, int batch, const int32_t dims[3], const float *x, const float *dy, float *grads);
/* One training batch: inputs x [batch][(L+7)^3][5] (normalized), targets
 * [batch][L^3][5] (normalized); loss = mean over cells and components of
 * (-div G - target)^2 with G on the (L+1)^3 faces; gradients kept on the device. */
int dg_cnn_train_batch(dg_cnn *m, int batch, int L, const float *x, const float *target, double *loss, float *grads);
/* Adam (fp32 master weights) with grads from the host, or the last computed
 * gradients when grads is NULL. */
int dg_cnn_adam_step(dg_cnn *m, const float *grads, float lr, float beta1, float beta2, float eps);

/* Training data from the solver: restricts `fine` into `coarse` (U^n), steps
 * the coarse grid once and the fine grid twice with a common dt (the smaller
 * of the coarse CFL dt and twice the fine one), and stores U^n and the missed
 * flux divergence R = (restricted fine - coarse) / (dt ds^2).  Both grids
 * FP32 and periodic; the fine state advances by two steps. */
int dg_cnn_add_sample(dg_cnn *m, dg_ctx *fine, dg_ctx *coarse);
/* dg_cnn_add_sample with two options.  keep_coarse = 1: U^n is the coarse
 * grid's own current state (its trajectory) instead of the restriction of
 * `fine`, so R also holds the coarse trajectory's accumulated departure from
 * the restricted fine one; the coarse grid then advances by its plain step.
 * store = 0: both grids advance (the fine by two steps, the coarse by one, the
 * common dt) and no sample is kept. */
int dg_cnn_add_sample_mode(dg_cnn *m, dg_ctx *fine, dg_ctx *coarse, int keep_coarse, int store);
int dg_cnn_sample_count(dg_cnn *m);
/* Standard deviation of Gaussian noise added to the normalized network inputs
 * during dg_cnn_train (0 = none; training-time augmentation). */
int dg_cnn_set_input_noise(dg_cnn *m, float std);
/* Normalization from the samples (mean/std of the primitives, RMS of R). */
int dg_cnn_fit_normalization(dg_cnn *m);
/* steps of Adam (beta 0.9/0.999, eps 1e-8) on random periodic crops of the
 * samples (batch crops of L^3 cells); losses[steps] (may be NULL). */
int dg_cnn_train(dg_cnn *m, int steps, int batch, int L, float lr, uint64_t seed, float *losses);

/* dg_step / dg_run with the learned correction (ctx: FP32, periodic);
 * dg_learned_step takes the CFL dt when dt <= 0, else the given dt. */
int dg_learned_step(dg_ctx *ctx, dg_cnn *m, double dt, double *dt_used);
int dg_learned_run(dg_ctx *ctx, dg_cnn *m, int64_t max_steps, double t_end, double *t_out, int64_t *steps_out);
/* A learned step with the correction G [n0*n1*n2][16] given by the host
 * (normalized units, as dg_learned_correction returns) instead of the
 * network's: for tests of the correction, the fallback and conservation. */
int dg_learned_step_given(dg_cnn *m, dg_ctx *ctx, const float *G, double *dt_used);
/* Timing harness: the network's forward on the state of ctx, nrep times, HIP
 * events around each launch; mean ms of [input preparation, layer 1, layer 2,
 * layer 3, total]. */
int dg_learned_time(dg_cnn *m, dg_ctx *ctx, int nrep, double ms[5]);
/* Cells that kept the plain flux in the last learned step. */
int dg_learned_flagged(dg_cnn *m, int64_t *count);
/* Network output for the current state of ctx: G [n0*n1*n2][16] (channel
 * 5*axis + component, channel 15 unused), normalized units. */
int dg_learned_correction(dg_cnn *m, dg_ctx *ctx, float *G);

#ifdef __cplusplus
}
#endif

#endif /* DEEPGREEN_H */
