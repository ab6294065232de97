/*
 * dg_oracle_impl.h -- body of the CPU oracle, instantiated once per real type
 * by dg_oracle.c (define R = float|double and SFX = f32|f64 before including).
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference Rust solver's
 * per-step numerics so the GPU library can be checked against it.  It is never
 * linked into, called by or shipped with the product library.
 *
 * Every function keeps the reference's loop ranges, operand order and
 * association, because the float instantiation is meant to reproduce the
 * reference f32 arithmetic bit for bit (compiled with -ffp-contract=off and no
 * fast-math).  Citations are to src/rs/src/ in the reference repository.
 *
 * Layout (grid.rs:79-118, 169-221): every field is a C-order array over the
 * ghosted grid (nx, ny, nz), k fastest.  Gradients and fluxes are
 * Array4(nx, ny, nz, 3) with the axis index last.
 */

#define CAT2_(a, b) a##_##b
#define CAT2(a, b) CAT2_(a, b)
#define FN(name) CAT2(name, SFX)
#define CTX FN(dgo_ctx)

typedef struct CTX {
    dgo_params p;
    int64_t n3;              /* nx*ny*nz */
    R *prim[5];              /* rho, vx, vy, vz, p          (grid.rs:95-101) */
    R *xtr[5];               /* primitives_xtr              (grid.rs:114)    */
    R *cons[5];              /* mass, mx, my, mz, energy    (grid.rs:87-93)  */
    R *grad[5];              /* drho, dvx, dvy, dvz, dp (x3) (grid.rs:103-109) */
    R *flux[5];              /* mass, mx, my, mz, energy (x3) (grid.rs:79-85) */
} CTX;

#define I3(c, i, j, k) ((((int64_t)(i)) * (c)->p.ny + (j)) * (c)->p.nz + (k))
#define I4(c, i, j, k, d) (I3(c, i, j, k) * 3 + (d))

static R FN(gm1)(const CTX *c) { return (R)c->p.gamma - (R)1.0; }

void FN(dgo_free)(CTX *c);

/* ---------------------------------------------------------------- setup -- */

CTX *FN(dgo_new)(const dgo_params *p)
{
    CTX *c = (CTX *)calloc(1, sizeof(CTX));
    if (!c) return NULL;
    c->p = *p;
    c->n3 = p->nx * p->ny * p->nz;
    for (int f = 0; f < 5; ++f) {
        c->prim[f] = (R *)calloc((size_t)c->n3, sizeof(R));
        c->xtr[f] = (R *)calloc((size_t)c->n3, sizeof(R));
        c->cons[f] = (R *)calloc((size_t)c->n3, sizeof(R));
        c->grad[f] = (R *)calloc((size_t)c->n3 * 3, sizeof(R));
        c->flux[f] = (R *)calloc((size_t)c->n3 * 3, sizeof(R));
        if (!c->prim[f] || !c->xtr[f] || !c->cons[f] || !c->grad[f] || !c->flux[f]) {
            FN(dgo_free)(c);
            return NULL;
        }
    }
    return c;
}

void FN(dgo_free)(CTX *c)
{
    if (!c) return;
    for (int f = 0; f < 5; ++f) {
        free(c->prim[f]); free(c->xtr[f]); free(c->cons[f]);
        free(c->grad[f]); free(c->flux[f]);
    }
    free(c);
}

/* which: 0-4 primitives, 5-9 xtr, 10-14 conserved, 15-19 gradients,
 * 20-24 fluxes.  Gradient and flux arrays hold 3*n3 values. */
R *FN(dgo_field)(CTX *c, int which)
{
    if (which < 0 || which >= 25) return NULL;
    int f = which % 5;
    switch (which / 5) {
    case 0: return c->prim[f];
    case 1: return c->xtr[f];
    case 2: return c->cons[f];
    case 3: return c->grad[f];
    default: return c->flux[f];
    }
}

/* ------------------------------------------------------------ boundaries -- */

/* Copy the plane at index `src` along `axis` onto the plane `dst`, for every
 * index of the other two axes (ghosts included) and every one of `ncomp`
 * trailing components.  This is one `mesh.slice_mut(...).assign(&...)` of
 * boundaries.rs.  `neg` multiplies by -1.0 (handle_walls' mapv_inplace). */
static void FN(copy_plane)(const CTX *c, R *a, int axis, int64_t dst, int64_t src,
                           int ncomp, int neg)
{
    const int64_t nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    if (axis == 0) {
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t k = 0; k < nz; ++k)
                for (int d = 0; d < ncomp; ++d) {
                    R v = a[(I3(c, src, j, k)) * ncomp + d];
                    a[(I3(c, dst, j, k)) * ncomp + d] = neg ? (R)-1.0 * v : v;
                }
    } else if (axis == 1) {
        for (int64_t i = 0; i < nx; ++i)
            for (int64_t k = 0; k < nz; ++k)
                for (int d = 0; d < ncomp; ++d) {
                    R v = a[(I3(c, i, src, k)) * ncomp + d];
                    a[(I3(c, i, dst, k)) * ncomp + d] = neg ? (R)-1.0 * v : v;
                }
    } else {
        for (int64_t i = 0; i < nx; ++i)
            for (int64_t j = 0; j < ny; ++j)
                for (int d = 0; d < ncomp; ++d) {
                    R v = a[(I3(c, i, j, src)) * ncomp + d];
                    a[(I3(c, i, j, dst)) * ncomp + d] = neg ? (R)-1.0 * v : v;
                }
    }
}

/* apply_boundary_conditions (boundaries.rs:5-99) for ncomp == 1 and
 * apply_boundary_conditions4 (boundaries.rs:101-195) for ncomp == 3.
 * Faces in the order x-, x+, y-, y+, z-, z+; WALL is a no-op here. */
static void FN(apply_bc)(const CTX *c, R *a, int ncomp)
{
    const int64_t n[3] = {c->p.nx, c->p.ny, c->p.nz};
    const int64_t g = c->p.nghosts;
    for (int face = 0; face < 6; ++face) {
        const int axis = face / 2;
        const int upper = face % 2;
        const int64_t m = n[axis];
        const int bc = c->p.bcs[face];
        if (bc == DGO_OUTFLOW) {
            if (!upper) {
                FN(copy_plane)(c, a, axis, 0, 2, ncomp, 0);
                FN(copy_plane)(c, a, axis, 1, 2, ncomp, 0);
            } else {
                FN(copy_plane)(c, a, axis, m - 2, m - g - 1, ncomp, 0);
                FN(copy_plane)(c, a, axis, m - 1, m - g - 1, ncomp, 0);
            }
        } else if (bc == DGO_PERIODIC) {
            if (!upper) {
                FN(copy_plane)(c, a, axis, 0, m - g - 2, ncomp, 0);
                FN(copy_plane)(c, a, axis, 1, m - g - 1, ncomp, 0);
            } else {
                FN(copy_plane)(c, a, axis, m - 2, 2, ncomp, 0);
                FN(copy_plane)(c, a, axis, m - 1, 3, ncomp, 0);
            }
        }
    }
}

/* handle_walls (boundaries.rs:197-407): mirror the first two interior layers
 * into the ghosts of every WALL face, negating the normal velocity. */
static void FN(handle_walls)(const CTX *c, R *const w[5])
{
    const int64_t n[3] = {c->p.nx, c->p.ny, c->p.nz};
    const int64_t g = c->p.nghosts;
    for (int face = 0; face < 6; ++face) {
        if (c->p.bcs[face] != DGO_WALL) continue;
        const int axis = face / 2;
        const int64_t m = n[axis];
        int64_t dst0, src0, dst1, src1;
        if (face % 2 == 0) { dst0 = 1; src0 = 2; dst1 = 0; src1 = 3; }
        else { dst0 = m - g; src0 = m - g - 1; dst1 = m - 1; src1 = m - g - 2; }
        for (int f = 0; f < 5; ++f)
            FN(copy_plane)(c, w[f], axis, dst0, src0, 1, f == 1 + axis);
        for (int f = 0; f < 5; ++f)
            FN(copy_plane)(c, w[f], axis, dst1, src1, 1, f == 1 + axis);
    }
}

/* grid.rs:121-135 */
void FN(dgo_update_primitive_ghost_cells)(CTX *c)
{
    FN(handle_walls)(c, c->prim);
    for (int f = 0; f < 5; ++f) FN(apply_bc)(c, c->prim[f], 1);
}

/* grid.rs:137-151 */
void FN(dgo_update_primitive_xtr_ghost_cells)(CTX *c)
{
    FN(handle_walls)(c, c->xtr);
    for (int f = 0; f < 5; ++f) FN(apply_bc)(c, c->xtr[f], 1);
}

/* grid.rs:153-159 */
void FN(dgo_update_gradients_ghost_cells)(CTX *c)
{
    for (int f = 0; f < 5; ++f) FN(apply_bc)(c, c->grad[f], 3);
}

/* grid.rs:161-167 */
void FN(dgo_update_fluxes_ghost_cells)(CTX *c)
{
    for (int f = 0; f < 5; ++f) FN(apply_bc)(c, c->flux[f], 3);
}

/* ----------------------------------------------------------- numerics -- */

/* calc_primitives (physics.rs:109-141), every cell. */
void FN(dgo_calc_primitives)(CTX *c)
{
    const R ds = (R)c->p.ds;
    const R V = ds * ds * ds; /* ds.powi(3) */
    const R gm1 = FN(gm1)(c);
    R *rho = c->prim[0], *vx = c->prim[1], *vy = c->prim[2], *vz = c->prim[3], *p = c->prim[4];
    const R *M = c->cons[0], *Mx = c->cons[1], *My = c->cons[2], *Mz = c->cons[3], *E = c->cons[4];
    const int64_t n3 = c->n3;
#pragma omp parallel for schedule(static) if (c->p.nthreads > 1) num_threads(c->p.nthreads)
    for (int64_t q = 0; q < n3; ++q) {
        const R r = M[q] / V;
        rho[q] = r;
        const R ux = Mx[q] / (r * V);
        const R uy = My[q] / (r * V);
        const R uz = Mz[q] / (r * V);
        vx[q] = ux; vy[q] = uy; vz[q] = uz;
        p[q] = (E[q] / V - (R)0.5 * r * (ux * ux + uy * uy + uz * uz)) * gm1;
    }
}

/* calc_conserved (physics.rs:143-175), every cell. */
void FN(dgo_calc_conserved)(CTX *c)
{
    const R ds = (R)c->p.ds;
    const R V = ds * ds * ds;
    const R gm1 = FN(gm1)(c);
    const R *rho = c->prim[0], *vx = c->prim[1], *vy = c->prim[2], *vz = c->prim[3], *p = c->prim[4];
    R *M = c->cons[0], *Mx = c->cons[1], *My = c->cons[2], *Mz = c->cons[3], *E = c->cons[4];
    const int64_t n3 = c->n3;
#pragma omp parallel for schedule(static) if (c->p.nthreads > 1) num_threads(c->p.nthreads)
    for (int64_t q = 0; q < n3; ++q) {
        const R r = rho[q];
        M[q] = r * V;
        Mx[q] = vx[q] * r * V;
        My[q] = vy[q] * r * V;
        Mz[q] = vz[q] * r * V;
        E[q] = V * (r * (vx[q] * vx[q] + vy[q] * vy[q] + vz[q] * vz[q]) / (R)2.0 + p[q] / gm1);
    }
}

/* calc_timestep (physics.rs:177-196).  The minimum runs over every cell,
 * ghosts included.  ndarray-stats' min() fails on any NaN, and the reference
 * then uses 1.0 (physics.rs:191-194); dt_nan_fallback = 0 propagates NaN. */
R FN(dgo_calc_timestep)(CTX *c)
{
    const R ds = (R)c->p.ds;
    const R gamma = (R)c->p.gamma;
    const R *rho = c->prim[0], *vx = c->prim[1], *vy = c->prim[2], *vz = c->prim[3], *p = c->prim[4];
    const int64_t n3 = c->n3;
    R mn = (R)INFINITY;
    int any_nan = 0;
    const int par = c->p.nthreads > 1 && !c->p.ref_threading;
#pragma omp parallel for schedule(static) reduction(min : mn) reduction(| : any_nan) if (par) num_threads(c->p.nthreads)
    for (int64_t q = 0; q < n3; ++q) {
        const R v = ds / (RSQRT(gamma * p[q] / rho[q]) +
                          RSQRT(vx[q] * vx[q] + vy[q] * vy[q] + vz[q] * vz[q]));
        if (v != v) any_nan = 1;
        else if (v < mn) mn = v;
    }
    R min_val;
    if (any_nan) min_val = c->p.dt_nan_fallback ? (R)1.0 : (R)NAN;
    else min_val = mn;
    return (R)c->p.cfl * min_val;
}

/* Limited slope from the backward and forward differences a, b.  The
 * reference has limiters only as dead code (physics.f90:205-238); these are
 * the textbook minmod and van Leer (harmonic-mean) slopes, an extension. */
static R FN(limited)(R a, R b, int limiter)
{
    if (!(a * b > (R)0)) return (R)0;
    if (limiter == DGO_LIM_MINMOD) return RFABS(a) <= RFABS(b) ? a : b;
    return (R)2 * a * b / (a + b); /* van Leer */
}

/* The limited slope of one (backward, forward) difference pair, exported for
 * the tests that pin it. */
R FN(dgo_limited)(R a, R b, int limiter) { return FN(limited)(a, b, limiter); }

/* calc_gradients (physics.rs:198-215) for all five primitives, on [1, n-1)^3.
 * With limiter != none the central difference is replaced by the limited
 * slope divided by ds. */
void FN(dgo_calc_gradients)(CTX *c)
{
    const R ds = (R)c->p.ds;
    const int64_t nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    if (c->p.limiter != DGO_LIM_NONE) {
        const int lim = c->p.limiter;
        for (int f = 0; f < 5; ++f) {
            const R *q = c->prim[f];
            R *g = c->grad[f];
#pragma omp parallel for schedule(static) if (c->p.nthreads > 1) num_threads(c->p.nthreads)
            for (int64_t i = 1; i < nx - 1; ++i)
                for (int64_t j = 1; j < ny - 1; ++j)
                    for (int64_t k = 1; k < nz - 1; ++k) {
                        const R q0 = q[I3(c, i, j, k)];
                        g[I4(c, i, j, k, 0)] = FN(limited)(q0 - q[I3(c, i - 1, j, k)], q[I3(c, i + 1, j, k)] - q0, lim) / ds;
                        g[I4(c, i, j, k, 1)] = FN(limited)(q0 - q[I3(c, i, j - 1, k)], q[I3(c, i, j + 1, k)] - q0, lim) / ds;
                        g[I4(c, i, j, k, 2)] = FN(limited)(q0 - q[I3(c, i, j, k - 1)], q[I3(c, i, j, k + 1)] - q0, lim) / ds;
                    }
        }
        return;
    }
    for (int f = 0; f < 5; ++f) {
        const R *q = c->prim[f];
        R *g = c->grad[f];
#pragma omp parallel for schedule(static) if (c->p.nthreads > 1) num_threads(c->p.nthreads)
        for (int64_t i = 1; i < nx - 1; ++i)
            for (int64_t j = 1; j < ny - 1; ++j)
                for (int64_t k = 1; k < nz - 1; ++k) {
                    g[I4(c, i, j, k, 0)] = (q[I3(c, i + 1, j, k)] - q[I3(c, i - 1, j, k)]) / ((R)2.0 * ds);
                    g[I4(c, i, j, k, 1)] = (q[I3(c, i, j + 1, k)] - q[I3(c, i, j - 1, k)]) / ((R)2.0 * ds);
                    g[I4(c, i, j, k, 2)] = (q[I3(c, i, j, k + 1)] - q[I3(c, i, j, k - 1)]) / ((R)2.0 * ds);
                }
    }
}

/* extrapolate_primitives_t (physics.rs:217-282), every cell. */
void FN(dgo_extrapolate)(CTX *c, R dt)
{
    const R gamma = (R)c->p.gamma;
    const R *rho = c->prim[0], *vx = c->prim[1], *vy = c->prim[2], *vz = c->prim[3], *pp = c->prim[4];
    const R *gr = c->grad[0], *gx = c->grad[1], *gy = c->grad[2], *gz = c->grad[3], *gp = c->grad[4];
    R *orho = c->xtr[0], *ovx = c->xtr[1], *ovy = c->xtr[2], *ovz = c->xtr[3], *op = c->xtr[4];
    const int64_t n3 = c->n3;
    const int par = c->p.nthreads > 1 && !c->p.ref_threading;
#pragma omp parallel for schedule(static) if (par) num_threads(c->p.nthreads)
    for (int64_t q = 0; q < n3; ++q) {
        const R r = rho[q], ux = vx[q], uy = vy[q], uz = vz[q], p = pp[q];
        const R drho_dx = gr[3 * q + 0], dvx_dx = gx[3 * q + 0], dvy_dx = gy[3 * q + 0], dvz_dx = gz[3 * q + 0];
        const R drho_dy = gr[3 * q + 1], dvx_dy = gx[3 * q + 1], dvy_dy = gy[3 * q + 1], dvz_dy = gz[3 * q + 1];
        const R drho_dz = gr[3 * q + 2], dvx_dz = gx[3 * q + 2], dvy_dz = gy[3 * q + 2], dvz_dz = gz[3 * q + 2];
        const R dp_dx = gp[3 * q + 0], dp_dy = gp[3 * q + 1], dp_dz = gp[3 * q + 2];
        orho[q] = r - (R)0.5 * dt *
                          (ux * drho_dx + r * dvx_dx + uy * drho_dy + r * dvy_dy + uz * drho_dz + r * dvz_dz);
        ovx[q] = ux - (R)0.5 * dt * (ux * dvx_dx + uy * dvx_dy + uz * dvx_dz + ((R)1.0 / r) * dp_dx);
        ovy[q] = uy - (R)0.5 * dt * (ux * dvy_dx + uy * dvy_dy + uz * dvy_dz + ((R)1.0 / r) * dp_dy);
        ovz[q] = uz - (R)0.5 * dt * (ux * dvz_dx + uy * dvz_dy + uz * dvz_dz + ((R)1.0 / r) * dp_dz);
        op[q] = p - (R)0.5 * dt *
                        (gamma * p * (dvx_dx + dvy_dy + dvz_dz) + ux * dp_dx + uy * dp_dy + uz * dp_dz);
    }
}

/* calc_flux_reconstruction (physics.rs:284-409) for one axis.  The argument
 * lists are already rotated by the caller exactly as in physics.rs:616-677:
 * slot 1 is the normal momentum / velocity, slots 2-3 the tangential ones. */
static void FN(flux_axis)(CTX *c, R *const F[5], R *const G[5], R *const W[5], int dim)
{
    const R ds = (R)c->p.ds;
    const R gamma = (R)c->p.gamma;
    const R gm1 = FN(gm1)(c);
    const R G_acc = (R)c->p.g;
    const int64_t nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const int64_t ghosts = c->p.nghosts;
    const int64_t oi = dim == 0, oj = dim == 1, ok = dim == 2;
    const int grav = ok == 1 && c->p.gravity;
    const int par = c->p.nthreads > 1 && !c->p.ref_threading;
#pragma omp parallel for schedule(static) if (par) num_threads(c->p.nthreads)
    for (int64_t i = 1; i < nx - 1; ++i)
        for (int64_t j = 1; j < ny - 1; ++j)
            for (int64_t k = 1; k < nz - 1; ++k) {
                const int64_t a = I3(c, i + oi, j + oj, k + ok); /* "l": cell c + e_d */
                const int64_t b = I3(c, i, j, k);                 /* "r": cell c       */
                const R rl = W[0][a] - (G[0][3 * a + dim]) * (ds / (R)2.);
                const R rr = W[0][b] + (G[0][3 * b + dim]) * (ds / (R)2.);
                const R vxl = W[1][a] - (G[1][3 * a + dim]) * (ds / (R)2.);
                const R vxr = W[1][b] + (G[1][3 * b + dim]) * (ds / (R)2.);
                const R vyl = W[2][a] - (G[2][3 * a + dim]) * (ds / (R)2.);
                const R vyr = W[2][b] + (G[2][3 * b + dim]) * (ds / (R)2.);
                const R vzl = W[3][a] - (G[3][3 * a + dim]) * (ds / (R)2.);
                const R vzr = W[3][b] + (G[3][3 * b + dim]) * (ds / (R)2.);
                const R pl = W[4][a] - (G[4][3 * a + dim]) * (ds / (R)2.);
                const R pr = W[4][b] + (G[4][3 * b + dim]) * (ds / (R)2.);

                const R en_left = (R)0.5 * (rl * (vxl * vxl + vyl * vyl + vzl * vzl)) + (pl / gm1);
                const R en_right = (R)0.5 * (rr * (vxr * vxr + vyr * vyr + vzr * vzr)) + (pr / gm1);

                const R rho_star = (rl + rr) / (R)2.0;
                const R mx_star = (rl * vxl + rr * vxr) / (R)2.0;
                const R my_star = (rl * vyl + rr * vyr) / (R)2.0;
                const R mz_star = (rl * vzl + rr * vzr) / (R)2.0;
                const R en_star = (R)0.5 * (en_left + en_right);
                const R p_star =
                    gm1 * (en_star - (R)0.5 * (mx_star * mx_star + my_star * my_star + mz_star * mz_star) / rho_star);

                const int64_t o = 3 * b + dim;
                R f0 = mx_star;
                R f1 = (mx_star * mx_star / rho_star) + p_star;
                R f2 = mx_star * my_star / rho_star;
                R f3 = mx_star * mz_star / rho_star;
                R f4 = (en_star + p_star) * (mx_star / rho_star);

                if (grav) { /* physics.rs:366-372 */
                    const R h = ((R)nz - (R)2.0 * (R)ghosts) * ds -
                                ((R)(k + c->p.gravity_k_shift) - (R)2.0) * ds;
                    f1 += (R)1.0 * rho_star * G_acc * h;
                    f4 += (R)0.5 * rho_star * (vxl + vxr) * G_acc * h;
                }

                const R c_l = RSQRT(gamma * pl / rl) + RFABS(vxl);
                const R c_r = RSQRT(gamma * pr / rr) + RFABS(vxr);
                const R c_star = RFMAX(c_l, c_r); /* f32::max ignores NaN */

                f0 = mx_star - (R)0.5 * c_star * (rl - rr);
                f1 -= c_star * (rl * vxl - rr * vxr) / (R)2.0;
                f2 -= c_star * (rl * vyl - rr * vyr) / (R)2.0;
                f3 -= c_star * (rl * vzl - rr * vzr) / (R)2.0;
                f4 -= c_star * (en_left - en_right) / (R)2.0;
                F[0][o] = f0 * ds;
                F[1][o] = f1 * ds;
                F[2][o] = f2 * ds;
                F[3][o] = f3 * ds;
                F[4][o] = f4 * ds;
            }
}

void FN(dgo_calc_fluxes)(CTX *c)
{
    /* physics.rs:616-677: the three calls with their rotated arguments. */
    R *const *Fs = c->flux, *const *Gs = c->grad, *const *Ws = c->xtr;
    {
        R *const F[5] = {Fs[0], Fs[1], Fs[2], Fs[3], Fs[4]};
        R *const G[5] = {Gs[0], Gs[1], Gs[2], Gs[3], Gs[4]};
        R *const W[5] = {Ws[0], Ws[1], Ws[2], Ws[3], Ws[4]};
        FN(flux_axis)(c, F, G, W, 0);
    }
    {
        R *const F[5] = {Fs[0], Fs[2], Fs[1], Fs[3], Fs[4]};
        R *const G[5] = {Gs[0], Gs[2], Gs[1], Gs[3], Gs[4]};
        R *const W[5] = {Ws[0], Ws[2], Ws[1], Ws[3], Ws[4]};
        FN(flux_axis)(c, F, G, W, 1);
    }
    {
        R *const F[5] = {Fs[0], Fs[3], Fs[1], Fs[2], Fs[4]};
        R *const G[5] = {Gs[0], Gs[3], Gs[1], Gs[2], Gs[4]};
        R *const W[5] = {Ws[0], Ws[3], Ws[1], Ws[2], Ws[4]};
        FN(flux_axis)(c, F, G, W, 2);
    }
}

/* add_fluxes (physics.rs:539-593) on [1, n-1)^3. */
void FN(dgo_add_fluxes)(CTX *c, R dt)
{
    const R ds = (R)c->p.ds;
    const int64_t nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const int par = c->p.nthreads > 1 && !c->p.ref_threading;
    for (int f = 0; f < 5; ++f) {
        R *U = c->cons[f];
        const R *F = c->flux[f];
#pragma omp parallel for schedule(static) if (par) num_threads(c->p.nthreads)
        for (int64_t i = 1; i < nx - 1; ++i)
            for (int64_t j = 1; j < ny - 1; ++j)
                for (int64_t k = 1; k < nz - 1; ++k) {
                    U[I3(c, i, j, k)] -= dt * ds *
                                         (F[I4(c, i, j, k, 0)] - F[I4(c, i - 1, j, k, 0)] +
                                          F[I4(c, i, j, k, 1)] - F[I4(c, i, j - 1, k, 1)] +
                                          F[I4(c, i, j, k, 2)] - F[I4(c, i, j, k - 1, 2)]);
                }
    }
}

/* Conducting-wall heat source of the CUDA variant (src/cu/include/kernels.h:
 * 561-661), restated: for a conducting lower wall along axis d, the energy
 * flux through the wall face of the first interior cell gains Q/V, with
 * t_in = p / (RS rho) of the wall ghost (index 1) in the state the CUDA flux
 * kernel reads, the half-step extrapolated one (physics.hpp:437-442 passes
 * rho_xtr and p_xtr as its rho and p): here the xtr arrays after their ghost
 * fill, where the ghost is the mirror of the first interior cell (handle_walls
 * on xtr, grid.rs:137-151), hc = 10 - v_out +
 * 10 sqrt(v_out), S = ds^2, V = ds^3, Q = (T_out - t_in) S / (d/kappa +
 * 1/hc), d = 0.008, kappa = 0.2.  Through the update (F scaled by ds,
 * U -= dt ds dF; physics.rs:401-405, 539-593) the first interior cell's
 * energy gains dt ds (ds Q/V); added after add_fluxes, axes in the order x,
 * y, z.  The CUDA code's upper-wall branches act on ghost cells only and
 * change no interior cell, so only lower walls carry the source.
 * PARITY UNPINNED: the CUDA build cannot run here. */
void FN(dgo_heat)(CTX *c, R dt)
{
    if (!c->p.heat_mask) return;
    const R ds = (R)c->p.ds;
    const R S = ds * ds, V = ds * ds * ds;
    const R vout = (R)c->p.v_out;
    const R hc = (R)10.0 - vout + (R)10.0 * RSQRT(vout);
    const R dk = (R)0.008 / (R)0.2;
    const int64_t n[3] = {c->p.nx, c->p.ny, c->p.nz};
    for (int d = 0; d < 3; ++d) {
        if (!(c->p.heat_mask & (1 << d))) continue;
        for (int64_t i = 2; i < n[0] - 2; ++i)
            for (int64_t j = 2; j < n[1] - 2; ++j)
                for (int64_t k = 2; k < n[2] - 2; ++k) {
                    const int64_t q[3] = {i, j, k};
                    if (q[d] != 2) continue;
                    int64_t gq[3] = {i, j, k};
                    gq[d] = 1;   /* the wall ghost */
                    const int64_t gi = I3(c, gq[0], gq[1], gq[2]);
                    const R t_in = c->xtr[4][gi] / ((R)c->p.rs * c->xtr[0][gi]);
                    const R Q = ((R)c->p.t_out - t_in) * (S / (dk + (R)1.0 / hc));
                    c->cons[4][I3(c, i, j, k)] += dt * ds * ((Q / V) * ds);
                }
    }
}

/* compute_step (physics.rs:595-681).  The incoming *dt is ignored. */
void FN(dgo_compute_step)(CTX *c, R *dt)
{
    FN(dgo_calc_primitives)(c);
    FN(dgo_update_primitive_ghost_cells)(c);
    *dt = FN(dgo_calc_timestep)(c);
    FN(dgo_calc_gradients)(c);
    FN(dgo_update_gradients_ghost_cells)(c);
    FN(dgo_extrapolate)(c, *dt);
    FN(dgo_update_primitive_xtr_ghost_cells)(c);
    FN(dgo_calc_fluxes)(c);
    FN(dgo_update_fluxes_ghost_cells)(c);
    FN(dgo_add_fluxes)(c, *dt);
    FN(dgo_heat)(c, *dt);
}

/* compute (main.rs:38-68) without the HDF5 dump: calc_conserved, then step
 * while !(tnow > total_time || step >= max_steps).  Time accumulates in R.
 * dt_hist (may be NULL) receives the dt of every step.  Returns the step
 * count; *t_out receives tnow. */
int64_t FN(dgo_compute)(CTX *c, double total_time, int64_t max_steps, double *t_out, R *dt_hist)
{
    R tnow = (R)0.0;
    int64_t step = 0;
    R dt = (R)0.5;
    FN(dgo_calc_conserved)(c);
    for (;;) {
        if (tnow > (R)total_time || step >= max_steps) break;
        FN(dgo_compute_step)(c, &dt);
        if (dt_hist) dt_hist[step] = dt;
        tnow += dt;
        step += 1;
    }
    if (t_out) *t_out = (double)tnow;
    return step;
}

/* Step loop without the leading calc_conserved: the caller's conserved
 * arrays are the state.  Used to continue a run and by the timing harness. */
int64_t FN(dgo_steps)(CTX *c, int64_t nsteps, R *dt_hist)
{
    R dt = (R)0.5;
    for (int64_t s = 0; s < nsteps; ++s) {
        FN(dgo_compute_step)(c, &dt);
        if (dt_hist) dt_hist[s] = dt;
    }
    return nsteps;
}

#undef I3
#undef I4
#undef CTX
#undef FN
#undef CAT2
#undef CAT2_
