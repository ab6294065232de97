// dg_dump.hip -- dg_step_dump: one step that also hands back the arrays the
// reference's SimulationGrid holds after compute_step and that its state dump
// writes (io.rs:5-196): the primitives of the start of the step, the
// gradients and the fluxes, each after its ghost fill.
//
// The fused step kernel never materializes gradients or fluxes, so a dump
// recomputes them from the state about to be stepped, pass by pass as the
// reference does (physics.rs:595-681): primitives on the ghosted grid with
// the boundary maps (calc_primitives + update_primitive_ghost_cells),
// calc_gradients on [1, m-1)^3 + update_gradients_ghost_cells,
// extrapolate_primitives_t on every cell + update_primitive_xtr_ghost_cells,
// calc_flux_reconstruction x3 on [1, m-1)^3 + update_fluxes_ghost_cells.
// Ghost updates are plane copies applied face by face in the reference order
// (x-, x+, y-, y+, z-, z+; boundaries.rs:5-407), so edge and corner ghosts
// are the same compositions.  Arithmetic keeps the reference's operation order
// and IEEE division and square root (the library is built without FMA
// contraction), whatever the context's math setting: the dumped arrays are
// the reference arithmetic's on the state of the step's start.  Only dumps
// use these kernels; they run one thread per ghosted cell.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>

#include "../../include/deepgreen.h"
#include "dg_ctx.h"
#include "dg_internal.h"

namespace dg {
namespace {

struct Grid {
    int m[3];        // ghosted sizes (n + 4)
    int n[3];        // interior sizes
    int bc[6];       // x-, x+, y-, y+, z-, z+ (conducting walls as walls)
    int limiter, gravity;
    double ds, gamma, g;
    long long G;     // ghosted cells
};

__device__ __forceinline__ long long gid(const Grid &g, int i, int j, int k)
{
    return ((long long)i * g.m[1] + j) * g.m[2] + k;
}

// Ghosted index gi (0..m-1) of one axis -> interior source index and flip
// (the boundary maps of update_primitive_ghost_cells).
__device__ __forceinline__ int src_index(int gi, int n, int lo, int hi, int *flip)
{
    const int p = gi - 2;
    *flip = 0;
    if (p < 0) {
        if (lo == BC_PERIODIC) return p + n;
        if (lo == BC_OUTFLOW) return 0;
        *flip = 1;
        return -1 - p;
    }
    if (p >= n) {
        if (hi == BC_PERIODIC) return p - n;
        if (hi == BC_OUTFLOW) return n - 1;
        *flip = 1;
        return 2 * n - 1 - p;
    }
    return p;
}

// calc_primitives (physics.rs:109-141) of every cell, ghosts taking the
// primitives of their mapped interior cell (normal velocity negated at walls).
template <typename T>
__global__ void __launch_bounds__(256) dump_prim_kernel(const Grid g, const Fields<const T> U, const Fields<T> P)
{
    const T ds = T(g.ds), gm1 = T(g.gamma) - T(1.0);
    const T V = ds * ds * ds;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.G; c += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(c % g.m[2]), j = (int)((c / g.m[2]) % g.m[1]), i = (int)(c / ((long long)g.m[1] * g.m[2]));
        int fx, fy, fz;
        const int si = src_index(i, g.n[0], g.bc[0], g.bc[1], &fx);
        const int sj = src_index(j, g.n[1], g.bc[2], g.bc[3], &fy);
        const int sk = src_index(k, g.n[2], g.bc[4], g.bc[5], &fz);
        const long long s = ((long long)si * g.n[1] + sj) * g.n[2] + sk;
        const T M = U.f[0][s], Mx = U.f[1][s], My = U.f[2][s], Mz = U.f[3][s], E = U.f[4][s];
        const T r = M / V;
        const T ux = Mx / (r * V), uy = My / (r * V), uz = Mz / (r * V);
        P.f[0][c] = r;
        P.f[1][c] = fx ? T(-1.0) * ux : ux;
        P.f[2][c] = fy ? T(-1.0) * uy : uy;
        P.f[3][c] = fz ? T(-1.0) * uz : uz;
        P.f[4][c] = (E / V - T(0.5) * r * (ux * ux + uy * uy + uz * uz)) * gm1;
    }
}

template <typename T> __device__ __forceinline__ T dump_fabs(T x);
template <> __device__ __forceinline__ float dump_fabs<float>(float x) { return fabsf(x); }
template <> __device__ __forceinline__ double dump_fabs<double>(double x) { return fabs(x); }

// The library's limited slopes (minmod, van Leer) of a difference pair.
template <typename T>
__device__ __forceinline__ T limited_slope(T a, T b, int lim)
{
    if (!(a * b > T(0))) return T(0);
    if (lim == LIM_MINMOD) return dump_fabs<T>(a) <= dump_fabs<T>(b) ? a : b;
    return T(2) * a * b / (a + b);
}

// calc_gradients (physics.rs:198-215) on [1, m-1)^3; zero elsewhere (the
// reference's arrays start at zero and those cells are only written by the
// ghost copies that follow).
template <typename T>
__global__ void __launch_bounds__(256) dump_grad_kernel(const Grid g, const Fields<const T> P, const Fields<T> D)
{
    const T ds = T(g.ds);
    const long long sx = (long long)g.m[1] * g.m[2], sy = g.m[2];
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.G; c += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(c % g.m[2]), j = (int)((c / g.m[2]) % g.m[1]), i = (int)(c / sx);
        const bool inner = i >= 1 && i < g.m[0] - 1 && j >= 1 && j < g.m[1] - 1 && k >= 1 && k < g.m[2] - 1;
        const long long st[3] = {sx, sy, 1};
        for (int f = 0; f < 5; ++f) {
            const T *q = P.f[f];
            for (int d = 0; d < 3; ++d) {
                T v = T(0);
                if (inner) {
                    if (g.limiter == LIM_NONE) v = (q[c + st[d]] - q[c - st[d]]) / (T(2.0) * ds);
                    else v = limited_slope<T>(q[c] - q[c - st[d]], q[c + st[d]] - q[c], g.limiter) / ds;
                }
                D.f[f][c * 3 + d] = v;
            }
        }
    }
}

// extrapolate_primitives_t (physics.rs:217-282) on every cell.
template <typename T>
__global__ void __launch_bounds__(256)
dump_xtr_kernel(const Grid g, const Fields<const T> P, const Fields<const T> D, const Fields<T> X, double dtd)
{
    const T gamma = T(g.gamma), dt = T(dtd);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < g.G; c += (long long)gridDim.x * blockDim.x) {
        const T r = P.f[0][c], ux = P.f[1][c], uy = P.f[2][c], uz = P.f[3][c], p = P.f[4][c];
        const T *gr = D.f[0] + 3 * c, *gx = D.f[1] + 3 * c, *gy = D.f[2] + 3 * c, *gz = D.f[3] + 3 * c,
                *gp = D.f[4] + 3 * c;
        X.f[0][c] = r - T(0.5) * dt * (ux * gr[0] + r * gx[0] + uy * gr[1] + r * gy[1] + uz * gr[2] + r * gz[2]);
        X.f[1][c] = ux - T(0.5) * dt * (ux * gx[0] + uy * gx[1] + uz * gx[2] + (T(1.0) / r) * gp[0]);
        X.f[2][c] = uy - T(0.5) * dt * (ux * gy[0] + uy * gy[1] + uz * gy[2] + (T(1.0) / r) * gp[1]);
        X.f[3][c] = uz - T(0.5) * dt * (ux * gz[0] + uy * gz[1] + uz * gz[2] + (T(1.0) / r) * gp[2]);
        X.f[4][c] = p - T(0.5) * dt * (gamma * p * (gx[0] + gy[1] + gz[2]) + ux * gp[0] + uy * gp[1] + uz * gp[2]);
    }
}

template <typename T> __device__ __forceinline__ T dump_sqrt(T x);
template <> __device__ __forceinline__ float dump_sqrt<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double dump_sqrt<double>(double x) { return sqrt(x); }
template <typename T> __device__ __forceinline__ T dump_fmax(T a, T b);
template <> __device__ __forceinline__ float dump_fmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __device__ __forceinline__ double dump_fmax<double>(double a, double b) { return fmax(a, b); }

// calc_flux_reconstruction (physics.rs:284-409) along axis `dim` on
// [1, m-1)^3, with the argument rotation of physics.rs:616-677 (slot 1 the
// normal momentum); zero elsewhere.
template <typename T>
__global__ void __launch_bounds__(256)
dump_flux_kernel(const Grid g, const Fields<const T> D, const Fields<const T> X, const Fields<T> F, int dim)
{
    const T ds = T(g.ds), gamma = T(g.gamma), gm1 = gamma - T(1.0), Gacc = T(g.g);
    // slot s of the rotated argument lists -> field index
    int rot[5] = {0, 1 + dim, dim == 0 ? 2 : 1, dim == 2 ? 2 : 3, 4};
    const long long sx = (long long)g.m[1] * g.m[2], sy = g.m[2];
    const long long st = dim == 0 ? sx : (dim == 1 ? sy : 1);
    const bool grav = dim == 2 && g.gravity;
    for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b < g.G; b += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(b % g.m[2]), j = (int)((b / g.m[2]) % g.m[1]), i = (int)(b / sx);
        const bool inner = i >= 1 && i < g.m[0] - 1 && j >= 1 && j < g.m[1] - 1 && k >= 1 && k < g.m[2] - 1;
        if (!inner) {
            for (int f = 0; f < 5; ++f) F.f[f][3 * b + dim] = T(0);
            continue;
        }
        const long long a = b + st;   // "l": cell c + e_d; "r": cell c
        T l[5], r[5];
        for (int s = 0; s < 5; ++s) {
            const int f = rot[s];
            l[s] = X.f[f][a] - (D.f[f][3 * a + dim]) * (ds / T(2.));
            r[s] = X.f[f][b] + (D.f[f][3 * b + dim]) * (ds / T(2.));
        }
        const T rl = l[0], vxl = l[1], vyl = l[2], vzl = l[3], pl = l[4];
        const T rr = r[0], vxr = r[1], vyr = r[2], vzr = r[3], pr = r[4];
        const T en_left = T(0.5) * (rl * (vxl * vxl + vyl * vyl + vzl * vzl)) + (pl / gm1);
        const T en_right = T(0.5) * (rr * (vxr * vxr + vyr * vyr + vzr * vzr)) + (pr / gm1);
        const T rho_star = (rl + rr) / T(2.0);
        const T mx = (rl * vxl + rr * vxr) / T(2.0);
        const T my = (rl * vyl + rr * vyr) / T(2.0);
        const T mz = (rl * vzl + rr * vzr) / T(2.0);
        const T en_star = T(0.5) * (en_left + en_right);
        const T p_star = gm1 * (en_star - T(0.5) * (mx * mx + my * my + mz * mz) / rho_star);
        T f0 = mx;
        T f1 = (mx * mx / rho_star) + p_star;
        T f2 = mx * my / rho_star;
        T f3 = mx * mz / rho_star;
        T f4 = (en_star + p_star) * (mx / rho_star);
        if (grav) {   // physics.rs:366-372, 0-based k
            const T h = (T(g.m[2]) - T(2.0) * T(2)) * ds - (T(k) - T(2.0)) * ds;
            f1 += T(1.0) * rho_star * Gacc * h;
            f4 += T(0.5) * rho_star * (vxl + vxr) * Gacc * h;
        }
        const T c_l = dump_sqrt<T>(gamma * pl / rl) + dump_fabs<T>(vxl);
        const T c_r = dump_sqrt<T>(gamma * pr / rr) + dump_fabs<T>(vxr);
        const T c = dump_fmax<T>(c_l, c_r);
        f0 = mx - T(0.5) * c * (rl - rr);
        f1 -= c * (rl * vxl - rr * vxr) / T(2.0);
        f2 -= c * (rl * vyl - rr * vyr) / T(2.0);
        f3 -= c * (rl * vzl - rr * vzr) / T(2.0);
        f4 -= c * (en_left - en_right) / T(2.0);
        const T out[5] = {f0 * ds, f1 * ds, f2 * ds, f3 * ds, f4 * ds};
        for (int s = 0; s < 5; ++s) F.f[rot[s]][3 * b + dim] = out[s];
    }
}

// One plane copy of a ghost update (a `slice_mut(..).assign(..)` of
// boundaries.rs) on five arrays of ncomp components: plane dst <- plane src
// along `axis`; component field `negf` (if >= 0) is negated (handle_walls).
template <typename T>
__global__ void __launch_bounds__(256)
plane_copy_kernel(const Grid g, const Fields<T> A, int ncomp, int axis, int dst, int src, int negf)
{
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    const long long cells = (long long)g.m[o1] * g.m[o2];
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < 5 * cells; e += (long long)gridDim.x * blockDim.x) {
        const int f = (int)(e / cells);
        const long long r = e % cells;
        int q[3];
        q[o1] = (int)(r / g.m[o2]);
        q[o2] = (int)(r % g.m[o2]);
        q[axis] = src;
        const long long s = gid(g, q[0], q[1], q[2]);
        q[axis] = dst;
        const long long d = gid(g, q[0], q[1], q[2]);
        for (int c = 0; c < ncomp; ++c) {
            const T v = A.f[f][s * ncomp + c];
            A.f[f][d * ncomp + c] = f == negf ? T(-1.0) * v : v;
        }
    }
}

template <typename T>
int copy_plane(const Grid &g, const Fields<T> &A, int ncomp, int axis, int dst, int src, int negf, hipStream_t s)
{
    const int o1 = axis == 0 ? 1 : 0, o2 = axis == 2 ? 1 : 2;
    const long long work = 5LL * g.m[o1] * g.m[o2];
    const int blocks = (int)std::min<long long>((work + 255) / 256, 4096);
    hipLaunchKernelGGL(plane_copy_kernel<T>, dim3(blocks), dim3(256), 0, s, g, A, ncomp, axis, dst, src, negf);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// apply_boundary_conditions (ncomp 1) / apply_boundary_conditions4 (ncomp 3)
// of boundaries.rs:5-195 on five arrays; walls are no-ops there.
template <typename T>
int apply_bc(const Grid &g, const Fields<T> &A, int ncomp, hipStream_t s)
{
    int rc = 0;
    for (int face = 0; face < 6 && !rc; ++face) {
        const int axis = face / 2, upper = face % 2, m = g.m[axis], bc = g.bc[face];
        if (bc == BC_OUTFLOW) {
            if (!upper) {
                rc |= copy_plane<T>(g, A, ncomp, axis, 0, 2, -1, s);
                rc |= copy_plane<T>(g, A, ncomp, axis, 1, 2, -1, s);
            } else {
                rc |= copy_plane<T>(g, A, ncomp, axis, m - 2, m - 3, -1, s);
                rc |= copy_plane<T>(g, A, ncomp, axis, m - 1, m - 3, -1, s);
            }
        } else if (bc == BC_PERIODIC) {
            if (!upper) {
                rc |= copy_plane<T>(g, A, ncomp, axis, 0, m - 4, -1, s);
                rc |= copy_plane<T>(g, A, ncomp, axis, 1, m - 3, -1, s);
            } else {
                rc |= copy_plane<T>(g, A, ncomp, axis, m - 2, 2, -1, s);
                rc |= copy_plane<T>(g, A, ncomp, axis, m - 1, 3, -1, s);
            }
        }
    }
    return rc;
}

// handle_walls (boundaries.rs:197-407) on five primitive-like arrays.
template <typename T>
int handle_walls(const Grid &g, const Fields<T> &A, hipStream_t s)
{
    int rc = 0;
    for (int face = 0; face < 6 && !rc; ++face) {
        if (g.bc[face] != BC_WALL) continue;
        const int axis = face / 2, m = g.m[axis];
        if (face % 2 == 0) {
            rc |= copy_plane<T>(g, A, 1, axis, 1, 2, 1 + axis, s);
            rc |= copy_plane<T>(g, A, 1, axis, 0, 3, 1 + axis, s);
        } else {
            rc |= copy_plane<T>(g, A, 1, axis, m - 2, m - 3, 1 + axis, s);
            rc |= copy_plane<T>(g, A, 1, axis, m - 1, m - 4, 1 + axis, s);
        }
    }
    return rc;
}

template <typename T>
int dump_pipeline(const Grid &g, const void *const u[5], T *base, double dt, hipStream_t s)
{
    // layout of the dump buffer: P[5][G], X[5][G], D[5][3G], F[5][3G]
    Fields<T> P, X, D, F;
    Fields<const T> U, Pc, Xc, Dc;
    for (int f = 0; f < 5; ++f) {
        P.f[f] = base + (size_t)f * g.G;
        X.f[f] = base + (size_t)(5 + f) * g.G;
        D.f[f] = base + (size_t)10 * g.G + (size_t)f * 3 * g.G;
        F.f[f] = base + (size_t)25 * g.G + (size_t)f * 3 * g.G;
        U.f[f] = static_cast<const T *>(u[f]);
        Pc.f[f] = P.f[f];
        Xc.f[f] = X.f[f];
        Dc.f[f] = D.f[f];
    }
    const int blocks = (int)std::min<long long>((g.G + 255) / 256, 8192);
    int rc = 0;
    hipLaunchKernelGGL(dump_prim_kernel<T>, dim3(blocks), dim3(256), 0, s, g, U, P);
    hipLaunchKernelGGL(dump_grad_kernel<T>, dim3(blocks), dim3(256), 0, s, g, Pc, D);
    rc |= apply_bc<T>(g, D, 3, s);
    hipLaunchKernelGGL(dump_xtr_kernel<T>, dim3(blocks), dim3(256), 0, s, g, Pc, Dc, X, dt);
    rc |= handle_walls<T>(g, X, s);
    rc |= apply_bc<T>(g, X, 1, s);
    for (int dim = 0; dim < 3; ++dim)
        hipLaunchKernelGGL(dump_flux_kernel<T>, dim3(blocks), dim3(256), 0, s, g, Dc, Xc, F, dim);
    rc |= apply_bc<T>(g, F, 3, s);
    return (rc || hipGetLastError() != hipSuccess) ? -2 : 0;
}

} // namespace
} // namespace dg

extern "C" int dg_step_dump(dg_ctx *c, double *dt_used, void *const prim[5], void *const grad[5],
                            void *const flux[5])
{
    if (!c) return DG_ERR_INVALID;
    if (!c->have_state) { c->err = "no state set"; return DG_ERR_STATE; }
    if (hipSetDevice(c->cfg.device) != hipSuccess) { c->err = "hipSetDevice"; return DG_ERR_DEVICE; }
    dg::Grid g{};
    for (int d = 0; d < 3; ++d) {
        g.n[d] = c->args.n[d];
        g.m[d] = c->args.n[d] + 4;
    }
    for (int i = 0; i < 6; ++i) g.bc[i] = c->args.bc[i];
    g.limiter = c->args.limiter;
    g.gravity = c->args.gravity;
    g.ds = c->args.ds;
    g.gamma = c->args.gamma;
    g.g = c->args.g;
    g.G = (long long)g.m[0] * g.m[1] * g.m[2];
    const size_t bytes = (size_t)40 * g.G * c->el;
    if (!c->dumpbuf) {
        if (hipMalloc(&c->dumpbuf, bytes) != hipSuccess) {
            c->dumpbuf = nullptr;
            c->err = "dump buffers: device allocation failed";
            return DG_ERR_ALLOC;
        }
    }
    // dt of the coming step (the step kernel reads the same device value)
    if (hipMemcpyAsync(c->sc_host, c->sc, sizeof(dg::Scalars), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        c->err = "reading the scalars";
        return DG_ERR_DEVICE;
    }
    const double dt = c->sc_host->dt;
    if (!c->cfg.dt_nan_fallback && !std::isfinite(dt)) { c->err = "dt is not finite"; return DG_ERR_NONFINITE; }
    const void *u[5];
    for (int f = 0; f < 5; ++f) u[f] = c->buf[c->cur][f];
    const int rc = c->prec == DG_FP64
                       ? dg::dump_pipeline<double>(g, u, static_cast<double *>(c->dumpbuf), dt, c->stream)
                       : dg::dump_pipeline<float>(g, u, static_cast<float *>(c->dumpbuf), dt, c->stream);
    if (rc) { c->err = "dump kernels"; return DG_ERR_DEVICE; }
    const size_t one = (size_t)g.G * c->el;
    char *base = static_cast<char *>(c->dumpbuf);
    for (int f = 0; f < 5; ++f) {
        hipError_t e = hipSuccess;
        if (prim && prim[f]) e = hipMemcpyAsync(prim[f], base + f * one, one, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && grad && grad[f])
            e = hipMemcpyAsync(grad[f], base + (10 + 3 * f) * one, 3 * one, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && flux && flux[f])
            e = hipMemcpyAsync(flux[f], base + (25 + 3 * f) * one, 3 * one, hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) { c->err = "dump download"; return DG_ERR_DEVICE; }
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) { c->err = "dump download"; return DG_ERR_DEVICE; }
    return dg_step(c, dt_used);
}
