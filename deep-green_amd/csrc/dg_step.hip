// dg_step.hip -- fused compressible-Euler step for gfx950 (MI355X).
//
// One kernel performs the whole reference step (physics.rs:595-681):
// primitives (calc_primitives :109-141), ghost fill (boundaries.rs as index
// maps), central or limited slopes (calc_gradients :198-215), the Hancock
// half-step (extrapolate_primitives_t :217-282), the Rusanov face flux
// (calc_flux_reconstruction :284-409, one evaluation per face), the
// conservative update (add_fluxes :539-593) and, in its epilogue, the
// wave-speed maximum for the next step's dt (calc_timestep :177-196) and the
// FP64 conservation partial sums.  State is read once and written once.
//
// Work decomposition: a workgroup owns a (TY x 64) tile of the (y, z) plane
// and marches along x over a chunk of planes.  Wave w is tile row w; lane l
// is tile column l, so every global access is a 64-wide coalesced row.
//   * primitives of three consecutive planes sit in an LDS ring (tile plus a
//     two-cell halo), filled by mapping ghost positions through the boundary
//     maps, so no ghost cells exist in HBM;
//   * each thread evaluates slopes and the extrapolated state of its cell,
//     then the flux through its +x (previous plane), +y and +z faces;
//   * y faces exchange through LDS, z faces through lane shuffles, x faces
//     through registers carried from the previous plane;
//   * halo states (row -1, row TY, column -1, column 64) are evaluated by
//     waves 0, 1 and 2.
// Boundaries: a ghost state is evaluated at its mapped interior cell
// (periodic: the wrapped cell; outflow: the boundary cell, which reproduces
// the reference's copies of xtr and gradients, boundaries.rs:5-195); a wall
// ghost state is the mirrored extrapolated state of the boundary cell plus
// the normal slope evaluated on the mirrored ghost primitives
// (handle_walls :197-407 applied to primitives and xtr, gradients computed on
// the ghost).  The outflow lower face copies the first interior face
// (apply_boundary_conditions4 on fluxes, boundaries.rs:107-111).
//
// Arithmetic: with FAST = false every expression keeps the reference's
// operation order (the library is built with -ffp-contract=off and IEEE
// division and square root), so results match the reference arithmetic bit
// for bit.  FAST = true is the same scheme reassociated around FMAs and
// hardware reciprocals.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <math.h>
#include <stdint.h>

#include "dg_internal.h"

namespace dg {

constexpr int TZ = 64;       // tile columns == wavefront width
constexpr int RZ = TZ + 4;   // LDS region columns (two-cell halo)

// ------------------------------------------------------------ math helpers --
template <typename T> __device__ __forceinline__ T dsqrt(T x);
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return sqrt(x); }
template <typename T> __device__ __forceinline__ T dmax(T a, T b);  // NaN-ignoring, as f32::max
template <> __device__ __forceinline__ float dmax<float>(float a, float b) { return fmaxf(a, b); }
template <> __device__ __forceinline__ double dmax<double>(double a, double b) { return fmax(a, b); }
template <typename T> __device__ __forceinline__ T dabs(T x) { return x < T(0) ? -x : (x == T(0) ? T(0) * x : x); }
template <> __device__ __forceinline__ float dabs<float>(float x) { return fabsf(x); }
template <> __device__ __forceinline__ double dabs<double>(double x) { return fabs(x); }
template <typename T> __device__ __forceinline__ T drcp(T x);
template <> __device__ __forceinline__ float drcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
// FP64 fast math: the hardware reciprocal estimate refined by one Newton
// step (which squares its relative error)
template <> __device__ __forceinline__ double drcp<double>(double x)
{
    const double r = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, r, 1.0), r, r);
}
// reciprocal square root: hardware estimate (FP32); FP64 refined by one
// Newton step y <- y + y (1 - x y^2) / 2
template <typename T> __device__ __forceinline__ T drsq(T x);
template <> __device__ __forceinline__ float drsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <> __device__ __forceinline__ double drsq<double>(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}
template <typename T> __device__ __forceinline__ T dsqrt_fast(T x);
template <> __device__ __forceinline__ float dsqrt_fast<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double dsqrt_fast<double>(double x) { return sqrt(x); }

// Whole-wavefront lane shifts (DPP wave_shl:1 / wave_shr:1): lane i receives
// lane i+1 (shl) or lane i-1 (shr); the lane without a source (63 for shl, 0
// for shr) keeps `edge` (bound_ctrl off: the destination is not written).
__device__ __forceinline__ int dpp_from_next(int edge, int v)
{ return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_from_prev(int edge, int v)
{ return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xf, 0xf, false); }
template <typename T> __device__ __forceinline__ T from_next_lane(T v, T edge);
template <typename T> __device__ __forceinline__ T from_prev_lane(T v, T edge);
template <> __device__ __forceinline__ float from_next_lane<float>(float v, float edge)
{ return __builtin_bit_cast(float, dpp_from_next(__builtin_bit_cast(int, edge), __builtin_bit_cast(int, v))); }
template <> __device__ __forceinline__ float from_prev_lane<float>(float v, float edge)
{ return __builtin_bit_cast(float, dpp_from_prev(__builtin_bit_cast(int, edge), __builtin_bit_cast(int, v))); }
template <> __device__ __forceinline__ double from_next_lane<double>(double v, double edge)
{
    const long long b = __builtin_bit_cast(long long, v), e = __builtin_bit_cast(long long, edge);
    const int lo = dpp_from_next((int)(e & 0xffffffffLL), (int)(b & 0xffffffffLL));
    const int hi = dpp_from_next((int)(e >> 32), (int)(b >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
template <> __device__ __forceinline__ double from_prev_lane<double>(double v, double edge)
{
    const long long b = __builtin_bit_cast(long long, v), e = __builtin_bit_cast(long long, edge);
    const int lo = dpp_from_prev((int)(e & 0xffffffffLL), (int)(b & 0xffffffffLL));
    const int hi = dpp_from_prev((int)(e >> 32), (int)(b >> 32));
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Element idx of a field addressed with a 32-bit byte offset from the field
// base (a field is < 4 GiB; the base stays in scalar registers).
template <typename T> __device__ __forceinline__ const T &at(const T *base, uint32_t idx)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <typename T> __device__ __forceinline__ T &at(T *base, uint32_t idx)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (uint32_t)(idx * (uint32_t)sizeof(T)));
}

template <typename T> __device__ __forceinline__ T dfma(T a, T b, T c);
template <> __device__ __forceinline__ float dfma<float>(float a, float b, float c) { return fmaf(a, b, c); }
template <> __device__ __forceinline__ double dfma<double>(double a, double b, double c) { return fma(a, b, c); }

// Position p (interior coordinates, -2..n+1 for ghosts) -> interior source
// cell and whether the normal velocity flips (wall).  Positions further out
// only feed cells that are never stored; they are clamped.
struct AxisMap { int idx; int flip; };
__device__ __forceinline__ AxisMap map_axis(int p, int n, int lo, int hi)
{
    if (p < 0) {
        if (p < -2) return {0, 0};
        if (lo == BC_PERIODIC) return {p + n, 0};
        if (lo == BC_OUTFLOW) return {0, 0};
        return {-1 - p, 1};           // wall: -1 -> 0, -2 -> 1
    }
    if (p >= n) {
        if (p > n + 1) return {n - 1, 0};
        if (hi == BC_PERIODIC) return {p - n, 0};
        if (hi == BC_OUTFLOW) return {n - 1, 0};
        return {2 * n - 1 - p, 1};    // wall: n -> n-1, n+1 -> n-2
    }
    return {p, 0};
}

// Tile-local position whose state a tile cell at local p takes (see the
// header): periodic keeps p (the LDS ring holds wrapped primitives), outflow
// and wall use the adjacent interior cell.  Result stays in [-1, tn].
__device__ __forceinline__ int eval_local(int p, int o, int n, int lo, int hi, int tn)
{
    int g = o + p;
    if (g < 0) { if (lo != BC_PERIODIC) g = 0; }
    else if (g >= n) { if (hi != BC_PERIODIC) g = n - 1; }
    int q = g - o;
    return q < -1 ? -1 : (q > tn ? tn : q);
}

__device__ __forceinline__ bool wall_ghost(int p, int o, int n, int lo, int hi)
{
    int g = o + p;
    return (g == -1 && lo == BC_WALL) || (g == n && hi == BC_WALL);
}

// calc_primitives (physics.rs:109-141) for one cell.
// pow2: V is a power of two, so u / V == u * invV exactly.
template <typename T, bool FAST>
__device__ __forceinline__ void to_prim(const T u[5], T V, T invV, T gm1, T w[5], bool pow2 = false)
{
    if constexpr (!FAST) {
        const T r = pow2 ? u[0] * invV : u[0] / V;
        const T ux = u[1] / (r * V), uy = u[2] / (r * V), uz = u[3] / (r * V);
        w[0] = r; w[1] = ux; w[2] = uy; w[3] = uz;
        w[4] = ((pow2 ? u[4] * invV : u[4] / V) - T(0.5) * r * (ux * ux + uy * uy + uz * uz)) * gm1;
    } else {
        const T r = u[0] * invV;
        const T im = drcp<T>(u[0]);
        const T ux = u[1] * im, uy = u[2] * im, uz = u[3] * im;
        w[0] = r; w[1] = ux; w[2] = uy; w[3] = uz;
        const T v2 = dfma<T>(uz, uz, dfma<T>(uy, uy, ux * ux));
        w[4] = dfma<T>(u[4], invV, -T(0.5) * r * v2) * gm1;
    }
}

// Wave-speed denominator of calc_timestep: sqrt(gamma p / rho) + |v|.
template <typename T, bool FAST>
__device__ __forceinline__ T wave_den(const T w[5], T gamma)
{
    static_assert(!FAST, "fast math takes the wave speed from the conserved state (cell_den)");
    return dsqrt<T>(gamma * w[4] / w[0]) + dsqrt<T>(w[1] * w[1] + w[2] * w[2] + w[3] * w[3]);
}

// The same from a conserved state u (extensive).  Fast math: with
// x = gamma (gamma - 1) (E - |m|^2 / (2 U_rho)), gamma p / rho = x / U_rho, so
// the sound speed is |x| / sqrt(x U_rho) (one reciprocal square root) and
// |v| = sqrt(|m|^2) / |U_rho|.
template <typename T, bool FAST>
__device__ __forceinline__ T cell_den(const T u[5], T V, T invV, T gm1, T gamma, bool pow2 = false)
{
    if constexpr (!FAST) {
        T w[5];
        to_prim<T, false>(u, V, invV, gm1, w, pow2);
        return wave_den<T, false>(w, gamma);
    } else {
        const T im = drcp<T>(u[0]);
        const T m2 = dfma<T>(u[3], u[3], dfma<T>(u[2], u[2], u[1] * u[1]));
        const T x = (gamma * gm1) * dfma<T>(T(-0.5) * m2, im, u[4]);
        return dfma<T>(dabs<T>(x), drsq<T>(x * u[0]), dsqrt_fast<T>(m2) * dabs<T>(im));
    }
}

// Limited slope (extension; the reference's limiters are dead code,
// physics.f90:205-238): textbook minmod and van Leer.
template <typename T>
__device__ __forceinline__ T limited(T a, T b, int lim)
{
    if (!(a * b > T(0))) return T(0);
    if (lim == LIM_MINMOD) return dabs<T>(a) <= dabs<T>(b) ? a : b;
    return T(2) * a * b / (a + b);
}

// Rusanov flux of calc_flux_reconstruction (physics.rs:338-405) for one face.
// l = state of the cell on the + side ("l" in the reference), r = state of the
// cell on the - side, both (rho, v_normal, v_t1, v_t2, p); F returns the
// rotated fluxes multiplied by ds.  FAST: the fifth component of l and r is
// P = 2 p / (gamma - 1) (so 2 E = rho |v|^2 + P), F is not multiplied by ds,
// hgm1 = (gamma - 1) / 2 and sK = sqrt(gamma (gamma - 1) / 2).
template <typename T, bool FAST>
__device__ __forceinline__ void rusanov(const T l[5], const T r[5], T F[5], T gamma, T gm1, T ds, bool grav, T G,
                                        T h, T hgm1, T sK)
{
    const T rl = l[0], vxl = l[1], vyl = l[2], vzl = l[3], pl = l[4];
    const T rr = r[0], vxr = r[1], vyr = r[2], vzr = r[3], pr = r[4];
    if constexpr (!FAST) {
        const T en_left = T(0.5) * (rl * (vxl * vxl + vyl * vyl + vzl * vzl)) + (pl / gm1);
        const T en_right = T(0.5) * (rr * (vxr * vxr + vyr * vyr + vzr * vzr)) + (pr / gm1);
        const T rho_star = (rl + rr) / T(2.0);
        const T mx = (rl * vxl + rr * vxr) / T(2.0);
        const T my = (rl * vyl + rr * vyr) / T(2.0);
        const T mz = (rl * vzl + rr * vzr) / T(2.0);
        const T en_star = T(0.5) * (en_left + en_right);
        const T p_star = gm1 * (en_star - T(0.5) * (mx * mx + my * my + mz * mz) / rho_star);
        T f1 = (mx * mx / rho_star) + p_star;
        T f2 = mx * my / rho_star;
        T f3 = mx * mz / rho_star;
        T f4 = (en_star + p_star) * (mx / rho_star);
        if (grav) {
            f1 += T(1.0) * rho_star * G * h;
            f4 += T(0.5) * rho_star * (vxl + vxr) * G * h;
        }
        const T c_l = dsqrt<T>(gamma * pl / rl) + dabs<T>(vxl);
        const T c_r = dsqrt<T>(gamma * pr / rr) + dabs<T>(vxr);
        const T c = dmax<T>(c_l, c_r);
        T f0 = mx - T(0.5) * c * (rl - rr);
        f1 -= c * (rl * vxl - rr * vxr) / T(2.0);
        f2 -= c * (rl * vyl - rr * vyr) / T(2.0);
        f3 -= c * (rl * vzl - rr * vzr) / T(2.0);
        f4 -= c * (en_left - en_right) / T(2.0);
        F[0] = f0 * ds; F[1] = f1 * ds; F[2] = f2 * ds; F[3] = f3 * ds; F[4] = f4 * ds;
    } else {
        // sums S = 2 * (star state); fluxes returned as 2 * flux / 1 (the
        // caller folds the factor 1/2 and ds into the update).  pl, pr hold
        // P = 2 p / (gamma - 1): the doubled total energies are
        // e2 = m . v + P, and Se2 = 2 Se.
        (void)gamma; (void)gm1;
        const T mxl = rl * vxl, myl = rl * vyl, mzl = rl * vzl;
        const T mxr = rr * vxr, myr = rr * vyr, mzr = rr * vzr;
        const T e2l = dfma<T>(mzl, vzl, dfma<T>(myl, vyl, dfma<T>(mxl, vxl, pl)));
        const T e2r = dfma<T>(mzr, vzr, dfma<T>(myr, vyr, dfma<T>(mxr, vxr, pr)));
        const T Sr = rl + rr, Sx = mxl + mxr, Sy = myl + myr, Sz = mzl + mzr, Se2 = e2l + e2r;
        const T iS = drcp<T>(Sr);
        const T m2 = dfma<T>(Sz, Sz, dfma<T>(Sy, Sy, Sx * Sx));
        const T P2 = hgm1 * dfma<T>(-m2, iS, Se2);          // 2 p* = (gamma - 1) (Se - m2 / (2 Sr))
        const T un = Sx * iS;                                // u* normal
        T f0 = Sx;
        T f1 = dfma<T>(Sx, un, P2);
        T f2 = Sy * un;
        T f3 = Sz * un;
        T f4 = dfma<T>(T(0.5), Se2, P2) * un;
        if (grav) {
            const T gh = G * h;
            f1 = dfma<T>(Sr, gh, f1);
            f4 = dfma<T>(T(0.5) * Sr * (vxl + vxr), gh, f4);
        }
        // sqrt(gamma p / rho) = sK |P| / sqrt(P rho): one reciprocal square
        // root; the absolute value keeps the identity for the negative face
        // pressures and densities the unlimited scheme produces at strong
        // jumps (p / rho > 0), and a negative p / rho gives NaN as sqrt does
        const T al = dfma<T>(sK * dabs<T>(pl), drsq<T>(pl * rl), dabs<T>(vxl));
        const T ar = dfma<T>(sK * dabs<T>(pr), drsq<T>(pr * rr), dabs<T>(vxr));
        const T c = dmax<T>(al, ar);
        F[0] = dfma<T>(-c, rl - rr, f0);
        F[1] = dfma<T>(-c, mxl - mxr, f1);
        F[2] = dfma<T>(-c, myl - myr, f2);
        F[3] = dfma<T>(-c, mzl - mzr, f3);
        F[4] = dfma<T>(T(-0.5) * c, e2l - e2r, f4);
        (void)ds;
    }
}

// Rotated slots per axis (physics.rs:616-677): field index of slot s.
__device__ __forceinline__ int rot_field(int axis, int slot)
{
    // x: (rho, vx, vy, vz, p); y: (rho, vy, vx, vz, p); z: (rho, vz, vx, vy, p)
    if (slot == 0 || slot == 4 || axis == 0) return slot;
    if (axis == 1) return slot == 1 ? 2 : (slot == 2 ? 1 : 3);
    return slot == 1 ? 3 : (slot == 2 ? 1 : 2);
}

template <typename T>
struct Consts {
    T ds, gamma, gm1, V, invV, two_ds, inv_two_ds, inv_ds, half_ds, half_dt, dtds, G;
    T hgm1, sK, irk;   // fast math: (gamma - 1) / 2, sqrt(gamma (gamma - 1) / 2), (gamma - 1) / 2
    bool pow2;   // ds a power of two: divisions by ds, 2 ds, ds^3 as exact products
};

// Constants of one step with time step dt.
template <typename T, bool FAST>
__device__ __forceinline__ Consts<T> make_consts(const StepArgs &A, T dt)
{
    Consts<T> C;
    C.ds = T(A.ds);
    C.gamma = T(A.gamma);
    C.gm1 = C.gamma - T(1.0);
    C.V = C.ds * C.ds * C.ds;
    C.invV = T(1) / C.V;
    C.two_ds = T(2.0) * C.ds;
    C.inv_two_ds = T(1) / C.two_ds;
    C.inv_ds = T(1) / C.ds;
    C.pow2 = A.pow2 != 0;
    // fast math: unlimited slopes are central differences (2 ds times the
    // gradient), limited ones one-sided (ds times the gradient)
    const bool one_sided = FAST && A.limiter != LIM_NONE;
    C.half_ds = FAST ? (one_sided ? T(0.5) : T(0.25)) : C.ds / T(2.);
    C.G = T(A.g);
    C.half_dt = FAST ? T(0.5) * dt * (one_sided ? C.inv_ds : C.inv_two_ds) : T(0.5) * dt;
    // fast math: the ring and the fluxes are extensive (times V), so the
    // update divides by V: dt ds^2 / 2 / ds^3
    C.dtds = FAST ? T(0.5) * dt * C.inv_ds : dt * C.ds;
    C.hgm1 = T(0.5) * C.gm1;
    C.sK = dsqrt<T>(C.gamma * C.hgm1);
    C.irk = C.hgm1;
    return C;
}

// Two-element vector for 8-byte (float) / 16-byte (double) LDS accesses.
template <typename T> using vec2 = T __attribute__((ext_vector_type(2)));

// LDS layout of the step kernel.
template <typename T, int TY>
struct StepLDS {
    // The primitives ring is cell-major: a cell's 6 words (ring_cell: the
    // 5 primitives and, in fast math, its scaled inverse density, the
    // pressure word holding 2 p / (gamma - 1)) are consecutive, read as three
    // pairs; every neighbour of a cell, in all three ring slots, is then
    // within the DS offset range of one address register.
    static constexpr int NC = (TY + 4) * RZ;  // region cells of one plane
    static constexpr int CS = 6;              // words per cell
    static constexpr int PS = NC * CS;        // one plane
    alignas(16) T W[3 * PS];  // [plane slot][row][col][ring_cell words]
    // Ym[.][r]: -y face state of row r (r = 0..TY) until the flux phase, in
    // which the one thread that reads it replaces it by the flux through
    // face (r-1, r) (r = 0: the tile's lower y face)
    T Ym[5][TY + 1][TZ];
    T Yp[5][TZ];             // +y face state of row -1
    T Zp[5][TY];             // +z face state of column -1
    T Zm[5][TY];             // -z face state of column 64
    T Z0[5][TY];             // -z face state of column 0, then the flux through the tile's lower z face
};

// The six ring words of a cell from its conserved state u.  Exact math:
// calc_primitives (rho, vx, vy, vz, p) and a pad.  Fast math, extensive
// (times V, so no division by V): the mass U_rho, the velocities, the
// pressure word P = 2 E - |m|^2 / U_rho = V 2 p / (gamma - 1), which the flux
// uses as is (twice the internal energy) and the Hancock step extrapolates
// like p (it is linear in p), and the sixth word 1 / (k U_rho) = irk / U_rho
// with k = 2 / (gamma - 1), so that the Hancock step's grad(p) / rho is that
// word times grad(P).  Every face state, and so every flux, is then V times
// the intensive one; the sound speed sqrt(gamma (gamma - 1) / 2) |P| /
// sqrt(P U_rho) does not change.  FP32 keeps P U_rho ~ V^2 normal and
// finite for 1e-5 <= ds <= 1e4 (checked at create).
template <typename T, bool FAST>
__device__ __forceinline__ void ring_cell(const T u[5], const Consts<T> &C, T w[6])
{
    if constexpr (!FAST) {
        to_prim<T, false>(u, C.V, C.invV, C.gm1, w, C.pow2);
        w[5] = T(0);
    } else {
        const T im = drcp<T>(u[0]);
        const T ux = u[1] * im, uy = u[2] * im, uz = u[3] * im;
        w[0] = u[0]; w[1] = ux; w[2] = uy; w[3] = uz;
        const T ke2 = dfma<T>(uz, u[3], dfma<T>(uy, u[2], ux * u[1]));   // |m|^2 / U_rho
        w[4] = dfma<T>(T(2), u[4], -ke2);
        w[5] = C.irk * im;
    }
}

// Slope of one field along one axis from the primitives m (at -1), c, p (at
// +1).  Exact math: the gradient of calc_gradients (physics.rs:206-208) or
// the limited slope / ds.  Fast math: the same scaled by 2 ds (unlimited) or
// ds (limited), the factor folded into the constants that consume it.
template <typename T, bool FAST, int LIM>
__device__ __forceinline__ T slope(T m, T c, T p, const Consts<T> &C)
{
    if constexpr (LIM == LIM_NONE) {
        if constexpr (!FAST) return C.pow2 ? (p - m) * C.inv_two_ds : (p - m) / C.two_ds;
        else return p - m;
    } else {
        if constexpr (!FAST) return C.pow2 ? limited<T>(c - m, p - c, LIM) * C.inv_ds : limited<T>(c - m, p - c, LIM) / C.ds;
        else if constexpr (LIM == LIM_MINMOD && sizeof(T) == sizeof(float))
            return __builtin_amdgcn_fmed3f(c - m, p - c, 0.0f);   // minmod(a, b) = median(a, b, 0)
        else return limited<T>(c - m, p - c, LIM);
    }
}

// Where a tile cell takes its state from (precomputed once per thread: the
// mapping does not depend on the plane).  r: flat region index (row*RZ+col)
// whose slopes and extrapolated state are used; for a wall ghost along y
// (mask bit 0) or z (bit 1) the normal velocity of xtr flips and the normal
// slope is evaluated at the ghost's own position ry_wall / rz_wall.
struct EvalDesc { int r, ry_wall, rz_wall, mask, walls; };   // walls: some lane of the wave has mask != 0

__device__ __forceinline__ EvalDesc make_desc(int py, int pz, int y0, int z0, const StepArgs &A, int ty_max)
{
    const int qy = eval_local(py, y0, A.n[1], A.bc[2], A.bc[3], ty_max);
    const int qz = eval_local(pz, z0, A.n[2], A.bc[4], A.bc[5], TZ);
    EvalDesc d;
    d.r = (qy + 2) * RZ + (qz + 2);
    d.mask = 0;
    d.walls = 1;
    d.ry_wall = d.rz_wall = d.r;
    if (wall_ghost(py, y0, A.n[1], A.bc[2], A.bc[3])) { d.mask |= 1; d.ry_wall = (py + 2) * RZ + (qz + 2); }
    if (wall_ghost(pz, z0, A.n[2], A.bc[4], A.bc[5])) { d.mask |= 2; d.rz_wall = (qy + 2) * RZ + (pz + 2); }
    return d;
}

// Slopes and Hancock-extrapolated state (extrapolate_primitives_t,
// physics.rs:217-282) of a cell described by `d`, from ring slots sm, s0, sp.
// g[f*3+d] is the slope of field f along axis d.
template <typename T>
__device__ __forceinline__ void ld_cell(const T *p, T v[6])
{
    const vec2<T> *q = reinterpret_cast<const vec2<T> *>(p);
    const vec2<T> a = q[0], b = q[1], c = q[2];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y;
}
// the 5 primitives only (a neighbour's 1/rho is not needed)
template <typename T>
__device__ __forceinline__ void ld_prims(const T *p, T v[6])
{
    const vec2<T> *q = reinterpret_cast<const vec2<T> *>(p);
    const vec2<T> a = q[0], b = q[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = p[4]; v[5] = T(0);
}

// Hancock half-step (extrapolate_primitives_t, physics.rs:217-282) of a
// cell with primitives w, slopes g (g[f*3+d]: field f along axis d) and, in
// fast math, the ring's sixth word ir (w[4] then holds P = 2 p / (gamma - 1),
// which the step extrapolates like p, and ir grad P = grad p / rho).
template <typename T, bool FAST>
__device__ __forceinline__ void hancock(const T w[5], T ir, const T g[15], const Consts<T> &C, T xtr[5])
{
    const T r = w[0], ux = w[1], uy = w[2], uz = w[3], p = w[4];
    const T drx = g[0], dry = g[1], drz = g[2];
    const T dxx = g[3], dxy = g[4], dxz = g[5];
    const T dyx = g[6], dyy = g[7], dyz = g[8];
    const T dzx = g[9], dzy = g[10], dzz = g[11];
    const T dpx = g[12], dpy = g[13], dpz = g[14];
    const T hd = C.half_dt;
    if constexpr (!FAST) {
        // physics.rs:251-278, same association
        (void)ir;
        xtr[0] = r - hd * (ux * drx + r * dxx + uy * dry + r * dyy + uz * drz + r * dzz);
        xtr[1] = ux - hd * (ux * dxx + uy * dxy + uz * dxz + (T(1.0) / r) * dpx);
        xtr[2] = uy - hd * (ux * dyx + uy * dyy + uz * dyz + (T(1.0) / r) * dpy);
        xtr[3] = uz - hd * (ux * dzx + uy * dzy + uz * dzz + (T(1.0) / r) * dpz);
        xtr[4] = p - hd * (C.gamma * p * (dxx + dyy + dzz) + ux * dpx + uy * dpy + uz * dpz);
    } else {
        // hd here is dt / 2 over the slopes' scale (2 ds or ds): slopes are undivided differences
        const T div = dxx + dyy + dzz;
        xtr[0] = dfma<T>(-hd, dfma<T>(uz, drz, dfma<T>(uy, dry, dfma<T>(ux, drx, r * div))), r);
        xtr[1] = dfma<T>(-hd, dfma<T>(ir, dpx, dfma<T>(uz, dxz, dfma<T>(uy, dxy, ux * dxx))), ux);
        xtr[2] = dfma<T>(-hd, dfma<T>(ir, dpy, dfma<T>(uz, dyz, dfma<T>(uy, dyy, ux * dyx))), uy);
        xtr[3] = dfma<T>(-hd, dfma<T>(ir, dpz, dfma<T>(uz, dzz, dfma<T>(uy, dzy, ux * dzx))), uz);
        xtr[4] = dfma<T>(-hd, dfma<T>(uz, dpz, dfma<T>(uy, dpy, dfma<T>(ux, dpx, C.gamma * p * div))), p);
    }
}

template <typename T, int TY, bool FAST, int LIM>
__device__ __forceinline__ void eval_cell(const StepLDS<T, TY> &L, int sm, int s0, int sp, const EvalDesc &d,
                                          const Consts<T> &C, T xtr[5], T g[15])
{
    constexpr int CS = StepLDS<T, TY>::CS, PS = StepLDS<T, TY>::PS;
    const T *W0 = L.W + s0 * PS;
    const T *c0 = W0 + d.r * CS;
    T cc[6], xm[6], xp[6], ym[6], yp[6], zm[6], zp[6];
    ld_cell<T>(c0, cc);
    ld_prims<T>(L.W + sm * PS + d.r * CS, xm);
    ld_prims<T>(L.W + sp * PS + d.r * CS, xp);
    ld_prims<T>(c0 - RZ * CS, ym);
    ld_prims<T>(c0 + RZ * CS, yp);
    ld_prims<T>(c0 - CS, zm);
    ld_prims<T>(c0 + CS, zp);
    T w[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        const T c = cc[f];
        w[f] = c;
        g[f * 3 + 0] = slope<T, FAST, LIM>(xm[f], c, xp[f], C);
        g[f * 3 + 1] = slope<T, FAST, LIM>(ym[f], c, yp[f], C);
        g[f * 3 + 2] = slope<T, FAST, LIM>(zm[f], c, zp[f], C);
    }
    hancock<T, FAST>(w, cc[5], g, C, xtr);
    if (!d.walls) return;   // wave-uniform: no lane evaluates a wall ghost
    if (d.mask & 1) {   // wall ghost along y: mirrored xtr, slope on the ghost
        xtr[2] = -xtr[2];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const int o = d.ry_wall * CS + f;
            g[f * 3 + 1] = slope<T, FAST, LIM>(W0[o - RZ * CS], W0[o], W0[o + RZ * CS], C);
        }
    }
    if (d.mask & 2) {
        xtr[3] = -xtr[3];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const int o = d.rz_wall * CS + f;
            g[f * 3 + 2] = slope<T, FAST, LIM>(W0[o - CS], W0[o], W0[o + CS], C);
        }
    }
}

// Face state of a cell along axis d: xtr + sign * slope * ds/2
// (physics.rs:313-336).  Slopes are read as g[f*stride + d*(stride==3)]:
// stride 3 for the full slope set, 1 for a saved single-axis slope.
template <typename T, bool FAST>
__device__ __forceinline__ void face_state(const T xtr[5], const T *g, int d, bool plus, const Consts<T> &C,
                                           T s[5], int stride = 3)
{
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        const T gv = stride == 3 ? g[f * 3 + d] : g[f];
        if constexpr (!FAST) {
            const T q = gv * C.half_ds;
            s[f] = plus ? xtr[f] + q : xtr[f] - q;
        } else {
            s[f] = dfma<T>(plus ? C.half_ds : -C.half_ds, gv, xtr[f]);
        }
    }
}

// Flux through a face along `axis` from unrotated states (l on the + side).
template <typename T, bool FAST>
__device__ __forceinline__ void face_flux(const T l[5], const T r[5], int axis, const Consts<T> &C, bool grav,
                                          T h, T out[5])
{
    T lr[5], rr[5], F[5];
#pragma unroll
    for (int s = 0; s < 5; ++s) { lr[s] = l[rot_field(axis, s)]; rr[s] = r[rot_field(axis, s)]; }
    rusanov<T, FAST>(lr, rr, F, C.gamma, C.gm1, C.ds, grav, C.G, h, C.hgm1, C.sK);
#pragma unroll
    for (int s = 0; s < 5; ++s) out[rot_field(axis, s)] = F[s];
}

// Gravity height of physics.rs:368 for the cell with interior z index z.
template <typename T>
__device__ __forceinline__ T grav_h(int z, int nz, T ds)
{
    const int kg = z + 2, nzg = nz + 4;
    return (T(nzg) - T(2.0) * T(2)) * ds - (T(kg) - T(2.0)) * ds;
}

// Region cells (tile plus two-cell halo) a thread loads for every plane:
// flat LDS index, (y, z) offset of the mapped source cell, and flips.
template <int TY>
struct LoadPlan {
    static constexpr int NT = TY * TZ;
    static constexpr int NREG = (TY + 4) * RZ;
    static constexpr int NS = (NREG + NT - 1) / NT;
    int lds[NS];    // flat LDS index | flips << 12 (bit 13: y, bit 14: z); -1: none
    int off[NS];
};

template <int TY>
__device__ __forceinline__ void make_plan(LoadPlan<TY> &P, int tid, int y0, int z0, const StepArgs &A)
{
    constexpr int NT = LoadPlan<TY>::NT, NREG = LoadPlan<TY>::NREG;
#pragma unroll
    for (int q = 0; q < LoadPlan<TY>::NS; ++q) {
        const int c = tid + q * NT;
        if (c >= NREG) { P.lds[q] = -1; P.off[q] = 0; continue; }
        int ry, rz;
        if (c < NT) { ry = (c >> 6) + 2; rz = (c & 63) + 2; }   // the thread's own cell first
        else {
            const int h = c - NT;
            if (h < 4 * RZ) {
                const int rr = h / RZ;
                ry = rr < 2 ? rr : TY + rr;   // rows 0,1, TY+2, TY+3
                rz = h % RZ;
            } else {
                const int hh = h - 4 * RZ;
                ry = 2 + (hh >> 2);
                const int cc = hh & 3;
                rz = cc < 2 ? cc : TZ + cc;   // cols 0,1, 66,67
            }
        }
        const AxisMap my = map_axis(y0 + ry - 2, A.n[1], A.bc[2], A.bc[3]);
        const AxisMap mz = map_axis(z0 + rz - 2, A.n[2], A.bc[4], A.bc[5]);
        P.lds[q] = (ry * RZ + rz) | (my.flip ? 1 << 13 : 0) | (mz.flip ? 1 << 14 : 0);
        P.off[q] = my.idx * A.n[2] + mz.idx;
    }
}

// Load plane xg's primitives into ring slot `slot` (calc_primitives on the
// mapped cells; ghost positions take the boundary maps).  Split in two so
// the global loads can be issued one phase before their values are needed.
template <typename T, int TY>
struct PlaneLoad {
    T u[LoadPlan<TY>::NS][5];
    int xflip;
};

template <typename T, int TY>
__device__ __forceinline__ void issue_plane(PlaneLoad<T, TY> &Q, int xg, const LoadPlan<TY> &P, const StepArgs &A,
                                            const Fields<const T> &U)
{
    const AxisMap mx = map_axis(xg, A.n[0], A.bc[0], A.bc[1]);
    const uint32_t plane = (uint32_t)mx.idx * (uint32_t)A.n[1] * (uint32_t)A.n[2];
    Q.xflip = mx.flip;
#pragma unroll
    for (int q = 0; q < LoadPlan<TY>::NS; ++q) {
        if (P.lds[q] < 0) continue;
        const uint32_t idx = plane + (uint32_t)P.off[q];
#pragma unroll
        for (int f = 0; f < 5; ++f) Q.u[q][f] = at(U.f[f], idx);
    }
}

template <typename T, int TY, bool FAST>
__device__ __forceinline__ void finish_plane(StepLDS<T, TY> &L, int slot, const PlaneLoad<T, TY> &Q,
                                             const LoadPlan<TY> &P, const Consts<T> &C)
{
    constexpr int CS = StepLDS<T, TY>::CS, PS = StepLDS<T, TY>::PS;
    T *Ws = L.W + slot * PS;
#pragma unroll
    for (int q = 0; q < LoadPlan<TY>::NS; ++q) {
        // wall mirrors: rare, so skipped by the whole wave when no lane has one
        const bool flips = __any(Q.xflip || (P.lds[q] >= 0 && (P.lds[q] & (3 << 13)) != 0));
        if (P.lds[q] < 0) continue;
        T w[6];
        ring_cell<T, FAST>(Q.u[q], C, w);
        if (flips) {
            if (Q.xflip) w[1] = -w[1];
            if (P.lds[q] & (1 << 13)) w[2] = -w[2];
            if (P.lds[q] & (1 << 14)) w[3] = -w[3];
        }
        vec2<T> *o = reinterpret_cast<vec2<T> *>(Ws + (P.lds[q] & 0xfff) * CS);
        o[0] = vec2<T>{w[0], w[1]};
        o[1] = vec2<T>{w[2], w[3]};
        o[2] = vec2<T>{w[4], w[5]};
    }
}

template <typename T, int TY, bool FAST>
__device__ __forceinline__ void load_plane(StepLDS<T, TY> &L, int slot, int xg, const LoadPlan<TY> &P,
                                           const StepArgs &A, const Fields<const T> &U, const Consts<T> &C)
{
    PlaneLoad<T, TY> Q;
    issue_plane<T, TY>(Q, xg, P, A, U);
    finish_plane<T, TY, FAST>(L, slot, Q, P, C);
}

// Deterministic block reduction of (max den, nan flag, 5 sums) into partials.
template <int NT>
__device__ __forceinline__ void block_partials(double den, int nanf, const double sums[5], double *red,
                                               double *partials, int blk, int tid)
{
    double v[7] = {den, (double)nanf, sums[0], sums[1], sums[2], sums[3], sums[4]};
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v[0] = fmax(v[0], __shfl_xor(v[0], off));
        v[1] = fmax(v[1], __shfl_xor(v[1], off));
#pragma unroll
        for (int q = 2; q < 7; ++q) v[q] += __shfl_xor(v[q], off);
    }
    constexpr int NW = NT / 64;
    const int w = tid >> 6, lane = tid & 63;
    if (lane == 0)
        for (int q = 0; q < 7; ++q) red[w * 8 + q] = v[q];
    __syncthreads();
    if (tid == 0) {
        double o[7];
        for (int q = 0; q < 7; ++q) o[q] = red[q];
        for (int ww = 1; ww < NW; ++ww) {
            o[0] = fmax(o[0], red[ww * 8 + 0]);
            o[1] = fmax(o[1], red[ww * 8 + 1]);
            for (int q = 2; q < 7; ++q) o[q] += red[ww * 8 + q];
        }
        double *P = partials + (size_t)blk * PARTIAL_STRIDE;
        for (int q = 0; q < 7; ++q) P[q] = o[q];
        P[7] = 0.0;
    }
}

// One tile (TY x 64 columns, planes [bx*lx, min(bx*lx+lx, nx))) of one
// grid: the whole fused step for those cells; writes the tile's partial
// (max wave-speed denominator, NaN flag) to partial_out.  Called by the
// single-grid kernel and by the persistent ensemble kernel.
template <typename T, int TY, bool FAST, bool BULK, int LIM>
__device__ __forceinline__ void step_tile(StepLDS<T, TY> &L, double *red, const StepArgs &A,
                                          const Fields<const T> &U, const Fields<T> &Vout,
                                          const Scalars *__restrict__ sc, double *__restrict__ partial_out,
                                          int bx, int by, int bz)
{
    constexpr int NT = TY * TZ;
    const int tid = threadIdx.x;
    // the tile row is the wave index: uniform across the wave (scalar register)
    const int ty = __builtin_amdgcn_readfirstlane(tid >> 6), tz = tid & 63;
    const int nx = A.n[0], ny = A.n[1], nz = A.n[2];
    const int x0 = bx * A.lx, y0 = by * TY, z0 = bz * TZ;
    const int xend = min(x0 + A.lx, nx);
    const int yc = y0 + ty, zc = z0 + tz;
    const bool valid = yc < ny && zc < nz;
    const uint32_t plane_sz = (uint32_t)ny * (uint32_t)nz;   // n^3 * sizeof(T) < 4 GiB (checked at create)
    const uint32_t col = (uint32_t)yc * (uint32_t)nz + (uint32_t)zc;

    const Consts<T> C = make_consts<T, FAST>(A, T(sc->dt));

    if (!sc->active) {
        // inactive step of a captured run: carry the state through unchanged
        if (valid)
            for (int x = x0; x < xend; ++x)
                for (int f = 0; f < 5; ++f)
                    Vout.f[f][(uint32_t)x * plane_sz + col] = U.f[f][(uint32_t)x * plane_sz + col];
        return;
    }

    // plane-independent addressing, evaluated once
    LoadPlan<TY> plan;
    make_plan<TY>(plan, tid, y0, z0, A);
    EvalDesc own = make_desc(ty, tz, y0, z0, A, TY);
    EvalDesc halo = own;
    if (ty == 0) halo = make_desc(-1, tz, y0, z0, A, TY);
    else if (ty == 1) halo = make_desc(TY, tz, y0, z0, A, TY);
    else if (ty == 2 && tz < 2 * TY) halo = make_desc(tz < TY ? tz : tz - TY, tz < TY ? -1 : TZ, y0, z0, A, TY);
    own.walls = __any(own.mask != 0);
    halo.walls = __any(halo.mask != 0);
    const bool zgrav = A.gravity != 0;
    const T hz = zgrav ? grav_h<T>(zc, nz, C.ds) : T(0);
    // periodic: the reference copies the lower z face from face (n-1, n)
    // (boundaries.rs:174-178 on fluxes), gravity height included
    const T hzb = zgrav ? grav_h<T>((z0 == 0 && A.bc[4] == BC_PERIODIC) ? nz - 1 : z0 - 1, nz, C.ds) : T(0);
    const bool ylo_out = yc == 0 && A.bc[2] == BC_OUTFLOW;
    const bool zlo_tile = z0 == 0 && A.bc[4] == BC_OUTFLOW;   // uniform
    const bool zlo_out = zlo_tile && tz == 0;

    // carried values of the own column
    T xprev[5], gxprev[5];                      // xtr and x-slope of plane t-1
    T fxlo[5], fyu[5], fyl[5], fzu[5], fzl[5];  // flux parts of plane t-1's cell
#pragma unroll
    for (int f = 0; f < 5; ++f) { xprev[f] = gxprev[f] = fxlo[f] = fyu[f] = fyl[f] = fzu[f] = fzl[f] = T(0); }
    T maxden = T(0);
    double dmaxden = 0.0;
    int nanf = 0;

    // the ring holds planes t-1, t, t+1 when plane t's step starts; plane
    // t+2 is loaded during the flux phase of plane t (its slot held t-1,
    // whose last reader is phase B of plane t).  Plane t sits in slot
    // (t - x0 + 1) mod 3, so the march is unrolled by three with
    // compile-time slots.
    load_plane<T, TY, FAST>(L, 2, x0 - 2, plan, A, U, C);
    load_plane<T, TY, FAST>(L, 0, x0 - 1, plan, A, U, C);
    load_plane<T, TY, FAST>(L, 1, x0, plan, A, U, C);
    __syncthreads();

    auto plane_step = [&](auto slot, const int t) __attribute__((always_inline)) {
        constexpr int s0 = decltype(slot)::value, sm = (s0 + 2) % 3, sp = (s0 + 1) % 3;
        const bool full = t >= x0 && t < xend;

        // ---- B: halo face states (waves 0-2), then the own cell
        if (full && (ty < 2 || (ty == 2 && tz < 2 * TY))) {
            T hx[5], hg[15], hs[5];
            eval_cell<T, TY, FAST, LIM>(L, sm, s0, sp, halo, C, hx, hg);
            if (ty == 0) {                     // row -1: +y state
                face_state<T, FAST>(hx, hg, 1, true, C, hs);
#pragma unroll
                for (int f = 0; f < 5; ++f) L.Yp[f][tz] = hs[f];
            } else if (ty == 1) {              // row TY: -y state
                face_state<T, FAST>(hx, hg, 1, false, C, hs);
#pragma unroll
                for (int f = 0; f < 5; ++f) L.Ym[f][TY][tz] = hs[f];
            } else if (tz < TY) {              // column -1: +z state
                face_state<T, FAST>(hx, hg, 2, true, C, hs);
#pragma unroll
                for (int f = 0; f < 5; ++f) L.Zp[f][tz] = hs[f];
            } else {                           // column 64: -z state
                face_state<T, FAST>(hx, hg, 2, false, C, hs);
#pragma unroll
                for (int f = 0; f < 5; ++f) L.Zm[f][tz - TY] = hs[f];
            }
        }
        T xtr[5], g[15];
        eval_cell<T, TY, FAST, LIM>(L, sm, s0, sp, own, C, xtr, g);
        T szm[5];
        face_state<T, FAST>(xtr, g, 2, false, C, szm);
        if (full) {
            T sym[5];
            face_state<T, FAST>(xtr, g, 1, false, C, sym);
#pragma unroll
            for (int f = 0; f < 5; ++f) L.Ym[f][ty][tz] = sym[f];
            if (tz == 0) {
#pragma unroll
                for (int f = 0; f < 5; ++f) L.Z0[f][ty] = szm[f];
            }
        }
        __syncthreads();

        // ---- C: fluxes through the own cell's +x (of plane t-1), +y, +z faces;
        // the loads of plane t+2 are issued first so their latency overlaps
        // the flux arithmetic
        const bool pre = t + 2 <= xend + 1;
        PlaneLoad<T, TY> next;
        if (pre) issue_plane<T, TY>(next, t + 2, plan, A, U);
        T fxu[5] = {T(0), T(0), T(0), T(0), T(0)};  // face (t-1, t) along x
        const bool xbnd = (t == 0 && A.bc[0] == BC_WALL) || (t == nx && A.bc[1] != BC_PERIODIC);
        if (t >= x0 && !xbnd) {
            T r[5], l[5];
            face_state<T, FAST>(xprev, gxprev, 0, true, C, r, 1);
            face_state<T, FAST>(xtr, g, 0, false, C, l);
            face_flux<T, FAST>(l, r, 0, C, false, T(0), fxu);
        } else if (t >= x0) {                        // a domain-boundary plane (uniform branch)
            T r[5], l[5];
            if (t == 0) {                            // wall: ghost plane -1 mirrors plane 0
                T m[5] = {xtr[0], -xtr[1], xtr[2], xtr[3], xtr[4]};
                face_state<T, FAST>(m, gxprev, 0, true, C, r, 1);
                face_state<T, FAST>(xtr, g, 0, false, C, l);
            } else if (A.bc[1] == BC_OUTFLOW) {      // ghost nx copies plane nx-1
                face_state<T, FAST>(xprev, gxprev, 0, true, C, r, 1);
                face_state<T, FAST>(xprev, gxprev, 0, false, C, l, 1);
            } else {                                 // wall: ghost nx mirrors plane nx-1
                T m[5] = {xprev[0], -xprev[1], xprev[2], xprev[3], xprev[4]};
                face_state<T, FAST>(xprev, gxprev, 0, true, C, r, 1);
                face_state<T, FAST>(m, g, 0, false, C, l);
            }
            face_flux<T, FAST>(l, r, 0, C, false, T(0), fxu);
        }
        T fyu_n[5], fzu_n[5];
        if (full) {
            T r[5], l[5];
            face_state<T, FAST>(xtr, g, 1, true, C, r);
#pragma unroll
            for (int f = 0; f < 5; ++f) l[f] = L.Ym[f][ty + 1][tz];
            face_flux<T, FAST>(l, r, 1, C, false, T(0), fyu_n);
#pragma unroll
            for (int f = 0; f < 5; ++f) L.Ym[f][ty + 1][tz] = fyu_n[f];   // read above by this thread only
            face_state<T, FAST>(xtr, g, 2, true, C, r);
#pragma unroll
            for (int f = 0; f < 5; ++f) l[f] = from_next_lane<T>(szm[f], L.Zm[f][ty]);
            face_flux<T, FAST>(l, r, 2, C, zgrav, hz, fzu_n);
            // the tile's lower y faces (wave 3) and lower z faces (wave 0):
            // waves on the SIMDs that evaluated no halo row in phase B
            if (ty == 3) {
                T rb[5], lb[5], Fb[5];
#pragma unroll
                for (int f = 0; f < 5; ++f) { rb[f] = L.Yp[f][tz]; lb[f] = L.Ym[f][0][tz]; }
                face_flux<T, FAST>(lb, rb, 1, C, false, T(0), Fb);
#pragma unroll
                for (int f = 0; f < 5; ++f) L.Ym[f][0][tz] = Fb[f];
            } else if (ty == 0 && tz < TY) {
                T rb[5], lb[5], Fb[5];
#pragma unroll
                for (int f = 0; f < 5; ++f) { rb[f] = L.Zp[f][tz]; lb[f] = L.Z0[f][tz]; }
                face_flux<T, FAST>(lb, rb, 2, C, zgrav, hzb, Fb);
#pragma unroll
                for (int f = 0; f < 5; ++f) L.Z0[f][tz] = Fb[f];
            }
        }
#pragma unroll
        for (int f = 0; f < 5; ++f) { xprev[f] = xtr[f]; gxprev[f] = g[f * 3 + 0]; }
        if (pre) finish_plane<T, TY, FAST>(L, sm, next, plan, C);
        const int xc = t - 1;
        const bool upd = xc >= x0 && xc < xend && valid;
        const uint32_t uidx = (uint32_t)xc * plane_sz + col;
        T uold[5];
        if (upd) {
#pragma unroll
            for (int f = 0; f < 5; ++f) uold[f] = at(U.f[f], uidx);
        }
        __syncthreads();

        // ---- D: lower y/z fluxes, finish cell t-1, carry plane t
        T fyl_n[5], fzl_n[5];
        if (full) {
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                fyl_n[f] = L.Ym[f][ty][tz];
                fzl_n[f] = from_prev_lane<T>(fzu_n[f], L.Z0[f][ty]);
            }
            if (ylo_out) {                           // boundaries.rs:137-141 on fluxes
#pragma unroll
                for (int f = 0; f < 5; ++f) fyl_n[f] = fyu_n[f];
            }
            if (zlo_tile) {                          // boundaries.rs:167-171 on fluxes
#pragma unroll
                for (int f = 0; f < 5; ++f) fzl_n[f] = zlo_out ? fzu_n[f] : fzl_n[f];
            }
        }
        if (upd) {
            // outflow at the lower x face: the first cell's lower face copies
            // its upper face (boundaries.rs:107-111 on fluxes)
            const bool xlo_out = xc == 0 && A.bc[0] == BC_OUTFLOW;
            if constexpr (!FAST) {
                if (xlo_out) {
#pragma unroll
                    for (int f = 0; f < 5; ++f) fxlo[f] = fxu[f];
                }
            }
            T un[5];
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const T u = uold[f];
                T acc;
                if constexpr (!FAST) acc = ((((fxu[f] - fxlo[f]) + fyu[f]) - fyl[f]) + fzu[f]) - fzl[f];
                else acc = dfma<T>(xlo_out ? T(0) : T(1), fxu[f], fyu[f]);   // fyu: y/z sum minus the lower x face
                un[f] = FAST ? dfma<T>(-C.dtds, acc, u) : u - C.dtds * acc;
                at(Vout.f[f], uidx) = un[f];
            }
            if constexpr (BULK) {
                T w[5];
                to_prim<T, FAST>(un, C.V, C.invV, C.gm1, w, C.pow2);
                double wd[5];
                for (int f = 0; f < 5; ++f) wd[f] = (double)w[f];
                const double den = wave_den<double, false>(wd, (double)C.gamma);
                nanf |= den != den;
                dmaxden = fmax(dmaxden, den);   // NaN-ignoring
            } else {
                const T den = cell_den<T, FAST>(un, C.V, C.invV, C.gm1, C.gamma, C.pow2);
                nanf |= den != den;
                maxden = dmax<T>(maxden, den);  // NaN-ignoring
            }
        }
        if (full) {
            const bool xlo_next = t == 0 && A.bc[0] == BC_OUTFLOW;
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                if constexpr (!FAST) {
                    fxlo[f] = fxu[f];
                    fyu[f] = fyu_n[f]; fyl[f] = fyl_n[f]; fzu[f] = fzu_n[f]; fzl[f] = fzl_n[f];
                } else {
                    const T yz = ((fyu_n[f] - fyl_n[f]) + fzu_n[f]) - fzl_n[f];
                    fyu[f] = dfma<T>(xlo_next ? T(0) : T(-1), fxu[f], yz);
                }
            }
        }
    };
    for (int t = x0 - 1;;) {
        plane_step(std::integral_constant<int, 0>{}, t);
        if (++t > xend) break;
        plane_step(std::integral_constant<int, 1>{}, t);
        if (++t > xend) break;
        plane_step(std::integral_constant<int, 2>{}, t);
        if (++t > xend) break;
    }

    if constexpr (!BULK) dmaxden = (double)maxden;
    const double nosums[5] = {0, 0, 0, 0, 0};
    block_partials<NT>(dmaxden, nanf, nosums, red, partial_out, 0, tid);
}

// Occupancy target: the FP32 fast kernels fit 128 VGPRs (4 waves per SIMD),
// the FP64 fast kernels 256 (2 waves per SIMD).
template <typename T, bool FAST> constexpr int waves_per_eu() { return FAST ? (sizeof(T) == 4 ? 4 : 2) : 1; }

template <typename T, int TY, bool FAST, bool BULK, int LIM>
__global__ void __launch_bounds__(TY * TZ) __attribute__((amdgpu_waves_per_eu(waves_per_eu<T, FAST>())))
step_kernel(const StepArgs A, const Fields<const T> U, const Fields<T> Vout, const Scalars *__restrict__ sc,
            double *__restrict__ partials)
{
    __shared__ StepLDS<T, TY> L;
    __shared__ double red[(TY * TZ / 64) * 8];
#ifdef DG_EMU_POISON
    {   // CPU-emulation builds only: make every unwritten LDS read visible
        T *raw = reinterpret_cast<T *>(&L);
        for (size_t q = threadIdx.x; q < sizeof(L) / sizeof(T); q += TY * TZ) raw[q] = T(NAN);
        __syncthreads();
    }
#endif
    // tile of this workgroup (optionally spread so each XCD gets a
    // contiguous range of neighbouring tiles)
    int b = blockIdx.x;
    if (A.swizzle) b = (b & 7) * (A.nblocks >> 3) + (b >> 3);
    const int bz = b % A.nbz;
    const int by = (b / A.nbz) % A.nby;
    const int bx = b / (A.nbz * A.nby);
    step_tile<T, TY, FAST, BULK, LIM>(L, red, A, U, Vout, sc, partials + (size_t)blockIdx.x * PARTIAL_STRIDE,
                                      bx, by, bz);
}

// Persistent ensemble step: workgroups take (instance, tile) items from a
// work list sorted by cost, largest first, through an atomic counter (dynamic
// load balance); every instance has its own grid, state buffers, scalars and
// partial slots.
template <typename T, int TY, bool FAST, bool BULK, int LIM, int W>
__global__ void __launch_bounds__(TY * TZ) __attribute__((amdgpu_waves_per_eu(W)))
ens_step_kernel(const EnsInstance *__restrict__ inst, const EnsItem *__restrict__ items, int nitems, int parity,
                double *__restrict__ partials, int *__restrict__ next)
{
    __shared__ StepLDS<T, TY> L;
    __shared__ double red[(TY * TZ / 64) * 8];
    __shared__ int s_item;
    for (;;) {
        // next item of the (heaviest-first) work list; the counter is reset
        // by the dt kernel that follows every step launch
        if (threadIdx.x == 0) s_item = atomicAdd(next, 1);
        __syncthreads();
        const int it = s_item;
        if (it >= nitems) break;
        const EnsItem w = items[it];
        const EnsInstance &I = inst[w.inst];
        const StepArgs A = I.args;
        Fields<const T> U;
        Fields<T> V;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            U.f[f] = static_cast<const T *>(I.buf[parity][f]);
            V.f[f] = static_cast<T *>(I.buf[1 - parity][f]);
        }
        step_tile<T, TY, FAST, BULK, LIM>(L, red, A, U, V, I.sc, partials + (size_t)w.slot * PARTIAL_STRIDE,
                                          w.bx, w.by, w.bz);
        __syncthreads();   // LDS and the reduction scratch are reused by the next item
    }
}

// --------------------------------------------------------- small kernels --
template <typename T, bool BULK, bool FAST>
__global__ void __launch_bounds__(256)
reduce_kernel(const StepArgs A, const Fields<const T> U, double *__restrict__ partials)
{
    __shared__ double red[4 * 8];
    const T ds = T(A.ds), gamma = T(A.gamma), gm1 = gamma - T(1.0);
    const T V = ds * ds * ds;
    const size_t n3 = (size_t)A.n[0] * A.n[1] * A.n[2];
    double sums[5] = {0, 0, 0, 0, 0};
    double dmaxden = 0.0;
    T maxden = T(0);
    int nanf = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        T u[5];
        for (int f = 0; f < 5; ++f) { u[f] = U.f[f][i]; sums[f] += (double)u[f]; }
        // same math as the step epilogue
        if constexpr (BULK) {
            T w[5];
            to_prim<T, FAST>(u, V, T(1) / V, gm1, w);
            double wd[5];
            for (int f = 0; f < 5; ++f) wd[f] = (double)w[f];
            const double den = wave_den<double, false>(wd, (double)gamma);
            if (den != den) nanf = 1; else dmaxden = fmax(dmaxden, den);
        } else {
            const T den = cell_den<T, FAST>(u, V, T(1) / V, gm1, gamma);
            if (den != den) nanf = 1; else maxden = dmax<T>(maxden, den);
        }
    }
    if constexpr (!BULK) dmaxden = (double)maxden;
    block_partials<256>(dmaxden, nanf, sums, red, partials, blockIdx.x, threadIdx.x);
}

// dt from the partials (calc_timestep, physics.rs:177-196: CFL * min over
// cells of ds / den == CFL * ds / max den), conservation sums, and the run
// clock (main.rs:64-66).  One workgroup of NT threads; fixed reduction
// order.  mode 0: dt and sums of a fresh state; 1: after a step (advance the
// clock, new dt); 2: sums only.
template <typename T, bool BULK, int NT>
__device__ __forceinline__ void finalize_body(double ds, Scalars *__restrict__ sc, const double *__restrict__ partials,
                                              int np, double cfl, int nan_fallback, int mode, double *red)
{
    const int tid = threadIdx.x;
    double den = 0.0, nanf = 0.0, s[5] = {0, 0, 0, 0, 0};
    for (int i = tid; i < np; i += NT) {
        const double *P = partials + (size_t)i * PARTIAL_STRIDE;
        den = fmax(den, P[0]);
        nanf = fmax(nanf, P[1]);
        for (int q = 0; q < 5; ++q) s[q] += P[2 + q];
    }
    double v[7] = {den, nanf, s[0], s[1], s[2], s[3], s[4]};
    for (int off = 32; off >= 1; off >>= 1) {
        v[0] = fmax(v[0], __shfl_xor(v[0], off));
        v[1] = fmax(v[1], __shfl_xor(v[1], off));
        for (int q = 2; q < 7; ++q) v[q] += __shfl_xor(v[q], off);
    }
    if ((tid & 63) == 0)
        for (int q = 0; q < 7; ++q) red[(tid >> 6) * 8 + q] = v[q];
    __syncthreads();
    if (tid != 0) return;
    double o[7];
    for (int q = 0; q < 7; ++q) o[q] = red[q];
    for (int w = 1; w < NT / 64; ++w) {
        o[0] = fmax(o[0], red[w * 8 + 0]);
        o[1] = fmax(o[1], red[w * 8 + 1]);
        for (int q = 2; q < 7; ++q) o[q] += red[w * 8 + q];
    }
    if (mode == 2) {
        for (int q = 0; q < 5; ++q) sc->sums[q] = o[2 + q];
        return;
    }
    if (mode == 1) {
        if (!sc->active) return;
        if constexpr (!BULK) sc->t = (double)(T(sc->t) + T(sc->dt));
        else sc->t = sc->t + sc->dt;
        sc->steps += 1;
    }
    const bool nan = o[1] != 0.0;
    double dt;
    if constexpr (!BULK) {
        const T minval = nan ? (nan_fallback ? T(1.0) : T(NAN)) : T(ds) / T(o[0]);
        dt = (double)(T(cfl) * minval);
    } else {
        const double minval = nan ? (nan_fallback ? 1.0 : NAN) : ds / o[0];
        dt = cfl * minval;
    }
    sc->dt = dt;
    sc->nonfinite = nan ? 1 : 0;
    sc->maxden = o[0];
    if (mode == 0)
        for (int q = 0; q < 5; ++q) sc->sums[q] = o[2 + q];
    sc->active = !((sc->t > sc->t_end) || (sc->steps >= sc->max_steps));
}

template <typename T, bool BULK>
__global__ void __launch_bounds__(1024)
finalize_kernel(const StepArgs A, Scalars *__restrict__ sc, const double *__restrict__ partials, int np,
                double cfl, int nan_fallback, int mode)
{
    __shared__ double red[16 * 8];
    finalize_body<T, BULK, 1024>(A.ds, sc, partials, np, cfl, nan_fallback, mode, red);
}

// One workgroup per ensemble instance: its partial slots -> its scalars.
template <typename T, bool BULK>
__global__ void __launch_bounds__(256)
ens_finalize_kernel(const EnsInstance *__restrict__ inst, const double *__restrict__ partials, int mode,
                    int *__restrict__ next)
{
    __shared__ double red[4 * 8];
    if (blockIdx.x == 0 && threadIdx.x == 0) *next = 0;   // work counter of the next step launch
    const EnsInstance &I = inst[blockIdx.x];
    finalize_body<T, BULK, 256>(I.args.ds, I.sc, partials + (size_t)I.part_base * PARTIAL_STRIDE, I.nparts, I.cfl,
                                I.nan_fallback, mode, red);
}

// calc_conserved (physics.rs:143-175), in place on interior arrays holding W.
template <typename T>
__global__ void __launch_bounds__(256) prim_to_cons_kernel(const StepArgs A, const Fields<T> U)
{
    const T ds = T(A.ds), gm1 = T(A.gamma) - T(1.0);
    const T V = ds * ds * ds;
    const size_t n3 = (size_t)A.n[0] * A.n[1] * A.n[2];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        const T r = U.f[0][i], vx = U.f[1][i], vy = U.f[2][i], vz = U.f[3][i], p = U.f[4][i];
        U.f[0][i] = r * V;
        U.f[1][i] = vx * r * V;
        U.f[2][i] = vy * r * V;
        U.f[3][i] = vz * r * V;
        U.f[4][i] = V * (r * (vx * vx + vy * vy + vz * vz) / T(2.0) + p / gm1);
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
cons_to_prim_kernel(const StepArgs A, const Fields<const T> U, const Fields<T> W)
{
    const T ds = T(A.ds), gm1 = T(A.gamma) - T(1.0);
    const T V = ds * ds * ds;
    const size_t n3 = (size_t)A.n[0] * A.n[1] * A.n[2];
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (size_t)gridDim.x * blockDim.x) {
        T u[5], w[5];
        for (int f = 0; f < 5; ++f) u[f] = U.f[f][i];
        to_prim<T, false>(u, V, T(1) / V, gm1, w);
        for (int f = 0; f < 5; ++f) W.f[f][i] = w[f];
    }
}

// ------------------------------------------ learned flux correction --
// The network's output G (fp32, 16 channels per cell, channel 5d+q = the
// correction of flux component q on the face c + e_d/2, in units of
// out_scale[q]) corrects the plain update of a periodic FP32 grid:
//   U[c] += -dt ds^2 out_scale[q] sum_d (m(c, c+e_d) G[c][5d+q] - m(c-e_d, c) G[c-e_d][5d+q]),
// the face flux correction telescopes, so the totals are unchanged.  A cell
// whose corrected state would have rho <= 0 or p <= 0 is flagged; every face
// of a flagged cell keeps the plain flux (m = 0), so each face still carries
// one flux for both of its cells.  Keeping those faces plain changes the
// updates of the flagged cells' neighbours, so two more passes (run only when
// the first flagged a cell) flag every cell whose state is invalid under the
// current masks.
struct CorrScale { float s[5]; };

__device__ __forceinline__ long long wrap_index(int x, int y, int z, const StepArgs &A)
{
    x = x < 0 ? x + A.n[0] : (x >= A.n[0] ? x - A.n[0] : x);
    y = y < 0 ? y + A.n[1] : (y >= A.n[1] ? y - A.n[1] : y);
    z = z < 0 ? z + A.n[2] : (z >= A.n[2] ? z - A.n[2] : z);
    return ((long long)x * A.n[1] + y) * A.n[2] + z;
}

// G holds fp32 (host-given corrections) or bf16 (the network's output layer).
__device__ __forceinline__ float gval(const float *G, long long i) { return G[i]; }
__device__ __forceinline__ float gval(const __bf16 *G, long long i) { return (float)G[i]; }

// Corrected increment of cell (x, y, z); flags == nullptr: all faces corrected.
template <typename GT>
__device__ __forceinline__ void corr_delta(const StepArgs &A, const GT *__restrict__ G, const CorrScale &S,
                                           float coef, const unsigned char *__restrict__ flags, int x, int y, int z,
                                           float d[5])
{
    const long long c = wrap_index(x, y, z, A);
    const long long lo[3] = {wrap_index(x - 1, y, z, A), wrap_index(x, y - 1, z, A), wrap_index(x, y, z - 1, A)};
    const long long hi[3] = {wrap_index(x + 1, y, z, A), wrap_index(x, y + 1, z, A), wrap_index(x, y, z + 1, A)};
    const bool fc = flags && flags[c];
    float mh[3], ml[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        mh[a] = (fc || (flags && flags[hi[a]])) ? 0.0f : 1.0f;
        ml[a] = (fc || (flags && flags[lo[a]])) ? 0.0f : 1.0f;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        float div = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) div += mh[a] * gval(G, c * 16 + 5 * a + q) - ml[a] * gval(G, lo[a] * 16 + 5 * a + q);
        d[q] = coef * S.s[q] * div;
    }
}

__global__ void zero_count_kernel(int *cnt) { *cnt = 0; }

// MASKED = false: every cell of U gets its increment with every face
// corrected, in place, and B (5 fields of n3) keeps the plain state;
// flags[i] = corrected state of i invalid, cnt counts the flagged cells, and
// each block writes the wave-speed partial of the corrected state (the
// final one when no cell is flagged; launched with reduce_blocks blocks).
// MASKED = true (skipped when cnt is 0; U and partials untouched):
// flags[i] = prev[i] or the plain state B of i plus its increment with the
// faces of the cells flagged in prev kept plain is invalid.
template <bool FAST, bool MASKED, typename GT>
__global__ void __launch_bounds__(256)
corr_flag_kernel(const StepArgs A, const Fields<float> U, float *__restrict__ B, const GT *__restrict__ G,
                 const CorrScale S, const Scalars *__restrict__ sc, const unsigned char *__restrict__ prev,
                 unsigned char *__restrict__ flags, int *__restrict__ cnt, double *__restrict__ partials)
{
    if (MASKED && *cnt == 0) return;
    __shared__ double red[4 * 8];
    const float ds = float(A.ds), V = ds * ds * ds, gm1 = float(A.gamma) - 1.0f, gamma = float(A.gamma);
    const float coef = sc->active ? -float(sc->dt) * ds * ds : 0.0f;   // an inactive step changes nothing
    const long long n3 = (long long)A.n[0] * A.n[1] * A.n[2];
    float maxden = 0.0f;
    int nanf = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % A.n[2]), y = (int)((i / A.n[2]) % A.n[1]), x = (int)(i / ((long long)A.n[1] * A.n[2]));
        float d[5], u[5], w[5];
        corr_delta(A, G, S, coef, MASKED ? prev : nullptr, x, y, z, d);
        for (int q = 0; q < 5; ++q) {
            const float up = MASKED ? B[q * n3 + i] : U.f[q][i];
            if (!MASKED) B[q * n3 + i] = up;
            u[q] = up + d[q];
        }
        to_prim<float, FAST>(u, V, 1.0f / V, gm1, w);
        const bool bad = !(w[0] > 0.0f && w[4] > 0.0f);
        if (MASKED) {
            flags[i] = (prev[i] || bad) ? 1 : 0;
        } else {
            flags[i] = bad ? 1 : 0;
            if (bad) atomicAdd(cnt, 1);
            for (int q = 0; q < 5; ++q) U.f[q][i] = u[q];
            const float den = cell_den<float, FAST>(u, V, 1.0f / V, gm1, gamma);
            if (den != den) nanf = 1; else maxden = fmaxf(maxden, den);
        }
    }
    if (!MASKED) {
        const double nosums[5] = {0, 0, 0, 0, 0};
        block_partials<256>((double)maxden, nanf, nosums, red, partials, blockIdx.x, threadIdx.x);
    }
}

// Only when a cell is flagged (cnt > 0): U = B + the increments under the
// final flags, and the wave-speed partials of that state.  With no cell
// flagged the flag kernel's corrected state and partials stand.
template <bool FAST, typename GT>
__global__ void __launch_bounds__(256)
corr_apply_kernel(const StepArgs A, const Fields<float> U, const float *__restrict__ B, const GT *__restrict__ G,
                  const CorrScale S, const Scalars *__restrict__ sc, const unsigned char *__restrict__ flags,
                  const int *__restrict__ cnt, double *__restrict__ partials)
{
    if (*cnt == 0) return;
    __shared__ double red[4 * 8];
    const float ds = float(A.ds), V = ds * ds * ds, gm1 = float(A.gamma) - 1.0f, gamma = float(A.gamma);
    const float coef = sc->active ? -float(sc->dt) * ds * ds : 0.0f;
    const long long n3 = (long long)A.n[0] * A.n[1] * A.n[2];
    float maxden = 0.0f;
    int nanf = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % A.n[2]), y = (int)((i / A.n[2]) % A.n[1]), x = (int)(i / ((long long)A.n[1] * A.n[2]));
        float d[5], u[5];
        corr_delta(A, G, S, coef, flags, x, y, z, d);
        for (int q = 0; q < 5; ++q) {
            u[q] = B[q * n3 + i] + d[q];
            U.f[q][i] = u[q];
        }
        const float den = cell_den<float, FAST>(u, V, 1.0f / V, gm1, gamma);
        if (den != den) nanf = 1; else maxden = fmaxf(maxden, den);
    }
    const double nosums[5] = {0, 0, 0, 0, 0};
    block_partials<256>((double)maxden, nanf, nosums, red, partials, blockIdx.x, threadIdx.x);
}

// 2x restriction: coarse[I] = sum of the 8 children (U is extensive).
__global__ void __launch_bounds__(256)
restrict_kernel(const StepArgs Af, const Fields<const float> F, const StepArgs Ac, const Fields<float> C)
{
    const long long n3 = (long long)Ac.n[0] * Ac.n[1] * Ac.n[2];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x) {
        const int z = (int)(i % Ac.n[2]), y = (int)((i / Ac.n[2]) % Ac.n[1]), x = (int)(i / ((long long)Ac.n[1] * Ac.n[2]));
        for (int q = 0; q < 5; ++q) {
            float s = 0.0f;
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b)
                    for (int c = 0; c < 2; ++c)
                        s += F.f[q][((long long)(2 * x + a) * Af.n[1] + (2 * y + b)) * Af.n[2] + (2 * z + c)];
            C.f[q][i] = s;
        }
    }
}

// R = (Ut - Up) / (dt ds^2): the flux divergence the coarse update misses.
__global__ void __launch_bounds__(256)
residual_kernel(long long n3, const Fields<const float> Ut, const Fields<const float> Up, float inv, const Fields<float> R)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += (long long)gridDim.x * blockDim.x)
        for (int q = 0; q < 5; ++q) R.f[q][i] = (Ut.f[q][i] - Up.f[q][i]) * inv;
}

__global__ void set_dt_kernel(Scalars *sc, double dt) { sc->dt = dt; }

// ---------------------------------------------------- conducting walls --
// Heat through conducting lower walls (the CUDA variant's source,
// src/cu/include/kernels.h:561-661): the first interior cell along a
// conducting lower wall gains dt ds (ds Q/V) of energy, Q = (T_out - t_in) S /
// (d/kappa + 1/hc), S = ds^2, V = ds^3, d 0.008, kappa 0.2, hc = 10 - v_out +
// 10 sqrt(v_out).  t_in = p / (RS rho) of the wall ghost in the state the
// CUDA flux kernel reads, the half-step extrapolated one (physics.hpp:437-442
// passes rho_xtr / p_xtr): the ghost's extrapolated state is the mirror of the
// cell's own (handle_walls on xtr), so t_in comes from the cell's Hancock
// state, evaluated here exactly as the step kernel evaluates it.  Added after
// the step, axes in the order x, y, z; one thread per cell of the union of
// the wall layers.
struct HeatArgs { double t_out, v_out, rs; int mask; };

// Slopes and Hancock state of interior cell (i, j, k) from the conserved
// state U (neighbours through the boundary maps), as eval_cell computes them.
template <typename T, bool FAST, int LIM>
__device__ __forceinline__ void cell_xtr(const StepArgs &A, const Fields<const T> &U, const Consts<T> &C, int i,
                                         int j, int k, T xtr[5])
{
    T w[7][6];   // centre, x-, x+, y-, y+, z-, z+
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        int p[3] = {i, j, k};
        if (q > 0) p[(q - 1) / 2] += (q & 1) ? -1 : 1;
        const AxisMap mx = map_axis(p[0], A.n[0], A.bc[0], A.bc[1]);
        const AxisMap my = map_axis(p[1], A.n[1], A.bc[2], A.bc[3]);
        const AxisMap mz = map_axis(p[2], A.n[2], A.bc[4], A.bc[5]);
        const uint32_t idx = ((uint32_t)mx.idx * (uint32_t)A.n[1] + (uint32_t)my.idx) * (uint32_t)A.n[2] + (uint32_t)mz.idx;
        T u[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) u[f] = at(U.f[f], idx);
        ring_cell<T, FAST>(u, C, w[q]);
        if (mx.flip) w[q][1] = -w[q][1];
        if (my.flip) w[q][2] = -w[q][2];
        if (mz.flip) w[q][3] = -w[q][3];
    }
    T g[15];
#pragma unroll
    for (int f = 0; f < 5; ++f)
#pragma unroll
        for (int a = 0; a < 3; ++a) g[f * 3 + a] = slope<T, FAST, LIM>(w[1 + 2 * a][f], w[0][f], w[2 + 2 * a][f], C);
    hancock<T, FAST>(w[0], w[0][5], g, C, xtr);
}

template <typename T, bool FAST>
__global__ void __launch_bounds__(256)
heat_kernel(const StepArgs A, const Fields<const T> Uold, const Fields<T> Unew, const Scalars *__restrict__ sc,
            const HeatArgs H)
{
    if (!sc->active) return;
    const T dt = T(sc->dt);
    const Consts<T> C = make_consts<T, FAST>(A, dt);
    const T ds = T(A.ds);
    const T S = ds * ds, V = ds * ds * ds;
    const T vout = T(H.v_out);
    const T hc = T(10.0) - vout + T(10.0) * dsqrt<T>(vout);
    const T dk = T(0.008) / T(0.2);
    const int n0 = A.n[0], n1 = A.n[1], n2 = A.n[2];
    // cells of the union: layer x (i = 0), then y (j = 0, i > 0 if x is a wall layer), then z
    const long long lx = (H.mask & 1) ? (long long)n1 * n2 : 0;
    const long long ly = (H.mask & 2) ? (long long)n0 * n2 : 0;
    const long long lz = (H.mask & 4) ? (long long)n0 * n1 : 0;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < lx + ly + lz;
         e += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        if (e < lx) { i = 0; j = (int)(e / n2); k = (int)(e % n2); }
        else if (e < lx + ly) { const long long r = e - lx; j = 0; i = (int)(r / n2); k = (int)(r % n2); }
        else { const long long r = e - lx - ly; k = 0; i = (int)(r / n1); j = (int)(r % n1); }
        // the cell is handled by the first layer (x, y, z order) that contains it
        if (e >= lx && (H.mask & 1) && i == 0) continue;
        if (e >= lx + ly && (H.mask & 2) && j == 0) continue;
        const long long c = ((long long)i * n1 + j) * n2 + k;
        T xtr[5];
        if (A.limiter == LIM_MINMOD) cell_xtr<T, FAST, LIM_MINMOD>(A, Uold, C, i, j, k, xtr);
        else if (A.limiter == LIM_VANLEER) cell_xtr<T, FAST, LIM_VANLEER>(A, Uold, C, i, j, k, xtr);
        else cell_xtr<T, FAST, LIM_NONE>(A, Uold, C, i, j, k, xtr);
        // fast math: the ring's V 2 p / (gamma - 1) and V rho (the V cancels)
        const T p_xtr = FAST ? xtr[4] * C.hgm1 : xtr[4];
        const T t_in = p_xtr / (T(H.rs) * xtr[0]);
        const T Q = (T(H.t_out) - t_in) * (S / (dk + T(1.0) / hc));
        const T dE = dt * ds * ((Q / V) * ds);
        T E = Unew.f[4][c];
        const bool on[3] = {(H.mask & 1) && i == 0, (H.mask & 2) && j == 0, (H.mask & 4) && k == 0};
        for (int d = 0; d < 3; ++d)
            if (on[d]) E += dE;
        Unew.f[4][c] = E;
    }
}

// ---------------------------------------------------------------- launch --
template <typename T> static Fields<const T> cfields(const void *const p[5])
{
    Fields<const T> F;
    for (int f = 0; f < 5; ++f) F.f[f] = static_cast<const T *>(p[f]);
    return F;
}
template <typename T> static Fields<T> mfields(void *const p[5])
{
    Fields<T> F;
    for (int f = 0; f < 5; ++f) F.f[f] = static_cast<T *>(p[f]);
    return F;
}

template <typename T, int TY, bool FAST, bool BULK>
static void launch_step_t(const StepArgs &a, const void *const in[5], void *const out[5], const Scalars *sc,
                          double *partials, hipStream_t s)
{
    const dim3 grid(a.nblocks), block(TY * TZ);
    if (a.limiter == LIM_MINMOD)
        hipLaunchKernelGGL((step_kernel<T, TY, FAST, BULK, LIM_MINMOD>), grid, block, 0, s, a, cfields<T>(in),
                           mfields<T>(out), sc, partials);
    else if (a.limiter == LIM_VANLEER)
        hipLaunchKernelGGL((step_kernel<T, TY, FAST, BULK, LIM_VANLEER>), grid, block, 0, s, a, cfields<T>(in),
                           mfields<T>(out), sc, partials);
    else
        hipLaunchKernelGGL((step_kernel<T, TY, FAST, BULK, LIM_NONE>), grid, block, 0, s, a, cfields<T>(in),
                           mfields<T>(out), sc, partials);
}

int launch_step(int prec, int fast, int tile_y, const StepArgs &a, const void *const in[5], void *const out[5],
                const Scalars *sc, double *partials, hipStream_t s)
{
#define DG_CASE(T, TYV, F, B) launch_step_t<T, TYV, F, B>(a, in, out, sc, partials, s)
    if (prec == 0) {
        if (tile_y == 4) { if (fast) DG_CASE(double, 4, true, false); else DG_CASE(double, 4, false, false); }
        else if (tile_y == 8) { if (fast) DG_CASE(double, 8, true, false); else DG_CASE(double, 8, false, false); }
        else return -1;
    } else {
        const bool bulk = prec == 2;
        if (tile_y == 4) {
            if (fast) { if (bulk) DG_CASE(float, 4, true, true); else DG_CASE(float, 4, true, false); }
            else { if (bulk) DG_CASE(float, 4, false, true); else DG_CASE(float, 4, false, false); }
        } else if (tile_y == 8) {
            if (fast) { if (bulk) DG_CASE(float, 8, true, true); else DG_CASE(float, 8, true, false); }
            else { if (bulk) DG_CASE(float, 8, false, true); else DG_CASE(float, 8, false, false); }
        } else if (tile_y == 16) {
            if (fast) { if (bulk) DG_CASE(float, 16, true, true); else DG_CASE(float, 16, true, false); }
            else { if (bulk) DG_CASE(float, 16, false, true); else DG_CASE(float, 16, false, false); }
        } else return -1;
    }
#undef DG_CASE
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <typename T, int TY, bool FAST, bool BULK, int W>
static void launch_ens_tw(const EnsInstance *inst, const EnsItem *items, int nitems, int grid, int parity,
                          int limiter, double *partials, int *next, hipStream_t s)
{
    const dim3 g(grid), b(TY * TZ);
    if (limiter == LIM_MINMOD)
        hipLaunchKernelGGL((ens_step_kernel<T, TY, FAST, BULK, LIM_MINMOD, W>), g, b, 0, s, inst, items, nitems, parity, partials, next);
    else if (limiter == LIM_VANLEER)
        hipLaunchKernelGGL((ens_step_kernel<T, TY, FAST, BULK, LIM_VANLEER, W>), g, b, 0, s, inst, items, nitems, parity, partials, next);
    else
        hipLaunchKernelGGL((ens_step_kernel<T, TY, FAST, BULK, LIM_NONE, W>), g, b, 0, s, inst, items, nitems, parity, partials, next);
}

// waves: occupancy target of the FP32 fast kernels (3: up to 168 VGPRs, 4:
// 128); the other variants are compiled without a target.
template <typename T, int TY, bool FAST, bool BULK>
static void launch_ens_t(const EnsInstance *inst, const EnsItem *items, int nitems, int grid, int parity,
                         int limiter, double *partials, int *next, int waves, hipStream_t s)
{
    if constexpr (FAST && sizeof(T) == 4) {
        if (waves == 3) launch_ens_tw<T, TY, FAST, BULK, 3>(inst, items, nitems, grid, parity, limiter, partials, next, s);
        else launch_ens_tw<T, TY, FAST, BULK, 4>(inst, items, nitems, grid, parity, limiter, partials, next, s);
    } else {
        launch_ens_tw<T, TY, FAST, BULK, 1>(inst, items, nitems, grid, parity, limiter, partials, next, s);
    }
}

int launch_ens_step(int prec, int fast, int tile_y, int limiter, int waves, const EnsInstance *inst,
                    const EnsItem *items, int nitems, int grid, int parity, double *partials, int *next, hipStream_t s)
{
#define DG_ENS(T, TYV, F, B) launch_ens_t<T, TYV, F, B>(inst, items, nitems, grid, parity, limiter, partials, next, waves, s)
    if (prec == 0) {
        if (tile_y == 4) DG_ENS(double, 4, false, false);
        else if (tile_y == 8) DG_ENS(double, 8, false, false);
        else return -1;
    } else {
        const bool bulk = prec == 2;
        if (tile_y == 4) {
            if (fast) { if (bulk) DG_ENS(float, 4, true, true); else DG_ENS(float, 4, true, false); }
            else { if (bulk) DG_ENS(float, 4, false, true); else DG_ENS(float, 4, false, false); }
        } else if (tile_y == 8) {
            if (fast) { if (bulk) DG_ENS(float, 8, true, true); else DG_ENS(float, 8, true, false); }
            else { if (bulk) DG_ENS(float, 8, false, true); else DG_ENS(float, 8, false, false); }
        } else if (tile_y == 16) {
            if (fast) { if (bulk) DG_ENS(float, 16, true, true); else DG_ENS(float, 16, true, false); }
            else { if (bulk) DG_ENS(float, 16, false, true); else DG_ENS(float, 16, false, false); }
        } else return -1;
    }
#undef DG_ENS
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <typename T, int TY, bool FAST, bool BULK, int W>
static int ens_occ_tw(int limiter)
{
    int n = 0;
    hipError_t e;
    if (limiter == LIM_MINMOD)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ens_step_kernel<T, TY, FAST, BULK, LIM_MINMOD, W>, TY * TZ, 0);
    else if (limiter == LIM_VANLEER)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ens_step_kernel<T, TY, FAST, BULK, LIM_VANLEER, W>, TY * TZ, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ens_step_kernel<T, TY, FAST, BULK, LIM_NONE, W>, TY * TZ, 0);
    return e == hipSuccess ? n : -1;
}

template <typename T, int TY, bool FAST, bool BULK>
static int ens_occ_t(int limiter, int waves)
{
    if constexpr (FAST && sizeof(T) == 4) {
        return waves == 3 ? ens_occ_tw<T, TY, FAST, BULK, 3>(limiter) : ens_occ_tw<T, TY, FAST, BULK, 4>(limiter);
    } else {
        return ens_occ_tw<T, TY, FAST, BULK, 1>(limiter);
    }
}

int ens_blocks_per_cu(int prec, int fast, int tile_y, int limiter, int waves)
{
    if (prec == 0) {
        if (tile_y == 4) return ens_occ_t<double, 4, false, false>(limiter, waves);
        if (tile_y == 8) return ens_occ_t<double, 8, false, false>(limiter, waves);
        return -1;
    }
    const bool bulk = prec == 2;
#define DG_OCC(TYV)                                                                                    \
    if (tile_y == TYV) {                                                                              \
        if (fast) return bulk ? ens_occ_t<float, TYV, true, true>(limiter, waves) : ens_occ_t<float, TYV, true, false>(limiter, waves); \
        return bulk ? ens_occ_t<float, TYV, false, true>(limiter, waves) : ens_occ_t<float, TYV, false, false>(limiter, waves); \
    }
    DG_OCC(4) DG_OCC(8) DG_OCC(16)
#undef DG_OCC
    return -1;
}

int launch_ens_finalize(int prec, const EnsInstance *inst, int ninst, const double *partials, int mode, int *next,
                        hipStream_t s)
{
    if (prec == 0)
        hipLaunchKernelGGL((ens_finalize_kernel<double, false>), dim3(ninst), dim3(256), 0, s, inst, partials, mode, next);
    else if (prec == 1)
        hipLaunchKernelGGL((ens_finalize_kernel<float, false>), dim3(ninst), dim3(256), 0, s, inst, partials, mode, next);
    else
        hipLaunchKernelGGL((ens_finalize_kernel<float, true>), dim3(ninst), dim3(256), 0, s, inst, partials, mode, next);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

static int cells_grid(long long n3)
{
    long long b = (n3 + 255) / 256;
    return (int)(b < 1 ? 1 : (b > 2048 ? 2048 : b));
}

// flags: the final flags in [0, n3), the second pass's in [n3, 2 n3), then
// the plain state B (5 n3 floats), then the count
static size_t corr_b_offset(long long n3) { return (size_t)((2 * n3 + 63) / 64) * 64; }
size_t corr_flags_bytes(long long n3) { return corr_b_offset(n3) + (size_t)n3 * 5 * sizeof(float) + 64; }

template <typename GT>
static void launch_corr_t(int fast, const StepArgs &a, const Fields<float> &U, float *B, const GT *G,
                          const CorrScale &S, const Scalars *sc, unsigned char *fa, unsigned char *fb, int *cnt,
                          double *partials, hipStream_t s)
{
    const long long n3 = (long long)a.n[0] * a.n[1] * a.n[2];
    const int nb = reduce_blocks(a);   // the flag and apply kernels write the finalize's partials
    const dim3 g(cells_grid(n3)), b(256);
    if (fast) {
        hipLaunchKernelGGL((corr_flag_kernel<true, false, GT>), dim3(nb), b, 0, s, a, U, B, G, S, sc, nullptr, fa, cnt,
                           partials);
        hipLaunchKernelGGL((corr_flag_kernel<true, true, GT>), g, b, 0, s, a, U, B, G, S, sc, fa, fb, cnt, partials);
        hipLaunchKernelGGL((corr_flag_kernel<true, true, GT>), g, b, 0, s, a, U, B, G, S, sc, fb, fa, cnt, partials);
        hipLaunchKernelGGL((corr_apply_kernel<true, GT>), dim3(nb), b, 0, s, a, U, B, G, S, sc, fa, cnt, partials);
    } else {
        hipLaunchKernelGGL((corr_flag_kernel<false, false, GT>), dim3(nb), b, 0, s, a, U, B, G, S, sc, nullptr, fa,
                           cnt, partials);
        hipLaunchKernelGGL((corr_flag_kernel<false, true, GT>), g, b, 0, s, a, U, B, G, S, sc, fa, fb, cnt, partials);
        hipLaunchKernelGGL((corr_flag_kernel<false, true, GT>), g, b, 0, s, a, U, B, G, S, sc, fb, fa, cnt, partials);
        hipLaunchKernelGGL((corr_apply_kernel<false, GT>), dim3(nb), b, 0, s, a, U, B, G, S, sc, fa, cnt, partials);
    }
}

int *corr_count(unsigned char *flags, long long n3)
{
    return reinterpret_cast<int *>(flags + corr_flags_bytes(n3) - 64);
}

int launch_corr(int fast, const StepArgs &a, void *const u[5], const void *G, int g_bf16, const float scale[5],
                const Scalars *sc, unsigned char *flags, double *partials, hipStream_t s, int count_zeroed)
{
    CorrScale S;
    for (int q = 0; q < 5; ++q) S.s[q] = scale[q];
    const long long n3 = (long long)a.n[0] * a.n[1] * a.n[2];
    unsigned char *fa = flags, *fb = flags + n3;
    float *B = reinterpret_cast<float *>(flags + corr_b_offset(n3));
    int *cnt = corr_count(flags, n3);
    const Fields<float> U = mfields<float>(u);
    if (!count_zeroed) hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(1), 0, s, cnt);
    if (g_bf16)
        launch_corr_t<__bf16>(fast, a, U, B, static_cast<const __bf16 *>(G), S, sc, fa, fb, cnt, partials, s);
    else
        launch_corr_t<float>(fast, a, U, B, static_cast<const float *>(G), S, sc, fa, fb, cnt, partials, s);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_restrict(const StepArgs &fine, const void *const f[5], const StepArgs &coarse, void *const c[5], hipStream_t s)
{
    const long long n3 = (long long)coarse.n[0] * coarse.n[1] * coarse.n[2];
    hipLaunchKernelGGL(restrict_kernel, dim3(cells_grid(n3)), dim3(256), 0, s, fine, cfields<float>(f), coarse,
                       mfields<float>(c));
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_residual(long long n3, const void *const ut[5], const void *const up[5], double dt, double ds,
                    void *const r[5], hipStream_t s)
{
    const float inv = (float)(1.0 / (dt * ds * ds));
    hipLaunchKernelGGL(residual_kernel, dim3(cells_grid(n3)), dim3(256), 0, s, n3, cfields<float>(ut), cfields<float>(up),
                       inv, mfields<float>(r));
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_set_dt(Scalars *sc, double dt, hipStream_t s)
{
    hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, s, sc, dt);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_heat(int prec, int fast, const StepArgs &a, const void *const uold[5], void *const unew[5],
                const Scalars *sc, double t_out, double v_out, double rs, int mask, hipStream_t s)
{
    HeatArgs H{t_out, v_out, rs, mask};
    const long long cells = (long long)a.n[1] * a.n[2] + (long long)a.n[0] * a.n[2] + (long long)a.n[0] * a.n[1];
    const int blocks = (int)std::min<long long>((cells + 255) / 256, 2048);
    if (prec == 0 && fast)
        hipLaunchKernelGGL((heat_kernel<double, true>), dim3(blocks), dim3(256), 0, s, a, cfields<double>(uold),
                           mfields<double>(unew), sc, H);
    else if (prec == 0)
        hipLaunchKernelGGL((heat_kernel<double, false>), dim3(blocks), dim3(256), 0, s, a, cfields<double>(uold),
                           mfields<double>(unew), sc, H);
    else if (fast)
        hipLaunchKernelGGL((heat_kernel<float, true>), dim3(blocks), dim3(256), 0, s, a, cfields<float>(uold),
                           mfields<float>(unew), sc, H);
    else
        hipLaunchKernelGGL((heat_kernel<float, false>), dim3(blocks), dim3(256), 0, s, a, cfields<float>(uold),
                           mfields<float>(unew), sc, H);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int step_lds_bytes(int prec, int tile_y)
{
    const size_t el = prec == 0 ? sizeof(double) : sizeof(float);
    const size_t ty = (size_t)tile_y;
    const size_t cells = 3 * 5 * (ty + 4) * RZ + 5 * (ty + 1) * TZ + 2 * 5 * TZ + 5 * ty * TZ + 4 * 5 * ty;
    return (int)(cells * el + (ty * TZ / 64) * 8 * sizeof(double));
}

int reduce_blocks(const StepArgs &a)
{
    const size_t n3 = (size_t)a.n[0] * a.n[1] * a.n[2];
    size_t nb = (n3 + 255) / 256;
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    return (int)nb;
}

int launch_reduce(int prec, int fast, const StepArgs &a, const void *const u[5], double *partials, int nblocks,
                  hipStream_t s)
{
    const dim3 g(nblocks), b(256);
    if (prec == 0 && !fast)
        hipLaunchKernelGGL((reduce_kernel<double, false, false>), g, b, 0, s, a, cfields<double>(u), partials);
    else if (prec == 0)
        hipLaunchKernelGGL((reduce_kernel<double, false, true>), g, b, 0, s, a, cfields<double>(u), partials);
    else if (prec == 1 && !fast)
        hipLaunchKernelGGL((reduce_kernel<float, false, false>), g, b, 0, s, a, cfields<float>(u), partials);
    else if (prec == 1)
        hipLaunchKernelGGL((reduce_kernel<float, false, true>), g, b, 0, s, a, cfields<float>(u), partials);
    else if (!fast)
        hipLaunchKernelGGL((reduce_kernel<float, true, false>), g, b, 0, s, a, cfields<float>(u), partials);
    else
        hipLaunchKernelGGL((reduce_kernel<float, true, true>), g, b, 0, s, a, cfields<float>(u), partials);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_finalize(int prec, const StepArgs &a, Scalars *sc, const double *partials, int nparts, double cfl,
                    int nan_fallback, int mode, hipStream_t s)
{
    if (prec == 0)
        hipLaunchKernelGGL((finalize_kernel<double, false>), dim3(1), dim3(1024), 0, s, a, sc, partials, nparts, cfl, nan_fallback, mode);
    else if (prec == 1)
        hipLaunchKernelGGL((finalize_kernel<float, false>), dim3(1), dim3(1024), 0, s, a, sc, partials, nparts, cfl, nan_fallback, mode);
    else
        hipLaunchKernelGGL((finalize_kernel<float, true>), dim3(1), dim3(1024), 0, s, a, sc, partials, nparts, cfl, nan_fallback, mode);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_prim_to_cons(int prec, const StepArgs &a, void *const u[5], hipStream_t s)
{
    const int nb = reduce_blocks(a);
    if (prec == 0) hipLaunchKernelGGL((prim_to_cons_kernel<double>), dim3(nb), dim3(256), 0, s, a, mfields<double>(u));
    else hipLaunchKernelGGL((prim_to_cons_kernel<float>), dim3(nb), dim3(256), 0, s, a, mfields<float>(u));
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_cons_to_prim(int prec, const StepArgs &a, const void *const u[5], void *const w[5], hipStream_t s)
{
    const int nb = reduce_blocks(a);
    if (prec == 0)
        hipLaunchKernelGGL((cons_to_prim_kernel<double>), dim3(nb), dim3(256), 0, s, a, cfields<double>(u), mfields<double>(w));
    else
        hipLaunchKernelGGL((cons_to_prim_kernel<float>), dim3(nb), dim3(256), 0, s, a, cfields<float>(u), mfields<float>(w));
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

} // namespace dg
