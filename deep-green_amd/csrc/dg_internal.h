// dg_internal.h -- types shared by the HIP kernels (dg_step.hip) and the
// C-ABI layer (dg_api.cpp).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace dg {

enum : int { BC_PERIODIC = 0, BC_OUTFLOW = 1, BC_WALL = 2 };
enum : int { LIM_NONE = 0, LIM_MINMOD = 1, LIM_VANLEER = 2 };

// Per-workgroup partial results written by the step and reduction kernels:
// [0] max wave-speed denominator, [1] NaN seen (0/1), [2..6] conserved sums (zero from
// the step kernel, which does not accumulate them).
constexpr int PARTIAL_STRIDE = 8;

struct StepArgs {
    int n[3];          // interior cells per axis
    int bc[6];         // x-, x+, y-, y+, z-, z+
    int limiter;
    int gravity;
    int lx;            // planes per workgroup along x
    int nbx, nby, nbz; // workgroups per axis
    int nblocks;
    int swizzle;       // 1: spread tiles so each XCD works on a contiguous range
    int pow2;          // 1: ds is a power of two, so x / ds, x / (2 ds) and x / ds^3
                       //    equal the products with the exact reciprocals bit for bit
    double ds, gamma, g;
};

// Device-resident scalars of a context.
struct Scalars {
    double dt;          // dt the next step uses (exactly representable in the state type)
    double t;           // run clock (float arithmetic for DG_FP32)
    long long steps;    // run step counter
    double t_end;
    long long max_steps;
    int active;         // 1: the next step advances the state
    int nonfinite;      // the last reduction saw a NaN wave speed
    double sums[5];     // FP64 conservation sums (of the state they were computed for)
    double maxden;      // max of sqrt(gamma p / rho) + |v| over the interior
};

template <typename T> struct Fields { T *f[5]; };

// HIP events of a timing harness, destroyed on every exit path.
struct EventSet {
    std::vector<hipEvent_t> ev;
    hipError_t create(size_t n)
    {
        ev.assign(n, nullptr);
        for (auto &e : ev) {
            const hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) { e = nullptr; return r; }
        }
        return hipSuccess;
    }
    ~EventSet()
    {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// One grid of a batched ensemble (device-resident table).
struct EnsInstance {
    StepArgs args;       // grid, boundaries, tiling (lx, nby, nbz) of this instance
    void *buf[2][5];     // state ping-pong buffers
    Scalars *sc;         // this instance's scalars
    int part_base;       // first partial slot
    int nparts;          // partial slots (= tiles)
    double cfl;
    int nan_fallback;
    int pad_;
};

// One (instance, tile) work item of the persistent ensemble kernel.
struct EnsItem { int inst, bx, by, bz, slot; };

// Kernel launch entry points (dg_step.hip).  `prec` is 0 FP64, 1 FP32,
// 2 FP32_BULK; `fast` selects the reassociated arithmetic.
int launch_step(int prec, int fast, int tile_y, const StepArgs &a, const void *const in[5],
                void *const out[5], const Scalars *sc, double *partials, hipStream_t s);
int launch_reduce(int prec, int fast, const StepArgs &a, const void *const u[5], double *partials,
                  int nblocks, hipStream_t s);
// mode 0: dt and sums of a fresh state; 1: after a step; 2: sums only
int launch_finalize(int prec, const StepArgs &a, Scalars *sc, const double *partials, int nparts,
                    double cfl, int nan_fallback, int mode, hipStream_t s);
int launch_prim_to_cons(int prec, const StepArgs &a, void *const u[5], hipStream_t s);
int launch_cons_to_prim(int prec, const StepArgs &a, const void *const u[5], void *const w[5],
                        hipStream_t s);
// `next`: the work counter (0 before a step launch; the finalize resets it).
// waves: occupancy target of the FP32 fast variants (3 or 4 waves per SIMD).
int launch_ens_step(int prec, int fast, int tile_y, int limiter, int waves, const EnsInstance *inst,
                    const EnsItem *items, int nitems, int grid, int parity, double *partials, int *next, hipStream_t s);
int launch_ens_finalize(int prec, const EnsInstance *inst, int ninst, const double *partials, int mode, int *next,
                        hipStream_t s);
// Learned flux correction (FP32 periodic grids): flags cells whose corrected
// state is not positive, applies the masked correction in place and writes
// reduce_blocks(a) wave-speed partials for the finalize (mode 1).  G: n3 x 16
// channels, fp32 (g_bf16 = 0) or bf16 (g_bf16 = 1).  flags has
// corr_flags_bytes(n3) bytes; its first n3 hold the final flags.
size_t corr_flags_bytes(long long n3);
// count_zeroed: the fallback count (corr_count) was already set to 0 on the
// stream since its last use; otherwise launch_corr clears it.
int launch_corr(int fast, const StepArgs &a, void *const u[5], const void *G, int g_bf16, const float scale[5],
                const Scalars *sc, unsigned char *flags, double *partials, hipStream_t s, int count_zeroed = 0);
int *corr_count(unsigned char *flags, long long n3);
// FP32: coarse = 2x restriction of fine (sum of the 8 children).
int launch_restrict(const StepArgs &fine, const void *const f[5], const StepArgs &coarse, void *const c[5], hipStream_t s);
// FP32: r = (ut - up) / (dt ds^2).
int launch_residual(long long n3, const void *const ut[5], const void *const up[5], double dt, double ds,
                    void *const r[5], hipStream_t s);
int launch_set_dt(Scalars *sc, double dt, hipStream_t s);
// Heat through conducting lower walls (mask bit 0 x-, 1 y-, 2 z-), added to unew.
int launch_heat(int prec, int fast, const StepArgs &a, const void *const uold[5], void *const unew[5],
                const Scalars *sc, double t_out, double v_out, double rs, int mask, hipStream_t s);
// Resident workgroups of the ensemble step kernel per CU (-1 on error).
int ens_blocks_per_cu(int prec, int fast, int tile_y, int limiter, int waves);
int step_lds_bytes(int prec, int tile_y);
int reduce_blocks(const StepArgs &a);

} // namespace dg
