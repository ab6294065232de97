// dg_noalloc.cpp -- dg_step_noalloc: one step on caller-owned device buffers
// that allocates nothing.
//
// The step and reduction kernels need device scratch beyond the state: one
// partial-result slot per workgroup (max wave-speed denominator, NaN flag,
// sums) and the scalar block (dt, clock, stop condition).  A context
// allocates these with hipMalloc at dg_create; callers that cannot allocate
// (code under HIP graph capture, serving loops, harnesses that forbid it)
// get them from `scratch`, a static device array of this library: each GPU
// holds its own copy for the life of the process.  Its device address is
// looked up once per device with hipGetSymbolAddress (the array's name is
// only a host-side handle) and kept in `g_scratch`.
//
// Every call on one GPU uses the same array and leaves its contents behind:
// the scalar block is re-initialized by a kernel at the start of each call
// (no memset, so the sequence can be captured into a graph), and every
// partial slot a call reads is written earlier in the same call.  Calls that
// overlap on one GPU (two host threads, or two streams) race on the array.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <mutex>

#include "../../include/deepgreen.h"
#include "dg_ctx.h"
#include "dg_internal.h"

namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kScalarBytes = 256;                                   // dg::Scalars, padded
constexpr size_t kPartialBytes = (size_t)DG_NOALLOC_MAX_SLOTS * dg::PARTIAL_STRIDE * sizeof(double);
static_assert(sizeof(dg::Scalars) <= kScalarBytes, "scalar block");

// [0, 256): scalars; [256, 256 + slots * 64): partial slots
__device__ __align__(256) unsigned char scratch[kScalarBytes + kPartialBytes];

std::mutex g_lock;
void *g_scratch[kMaxDevices] = {};

// Device address of this GPU's copy of `scratch` (the device must be current).
hipError_t scratch_ptr(int device, void **out)
{
    std::lock_guard<std::mutex> lk(g_lock);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

// Scalars of a fresh single step: clock 0, no stop condition.
__global__ void noalloc_arm_kernel(dg::Scalars *sc)
{
    sc->t = 0.0;
    sc->steps = 0;
    sc->t_end = INFINITY;
    sc->max_steps = LLONG_MAX;
    sc->active = 1;
    sc->nonfinite = 0;
    sc->dt = 0.0;
}

} // namespace

extern "C" int dg_step_noalloc(const dg_config *cfg, const void *const in[5], void *const out[5], void *stream,
                               double *dt_used)
{
    if (!cfg || !in || !out) return DG_ERR_INVALID;
    for (int f = 0; f < 5; ++f) {
        if (!in[f] || !out[f]) return DG_ERR_INVALID;
        for (int g = 0; g < 5; ++g)
            if (in[f] == out[g]) return DG_ERR_INVALID;   // the step reads neighbours of cells it writes
    }
    dg::StepArgs a{};
    int tile_y = 8, heat_mask = 0;
    if (dg::make_args(cfg, &a, &tile_y, &heat_mask) != DG_OK) return DG_ERR_INVALID;
    if (heat_mask) return DG_ERR_INVALID;   // the heat source needs a context (dg_set_heat)
    const int nred = dg::reduce_blocks(a);
    if (a.nblocks > DG_NOALLOC_MAX_SLOTS || nred > DG_NOALLOC_MAX_SLOTS) return DG_ERR_INVALID;
    if (cfg->device < 0 || cfg->device >= kMaxDevices) return DG_ERR_INVALID;
    if (hipSetDevice(cfg->device) != hipSuccess) return DG_ERR_DEVICE;
    void *base = nullptr;
    if (scratch_ptr(cfg->device, &base) != hipSuccess) return DG_ERR_DEVICE;
    dg::Scalars *sc = static_cast<dg::Scalars *>(base);
    double *partials = reinterpret_cast<double *>(static_cast<unsigned char *>(base) + kScalarBytes);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int prec = cfg->precision, fast = cfg->math == DG_MATH_FAST;

    // dt from the input state (as dg_set_state computes it), then the step
    hipLaunchKernelGGL(noalloc_arm_kernel, dim3(1), dim3(1), 0, s, sc);
    if (hipGetLastError() != hipSuccess) return DG_ERR_DEVICE;
    if (dg::launch_reduce(prec, fast, a, in, partials, nred, s)) return DG_ERR_DEVICE;
    if (dg::launch_finalize(prec, a, sc, partials, nred, cfg->cfl, cfg->dt_nan_fallback, 0, s)) return DG_ERR_DEVICE;
    double dt = 0.0;
    if (dt_used) {
        if (hipMemcpyAsync(&dt, &sc->dt, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) return DG_ERR_DEVICE;
        if (hipStreamSynchronize(s) != hipSuccess) return DG_ERR_DEVICE;
        if (!cfg->dt_nan_fallback && !std::isfinite(dt)) return DG_ERR_NONFINITE;   // as dg_step: no step taken
    }
    if (dg::launch_step(prec, fast, tile_y, a, in, out, sc, partials, s)) return DG_ERR_DEVICE;
    if (dt_used) {
        if (hipStreamSynchronize(s) != hipSuccess) return DG_ERR_DEVICE;
        *dt_used = dt;
    }
    return DG_OK;
}
