// dg_api.cpp -- C ABI (include/deepgreen.h) over the fused HIP step.
//
// A context owns two interior-only SoA conserved buffers on the device (the
// state ping-pongs between them), the per-workgroup partial sums, the
// device-resident scalars (dt, run clock, conservation sums) and one HIP
// stream.  Host arrays crossing the ABI use the reference's ghosted layout
// (grid.rs:79-109); the device never stores ghosts: the kernels map ghost
// positions through the boundary conditions.
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <thread>
#include <algorithm>

#include "../../include/deepgreen.h"
#include "dg_ctx.h"
#include "dg_internal.h"

namespace {

const int kGraphSteps = 16;  // even: the buffer parity is unchanged by a replay

// why the calling thread's last dg_create failed (dg_last_error(NULL))
thread_local std::string t_create_err;

int fail(dg_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

int hip_fail(dg_ctx *c, hipError_t e, const char *what)
{
    return fail(c, DG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCHK(c, expr)                                                                \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return hip_fail((c), e_, #expr);                         \
    } while (0)

__global__ void arm_kernel(dg::Scalars *sc, double t_end, long long max_steps_abs, int reset)
{
    if (reset) { sc->t = 0.0; sc->steps = 0; }
    sc->t_end = t_end;
    sc->max_steps = max_steps_abs;
    sc->active = !((sc->t > sc->t_end) || (sc->steps >= sc->max_steps));
}

int read_scalars(dg_ctx *c)
{
    HIPCHK(c, hipMemcpyAsync(c->sc_host, c->sc, sizeof(dg::Scalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

// Boundary map of one axis on the host (same rules as the kernels):
// ghosted index gi (0..n+3) -> interior index and flip.
inline void host_map(int64_t gi, int64_t n, int lo, int hi, int64_t *src, int *flip)
{
    int64_t p = gi - 2;
    *flip = 0;
    if (p < 0) {
        if (lo == DG_PERIODIC) *src = p + n;
        else if (lo == DG_OUTFLOW) *src = 0;
        else { *src = -1 - p; *flip = 1; }
    } else if (p >= n) {
        if (hi == DG_PERIODIC) *src = p - n;
        else if (hi == DG_OUTFLOW) *src = n - 1;
        else { *src = 2 * n - 1 - p; *flip = 1; }
    } else {
        *src = p;
    }
}

// fn(lo, hi) over [0, n) split across up to 16 host threads (the copies
// between the caller's ghosted arrays and the pinned staging buffers)
template <typename F> void parallel_for(int64_t n, F fn)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const int64_t nt = std::min<int64_t>(std::min<int64_t>(16, hw ? hw : 1), n);
    if (nt <= 1) {
        fn(int64_t(0), n);
        return;
    }
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo < hi) th.emplace_back([=] { fn(lo, hi); });
    }
    for (auto &t : th) t.join();
}

// One field of the interior state into the caller's ghosted array, ghost
// cells by the boundary maps (velocity component f = 1, 2, 3 negated across
// a wall of its own axis).
template <typename T>
void scatter_ghosted(const dg_ctx *c, const T *interior, T *out, int f)
{
    const int64_t nx = c->cfg.n[0], ny = c->cfg.n[1], nz = c->cfg.n[2];
    const int64_t gy = ny + 4, gz = nz + 4;
    std::vector<int64_t> sy(gy), sz(gz);
    std::vector<int> fy(gy), fz(gz);
    for (int64_t j = 0; j < gy; ++j) host_map(j, ny, c->cfg.bc[2], c->cfg.bc[3], &sy[j], &fy[j]);
    for (int64_t k = 0; k < gz; ++k) host_map(k, nz, c->cfg.bc[4], c->cfg.bc[5], &sz[k], &fz[k]);
    parallel_for(nx + 4, [&](int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; ++i) {
            int64_t sx; int fx;
            host_map(i, nx, c->cfg.bc[0], c->cfg.bc[1], &sx, &fx);
            for (int64_t j = 0; j < gy; ++j)
                for (int64_t k = 0; k < gz; ++k) {
                    T v = interior[(sx * ny + sy[j]) * nz + sz[k]];
                    if ((f == 1 && fx) || (f == 2 && fy[j]) || (f == 3 && fz[k])) v = -v;
                    out[(i * gy + j) * gz + k] = v;
                }
        }
    });
}

// The two pinned staging buffers and their events, at the first transfer.
int ensure_staging(dg_ctx *c)
{
    for (int b = 0; b < 2; ++b) {
        if (!c->stage[b] && hipHostMalloc(&c->stage[b], c->n3 * c->el, hipHostMallocDefault) != hipSuccess) {
            c->stage[b] = nullptr;
            (void)hipGetLastError();
            return fail(c, DG_ERR_ALLOC, "pinned staging buffer");
        }
        if (!c->stage_ev[b] && hipEventCreateWithFlags(&c->stage_ev[b], hipEventDisableTiming) != hipSuccess) {
            c->stage_ev[b] = nullptr;
            return fail(c, DG_ERR_DEVICE, "staging event");
        }
    }
    return DG_OK;
}

} // namespace

extern "C" {

void dg_config_default(dg_config *cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->n[0] = cfg->n[1] = cfg->n[2] = 64;
    cfg->nghosts = 2;
    for (int i = 0; i < 6; ++i) cfg->bc[i] = DG_PERIODIC;
    cfg->ds = 1.0 / 64.0;
    cfg->gamma = 5.0 / 3.0;
    cfg->cfl = 0.2;
    cfg->g = -0.1;
    cfg->gravity = 0;
    cfg->precision = DG_FP32;
    cfg->math = DG_MATH_EXACT;
    cfg->limiter = DG_LIMITER_NONE;
    cfg->dt_nan_fallback = 1;
}

const char *dg_last_error(const dg_ctx *ctx)
{
    if (ctx) return ctx->err.c_str();
    return t_create_err.empty() ? "null context" : t_create_err.c_str();
}

int dg_device_count(int *count)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
    return DG_OK;
}

int dg_set_heat(dg_ctx *c, double t_out_kelvin, double v_out, double rs)
{
    if (!c || !(t_out_kelvin >= 0) || !(v_out >= 0) || !(rs > 0)) return DG_ERR_INVALID;
    c->heat_t_out = t_out_kelvin;
    c->heat_v_out = v_out;
    c->heat_rs = rs;
    for (int b = 0; b < 2; ++b)
        if (c->graph[b]) {   // the captured steps hold the constants
            (void)hipGraphExecDestroy(c->graph[b]);
            c->graph[b] = nullptr;
        }
    return DG_OK;
}

int dg_device_info(int device, char *arch, int len, int *compute_units)
{
    if (!arch || len <= 0) return DG_ERR_INVALID;
    arch[0] = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return DG_ERR_DEVICE;
    std::string name(prop.gcnArchName);
    name = name.substr(0, name.find(':'));  // drop feature suffixes (":sramecc+:xnack-")
    std::snprintf(arch, (size_t)len, "%s", name.c_str());
    if (compute_units) *compute_units = prop.multiProcessorCount;
    return DG_OK;
}

int dg_create(const dg_config *cfg, dg_ctx **out)
{
    std::string why;
    const int rc = dg::create_ctx(cfg, nullptr, out, &why);
    t_create_err = rc ? why : std::string();
    return rc;
}

} // extern "C"

namespace dg {

// Workgroup tiling of a grid: tile_y rows (4, 8, 16; FP64 4, 8), chunk_x
// x-planes per workgroup (0: the longest chunks that still give ~512
// workgroups -- two resident per CU of a 256-CU device, one wave of
// workgroups -- and at least 8 planes; long chunks re-read fewer halo
// planes).
int tile_args(StepArgs &a, int prec, int tile_y, int chunk_x, int *chunk_out)
{
    if (tile_y != 4 && tile_y != 8 && tile_y != 16) return DG_ERR_INVALID;
    if (prec == DG_FP64 && tile_y > 8) return DG_ERR_INVALID;
    if (chunk_x < 0) return DG_ERR_INVALID;
    a.nby = (a.n[1] + tile_y - 1) / tile_y;
    a.nbz = (a.n[2] + 63) / 64;
    int lx = chunk_x;
    if (lx <= 0) {
        const long long yz = (long long)a.nby * a.nbz;
        const long long want = (long long)a.n[0] * yz / 512;
        lx = (int)(want < 8 ? 8 : want);
    }
    if (lx > a.n[0]) lx = a.n[0];
    a.lx = lx;
    a.nbx = (a.n[0] + lx - 1) / lx;
    const long long nb = (long long)a.nbx * a.nby * a.nbz;
    if (nb > INT_MAX / 2) return DG_ERR_INVALID;
    a.nblocks = (int)nb;
    a.swizzle = (a.nblocks % 8 == 0) ? 1 : 0;
    if (chunk_out) *chunk_out = lx;
    return DG_OK;
}

// Tiling of a context; drops captured graphs (they bake in the grid).
int set_tiling(dg_ctx *c, int tile_y, int chunk_x)
{
    dg::StepArgs a = c->args;
    int lx = 0;
    const int rc = tile_args(a, c->prec, tile_y, chunk_x, &lx);
    if (rc) return rc;
    c->args = a;
    c->tile_y = tile_y;
    c->chunk_x = lx;
    for (int b = 0; b < 2; ++b)
        if (c->graph[b]) {
            (void)hipGraphExecDestroy(c->graph[b]);
            c->graph[b] = nullptr;
        }
    return DG_OK;
}

// Validates a configuration and derives the kernel arguments (grid,
// boundaries, constants, default tiling) from it.
int make_args(const dg_config *cfg, StepArgs *out, int *tile_y_out, int *heat_mask_out, std::string *why)
{
    auto bad = [why](const char *msg) {
        if (why) *why = msg;
        return DG_ERR_INVALID;
    };
    if (!cfg || !out) return bad("null argument");
    if (cfg->nghosts != 2) return bad("nghosts must be 2");
    for (int d = 0; d < 3; ++d)
        if (cfg->n[d] < 2 || cfg->n[d] > (1 << 20)) return bad("interior size out of range [2, 2^20]");
    for (int i = 0; i < 6; ++i)
        if (cfg->bc[i] < 0 || cfg->bc[i] > 3) return bad("unknown boundary type");
    for (int d = 0; d < 3; ++d)  // periodic faces pair up (grid.rs:55-63)
        if ((cfg->bc[2 * d] == DG_PERIODIC) != (cfg->bc[2 * d + 1] == DG_PERIODIC))
            return bad("a periodic face needs a periodic opposite face");
    if (!(cfg->ds > 0) || !(cfg->gamma > 1) || !(cfg->cfl > 0)) return bad("ds, cfl must be > 0 and gamma > 1");
    if (cfg->precision < 0 || cfg->precision > 2 || cfg->math < 0 || cfg->math > 1)
        return bad("unknown precision or math");
    if (cfg->limiter < 0 || cfg->limiter > 2) return bad("unknown limiter");
    // fast math keeps extensive (times ds^3) face states; in FP32 their
    // products stay normal, finite numbers for 1e-5 <= ds <= 1e4
    if (cfg->math == DG_MATH_FAST && cfg->precision != DG_FP64 && !(cfg->ds >= 1e-5 && cfg->ds <= 1e4))
        return bad("fast math in FP32 needs 1e-5 <= ds <= 1e4 (exact math has no such limit)");
    const size_t n3 = (size_t)cfg->n[0] * cfg->n[1] * cfg->n[2];
    // the kernels address a field with 32-bit byte offsets
    const size_t el_bytes = cfg->precision == DG_FP64 ? sizeof(double) : sizeof(float);
    if (n3 * el_bytes >= ((size_t)1 << 32)) return bad("a field of this grid exceeds 4 GiB");
    const int prec = cfg->precision, fast = cfg->math == DG_MATH_FAST;
    // default tile height: 8 rows (two workgroups per CU); FP32 exact math
    // keeps one 16-row workgroup per CU (its register footprint allows no
    // second one, and the taller tile re-reads fewer halo rows): 29.8 against
    // 25.6 Gcell/s on a 512^3 periodic grid
    int ty = cfg->tile_y ? cfg->tile_y : ((prec == DG_FP32 && !fast) ? 16 : 8);
    if (prec == DG_FP64 && ty > 8) ty = 8;

    StepArgs a{};
    for (int d = 0; d < 3; ++d) a.n[d] = (int)cfg->n[d];
    // a conducting wall is a wall for the flow; a lower one carries the heat source
    int heat_mask = 0;
    for (int i = 0; i < 6; ++i) a.bc[i] = cfg->bc[i] == DG_CONDUCTING_WALL ? DG_WALL : cfg->bc[i];
    for (int d = 0; d < 3; ++d)
        if (cfg->bc[2 * d] == DG_CONDUCTING_WALL) heat_mask |= 1 << d;
    a.limiter = cfg->limiter;
    a.gravity = cfg->gravity;
    a.ds = cfg->ds;
    {
        // a power of two whose cube and halves stay normal in float: dividing by
        // ds, 2 ds or ds^3 is then exactly a multiplication by the reciprocal
        // (both round the same real number), which the exact kernels use
        int ex = 0;
        const double mant = std::frexp(cfg->ds, &ex);
        a.pow2 = (mant == 0.5 && ex > -40 && ex < 40) ? 1 : 0;
    }
    a.gamma = cfg->gamma;
    a.g = cfg->g;
    if (tile_args(a, prec, ty, cfg->chunk_x, nullptr) != DG_OK) return bad("invalid tile_y or chunk_x");
    *out = a;
    if (tile_y_out) *tile_y_out = ty;
    if (heat_mask_out) *heat_mask_out = heat_mask;
    return DG_OK;
}

// Partial-slot buffer large enough for the step and reduction grids.
int reserve_partials(dg_ctx *c)
{
    const int need = std::max(c->args.nblocks, c->nparts_reduce);
    if (c->partials && c->partials_cap >= need) return DG_OK;
    if (c->partials) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) return DG_ERR_DEVICE;
        (void)hipFree(c->partials);
        c->partials = nullptr;
        c->partials_cap = 0;
    }
    if (hipMalloc((void **)&c->partials, (size_t)need * dg::PARTIAL_STRIDE * sizeof(double)) != hipSuccess) {
        c->partials = nullptr;
        return DG_ERR_ALLOC;
    }
    c->partials_cap = need;
    return DG_OK;
}

// One step (the step kernel, then dt / clock) on the context's stream.
int launch_one_step(dg_ctx *c)
{
    const void *in[5];
    void *out[5];
    for (int f = 0; f < 5; ++f) { in[f] = c->buf[c->cur][f]; out[f] = c->buf[1 - c->cur][f]; }
    int rc = dg::launch_step(c->prec, c->fast, c->tile_y, c->args, in, out, c->sc, c->partials, c->stream);
    if (rc) return fail(c, DG_ERR_DEVICE, std::string("step launch: ") + hipGetErrorString(hipGetLastError()));
    int nparts = c->args.nblocks;
    if (c->heat_mask) {
        // heat through conducting walls, then the next dt from the heated state
        if (dg::launch_heat(c->prec, c->fast, c->args, in, out, c->sc, c->heat_t_out, c->heat_v_out, c->heat_rs,
                            c->heat_mask, c->stream))
            return fail(c, DG_ERR_DEVICE, "heat launch");
        const void *u[5] = {out[0], out[1], out[2], out[3], out[4]};
        if (dg::launch_reduce(c->prec, c->fast, c->args, u, c->partials, c->nparts_reduce, c->stream))
            return fail(c, DG_ERR_DEVICE, "reduce launch");
        nparts = c->nparts_reduce;
    }
    rc = dg::launch_finalize(c->prec, c->args, c->sc, c->partials, nparts, c->cfg.cfl,
                             c->cfg.dt_nan_fallback, 1, c->stream);
    if (rc) return fail(c, DG_ERR_DEVICE, "finalize launch");
    c->cur = 1 - c->cur;
    c->sums_stale = true;
    return DG_OK;
}

// Lift any stop condition of a previous dg_run (a step launched while the
// scalars say "inactive" carries the state through unchanged).
int arm_unbounded(dg_ctx *c)
{
    hipLaunchKernelGGL(arm_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, (double)INFINITY, (long long)LLONG_MAX, 0);
    return hipGetLastError() == hipSuccess ? DG_OK : fail(c, DG_ERR_DEVICE, "arm launch");
}

// mode 0: dt and sums of a freshly set state; 2: sums only
int reduce_current(dg_ctx *c, int mode)
{
    const void *u[5];
    for (int f = 0; f < 5; ++f) u[f] = c->buf[c->cur][f];
    if (dg::launch_reduce(c->prec, c->fast, c->args, u, c->partials, c->nparts_reduce, c->stream))
        return fail(c, DG_ERR_DEVICE, "reduce launch");
    if (dg::launch_finalize(c->prec, c->args, c->sc, c->partials, c->nparts_reduce, c->cfg.cfl,
                            c->cfg.dt_nan_fallback, mode, c->stream))
        return fail(c, DG_ERR_DEVICE, "finalize launch");
    c->sums_stale = false;
    return DG_OK;
}

// Contexts of an ensemble share the ensemble's stream (shared_stream != null).
int create_ctx(const dg_config *cfg, hipStream_t shared_stream, dg_ctx **out, std::string *why)
{
    auto failed = [why](int code, const std::string &msg) {
        if (why) *why = msg;
        return code;
    };
    if (!cfg || !out) return failed(DG_ERR_INVALID, "null argument");
    *out = nullptr;
    dg::StepArgs args{};
    int ty = 8, heat_mask = 0;
    std::string reason;
    if (make_args(cfg, &args, &ty, &heat_mask, &reason) != DG_OK)
        return failed(DG_ERR_INVALID, "invalid configuration: " + reason);

    dg_ctx *c = new (std::nothrow) dg_ctx();
    if (!c) return failed(DG_ERR_ALLOC, "host allocation of the context failed");
    static std::atomic<uint64_t> next_uid{1};
    c->uid = next_uid.fetch_add(1);
    c->cfg = *cfg;
    c->prec = cfg->precision;
    c->fast = cfg->math == DG_MATH_FAST;
    c->el = c->prec == DG_FP64 ? sizeof(double) : sizeof(float);
    c->n3 = (size_t)cfg->n[0] * cfg->n[1] * cfg->n[2];
    c->args = args;
    c->heat_mask = heat_mask;
    c->nparts_reduce = dg::reduce_blocks(args);
    if (set_tiling(c, ty, cfg->chunk_x) != DG_OK) { delete c; return failed(DG_ERR_INVALID, "invalid tiling"); }
    const size_t n3 = c->n3;

    hipError_t e = hipSetDevice(cfg->device);
    if (e != hipSuccess) { delete c; return failed(DG_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e)); }
    if (shared_stream) {
        c->stream = shared_stream;
        c->owns_stream = false;
    } else {
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return failed(DG_ERR_DEVICE, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e));
        }
    }
    int rc = DG_OK;
    std::string what;
    for (int b = 0; b < 2 && rc == DG_OK; ++b)
        for (int f = 0; f < 5 && rc == DG_OK; ++f)
            if (hipMalloc(&c->buf[b][f], n3 * c->el) != hipSuccess) {
                c->buf[b][f] = nullptr;
                rc = DG_ERR_ALLOC;
                what = "device allocation of the state (10 arrays of " + std::to_string(n3 * c->el) + " bytes) failed";
            }
    if (rc == DG_OK && (rc = reserve_partials(c)) != DG_OK) what = "device allocation of the partial results failed";
    if (rc == DG_OK && hipMalloc((void **)&c->sc, sizeof(dg::Scalars)) != hipSuccess) {
        c->sc = nullptr;
        rc = DG_ERR_ALLOC;
        what = "device allocation of the scalars failed";
    }
    if (rc == DG_OK && hipHostMalloc((void **)&c->sc_host, sizeof(dg::Scalars), hipHostMallocDefault) != hipSuccess) {
        c->sc_host = nullptr;
        rc = DG_ERR_ALLOC;
        what = "pinned host allocation of the scalars failed";
    }
    if (rc != DG_OK) {
        (void)hipGetLastError();   // the failed allocation's error is reported here, not by a later call
        dg_destroy(c);             // releases whatever was allocated
        return failed(rc, what);
    }
    std::memset(c->sc_host, 0, sizeof(dg::Scalars));
    c->sc_host->t_end = INFINITY;
    c->sc_host->max_steps = LLONG_MAX;
    c->sc_host->active = 1;
    if (hipMemcpy(c->sc, c->sc_host, sizeof(dg::Scalars), hipMemcpyHostToDevice) != hipSuccess) {
        dg_destroy(c);
        return failed(DG_ERR_DEVICE, "initializing the scalars");
    }
    *out = c;
    return DG_OK;
}
} // namespace dg

extern "C" {

void dg_destroy(dg_ctx *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int b = 0; b < 2; ++b)
        if (c->graph[b]) (void)hipGraphExecDestroy(c->graph[b]);
    for (int b = 0; b < 2; ++b)
        for (int f = 0; f < 5; ++f)
            if (c->buf[b][f]) (void)hipFree(c->buf[b][f]);
    if (c->partials) (void)hipFree(c->partials);
    if (c->sc) (void)hipFree(c->sc);
    if (c->scratch) (void)hipFree(c->scratch);
    if (c->dumpbuf) (void)hipFree(c->dumpbuf);
    if (c->sc_host) (void)hipHostFree(c->sc_host);
    for (int b = 0; b < 2; ++b) {
        if (c->stage[b]) (void)hipHostFree(c->stage[b]);
        if (c->stage_ev[b]) (void)hipEventDestroy(c->stage_ev[b]);
    }
    if (c->stream && c->owns_stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int dg_set_state(dg_ctx *c, const void *const f[5], int kind)
{
    if (!c || !f || (kind != DG_PRIMITIVE && kind != DG_CONSERVED)) return fail(c, DG_ERR_INVALID, "bad argument");
    for (int q = 0; q < 5; ++q)
        if (!f[q]) return fail(c, DG_ERR_INVALID, "null field");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const int64_t nx = c->cfg.n[0], ny = c->cfg.n[1], nz = c->cfg.n[2];
    const int64_t gy = ny + 4, gz = nz + 4;
    if (int rc = ensure_staging(c)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // field q is packed into stage[q % 2] while field q - 1 crosses to the device
    for (int q = 0; q < 5; ++q) {
        const int b = q & 1;
        if (q >= 2) HIPCHK(c, hipEventSynchronize(c->stage_ev[b]));   // field q - 2 has left stage[b]
        const char *src = static_cast<const char *>(f[q]);
        char *dst = static_cast<char *>(c->stage[b]);
        const size_t el = c->el;
        parallel_for(nx, [&](int64_t i0, int64_t i1) {
            for (int64_t i = i0; i < i1; ++i)
                for (int64_t j = 0; j < ny; ++j)
                    std::memcpy(dst + (i * ny + j) * nz * el, src + (((i + 2) * gy + (j + 2)) * gz + 2) * el,
                                nz * el);
        });
        HIPCHK(c, hipMemcpyAsync(c->buf[c->cur][q], c->stage[b], c->n3 * c->el, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipEventRecord(c->stage_ev[b], c->stream));
    }
    if (kind == DG_PRIMITIVE) {
        void *u[5];
        for (int q = 0; q < 5; ++q) u[q] = c->buf[c->cur][q];
        if (dg::launch_prim_to_cons(c->prec, c->args, u, c->stream)) return fail(c, DG_ERR_DEVICE, "prim_to_cons");
    }
    int rc = dg::reduce_current(c, 0);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_state = true;
    return DG_OK;
}

int dg_get_state(dg_ctx *c, void *const f[5], int kind)
{
    if (!c || !f || (kind != DG_PRIMITIVE && kind != DG_CONSERVED)) return fail(c, DG_ERR_INVALID, "bad argument");
    if (!c->have_state) return fail(c, DG_ERR_STATE, "no state set");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    const void *const *srcs = (const void *const *)c->buf[c->cur];
    void *prim[5];
    if (kind == DG_PRIMITIVE) {
        if (!c->scratch) HIPCHK(c, hipMalloc(&c->scratch, 5 * c->n3 * c->el));
        for (int q = 0; q < 5; ++q) prim[q] = static_cast<char *>(c->scratch) + q * c->n3 * c->el;
        if (dg::launch_cons_to_prim(c->prec, c->args, srcs, prim, c->stream))
            return fail(c, DG_ERR_DEVICE, "cons_to_prim");
        srcs = prim;
    }
    if (int rc = ensure_staging(c)) return rc;
    // field q crosses into stage[q % 2] while field q - 1 is scattered on the host
    auto scatter = [&](int q) {
        if (c->prec == DG_FP64)
            scatter_ghosted<double>(c, static_cast<const double *>(c->stage[q & 1]), static_cast<double *>(f[q]), q);
        else
            scatter_ghosted<float>(c, static_cast<const float *>(c->stage[q & 1]), static_cast<float *>(f[q]), q);
    };
    for (int q = 0; q < 5; ++q) {
        HIPCHK(c, hipMemcpyAsync(c->stage[q & 1], srcs[q], c->n3 * c->el, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipEventRecord(c->stage_ev[q & 1], c->stream));
        if (q >= 1) {
            HIPCHK(c, hipEventSynchronize(c->stage_ev[(q - 1) & 1]));
            scatter(q - 1);
        }
    }
    HIPCHK(c, hipEventSynchronize(c->stage_ev[0]));
    scatter(4);
    return DG_OK;
}

int dg_step(dg_ctx *c, double *dt_used)
{
    if (!c) return DG_ERR_INVALID;
    if (!c->have_state) return fail(c, DG_ERR_STATE, "no state set");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    hipLaunchKernelGGL(arm_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, (double)INFINITY, (long long)LLONG_MAX, 0);
    int rc = read_scalars(c);
    if (rc) return rc;
    const double dt = c->sc_host->dt;
    if (!c->cfg.dt_nan_fallback && !std::isfinite(dt)) return fail(c, DG_ERR_NONFINITE, "dt is not finite");
    rc = dg::launch_one_step(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (dt_used) *dt_used = dt;
    return DG_OK;
}

int dg_reset_clock(dg_ctx *c)
{
    if (!c) return DG_ERR_INVALID;
    HIPCHK(c, hipSetDevice(c->cfg.device));
    hipLaunchKernelGGL(arm_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, (double)INFINITY, (long long)LLONG_MAX, 1);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DG_OK;
}

// The captured steps bake in the buffer pointers, so there is one graph per
// parity of the current buffer.
static int ensure_graph(dg_ctx *c)
{
    if (c->graph[c->cur]) return DG_OK;
    hipGraph_t g = nullptr;
    const int cur0 = c->cur;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = DG_OK;
    for (int s = 0; s < kGraphSteps && rc == DG_OK; ++s) rc = dg::launch_one_step(c);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->cur = cur0;  // capture launched nothing
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&c->graph[cur0], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { c->graph[cur0] = nullptr; return hip_fail(c, e, "hipGraphInstantiate"); }
    c->graph_steps = kGraphSteps;
    return DG_OK;
}

int dg_run(dg_ctx *c, int64_t max_steps, double t_end, double *t_out, int64_t *steps_out)
{
    if (!c || max_steps < 0) return fail(c, DG_ERR_INVALID, "bad argument");
    if (!c->have_state) return fail(c, DG_ERR_STATE, "no state set");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    int rc = read_scalars(c);
    if (rc) return rc;
    const long long start = c->sc_host->steps;
    const long long cap = (max_steps > LLONG_MAX - start) ? LLONG_MAX : start + max_steps;
    hipLaunchKernelGGL(arm_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, t_end, cap, 0);
    const bool timed = std::isfinite(t_end);
    int64_t left = max_steps;
    while (left > 0) {
        if (left >= kGraphSteps) {
            rc = ensure_graph(c);
            if (rc) return rc;
            HIPCHK(c, hipGraphLaunch(c->graph[c->cur], c->stream));
            left -= c->graph_steps;
        } else {
            rc = dg::launch_one_step(c);
            if (rc) return rc;
            left -= 1;
        }
        if (timed) {
            rc = read_scalars(c);
            if (rc) return rc;
            if (!c->sc_host->active) break;
        }
    }
    rc = read_scalars(c);
    if (rc) return rc;
    if (t_out) *t_out = c->sc_host->t;
    if (steps_out) *steps_out = c->sc_host->steps;
    if (!c->cfg.dt_nan_fallback && c->sc_host->nonfinite) return fail(c, DG_ERR_NONFINITE, "dt is not finite");
    return DG_OK;
}

int dg_conservation_sums(dg_ctx *c, double out[5])
{
    if (!c || !out) return DG_ERR_INVALID;
    if (!c->have_state) return fail(c, DG_ERR_STATE, "no state set");
    int rc = DG_OK;
    if (c->sums_stale && (rc = dg::reduce_current(c, 2))) return rc;
    if ((rc = read_scalars(c))) return rc;
    for (int q = 0; q < 5; ++q) out[q] = c->sc_host->sums[q];
    return DG_OK;
}

int dg_next_dt(dg_ctx *c, double *dt)
{
    if (!c || !dt) return DG_ERR_INVALID;
    if (!c->have_state) return fail(c, DG_ERR_STATE, "no state set");
    int rc = read_scalars(c);
    if (rc) return rc;
    *dt = c->sc_host->dt;
    return DG_OK;
}

int dg_device_state(dg_ctx *c, void *ptrs[5])
{
    if (!c || !ptrs) return DG_ERR_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < 5; ++q) ptrs[q] = c->buf[c->cur][q];
    return DG_OK;
}

int dg_time_steps(dg_ctx *c, int64_t nsteps, double *ms_total, double *ms_step_kernel, double *ms_reduce_kernel)
{
    if (!c || nsteps <= 0) return fail(c, DG_ERR_INVALID, "bad argument");
    if (!c->have_state) return fail(c, DG_ERR_STATE, "no state set");
    HIPCHK(c, hipSetDevice(c->cfg.device));
    hipLaunchKernelGGL(arm_kernel, dim3(1), dim3(1), 0, c->stream, c->sc, (double)INFINITY, (long long)LLONG_MAX, 0);
    dg::EventSet es;
    HIPCHK(c, es.create((size_t)nsteps * 2 + 1));
    std::vector<hipEvent_t> &ev = es.ev;
    int rc = DG_OK;
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    for (int64_t s = 0; s < nsteps && rc == DG_OK; ++s) {
        const void *in[5];
        void *out[5];
        for (int f = 0; f < 5; ++f) { in[f] = c->buf[c->cur][f]; out[f] = c->buf[1 - c->cur][f]; }
        if (dg::launch_step(c->prec, c->fast, c->tile_y, c->args, in, out, c->sc, c->partials, c->stream))
            rc = fail(c, DG_ERR_DEVICE, "step launch");
        if (hipEventRecord(ev[2 * s + 1], c->stream) != hipSuccess) rc = DG_ERR_DEVICE;
        int nparts = c->args.nblocks;
        if (c->heat_mask) {   // conducting walls: heat, then the dt of the heated state
            const void *u2[5] = {out[0], out[1], out[2], out[3], out[4]};
            if (dg::launch_heat(c->prec, c->fast, c->args, in, out, c->sc, c->heat_t_out, c->heat_v_out, c->heat_rs,
                                c->heat_mask, c->stream) ||
                dg::launch_reduce(c->prec, c->fast, c->args, u2, c->partials, c->nparts_reduce, c->stream))
                rc = fail(c, DG_ERR_DEVICE, "heat launch");
            nparts = c->nparts_reduce;
        }
        if (dg::launch_finalize(c->prec, c->args, c->sc, c->partials, nparts, c->cfg.cfl,
                                c->cfg.dt_nan_fallback, 1, c->stream))
            rc = fail(c, DG_ERR_DEVICE, "finalize launch");
        if (hipEventRecord(ev[2 * s + 2], c->stream) != hipSuccess) rc = DG_ERR_DEVICE;
        c->cur = 1 - c->cur;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) rc = DG_ERR_DEVICE;
    double tk = 0, tr = 0;
    float ms = 0;
    for (int64_t s = 0; s < nsteps && rc == DG_OK; ++s) {
        (void)hipEventElapsedTime(&ms, ev[2 * s], ev[2 * s + 1]);
        tk += ms;
        (void)hipEventElapsedTime(&ms, ev[2 * s + 1], ev[2 * s + 2]);
        tr += ms;
    }
    (void)hipEventElapsedTime(&ms, ev[0], ev[2 * nsteps]);
    if (rc) return rc;
    if (ms_total) *ms_total = ms;
    if (ms_step_kernel) *ms_step_kernel = tk / nsteps;
    if (ms_reduce_kernel) *ms_reduce_kernel = tr / nsteps;
    return DG_OK;
}

int dg_kernel_info(dg_ctx *c, int *tile_y, int *chunk_x, int64_t *workgroups, int *lds_bytes, int *threads)
{
    if (!c) return DG_ERR_INVALID;
    if (tile_y) *tile_y = c->tile_y;
    if (chunk_x) *chunk_x = c->chunk_x;
    if (workgroups) *workgroups = c->args.nblocks;
    if (lds_bytes) *lds_bytes = dg::step_lds_bytes(c->prec, c->tile_y);
    if (threads) *threads = c->tile_y * 64;
    return DG_OK;
}

} // extern "C"
