// dg_ensemble.cpp -- C ABI for many independent grids stepped together
// (include/deepgreen.h, dg_ensemble_*).
//
// Every instance is an ordinary context (its own state buffers, scalars and
// reduction slots) created on the ensemble's stream, so state upload and
// download, the fresh-state dt and the conservation sums are the single-grid
// code.  A step of the ensemble is two launches:
//   ens_step_kernel      persistent workgroups taking (instance, tile) items
//                        from the work list (heaviest first) through an
//                        atomic counter; each tile writes one partial slot of
//                        its instance;
//   ens_finalize_kernel  one workgroup per instance: its partial slots -> its
//                        next dt and run clock (the finalize of dg_run); it
//                        also resets the work counter for the next step.
// The per-tile arithmetic is the single-grid step's (same device function),
// the max wave speed is an order-free max, so each instance's trajectory is
// bitwise the one dg_run produces, whatever the tile shape or the order of the
// work list.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/deepgreen.h"
#include "dg_ctx.h"
#include "dg_internal.h"

namespace {

const int kEnsGraphSteps = 16;    // even: the buffer parity is unchanged by a replay
const int kWholeRunSteps = 1024;  // runs up to this long (no stop time) are one captured graph

}  // namespace

struct dg_ensemble {
    std::vector<dg_ctx *> inst;
    int prec = 1, fast = 0, tile_y = 8, limiter = 0, device = 0;
    int waves = 4;             // occupancy target of the step kernel (FP32 fast: 3 or 4)
    hipStream_t stream = nullptr;
    dg::EnsInstance *d_inst = nullptr;
    dg::EnsItem *d_items = nullptr;
    double *partials = nullptr;
    int *d_count = nullptr;    // device: active instances
    int *d_next = nullptr;     // device: work counter of the step launch
    int *h_count = nullptr;    // pinned mirror
    int nitems = 0, grid = 0;
    int items_cap = 0, slots_cap = 0;  // allocated entries of d_items / partials
    int parity = 0;            // state buffer of every instance
    int64_t cells = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // captured steps, by state buffer parity
    struct RunGraph { int parity; int64_t steps; hipGraphExec_t exec; };
    std::vector<RunGraph> runs;   // whole runs captured as one graph
    std::string err;
};

namespace {

int efail(dg_ensemble *e, int code, const std::string &msg)
{
    if (e) e->err = msg;
    return code;
}

#define EHIP(e, expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return efail((e), DG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Arms every instance for a run of up to max_steps more steps (the test of
// dg_run: active while !(t > t_end || steps >= cap)).
__global__ void ens_arm_kernel(const dg::EnsInstance *inst, int n, long long max_steps, double t_end)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dg::Scalars *sc = inst[i].sc;
    const long long start = sc->steps;
    const long long cap = (max_steps > LLONG_MAX - start) ? LLONG_MAX : start + max_steps;
    sc->t_end = t_end;
    sc->max_steps = cap;
    sc->active = !((sc->t > sc->t_end) || (sc->steps >= sc->max_steps));
}

// Number of instances still active; one wave, fixed order.
__global__ void ens_count_active_kernel(const dg::EnsInstance *inst, int n, int *count)
{
    int c = 0;
    for (int i = threadIdx.x; i < n; i += 64) c += inst[i].sc->active;
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
    if (threadIdx.x == 0) *count = c;
}

int arm(dg_ensemble *e, long long max_steps, double t_end)
{
    const int n = (int)e->inst.size();
    hipLaunchKernelGGL(ens_arm_kernel, dim3((n + 255) / 256), dim3(256), 0, e->stream, e->d_inst, n, max_steps,
                       t_end);
    EHIP(e, hipGetLastError());
    return DG_OK;
}

int launch_one(dg_ensemble *e, int parity)
{
    if (dg::launch_ens_step(e->prec, e->fast, e->tile_y, e->limiter, e->waves, e->d_inst, e->d_items, e->nitems,
                            e->grid, parity, e->partials, e->d_next, e->stream))
        return efail(e, DG_ERR_DEVICE, std::string("ensemble step launch: ") + hipGetErrorString(hipGetLastError()));
    if (dg::launch_ens_finalize(e->prec, e->d_inst, (int)e->inst.size(), e->partials, 1, e->d_next, e->stream))
        return efail(e, DG_ERR_DEVICE, "ensemble finalize launch");
    return DG_OK;
}

void after_steps(dg_ensemble *e, int64_t nsteps)
{
    e->parity = (int)((e->parity + nsteps) & 1);
    for (dg_ctx *c : e->inst) {
        c->cur = e->parity;
        c->sums_stale = true;
    }
}

// nsteps steps from the current parity, captured into one executable graph
int capture(dg_ensemble *e, int64_t nsteps, hipGraphExec_t *out)
{
    *out = nullptr;
    hipGraph_t g = nullptr;
    EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = DG_OK;
    for (int64_t s = 0; s < nsteps && rc == DG_OK; ++s) rc = launch_one(e, e->parity ^ (int)(s & 1));
    hipError_t he = hipStreamEndCapture(e->stream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (he != hipSuccess) return efail(e, DG_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(he));
    he = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (he != hipSuccess) {
        *out = nullptr;
        return efail(e, DG_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(he));
    }
    return DG_OK;
}

int ensure_graph(dg_ensemble *e)
{
    if (e->graph[e->parity]) return DG_OK;
    return capture(e, kEnsGraphSteps, &e->graph[e->parity]);
}

// The graph of a whole run of nsteps from the current parity (cached).
int run_graph(dg_ensemble *e, int64_t nsteps, hipGraphExec_t *out)
{
    for (const auto &r : e->runs)
        if (r.parity == e->parity && r.steps == nsteps) { *out = r.exec; return DG_OK; }
    int rc = capture(e, nsteps, out);
    if (rc) return rc;
    e->runs.push_back({e->parity, nsteps, *out});
    return DG_OK;
}

void drop_graphs(dg_ensemble *e)
{
    for (int b = 0; b < 2; ++b)
        if (e->graph[b]) {
            (void)hipGraphExecDestroy(e->graph[b]);
            e->graph[b] = nullptr;
        }
    for (auto &r : e->runs) (void)hipGraphExecDestroy(r.exec);
    e->runs.clear();
}

void drop_graphs(dg_ensemble *e);

// Instance table and work list from the instances' current tiling: every
// (instance, tile) is one item with its own partial slot; items are ordered
// heaviest first (x-planes marched, halo included), ties in instance and tile
// order, so the persistent workgroups' round-robin share is balanced.
int build_worklist(dg_ensemble *e)
{
    const int count = (int)e->inst.size();
    std::vector<dg::EnsInstance> tab(count);
    std::vector<dg::EnsItem> items;
    std::vector<int> cost;
    long long slot = 0;
    for (int i = 0; i < count; ++i) {
        dg_ctx *c = e->inst[i];
        const dg::StepArgs &a = c->args;
        dg::EnsInstance &I = tab[i];
        std::memset(&I, 0, sizeof(I));
        I.args = a;
        for (int b = 0; b < 2; ++b)
            for (int q = 0; q < 5; ++q) I.buf[b][q] = c->buf[b][q];
        I.sc = c->sc;
        I.part_base = (int)slot;
        I.nparts = a.nblocks;
        I.cfl = c->cfg.cfl;
        I.nan_fallback = c->cfg.dt_nan_fallback;
        if (slot + a.nblocks > INT_MAX / 2) return efail(e, DG_ERR_INVALID, "too many work items");
        for (int bx = 0; bx < a.nbx; ++bx) {
            const int planes = std::min(a.lx, a.n[0] - bx * a.lx);
            for (int by = 0; by < a.nby; ++by)
                for (int bz = 0; bz < a.nbz; ++bz) {
                    items.push_back(dg::EnsItem{i, bx, by, bz, (int)slot++});
                    cost.push_back(planes + 4);
                }
        }
    }
    std::vector<int> order(items.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return cost[p] > cost[q]; });
    std::vector<dg::EnsItem> sorted(items.size());
    for (size_t k = 0; k < order.size(); ++k) sorted[k] = items[order[k]];

    EHIP(e, hipStreamSynchronize(e->stream));
    if (!e->d_inst) EHIP(e, hipMalloc((void **)&e->d_inst, tab.size() * sizeof(dg::EnsInstance)));
    if ((int)sorted.size() > e->items_cap) {
        if (e->d_items) EHIP(e, hipFree(e->d_items));
        e->d_items = nullptr;
        e->items_cap = 0;
        EHIP(e, hipMalloc((void **)&e->d_items, sorted.size() * sizeof(dg::EnsItem)));
        e->items_cap = (int)sorted.size();
    }
    if (slot > e->slots_cap) {
        if (e->partials) EHIP(e, hipFree(e->partials));
        e->partials = nullptr;
        e->slots_cap = 0;
        EHIP(e, hipMalloc((void **)&e->partials, (size_t)slot * dg::PARTIAL_STRIDE * sizeof(double)));
        e->slots_cap = (int)slot;
    }
    EHIP(e, hipMemcpy(e->d_inst, tab.data(), tab.size() * sizeof(dg::EnsInstance), hipMemcpyHostToDevice));
    EHIP(e, hipMemcpy(e->d_items, sorted.data(), sorted.size() * sizeof(dg::EnsItem), hipMemcpyHostToDevice));
    e->nitems = (int)sorted.size();
    drop_graphs(e);
    return DG_OK;
}

int choose_grid(dg_ensemble *e, int grid)
{
    if (grid == 0) {
        int cus = 0;
        const int per_cu = dg::ens_blocks_per_cu(e->prec, e->fast, e->tile_y, e->limiter, e->waves);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device) != hipSuccess || per_cu <= 0)
            return efail(e, DG_ERR_DEVICE, "occupancy query");
        grid = per_cu * cus;
    }
    e->grid = std::min(grid, e->nitems);
    return DG_OK;
}

int check_ready(dg_ensemble *e)
{
    for (size_t i = 0; i < e->inst.size(); ++i)
        if (!e->inst[i]->have_state) return efail(e, DG_ERR_STATE, "instance " + std::to_string(i) + " has no state");
    return DG_OK;
}

}  // namespace

extern "C" {

int dg_ensemble_create(const dg_config *cfgs, int count, int grid, dg_ensemble **out)
{
    if (!cfgs || !out || count <= 0 || grid < 0) return DG_ERR_INVALID;
    *out = nullptr;
    const dg_config &c0 = cfgs[0];
    const int ty0 = c0.tile_y ? c0.tile_y : 8;
    for (int i = 1; i < count; ++i) {
        const dg_config &c = cfgs[i];
        const int ty = c.tile_y ? c.tile_y : 8;
        if (c.precision != c0.precision || c.math != c0.math || c.limiter != c0.limiter || c.device != c0.device ||
            ty != ty0)
            return DG_ERR_INVALID;
    }
    if (c0.precision == DG_FP64 && ty0 > 8) return DG_ERR_INVALID;
    if (c0.precision == DG_FP64 && c0.math == DG_MATH_FAST) return DG_ERR_INVALID;   // FP32 fast math only

    dg_ensemble *e = new (std::nothrow) dg_ensemble();
    if (!e) return DG_ERR_ALLOC;
    e->device = c0.device;
    if (hipSetDevice(e->device) != hipSuccess ||
        hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return DG_ERR_DEVICE;
    }
    int rc = DG_OK;
    e->inst.reserve(count);
    for (int i = 0; i < count && rc == DG_OK; ++i) {
        dg_config c = cfgs[i];
        c.tile_y = ty0;
        dg_ctx *ctx = nullptr;
        rc = dg::create_ctx(&c, e->stream, &ctx);
        if (rc == DG_OK) e->inst.push_back(ctx);
    }
    if (rc != DG_OK) {
        dg_ensemble_destroy(e);
        return rc;
    }
    for (dg_ctx *c : e->inst)
        if (c->heat_mask) {   // the persistent step kernel has no heat source
            dg_ensemble_destroy(e);
            return DG_ERR_INVALID;
        }
    dg_ctx *f = e->inst[0];
    e->prec = f->prec;
    e->fast = f->fast;
    e->tile_y = f->tile_y;
    e->limiter = f->cfg.limiter;

    if (hipMalloc((void **)&e->d_count, sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&e->d_next, sizeof(int)) != hipSuccess ||
        hipMemset(e->d_next, 0, sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&e->h_count, sizeof(int), hipHostMallocDefault) != hipSuccess) {
        dg_ensemble_destroy(e);
        return DG_ERR_ALLOC;
    }
    for (dg_ctx *c : e->inst) e->cells += (int64_t)c->args.n[0] * c->args.n[1] * c->args.n[2];
    if ((rc = build_worklist(e)) || (rc = choose_grid(e, grid))) {
        dg_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return DG_OK;
}

void dg_ensemble_destroy(dg_ensemble *e)
{
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    for (dg_ctx *c : e->inst) dg_destroy(c);
    if (e->d_inst) (void)hipFree(e->d_inst);
    if (e->d_items) (void)hipFree(e->d_items);
    if (e->partials) (void)hipFree(e->partials);
    if (e->d_count) (void)hipFree(e->d_count);
    if (e->d_next) (void)hipFree(e->d_next);
    if (e->h_count) (void)hipHostFree(e->h_count);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *dg_ensemble_last_error(const dg_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

int dg_ensemble_set_state(dg_ensemble *e, int i, const void *const f[5], int kind)
{
    if (!e || i < 0 || i >= (int)e->inst.size()) return efail(e, DG_ERR_INVALID, "bad instance");
    const int rc = dg_set_state(e->inst[i], f, kind);
    if (rc) e->err = e->inst[i]->err;
    return rc;
}

int dg_ensemble_get_state(dg_ensemble *e, int i, void *const f[5], int kind)
{
    if (!e || i < 0 || i >= (int)e->inst.size()) return efail(e, DG_ERR_INVALID, "bad instance");
    const int rc = dg_get_state(e->inst[i], f, kind);
    if (rc) e->err = e->inst[i]->err;
    return rc;
}

int dg_ensemble_run(dg_ensemble *e, int64_t max_steps, double t_end)
{
    if (!e || max_steps < 0) return efail(e, DG_ERR_INVALID, "bad argument");
    int rc = check_ready(e);
    if (rc) return rc;
    EHIP(e, hipSetDevice(e->device));
    if ((rc = arm(e, max_steps, t_end))) return rc;
    const bool timed = std::isfinite(t_end);
    const int n = (int)e->inst.size();
    int64_t left = max_steps;
    if (!timed && max_steps > 1 && max_steps <= kWholeRunSteps) {
        // the whole run is one graph launch
        hipGraphExec_t g = nullptr;
        if ((rc = run_graph(e, max_steps, &g))) return rc;
        EHIP(e, hipGraphLaunch(g, e->stream));
        after_steps(e, max_steps);
        left = 0;
    }
    while (left > 0) {
        int64_t did;
        if (left >= kEnsGraphSteps) {
            if ((rc = ensure_graph(e))) return rc;
            EHIP(e, hipGraphLaunch(e->graph[e->parity], e->stream));
            did = kEnsGraphSteps;
        } else {
            if ((rc = launch_one(e, e->parity))) return rc;
            did = 1;
        }
        after_steps(e, did);
        left -= did;
        if (timed) {
            hipLaunchKernelGGL(ens_count_active_kernel, dim3(1), dim3(64), 0, e->stream, e->d_inst, n, e->d_count);
            EHIP(e, hipMemcpyAsync(e->h_count, e->d_count, sizeof(int), hipMemcpyDeviceToHost, e->stream));
            EHIP(e, hipStreamSynchronize(e->stream));
            if (*e->h_count == 0) break;
        }
    }
    EHIP(e, hipStreamSynchronize(e->stream));
    for (dg_ctx *c : e->inst) {
        EHIP(e, hipMemcpyAsync(c->sc_host, c->sc, sizeof(dg::Scalars), hipMemcpyDeviceToHost, e->stream));
    }
    EHIP(e, hipStreamSynchronize(e->stream));
    for (size_t i = 0; i < e->inst.size(); ++i) {
        dg_ctx *c = e->inst[i];
        if (!c->cfg.dt_nan_fallback && c->sc_host->nonfinite)
            return efail(e, DG_ERR_NONFINITE, "instance " + std::to_string(i) + ": dt is not finite");
    }
    return DG_OK;
}

int dg_ensemble_clock(dg_ensemble *e, int i, double *t, int64_t *steps, double *next_dt)
{
    if (!e || i < 0 || i >= (int)e->inst.size()) return efail(e, DG_ERR_INVALID, "bad instance");
    dg_ctx *c = e->inst[i];
    EHIP(e, hipMemcpyAsync(c->sc_host, c->sc, sizeof(dg::Scalars), hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    if (t) *t = c->sc_host->t;
    if (steps) *steps = c->sc_host->steps;
    if (next_dt) *next_dt = c->sc_host->dt;
    return DG_OK;
}

int dg_ensemble_conservation_sums(dg_ensemble *e, int i, double out[5])
{
    if (!e || i < 0 || i >= (int)e->inst.size()) return efail(e, DG_ERR_INVALID, "bad instance");
    const int rc = dg_conservation_sums(e->inst[i], out);
    if (rc) e->err = e->inst[i]->err;
    return rc;
}

int dg_ensemble_time_steps(dg_ensemble *e, int64_t nsteps, double *ms_step_kernel, double *ms_dt_kernel)
{
    if (!e || nsteps <= 0) return efail(e, DG_ERR_INVALID, "bad argument");
    int rc = check_ready(e);
    if (rc) return rc;
    EHIP(e, hipSetDevice(e->device));
    if ((rc = arm(e, LLONG_MAX, INFINITY))) return rc;
    dg::EventSet es;
    EHIP(e, es.create((size_t)nsteps * 3));
    std::vector<hipEvent_t> &ev = es.ev;
    for (int64_t s = 0; s < nsteps && rc == DG_OK; ++s) {
        if (hipEventRecord(ev[3 * s], e->stream) != hipSuccess) rc = DG_ERR_DEVICE;
        if (dg::launch_ens_step(e->prec, e->fast, e->tile_y, e->limiter, e->waves, e->d_inst, e->d_items, e->nitems,
                                e->grid, e->parity, e->partials, e->d_next, e->stream))
            rc = efail(e, DG_ERR_DEVICE, "ensemble step launch");
        if (hipEventRecord(ev[3 * s + 1], e->stream) != hipSuccess) rc = DG_ERR_DEVICE;
        if (dg::launch_ens_finalize(e->prec, e->d_inst, (int)e->inst.size(), e->partials, 1, e->d_next, e->stream))
            rc = efail(e, DG_ERR_DEVICE, "ensemble finalize launch");
        if (hipEventRecord(ev[3 * s + 2], e->stream) != hipSuccess) rc = DG_ERR_DEVICE;
        after_steps(e, 1);
    }
    if (hipStreamSynchronize(e->stream) != hipSuccess && rc == DG_OK) rc = efail(e, DG_ERR_DEVICE, "synchronize");
    double tk = 0, tf = 0;
    float ms = 0;
    for (int64_t s = 0; s < nsteps && rc == DG_OK; ++s) {
        (void)hipEventElapsedTime(&ms, ev[3 * s], ev[3 * s + 1]);
        tk += ms;
        (void)hipEventElapsedTime(&ms, ev[3 * s + 1], ev[3 * s + 2]);
        tf += ms;
    }
    if (rc) return rc;
    if (ms_step_kernel) *ms_step_kernel = tk / nsteps;
    if (ms_dt_kernel) *ms_dt_kernel = tf / nsteps;
    return DG_OK;
}

int dg_ensemble_retile(dg_ensemble *e, int tile_y, const int32_t *chunk_x, int grid, int waves)
{
    if (!e || tile_y < 0 || grid < 0 || (waves != 0 && waves != 3 && waves != 4))
        return efail(e, DG_ERR_INVALID, "bad argument");
    if (waves) e->waves = waves;
    EHIP(e, hipSetDevice(e->device));
    EHIP(e, hipStreamSynchronize(e->stream));
    const int ty = tile_y ? tile_y : e->tile_y;
    for (size_t i = 0; i < e->inst.size(); ++i) {
        dg_ctx *c = e->inst[i];
        int rc = dg::set_tiling(c, ty, chunk_x ? chunk_x[i] : c->chunk_x);
        if (rc == DG_OK) rc = dg::reserve_partials(c);
        if (rc) return efail(e, rc, "instance " + std::to_string(i) + ": bad tiling");
    }
    e->tile_y = ty;
    int rc = build_worklist(e);
    if (rc == DG_OK) rc = choose_grid(e, grid);
    return rc;
}

int dg_ensemble_info(dg_ensemble *e, int64_t *cells, int *items, int *grid, int *tile_y)
{
    if (!e) return DG_ERR_INVALID;
    if (cells) *cells = e->cells;
    if (items) *items = e->nitems;
    if (grid) *grid = e->grid;
    if (tile_y) *tile_y = e->tile_y;
    return DG_OK;
}

}  // extern "C"
