// dg_ctx.h -- the context object behind the C ABI (shared by dg_api.cpp and
// dg_ensemble.cpp).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include <string>

#include "../../include/deepgreen.h"
#include "dg_internal.h"

struct dg_ctx {
    uint64_t uid = 0;     // unique per created context (never reused, unlike the address)
    dg_config cfg{};
    int prec = 1, fast = 0, tile_y = 8, chunk_x = 8;
    dg::StepArgs args{};
    size_t n3 = 0, el = 4;
    void *buf[2][5] = {};
    int cur = 0;
    double *partials = nullptr;
    int partials_cap = 0;     // slots allocated
    int nparts_reduce = 0;
    dg::Scalars *sc = nullptr;      // device
    dg::Scalars *sc_host = nullptr; // pinned mirror
    void *stage[2] = {nullptr, nullptr};     // pinned staging, one field each (double-buffered host transfers)
    hipEvent_t stage_ev[2] = {nullptr, nullptr};  // copy out of / into stage[b] done
    void *scratch = nullptr;        // device scratch (one state) for primitive downloads
    void *dumpbuf = nullptr;        // ghosted arrays of dg_step_dump (allocated at the first dump)
    hipStream_t stream = nullptr;
    bool owns_stream = true;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // dg_run replay, by state buffer parity
    int graph_steps = 0;
    int heat_mask = 0;        // conducting lower walls (bit 0 x-, 1 y-, 2 z-)
    double heat_t_out = 273.0, heat_v_out = 0.0, heat_rs = 287.0;
    bool have_state = false;
    bool sums_stale = true;   // the step kernel does not accumulate sums
    std::string err;
};


namespace dg {
// Validates a configuration and derives the kernel arguments (default tiling
// included); tile_y_out / heat_mask_out may be null.
int make_args(const dg_config *cfg, StepArgs *out, int *tile_y_out, int *heat_mask_out, std::string *why = nullptr);
// Workgroup tiling of a grid (tile_y rows, chunk_x planes; 0 = automatic).
int tile_args(StepArgs &a, int prec, int tile_y, int chunk_x, int *chunk_out);
// why (may be null) receives the reason of a failure
int create_ctx(const dg_config *cfg, hipStream_t shared_stream, dg_ctx **out, std::string *why = nullptr);
int set_tiling(dg_ctx *c, int tile_y, int chunk_x);
int reserve_partials(dg_ctx *c);
// Clears the stop condition of a previous dg_run so that the next steps advance.
int arm_unbounded(dg_ctx *c);
// One step (step kernel, then dt and clock: finalize mode 1); flips c->cur.
int launch_one_step(dg_ctx *c);
// Reduction of the current state: mode 0 dt and sums of a fresh state, 2 sums only.
int reduce_current(dg_ctx *c, int mode);
}
