"""Benchmark: fused compressible-Euler step on one MI355X.

Workload (default): 3D Kelvin-Helmholtz shear layer, 512^3 periodic, FP32 SoA
state -- the configuration the HBM-roofline figure is quoted on.  One "step"
is one full reference step (primitives, dt reduction, slopes, Hancock
half-step, Rusanov fluxes on all faces, conservative update).  Inputs are
resident in HBM before the timed region; the timed region runs K steps from
the library's captured HIP graph with the dt kept on the device.

Prints one JSON line (rank 0).  Fields beyond the contract:
  roofline      dominant kernel (the fused step) against HBM: algorithmic
                bytes per launch (read U + write U: 10 * 4 B per interior
                cell) / mean launch time from HIP events on the library's
                stream.
  cpu_baseline  the reference CPU path on this host, on a bounded sample.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "deep-green_amd"))

HBM_PEAK_GBS = 8000.0


def khi_state(n, dtype, seed=1):
    """Shear layer along y (|y - 1/2| < 1/4: rho 2, u +0.5; else rho 1,
    u -0.5), p 2.5, v = 0.01 sin(4 pi x) + U(-1e-3, 1e-3) from the seed,
    cell-centred coordinates on the unit cube; ghosted C-order arrays."""
    ds = 1.0 / n
    c = (np.arange(n + 4) - 2 + 0.5) * ds
    shape = (n + 4,) * 3
    inner = (np.abs(c - 0.5) < 0.25)[None, :, None]
    rho = np.empty(shape, dtype)
    rho[...] = np.where(inner, 2.0, 1.0)
    u = np.empty(shape, dtype)
    u[...] = np.where(inner, 0.5, -0.5)
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1e-3, 1e-3, size=shape).astype(dtype)
    v += (0.01 * np.sin(4 * np.pi * c)).astype(dtype)[:, None, None]
    w = np.zeros(shape, dtype)
    p = np.full(shape, 2.5, dtype)
    return [rho, u, v, w, p], ds


def cpu_baseline(n_sample=128, steps=3):
    """Reference CPU path: the reference's own Fortran physics module
    (oracle/_ref harness, real(4) build, gamma 1.4 / CFL 0.4) on a KHI sample
    of n_sample^3 cells, all host cores; falls back to the C oracle (a port)
    when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import dg_oracle
    cores = os.cpu_count() or 1
    w, ds = khi_state(n_sample, np.float64)
    cells = n_sample ** 3 * steps
    if dg_oracle.ref_binary("sp", "cfg"):
        _, _, wall = dg_oracle.run_reference_fortran(w, ds, [0] * 6, steps, precision="sp", params="cfg",
                                                     nthreads=cores)
        kind = "reference"
        what = "reference Fortran physics.f90 (real(4)), OpenMP"
    else:
        o = dg_oracle.Oracle((n_sample,) * 3, ds, [0] * 6, precision="f32", nthreads=cores)
        o.set_primitives(w)
        o.calc_conserved()
        t0 = time.perf_counter()
        o.steps(steps)
        wall = time.perf_counter() - t0
        kind = "port"
        what = "C oracle restating physics.rs, OpenMP"
    return {"value": cells / wall / 1e9, "unit": "Gcell-updates/s", "cores": cores, "kind": kind,
            "sample": f"KHI {n_sample}^3, {steps} steps, {what}, {wall:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--math", default="fast", choices=["fast", "exact"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp32_bulk", "fp64"])
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (scale for rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nb = args.copy_calibration * (1 << 20)
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": nb}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import deepgreen as dg

    n = args.n
    dtype = np.float64 if args.precision == "fp64" else np.float32
    w, ds = khi_state(n, dtype)
    sim = dg.Simulation(n, n, n, ds, [dg.PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0, precision=args.precision,
                        math=args.math, tile_y=args.tile_y)
    solver = dg.Solver(sim)
    solver.set_primitives(w)
    del w
    info = solver.kernel_info()
    print(f"[bench] KHI {n}^3 {args.precision}/{args.math} tile {info}", file=sys.stderr, flush=True)

    solver.run(args.warmup)
    # timed region: K steps from the captured graph, bracketed by device syncs
    t0 = time.perf_counter()
    solver.run(args.steps)
    solver.next_dt()  # synchronizes the stream
    elapsed = time.perf_counter() - t0
    cells = n ** 3
    gcells = cells * args.steps / elapsed / 1e9
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    # per-kernel device times (HIP events on the library's stream)
    nk = min(args.steps, 20)
    _, ms_step, ms_red = solver.time_steps(nk)
    el = 8 if args.precision == "fp64" else 4
    alg_bytes = 10 * el * cells
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    finite = bool(np.isfinite(solver.conservation_sums()).all())
    solver.close()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_sample)

    out = {
        "metric": "Gcell-updates/s + % HBM peak",
        "value": round(gcells, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": f"khi{n}_{args.precision}", "cells": cells, "periodic": True,
                   "gamma": 1.4, "cfl": 0.4, "math": args.math, "limiter": "none",
                   "pct_hbm_peak": round(100.0 * gcells * 1e9 * 10 * el / (HBM_PEAK_GBS * 1e9), 2),
                   "tile": info, "state_finite": finite},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "step_kernel", "kernel_ms": round(ms_step, 4),
                     "reduce_kernel_ms": round(ms_red, 4), "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
