"""Benchmark: fused compressible-Euler step on one MI355X.

Workload (default, --workload khi): 3D Kelvin-Helmholtz shear layer, 512^3
periodic, FP32 SoA state -- the configuration the HBM-roofline figure is
quoted on.  One "step" is one full reference step (primitives, dt reduction,
slopes, Hancock half-step, Rusanov fluxes on all faces, conservative update).
Inputs are resident in HBM before the timed region; the timed region runs K
steps from the library's captured HIP graph with the dt kept on the device.

--workload ensemble: 256 cubic grids with edges drawn by seed from
{32, 48, 64, 96, 128}, random planar Riemann problems, outflow boundaries,
minmod slopes, FP32; one step advances every grid (one persistent step
launch + one dt launch, replayed from a captured HIP graph); tiles from the
autotuner's cache (tuned first if the cache has no entry).

Prints one JSON line (rank 0).  Fields beyond the contract:
  roofline      dominant kernel (the fused step) against HBM: algorithmic
                bytes per launch (read U + write U: 10 * 4 B per interior
                cell) / mean launch time from HIP events on the library's
                stream; traffic = HBM bytes per launch measured by rocprofv3
                (FETCH_SIZE + WRITE_SIZE, copy-calibrated) where a committed
                profile summary holds it.
  cpu_baseline  the reference CPU path on this host, on a bounded sample.
  secondary     (default KHI run) short runs of the other configurations:
                KHI 512^3 FP64 fast and FP32 exact, the ensemble, the blast.
  pcie_inclusive  (default KHI run) the rate of a host that hands its arrays
                over every step (set state, step, get state); never `value`.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "deep-green_amd"))

HBM_PEAK_GBS = 8000.0


def traffic_from_profile(name, kernel_ms):
    """HBM bytes per launch recorded in profiles/<name> (the line
    'hbm_bytes_per_launch: <int>'), or None.  The summary also records the
    kernel's mean duration in the profiled run ('kernel_ms: <float>'); a
    summary whose duration differs from this run's by more than 10 % was
    taken of another build of the kernel and is not used."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            txt = f.read()
    except OSError:
        return None
    m = re.search(r"^hbm_bytes_per_launch:\s*(\d+)", txt, re.M)
    t = re.search(r"^kernel_ms:\s*([0-9.eE+-]+)", txt, re.M)
    if not m or not t or abs(float(t.group(1)) - kernel_ms) > 0.1 * kernel_ms:
        return None
    return int(m.group(1))


GRAPH_STEPS = 16   # dg_run replays captured graphs of this many steps


def warmup_steps(requested):
    """Untimed warm-up length: at least one captured graph's worth of steps
    and a multiple of it, so the graph the timed run replays exists (and is
    instantiated) before the clock starts."""
    return max(GRAPH_STEPS, -(-int(requested) // GRAPH_STEPS) * GRAPH_STEPS)


def cpu_baseline(n_sample=256, steps=12):
    """Reference CPU path: the reference's own Fortran physics module
    (oracle/_ref harness, real(4) build, gamma 1.4 / CFL 0.4) on a KHI sample
    of n_sample^3 cells, all host cores; falls back to the C oracle (a port)
    when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import dg_oracle
    cores = os.cpu_count() or 1
    from deepgreen.setups import khi
    w, ds = khi(n_sample, np.float64)
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


def cpu_baseline_ensemble(members=64, steps=20, seed=0):
    """CPU path for the ensemble workload: the C oracle (a port of
    physics.rs with the minmod slopes the ensemble uses; the reference has no
    limited scheme to time) on the first `members` members of the seeded
    ensemble, all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import dg_oracle
    from deepgreen.setups import ensemble_members, planar_riemann
    cores = os.cpu_count() or 1
    cells, wall = 0, 0.0
    for n, rng in ensemble_members(256, seed)[:members]:
        w, ds, _ = planar_riemann(n, np.random.default_rng(rng.bit_generator.seed_seq), np.float64)
        o = dg_oracle.Oracle((n,) * 3, ds, [1] * 6, precision="f32", limiter="minmod", nthreads=cores)
        o.set_primitives(w)
        o.calc_conserved()
        t0 = time.perf_counter()
        o.steps(steps)
        wall += time.perf_counter() - t0
        cells += n ** 3 * steps
    return {"value": cells / wall / 1e9, "unit": "Gcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"first {members} ensemble members, {steps} steps each, C oracle (physics.rs port, minmod), "
                      f"OpenMP, {wall:.2f} s"}


def bench_khi(args):
    import deepgreen as dg
    from deepgreen.setups import khi

    n = args.n
    dtype = np.float64 if args.precision == "fp64" else np.float32
    w, ds = khi(n, dtype)
    sim = dg.Simulation(n, n, n, ds, [dg.PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0, precision=args.precision,
                        math=args.math, tile_y=args.tile_y)
    solver = dg.Solver(sim)
    solver.set_primitives(w)
    del w
    info = solver.kernel_info()
    print(f"[bench] KHI {n}^3 {args.precision}/{args.math} tile {info}", file=sys.stderr, flush=True)

    solver.run(warmup_steps(args.warmup))
    # timed region: K steps from the captured graph, bracketed by device syncs
    solver.next_dt()
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
    pcie = None
    if getattr(args, "pcie", False):
        # a host that hands its arrays over every step (dg_set_state, dg_step,
        # dg_get_state with the reference's ghosted layout in pageable memory)
        U = solver.get_state()
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            solver.set_conserved(U)
            solver.compute_step()
            U = solver.get_state()
        el_s = (time.perf_counter() - t0) / reps
        pcie = {"value": round(cells / el_s / 1e9, 3), "unit": "Gcell-updates/s", "ms_per_step": round(el_s * 1e3, 1),
                "note": f"per step: dg_set_state(conserved) + dg_step + dg_get_state(conserved) of ghosted "
                        f"{n + 4}^3 arrays in pageable host memory; never the headline value"}
        del U
    solver.close()
    workload = f"khi{n}_{args.precision}_{args.math}"
    traffic = traffic_from_profile(f"step_kernel_{workload}_hbm.md", ms_step)
    return {
        "pcie_inclusive": pcie,
        "s_per_step": elapsed / args.steps, "value": gcells, "dtype": "f64" if args.precision == "fp64" else "f32",
        "config": {"workload": f"khi{n}_{args.precision}", "cells": cells, "periodic": True,
                   "gamma": 1.4, "cfl": 0.4, "math": args.math, "limiter": "none",
                   "pct_hbm_peak": round(100.0 * gcells * 1e9 * 10 * el / (HBM_PEAK_GBS * 1e9), 2),
                   "tile": info, "state_finite": finite},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "step_kernel", "kernel_ms": round(ms_step, 4),
                     "reduce_kernel_ms": round(ms_red, 4), "alg_bytes_per_launch": alg_bytes},
    }


def bench_ensemble(args):
    import deepgreen as dg
    from deepgreen import autotune
    from deepgreen.setups import ensemble_members, ensemble_workload

    members = ensemble_members(args.members, args.seed)
    sizes = sorted({n for n, _ in members})
    counts = [sum(1 for m, _ in members if m == n) for n in sizes]
    t0 = time.perf_counter()
    choice = autotune.lookup(sizes, math=args.math)
    tuned_now = choice is None
    if tuned_now:
        choice = autotune.tune(sizes, counts, math=args.math,
                               log=lambda s: print(s, file=sys.stderr, flush=True))
    tune_s = time.perf_counter() - t0
    sims, states = ensemble_workload(args.members, args.seed, math=args.math, choice=choice)
    ens = dg.Ensemble(sims)
    autotune.apply(ens, choice)
    for i, w in states():
        ens.set_state(i, w)
    info = ens.info()
    print(f"[bench] ensemble of {len(sims)} ({dict(zip(sizes, counts))}), {info}, tiles {choice['tile_y']} "
          f"{choice['chunk_x']} waves {choice.get('waves', 4)} blocks/CU {choice['blocks_per_cu']}",
          file=sys.stderr, flush=True)
    # the timed run replays one captured graph of exactly args.steps steps;
    # the warm-up runs that graph once, so its capture is not timed
    ens.run(args.steps)
    ens.clock(0)
    t0 = time.perf_counter()
    ens.run(args.steps)
    ens.clock(0)   # synchronizes the stream
    elapsed = time.perf_counter() - t0
    cells = info["cells"]
    gcells = cells * args.steps / elapsed / 1e9
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)
    ms_step, ms_dt = ens.time_steps(min(args.steps, 20))
    alg_bytes = 40 * cells
    achieved = alg_bytes / (ms_step * 1e-3) / 1e9
    finite = all(bool(np.isfinite(ens.conservation_sums(i)).all()) for i in range(0, len(sims), 16))
    ens.close()
    traffic = traffic_from_profile("ens_step_kernel_ensemble256_fp32_fast_hbm.md", ms_step)
    return {
        "s_per_step": elapsed / args.steps, "value": gcells, "dtype": "f32",
        "config": {"workload": f"ensemble{args.members}_fp32", "members": len(sims), "cells": cells,
                   "sizes": dict(zip(map(str, sizes), counts)), "seed": args.seed, "boundaries": "outflow",
                   "gamma": 1.4, "cfl": 0.4, "math": args.math, "limiter": "minmod",
                   "pct_hbm_peak": round(100.0 * gcells * 1e9 * 40 / (HBM_PEAK_GBS * 1e9), 2),
                   "tile_y": choice["tile_y"], "waves": choice.get("waves", 4), "chunk_x": choice["chunk_x"],
                   "blocks_per_cu": choice["blocks_per_cu"], "work_items": info["items"], "grid": info["grid"],
                   "autotune_s": round(tune_s, 2), "tuned_in_run": tuned_now, "state_finite": finite},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "ens_step_kernel", "kernel_ms": round(ms_step, 4),
                     "dt_kernel_ms": round(ms_dt, 4), "alg_bytes_per_launch": alg_bytes},
    }


def bench_learned(args):
    """Config: learned flux correction at 128^3 (FP32, minmod, periodic).
    Training data from plain 256^3 runs (seeded random ICs) restricted to
    128^3; Adam lr 1e-3, batch 8 of 32^3 crops; evaluation against the plain
    512^3 solution restricted to 128^3, all runs from restrictions of the
    same 512^3 initial condition, to the same final time."""
    from deepgreen.learned import FluxCNN, evaluate, make_training_set

    def log(s):
        print(s, file=sys.stderr, flush=True)

    m = FluxCNN(seed=args.seed)
    t0 = time.perf_counter()
    make_training_set(m, seeds=range(1, 1 + args.train_ics), n_fine=2 * args.n_learned, t_end=args.t_end,
                      trajectory=args.train_mode == "trajectory",
                      every=args.sample_every, log=log)
    gen_s = time.perf_counter() - t0
    m.fit_normalization()
    t0 = time.perf_counter()
    losses = m.train(args.train_steps, batch=8, crop=32, lr=1e-3, seed=args.seed, log=log)
    train_s = time.perf_counter() - t0
    log(f"[learned] {m.samples} samples in {gen_s:.1f} s; {args.train_steps} Adam steps in {train_s:.1f} s, "
        f"loss {losses[:5].mean():.4g} -> {losses[-50:].mean():.4g}")
    res = evaluate(m, seed=1000 + args.seed, n_truth=4 * args.n_learned, T=args.t_end,
                   warmup_steps=args.learned_warmup, log=log)
    e, r = res["l1_rho"], res["runs"]
    speedup = r["plain_half"]["wall_s"] / r["learned_quarter"]["wall_s"]
    matched = e["learned_quarter"] <= e["plain_half"]
    n = args.n_learned
    # dominant kernel of the learned step: the 32 -> 32 convolution (MFMA)
    import deepgreen as dg
    from deepgreen.learned import _sim
    from deepgreen.setups import learned_ic
    w, ds = learned_ic(n, 1000 + args.seed)
    s = dg.Solver(_sim(n, ds))
    s.set_primitives(w)
    ms = m.time_forward(s, nrep=20)
    s.close()
    flops = (n + 2) ** 3 * 2 * 27 * 32 * 32
    tflops = flops / (ms[2] * 1e-3) / 1e12
    return {
        "metric": "error-matched speedup learned vs plain", "unit": "x", "value": speedup, "dtype": "f32",
        "s_per_step": r["learned_quarter"]["wall_s"] / max(r["learned_quarter"]["steps"], 1),
        "config": {"workload": f"learned{n}_fp32", "error_matched": matched,
                   "l1_rho": e, "runs": r, "T": args.t_end, "train_ics": args.train_ics,
                   "samples": m.samples, "train_steps": args.train_steps, "batch": 8, "crop": 32, "lr": 1e-3,
                   "train_mode": args.train_mode,
                   "loss_first": float(losses[:5].mean()), "loss_last": float(losses[-50:].mean()),
                   "data_gen_s": round(gen_s, 2), "train_s": round(train_s, 2),
                   "mass_learned_vs_plain": res["mass_learned_vs_plain"], "warmup_steps": args.learned_warmup,
                   "forward_ms": {"prep": round(ms[0], 4), "conv1": round(ms[1], 4), "conv2": round(ms[2], 4),
                                  "conv3": round(ms[3], 4), "total": round(ms[4], 4)}},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(tflops / 2500.0, 4), "traffic": None, "kernel": "conv_ring_kernel<32, 32, 1, 0>",
                     "kernel_ms": round(ms[2], 4), "flop_per_launch": flops},
    }


def bench_blast(args):
    """Config: 3D blast wave, 256^3 periodic unit cube, gamma 1.4, CFL 0.4,
    background rho 1, p 0.1, p 10 inside r < 0.1 about the centre, 500
    steps, minmod slopes (the reference's unlimited slopes turn the blast
    non-finite in step 2), FP64 and FP32-bulk (exact math, and fast math as a
    second line) from the same initial state; reports the FP32-bulk errors
    against FP64 (relative L1 and Linf per conserved field) and the
    throughputs."""
    import deepgreen as dg
    from deepgreen.setups import blast

    n = args.n_blast
    out = {}
    U = {}
    for key, prec, math_ in (("fp64", "fp64", "exact"), ("fp32_bulk", "fp32_bulk", "exact"),
                             ("fp32_bulk_fast", "fp32_bulk", "fast")):
        w, ds = blast(n, np.float64 if prec == "fp64" else np.float32)
        sim = dg.Simulation(n, n, n, ds, [dg.PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0, precision=prec,
                            math=math_, limiter="minmod")
        s = dg.Solver(sim)
        s.set_primitives(w)
        del w
        s.next_dt()
        t0 = time.perf_counter()
        t, steps = s.run(args.steps)
        s.next_dt()
        el = time.perf_counter() - t0
        out[key] = {"gcells": n ** 3 * args.steps / el / 1e9, "t": t, "steps": steps}
        U[key] = np.stack([np.asarray(a[2:-2, 2:-2, 2:-2], np.float64) for a in s.get_state()])
        s.close()
        print(f"[bench] blast {n}^3 {key}: {args.steps} steps in {el:.3f} s, t = {t:.6f}", file=sys.stderr,
              flush=True)
    names = ("mass", "momentum_x", "momentum_y", "momentum_z", "energy")
    errs = {}
    for key in ("fp32_bulk", "fp32_bulk_fast"):
        err = {}
        for q, nm in enumerate(names):
            d = np.abs(U[key][q] - U["fp64"][q])
            ref_l1 = np.abs(U["fp64"][q]).mean()
            ref_max = np.abs(U["fp64"][q]).max()
            err[nm] = {"rel_l1": float(d.mean() / ref_l1), "rel_linf": float(d.max() / ref_max)}
        errs[key] = err
    return {
        "s_per_step": 1.0 / (out["fp32_bulk"]["gcells"] * 1e9 / n ** 3), "value": out["fp32_bulk"]["gcells"],
        "dtype": "f32",
        "config": {"workload": f"blast{n}_fp64_vs_fp32bulk", "cells": n ** 3, "limiter": "minmod",
                   "fp64_gcells": round(out["fp64"]["gcells"], 3), "fp32_bulk_gcells": round(out["fp32_bulk"]["gcells"], 3),
                   "fp32_bulk_fast_gcells": round(out["fp32_bulk_fast"]["gcells"], 3),
                   "t_fp64": out["fp64"]["t"], "t_fp32_bulk": out["fp32_bulk"]["t"],
                   "t_fp32_bulk_fast": out["fp32_bulk_fast"]["t"],
                   "fp32_bulk_error_vs_fp64": errs["fp32_bulk"],
                   "fp32_bulk_fast_error_vs_fp64": errs["fp32_bulk_fast"]},
        "roofline": None,
    }


def secondary(args):
    """Short runs of the other benchmark configurations after the headline
    one (the default run only): KHI 512^3 in FP64 fast math and in FP32
    exact math, the 256-grid ensemble and the 256^3 blast (FP64 against
    FP32-bulk).  Each entry holds its own throughput; a configuration that
    fails records the error instead of ending the run."""
    import copy
    out = {}

    def run(name, fn, **kw):
        a = copy.copy(args)
        for k, v in kw.items():
            setattr(a, k, v)
        t0 = time.perf_counter()
        try:
            r = fn(a)
        except Exception as e:  # noqa: BLE001 - recorded, the headline line still prints
            out[name] = {"error": f"{type(e).__name__}: {e}"}
            return
        entry = {"value": round(r["value"], 3), "unit": "Gcell-updates/s", "steps": a.steps,
                 "ms_per_step": round(r["s_per_step"] * 1e3, 4), "wall_s": round(time.perf_counter() - t0, 1)}
        if r.get("roofline"):
            entry["kernel_ms"] = r["roofline"]["kernel_ms"]
            entry["frac_hbm"] = r["roofline"]["frac"]
            entry["traffic"] = r["roofline"]["traffic"]
        cfg = r["config"]
        for k in ("fp64_gcells", "fp32_bulk_gcells", "fp32_bulk_fast_gcells", "fp32_bulk_error_vs_fp64", "members",
                  "cells"):
            if k in cfg:
                entry[k] = cfg[k]
        out[name] = entry
        print(f"[bench] secondary {name}: {entry['value']} Gcell-updates/s", file=sys.stderr, flush=True)

    run("khi512_fp64_fast", bench_khi, precision="fp64", math="fast", steps=48, warmup=16, pcie=False)
    run("khi512_fp32_exact", bench_khi, precision="fp32", math="exact", steps=48, warmup=16, pcie=False)
    run("ensemble256_fp32_fast", bench_ensemble, workload="ensemble", math="fast", steps=40)
    run("blast256_fp64_vs_fp32bulk", bench_blast, workload="blast", steps=500)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (khi: 200, ensemble: 100)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="khi", choices=["khi", "ensemble", "learned", "blast"])
    ap.add_argument("--n-blast", type=int, default=256)
    ap.add_argument("--n-learned", type=int, default=128)
    ap.add_argument("--train-ics", type=int, default=32)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--sample-every", type=int, default=4)
    ap.add_argument("--train-mode", default="reset", choices=["reset", "trajectory"],
                    help="learned workload: samples of restricted fine states (reset) or of the coarse grid's own "
                         "plain trajectory (trajectory)")
    ap.add_argument("--t-end", type=float, default=0.03)
    ap.add_argument("--learned-warmup", type=int, default=0, help="plain steps before the correction starts")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--math", default="fast", choices=["fast", "exact"])
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp32_bulk", "fp64"])
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="default KHI run: skip the short runs of the other configurations")
    ap.add_argument("--cpu-sample", type=int, default=256, help="edge of the KHI sample the CPU path runs")
    ap.add_argument("--cpu-steps", type=int, default=12, help="steps of the CPU sample (~10 s on 16 cores)")
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
    if args.steps is None:
        args.steps = {"ensemble": 100, "blast": 500}.get(args.workload, 200)

    if args.workload == "ensemble":
        r = bench_ensemble(args)
    elif args.workload == "learned":
        r = bench_learned(args)
    elif args.workload == "blast":
        r = bench_blast(args)
    else:
        args.pcie = not args.no_secondary
        r = bench_khi(args)
    sec = None
    if args.workload == "khi" and not args.no_secondary and args.n == 512 and args.precision == "fp32" \
            and args.math == "fast":
        sec = secondary(args)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_ensemble() if args.workload == "ensemble" else cpu_baseline(args.cpu_sample,
                                                                                      args.cpu_steps)
    out = {
        "metric": r.get("metric", "Gcell-updates/s + % HBM peak"),
        "value": round(r["value"], 3),
        "unit": r.get("unit", "Gcell-updates/s"),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": {"khi": warmup_steps(args.warmup), "ensemble": args.steps}.get(args.workload, args.warmup),
        "ms_per_step": round(r["s_per_step"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": r["dtype"],
        "data": "synthetic",
        "config": r["config"],
        "roofline": r["roofline"],
        "cpu_baseline": cpu,
    }
    if sec is not None:
        out["secondary"] = sec
    if r.get("pcie_inclusive"):
        out["pcie_inclusive"] = r["pcie_inclusive"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
