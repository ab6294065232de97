"""Tile autotuner of the ensemble step.

Knobs of one persistent ensemble launch:
  tile_y         rows of a workgroup (4, 8 or 16; one kernel, so one value
                 for the whole ensemble),
  chunk_x        x-planes a workgroup marches, per size class (long chunks
                 re-read fewer halo planes, short ones give more, smaller
                 work items to balance),
  waves          occupancy target the FP32 fast step kernel is compiled for
                 (3 or 4 waves per SIMD: 168 or 128 VGPRs per lane),
  blocks_per_cu  resident workgroups per CU of the persistent grid.

The tuner builds one probe ensemble with the class mix of the real one
(scaled down to at most PROBE_CELLS cells, at least one member per class,
random planar Riemann problems), then re-tiles it in place and times the
persistent step kernel with HIP events: for every (tile_y, waves),
coordinate descent over the classes' chunk_x (largest class first); the best
(tile_y, waves, chunks); then the grid.  Results go to a JSON cache keyed by device architecture,
arithmetic (precision/math/limiter) and the set of size classes.  The choice
never changes results: each instance is bitwise the same under every tiling.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import Ensemble, Simulation, device_info, load_library
from .setups import planar_riemann

_HERE = os.path.dirname(os.path.abspath(__file__))
CACHE_PATH = os.environ.get("DG_AUTOTUNE_CACHE", os.path.join(_HERE, "autotune_cache.json"))

TILE_Y = (4, 8, 16)
WAVES = (4, 3)
BLOCKS_PER_CU = (1, 2, 0)           # 0: as many as fit
PROBE_CELLS = 32 * 1024 * 1024


def arith_key(precision, math, limiter):
    return f"{precision}/{math}/{limiter}"


def classes_key(sizes):
    return ",".join(str(n) for n in sorted(set(int(n) for n in sizes)))


def load_cache(path=CACHE_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=CACHE_PATH):
    try:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
        return path
    except OSError:
        return None


def lookup(sizes, precision="fp32", math="fast", limiter="minmod", device=0, cache_path=CACHE_PATH):
    """Cached choice for these size classes on this device, or None."""
    arch, _ = device_info(device)
    return load_cache(cache_path).get(arch, {}).get(arith_key(precision, math, limiter), {}).get(classes_key(sizes))


def chunk_candidates(n):
    return sorted({c for c in (8, 16, 32, 64, n) if c <= n})


def grid_for(blocks_per_cu, device=0):
    """Grid argument of an Ensemble for a blocks-per-CU choice (0: as many as fit)."""
    return 0 if not blocks_per_cu else int(blocks_per_cu) * device_info(device)[1]


def probe_mix(sizes, counts, max_cells=None):
    """Members per class of the probe: the real proportions, scaled to at most
    max_cells (default PROBE_CELLS) cells, at least one per class."""
    max_cells = PROBE_CELLS if max_cells is None else max_cells
    total = sum(c * n ** 3 for n, c in zip(sizes, counts))
    scale = min(1.0, max_cells / total)
    return [max(1, int(round(c * scale))) for c in counts]


def tune(sizes, counts=None, precision="fp32", math="fast", limiter="minmod", device=0, cache_path=CACHE_PATH,
         retune=False, steps=5, log=None):
    """Tile choice for an ensemble whose members have the edge lengths
    `sizes` (`counts` members each; default one each).  Returns the cache
    entry: {"tile_y", "chunk_x": {str(n): lx}, "blocks_per_cu", "ms_probe",
    "probe": {...}, "trials": [...]}."""
    load_library()
    arch, cus = device_info(device)
    sizes = [int(n) for n in sizes]
    counts = [int(c) for c in counts] if counts is not None else [1] * len(sizes)
    cache = load_cache(cache_path)
    slot = cache.setdefault(arch, {}).setdefault(arith_key(precision, math, limiter), {})
    ckey = classes_key(sizes)
    if not retune and ckey in slot:
        return slot[ckey]

    t0 = time.perf_counter()
    per = probe_mix(sizes, counts)
    dtype = np.float64 if precision == "fp64" else np.float32
    rng = np.random.default_rng(12345)
    sims, states, member_n = [], [], []
    for n, k in zip(sizes, per):
        for _ in range(k):
            sims.append(Simulation(n, n, n, 1.0 / n, [1] * 6, gamma=1.4, cfl=0.4, g=0.0, precision=precision,
                                   math=math, limiter=limiter, device=device, tile_y=8))
            states.append(planar_riemann(n, rng, dtype)[0])
            member_n.append(n)
    trials = []
    with Ensemble(sims) as ens:
        for i, w in enumerate(states):
            ens.set_state(i, w)

        def measure(ty, wv, chunk, bpc):
            ens.retile(ty, [chunk[n] for n in member_n], grid_for(bpc, device), wv)
            ens.time_steps(2)
            ms, _ = ens.time_steps(steps)
            trials.append({"tile_y": ty, "waves": wv, "chunk_x": {str(n): chunk[n] for n in sizes},
                           "blocks_per_cu": bpc, "ms": round(ms, 5)})
            if log:
                log(f"[autotune] tile_y={ty} waves={wv} chunk_x={chunk} blocks/CU={bpc or 'max'}: {ms:.4f} ms")
            return ms

        waves = WAVES if (precision != "fp64" and math == "fast") else (4,)
        best = None
        for ty in TILE_Y:
            if precision == "fp64" and ty > 8:
                continue
            for wv in waves:
                chunk = {n: min(32, n) for n in sizes}
                ms_cur = measure(ty, wv, chunk, 0)
                for n in sorted(set(sizes), reverse=True):
                    for lx in chunk_candidates(n):
                        if lx == chunk[n]:
                            continue
                        trial = dict(chunk)
                        trial[n] = lx
                        ms = measure(ty, wv, trial, 0)
                        if ms < ms_cur:
                            chunk, ms_cur = trial, ms
                if best is None or ms_cur < best[2]:
                    best = (ty, chunk, ms_cur, 0, wv)
        ty, chunk, _, _, wv = best
        for bpc in BLOCKS_PER_CU:
            ms = measure(ty, wv, chunk, bpc)
            if ms < best[2]:
                best = (ty, chunk, ms, bpc, wv)
    entry = {"tile_y": best[0], "waves": best[4], "chunk_x": {str(n): best[1][n] for n in sizes},
             "blocks_per_cu": best[3],
             "ms_probe": round(best[2], 5), "compute_units": cus,
             "probe": {"members": {str(n): k for n, k in zip(sizes, per)},
                       "cells": int(sum(k * n ** 3 for n, k in zip(sizes, per))), "steps_timed": steps},
             "tuned_s": round(time.perf_counter() - t0, 2), "trials": trials}
    slot[ckey] = entry
    save_cache(cache, cache_path)
    return entry


def apply(ens, choice, device=0):
    """Re-tile an Ensemble (its members cubic, keyed by edge length) to an
    autotuner choice."""
    chunks = [int(choice["chunk_x"].get(str(s.interior[0]), 0)) for s in ens.sims]
    ens.retile(int(choice["tile_y"]), chunks, grid_for(choice["blocks_per_cu"], device), int(choice.get("waves", 4)))
