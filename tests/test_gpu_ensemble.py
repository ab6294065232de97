"""GPU: the batched ensemble (dg_ensemble_*) against standalone contexts and
the oracle.

Each instance of an ensemble must follow bit for bit the trajectory that a
standalone context (dg_step / dg_run) of the same configuration and state
follows -- whatever the tiling (tile_y, per-instance chunk_x, persistent
grid) -- and, with exact math, the oracle's.
"""
import sys

import numpy as np
import pytest

import cases
from dg_oracle import Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def interior(U):
    return np.stack([np.array(a[2:-2, 2:-2, 2:-2]) for a in U])


def riemann_members(shapes, seed, precision="fp32", math="exact", limiter="minmod", tile_y=0, chunks=None,
                    bcs=None):
    from deepgreen import OUTFLOW, Simulation
    from deepgreen.setups import planar_riemann
    sims, states = [], []
    for k, shape in enumerate(shapes):
        n = max(shape)
        rng = np.random.default_rng([seed, k])
        w, _, _ = planar_riemann(n, rng, np.float64 if precision == "fp64" else np.float32)
        w = [np.ascontiguousarray(a[: shape[0] + 4, : shape[1] + 4, : shape[2] + 4]) for a in w]
        sims.append(Simulation(*shape, 1.0 / n, bcs[k] if bcs else [OUTFLOW] * 6, gamma=1.4, cfl=0.4, g=0.0,
                               precision=precision, math=math, limiter=limiter, tile_y=tile_y,
                               chunk_x=chunks[k] if chunks else 0))
        states.append(w)
    return sims, states


def standalone(sim, w, runs):
    from deepgreen import Solver
    with Solver(sim) as s:
        s.set_primitives(w)
        for r in runs:
            t, n = s.run(r)
        return interior(s.get_state()), (t, n, s.next_dt()), s.conservation_sums()


def test_ensemble_exact_matches_standalone_and_oracle():
    from deepgreen import Ensemble, OUTFLOW, PERIODIC, WALL
    shapes = [(16, 16, 16), (24, 20, 70), (9, 33, 12), (40, 8, 64), (12, 12, 130)]
    bcs = [[OUTFLOW] * 6, [PERIODIC, PERIODIC, OUTFLOW, OUTFLOW, WALL, WALL], [WALL] * 6,
           [OUTFLOW, OUTFLOW, PERIODIC, PERIODIC, OUTFLOW, OUTFLOW], [PERIODIC] * 6]
    sims, states = riemann_members(shapes, seed=3, chunks=[0, 5, 3, 16, 0], bcs=bcs)
    _say(f"  ensemble exact {shapes}")
    with Ensemble(sims) as ens:
        for i, w in enumerate(states):
            ens.set_state(i, w)
        ens.run(17)
        ens.run(16)
        for i, (sim, w) in enumerate(zip(sims, states)):
            got = interior(ens.get_state(i))
            U, clock, sums = standalone(sim, w, [33])
            np.testing.assert_array_equal(got, U, err_msg=f"instance {i} {shapes[i]}")
            assert ens.clock(i) == clock
            np.testing.assert_array_equal(ens.conservation_sums(i), sums)
            if i < 3:
                o = Oracle(shapes[i], sim.ds, sim.bcs, precision="f32", limiter="minmod", nthreads=4,
                           **{**cases.PARAMS["cfg"], "gravity_k_shift": 0})
                o.set_primitives(w)
                o.compute(1e30, 33)
                np.testing.assert_array_equal(got, np.stack([np.array(o.interior_view(a)) for a in o.cons]))


@pytest.mark.parametrize("precision", ["fp32", "fp32_bulk", "fp64"])
def test_ensemble_tiling_invariance(precision):
    """Fast math (FP64: exact): re-tiling in the middle of a run (tile_y,
    chunk_x per instance, grid) changes nothing; every instance equals its
    standalone run with the default tiling."""
    from deepgreen import Ensemble
    shapes = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128), (37, 29, 41)]
    math = "exact" if precision == "fp64" else "fast"
    sims, states = riemann_members(shapes, seed=5, precision=precision, math=math)
    _say(f"  ensemble tiling {precision} {shapes}")
    with Ensemble(sims) as ens:
        for i, w in enumerate(states):
            ens.set_state(i, w)
        ens.run(9)
        ens.retile(tile_y=4, chunk_x=[3, 48, 16, 7, 128, 0], grid=37)
        ens.run(16)
        ens.retile(tile_y=8 if precision == "fp64" else 16, chunk_x=[32, 8, 64, 96, 9, 5], grid=0)
        ens.run(7)
        for i, (sim, w) in enumerate(zip(sims, states)):
            U, clock, _ = standalone(sim, w, [32])
            np.testing.assert_array_equal(interior(ens.get_state(i)), U, err_msg=f"instance {i}")
            assert ens.clock(i) == clock


def test_ensemble_timed_stop():
    """A common t_end: each instance stops on its own clock, exactly where a
    standalone dg_run with the same t_end stops."""
    from deepgreen import Ensemble, Solver
    shapes = [(20, 20, 20), (40, 40, 40), (24, 24, 24)]
    sims, states = riemann_members(shapes, seed=9, math="fast")
    with Ensemble(sims) as ens:
        for i, w in enumerate(states):
            ens.set_state(i, w)
        ens.run(1000, 0.05)
        for i, (sim, w) in enumerate(zip(sims, states)):
            with Solver(sim) as s:
                s.set_primitives(w)
                t, n = s.run(1000, 0.05)
                U = interior(s.get_state())
            assert ens.clock(i)[:2] == (t, n)
            assert t > 0.05 and n < 1000
            np.testing.assert_array_equal(interior(ens.get_state(i)), U)


def test_ensemble_workload_sampled():
    """The 256-member workload (seeded sizes, planar Riemann, outflow,
    minmod, FP32 fast) for 100 steps; one member per size class checked bit
    for bit against its standalone run."""
    from deepgreen import Ensemble
    from deepgreen.setups import ensemble_workload
    sims, states = ensemble_workload(256, seed=0)
    pick, seen = {}, set()
    for i, s in enumerate(sims):
        if s.interior[0] not in seen:
            seen.add(s.interior[0])
            pick[i] = None
    _say(f"  ensemble workload: {len(sims)} members, checking {sorted(pick)}")
    with Ensemble(sims) as ens:
        for i, w in states():
            ens.set_state(i, w)
            if i in pick:
                pick[i] = w
        ens.run(100)
        for i, w in pick.items():
            U, clock, _ = standalone(sims[i], w, [100])
            np.testing.assert_array_equal(interior(ens.get_state(i)), U, err_msg=f"member {i}")
            assert ens.clock(i) == clock
            assert np.isfinite(U).all()


def test_autotune_small(tmp_path):
    from deepgreen import autotune
    path = str(tmp_path / "tune.json")
    old = autotune.PROBE_CELLS
    autotune.PROBE_CELLS = 64 ** 3
    try:
        e = autotune.tune([16, 24], counts=[3, 1], cache_path=path, steps=2, log=_say)
    finally:
        autotune.PROBE_CELLS = old
    assert e["tile_y"] in (4, 8, 16) and e["blocks_per_cu"] in (0, 1, 2)
    assert set(e["chunk_x"]) == {"16", "24"} and len(e["trials"]) >= 10
    assert autotune.lookup([24, 16], cache_path=path) == e
