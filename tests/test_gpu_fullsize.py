"""The benchmark workload at its full size (KHI 512^3, FP32 fast math, the
configuration `bench.py` times): properties that do not depend on the size.

* two steps against the oracle's FP32 arithmetic at 512^3 (fast math is the
  same scheme reassociated, so the difference is rounding);
* periodic boundaries conserve mass, momentum and energy to rounding over 20
  steps (FP64 sums of the FP32 state);
* the result is deterministic and independent of the tiling (tile height and
  x-chunk) bit for bit.
"""
import os
import sys

import numpy as np
import pytest

from dg_oracle import Oracle

pytestmark = pytest.mark.gpu

N = 512


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def khi512():
    from deepgreen.setups import khi
    _say(f"  building the KHI {N}^3 initial condition")
    return khi(N, np.float32)


def _solver(w, ds, tile_y=0, chunk_x=0):
    from deepgreen import PERIODIC, Simulation, Solver
    sim = Simulation(N, N, N, ds, [PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0, precision="fp32", math="fast",
                     tile_y=tile_y, chunk_x=chunk_x)
    s = Solver(sim)
    s.set_primitives(w)
    return s


def _interior_state(s):
    return np.stack([np.array(s.interior(a)) for a in s.get_state()])


def test_khi512_fast_two_steps_match_oracle(khi512):
    w, ds = khi512
    _say("  GPU: 2 steps")
    with _solver(w, ds) as s:
        dts = [s.compute_step() for _ in range(2)]
        U = _interior_state(s)
    _say(f"  oracle (FP32, {os.cpu_count()} threads): 2 steps")
    o = Oracle((N,) * 3, ds, [0] * 6, precision="f32", nthreads=min(16, os.cpu_count() or 1))
    o.set_primitives(w)
    n, _, dto = o.compute(1e30, 2)
    assert n == 2
    Uo = np.stack([np.array(o.interior_view(a)) for a in o.cons])
    del o
    np.testing.assert_allclose(dts, dto, rtol=1e-6)
    d = np.abs(U.astype(np.float64) - Uo)
    # scales: |rho|, |E| and, for each momentum component, the momentum
    # magnitude (z momentum is ~0 in this flow, so its own sum is no scale)
    mag = np.sqrt((Uo[1:4].astype(np.float64) ** 2).sum(0))
    l1_scale = np.array([np.abs(Uo[0]).sum()] + [mag.sum()] * 3 + [np.abs(Uo[4]).sum()])
    linf_scale = np.array([np.abs(Uo[0]).max()] + [mag.max()] * 3 + [np.abs(Uo[4]).max()])
    l1 = d.reshape(5, -1).sum(1) / l1_scale
    linf = d.reshape(5, -1).max(1) / linf_scale
    _say(f"  fast vs oracle FP32: rel L1 {l1}, rel Linf {linf}")
    # rounding of the reassociated arithmetic only (FP32 epsilon 6e-8)
    assert np.all(l1 < 1e-6), l1
    assert np.all(linf < 1e-5), linf


def test_khi512_fast_conserves_and_is_tiling_invariant(khi512):
    w, ds = khi512
    steps = 20
    with _solver(w, ds) as s:
        s0 = np.array(s.conservation_sums())
        s.run(steps)
        s1 = np.array(s.conservation_sums())
        Ua = _interior_state(s)
    _say(f"  sums before {s0}, after {steps} steps {s1}")
    scale = np.abs(Ua.astype(np.float64)).reshape(5, -1).sum(1)   # sum of |U|: the rounding scale
    drift = np.abs(s1 - s0) / scale
    _say(f"  conservation drift / sum|U|: {drift}")
    assert np.all(drift < 1e-6), drift
    # same run again (determinism) and with other tilings
    for ty, lx in ((0, 0), (4, 64), (16, 100)):
        with _solver(w, ds, tile_y=ty, chunk_x=lx) as s:
            s.run(steps)
            Ub = _interior_state(s)
        _say(f"  tile_y={ty} chunk_x={lx}: equal {np.array_equal(Ua, Ub)}")
        np.testing.assert_array_equal(Ua, Ub)
