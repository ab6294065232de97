"""dg_step_noalloc (the entry point that allocates nothing) against dg_step.

Inputs and outputs are device buffers owned by the caller: here the state
buffers of two contexts (dg_device_state), so the test needs no other device
allocator.  Each call must give the same state, bit for bit, and the same dt
as dg_step on a context holding the same state; two calls in a row share the
library's static scratch array, the second on a smaller grid.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def smooth_state(interior, seed, dtype):
    rng = np.random.default_rng(seed)
    shape = tuple(n + 4 for n in interior)
    x, y, z = np.meshgrid(*[np.arange(s) / s for s in shape], indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, size=5)
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x + ph[0]) * np.cos(2 * np.pi * y + ph[1])
    vx = 0.3 * np.sin(2 * np.pi * y + ph[2])
    vy = 0.2 * np.cos(2 * np.pi * z + ph[3])
    vz = 0.1 * np.sin(2 * np.pi * x + ph[4])
    p = 1.0 + 0.1 * np.cos(2 * np.pi * (x + y + z))
    return [a.astype(dtype) for a in (rho, vx, vy, vz, p)]


CASES = [
    # (interior, bcs, precision, math, limiter, gravity), second smaller than the first
    ((48, 40, 72), [0] * 6, "fp32", "fast", "none", False),
    ((20, 24, 16), [1, 1, 2, 2, 0, 0], "fp64", "exact", "minmod", True),
    ((12, 10, 14), [1, 2, 0, 0, 1, 1], "fp32_bulk", "exact", "vanleer", False),
]


def _sim(interior, bcs, precision, math, limiter, gravity):
    from deepgreen import Simulation
    return Simulation(*interior, 1.0 / 64, bcs, gravity=gravity, gamma=1.4, cfl=0.4, g=-1.0,
                      precision=precision, math=math, limiter=limiter)


def _noalloc(L, sim, src, dst, want_dt=True):
    cfg = sim.config(L)
    ins = (ctypes.c_void_p * 5)(*src.device_pointers())
    outs = (ctypes.c_void_p * 5)(*dst.device_pointers())
    dt = ctypes.c_double(-1.0)
    rc = L.dg_step_noalloc(ctypes.byref(cfg), ins, outs, None, ctypes.byref(dt) if want_dt else None)
    assert rc == 0, rc
    return dt.value


def test_noalloc_matches_dg_step_on_consecutive_calls():
    import deepgreen
    from deepgreen import Solver
    L = deepgreen.load_library()
    for case in CASES:
        sim = _sim(*case)
        _say(f"  noalloc {case}")
        w = smooth_state(sim.interior, 7, sim.dtype)
        with Solver(sim) as ref, Solver(sim) as a, Solver(sim) as b:
            for s in (ref, a, b):
                s.set_primitives(w)
            dt_ref = ref.compute_step()
            U_ref = ref.get_state()
            dt = _noalloc(L, sim, a, b)
            U = b.get_state()    # b's current buffer now holds the advanced state
            assert dt == dt_ref, (dt, dt_ref)
            for f in range(5):
                np.testing.assert_array_equal(b.interior(U[f]), ref.interior(U_ref[f]))
            # the input is unchanged
            U_in = a.get_state()
            a2 = Solver(sim)
            a2.set_primitives(w)
            U_in_ref = a2.get_state()
            a2.close()
            for f in range(5):
                np.testing.assert_array_equal(a.interior(U_in[f]), a.interior(U_in_ref[f]))


def test_noalloc_enqueue_only_then_synchronized_call():
    """A call without dt_used only enqueues; the next call on the same stream
    (with dt_used) synchronizes it, after which both outputs are complete."""
    import deepgreen
    from deepgreen import Solver
    L = deepgreen.load_library()
    sim = _sim(*CASES[0])
    w = smooth_state(sim.interior, 3, sim.dtype)
    with Solver(sim) as ref, Solver(sim) as a, Solver(sim) as b, Solver(sim) as c:
        for s in (ref, a, b, c):
            s.set_primitives(w)
        ref.compute_step()
        U_ref = ref.get_state()
        _noalloc(L, sim, a, b, want_dt=False)
        _noalloc(L, sim, a, c, want_dt=True)
        for s in (b, c):
            U = s.get_state()
            for f in range(5):
                np.testing.assert_array_equal(s.interior(U[f]), ref.interior(U_ref[f]))


def test_noalloc_rejects_what_it_cannot_do():
    import deepgreen
    from deepgreen import Simulation, Solver
    L = deepgreen.load_library()
    sim = _sim(*CASES[2])
    with Solver(sim) as a:
        a.set_primitives(smooth_state(sim.interior, 1, sim.dtype))
        cfg = sim.config(L)
        p = (ctypes.c_void_p * 5)(*a.device_pointers())
        assert L.dg_step_noalloc(ctypes.byref(cfg), p, p, None, None) == -1          # aliasing
        cw = Simulation(*sim.interior, sim.ds, [3, 1, 1, 1, 1, 1], precision="fp32_bulk").config(L)
        q = (ctypes.c_void_p * 5)(*[v + 64 for v in a.device_pointers()])
        assert L.dg_step_noalloc(ctypes.byref(cw), p, q, None, None) == -1            # conducting wall
        # a grid whose step needs more workgroups than the scratch array holds:
        # rejected from the configuration alone, before any device work
        big = Simulation(2, 2048, 2048, 1.0 / 2048, [0] * 6, tile_y=4, precision="fp32").config(L)
        assert L.dg_step_noalloc(ctypes.byref(big), p, q, None, None) == -1
