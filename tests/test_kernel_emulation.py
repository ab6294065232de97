"""The HIP step kernels compiled for the CPU (tests/cpu_emu: each workgroup
runs as cooperative host fibers with real barriers; LDS is poisoned with NaN at kernel
start so a read of an unwritten slot shows) and checked bit for bit against
the oracle on small grids.  Covers the kernel logic -- tiling, halos,
boundary maps, slot rotation, chunk seams -- without a GPU; the same source
then runs on the MI355X in test_gpu_parity.py.
"""
import os
import subprocess

import numpy as np
import pytest

import cases
from dg_oracle import Oracle

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
EMU = os.path.join(ROOT, "tests", "cpu_emu")
EMU_LIB = os.path.join(EMU, "libdeepgreen_emu.so")


@pytest.fixture(scope="module")
def emu_lib():
    subprocess.run(["make", "-s", "-C", EMU, "EXTRA=-DDG_EMU_POISON"], check=True)
    import deepgreen
    return deepgreen.load_library(EMU_LIB)


def run_both(lib, interior, w, ds, bcs, nsteps, precision, params="cfg", limiter="none", math="exact",
             tile_y=4, chunk_x=0):
    import deepgreen
    prm = dict(cases.PARAMS[params])
    prm["gravity_k_shift"] = 0
    sim = deepgreen.Simulation(*interior, ds, bcs, precision=precision, math=math, limiter=limiter,
                               tile_y=tile_y, chunk_x=chunk_x,
                               **{k: v for k, v in prm.items() if k != "gravity_k_shift"})
    with deepgreen.Solver(sim, lib=lib) as s:
        s.set_primitives(w)
        dts = np.array([s.compute_step() for _ in range(nsteps)])
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    o = Oracle(interior, ds, bcs, precision="f64" if precision == "fp64" else "f32", limiter=limiter, **prm)
    o.set_primitives(w)
    _, _, dto = o.compute(1e30, nsteps)
    Uo = np.stack([np.array(o.interior_view(a)) for a in o.cons])
    return U, dts, Uo, dto


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_emulated_kernel_smooth_mixed(emu_lib, precision):
    interior = (7, 6, 9)
    w, ds = cases.smooth_random(interior, seed=5)
    bcs = [cases.PERIODIC, cases.PERIODIC, cases.OUTFLOW, cases.OUTFLOW, cases.PERIODIC, cases.PERIODIC]
    U, dts, Uo, dto = run_both(emu_lib, interior, w, ds, bcs, 2, precision, chunk_x=3)
    np.testing.assert_array_equal(dts, dto.astype(np.float64))
    np.testing.assert_array_equal(U, Uo)


def test_emulated_kernel_walls_gravity(emu_lib):
    interior = (6, 5, 7)
    w, ds = cases.smooth_random(interior, seed=2, amp_v=0.1, p_base=100.0)
    U, dts, Uo, dto = run_both(emu_lib, interior, w, ds, [cases.WALL] * 6, 2, "fp32", params="native")
    np.testing.assert_array_equal(dts, dto.astype(np.float64))
    np.testing.assert_array_equal(U, Uo)


def test_emulated_kernel_outflow_limited_partial_tiles(emu_lib):
    interior = (5, 7, 66)   # partial y tile (7 of 8 rows) and a second z tile
    w, ds = cases.smooth_random(interior, seed=3)
    U, dts, Uo, dto = run_both(emu_lib, interior, w, ds, [cases.OUTFLOW] * 6, 1, "fp32", limiter="minmod",
                               tile_y=8)
    np.testing.assert_array_equal(dts, dto.astype(np.float64))
    np.testing.assert_array_equal(U, Uo)


def test_emulated_fast_math_close_to_exact(emu_lib):
    interior = (6, 5, 6)
    w, ds = cases.smooth_random(interior, seed=4)
    U, _, Uo, _ = run_both(emu_lib, interior, w, ds, [cases.PERIODIC] * 6, 2, "fp32", math="fast")
    np.testing.assert_allclose(U, Uo, rtol=2e-5, atol=1e-6 * np.abs(Uo).max())


def test_emulated_fast_minmod_close_to_exact(emu_lib):
    """Fast math with minmod slopes (the median-of-three form) on a state
    with jumps, where the limiter clips: within rounding of the oracle's
    minmod step."""
    interior = (7, 6, 8)
    w, ds = cases.sod(interior, axis=1)
    w = [np.asarray(a, np.float32) + np.float32(0.05) * np.asarray(b, np.float32)
         for a, b in zip(w, cases.smooth_random(interior, seed=9)[0])]
    U, dts, Uo, dto = run_both(emu_lib, interior, w, ds, [cases.OUTFLOW] * 6, 3, "fp32", limiter="minmod",
                               math="fast")
    np.testing.assert_allclose(dts, dto.astype(np.float64), rtol=2e-5)
    np.testing.assert_allclose(U, Uo, rtol=2e-5, atol=1e-6 * np.abs(Uo).max())


def test_emulated_periodic_gravity(emu_lib):
    """Periodic z with gravity: the lower boundary z-flux is the copy of the
    face (n-1, n), whose gravity height belongs to z = n-1."""
    interior = (5, 6, 8)
    w, ds = cases.smooth_random(interior, seed=6, amp_v=0.1, p_base=100.0)
    U, dts, Uo, dto = run_both(emu_lib, interior, w, ds, [cases.PERIODIC] * 6, 2, "fp64", params="native")
    np.testing.assert_array_equal(dts, dto.astype(np.float64))
    np.testing.assert_array_equal(U, Uo)


def oracle_run(interior, w, ds, bcs, nsteps, limiter="none"):
    o = Oracle(interior, ds, bcs, precision="f32", limiter=limiter, **cases.PARAMS["cfg"])
    o.set_primitives(w)
    n, t, _ = o.compute(1e30, nsteps)
    U = np.stack([np.array(o.interior_view(a)) for a in o.cons])
    o.calc_primitives()
    o.update_primitive_ghost_cells()
    return U, t, n, o.calc_timestep()


def test_emulated_ensemble_matches_oracle(emu_lib):
    """Three uneven grids in one ensemble (persistent kernel striding over the
    work list of all instances, four workgroups in the emulation), run for
    17 + 16 steps: a captured graph, a single step that flips the buffer
    parity, then the graph of the other parity.  Each instance is bitwise the
    oracle's trajectory (state, clock, next dt), which the standalone
    context also reproduces bit for bit."""
    import deepgreen
    prm = {k: v for k, v in cases.PARAMS["cfg"].items() if k != "gravity_k_shift"}
    members = [((6, 5, 7), [cases.OUTFLOW] * 6, 2),
               ((5, 7, 4), [cases.PERIODIC, cases.PERIODIC, cases.OUTFLOW, cases.OUTFLOW, cases.WALL, cases.WALL], 0),
               ((4, 4, 9), [cases.PERIODIC] * 6, 3)]
    sims, states = [], []
    for k, (interior, bcs, chunk) in enumerate(members):
        w, ds = cases.smooth_random(interior, seed=10 + k)
        sims.append(deepgreen.Simulation(*interior, ds, bcs, precision="fp32", math="exact", limiter="minmod",
                                         tile_y=4, chunk_x=chunk, **prm))
        states.append(w)
    with deepgreen.Ensemble(sims, lib=emu_lib) as ens:
        assert ens.info()["grid"] == 4
        for i, w in enumerate(states):
            ens.set_state(i, w)
        ens.run(17)
        ens.run(16)
        for i, (interior, bcs, _) in enumerate(members):
            got = np.stack([a[2:-2, 2:-2, 2:-2] for a in ens.get_state(i)])
            U, t, n, dt_next = oracle_run(interior, states[i], sims[i].ds, bcs, 33, limiter="minmod")
            np.testing.assert_array_equal(got, U)
            assert ens.clock(i) == (t, 33, dt_next) and n == 33


def test_emulated_run_graph_parity(emu_lib):
    """dg_run: 17 steps (one captured graph and one single step) then 16 (the
    graph captured for the other buffer parity) equal 33 oracle steps."""
    import deepgreen
    prm = {k: v for k, v in cases.PARAMS["cfg"].items() if k != "gravity_k_shift"}
    interior, bcs = (4, 5, 3), [cases.OUTFLOW] * 6
    w, ds = cases.smooth_random(interior, seed=21)
    sim = deepgreen.Simulation(*interior, ds, bcs, precision="fp32", tile_y=4, **prm)
    with deepgreen.Solver(sim, lib=emu_lib) as s:
        s.set_primitives(w)
        s.run(17)
        t, n = s.run(16)
        got = np.stack([a[2:-2, 2:-2, 2:-2] for a in s.get_state()])
    U, to, no, _ = oracle_run(interior, w, ds, bcs, 33)
    np.testing.assert_array_equal(got, U)
    assert (t, n) == (to, no)


def test_emulated_ensemble_retile_invariance(emu_lib):
    """Re-tiling an ensemble between runs (tile_y, per-instance chunk_x, grid)
    leaves every instance bitwise on the trajectory of its standalone run."""
    import deepgreen
    from deepgreen.setups import planar_riemann
    sims, states = [], []
    for k, n in enumerate([6, 9, 5]):
        w, _, _ = planar_riemann(n, np.random.default_rng(k))
        sims.append(deepgreen.Simulation(n, n, n, 1.0 / n, [cases.OUTFLOW] * 6, gamma=1.4, cfl=0.4, g=0.0,
                                         math="fast", limiter="minmod", tile_y=4))
        states.append(w)
    with deepgreen.Ensemble(sims, lib=emu_lib) as ens:
        for i, w in enumerate(states):
            ens.set_state(i, w)
        ens.run(3)
        ens.retile(tile_y=8, chunk_x=[2, 3, 1], grid=3)
        ens.run(2)
        got = [ens.get_state(i) for i in range(3)]
    for i, (sim, w) in enumerate(zip(sims, states)):
        with deepgreen.Solver(sim, lib=emu_lib) as s:
            s.set_primitives(w)
            s.run(5)
            ref = s.get_state()
        for a, b in zip(got[i], ref):
            np.testing.assert_array_equal(a[2:-2, 2:-2, 2:-2], b[2:-2, 2:-2, 2:-2])


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_emulated_conducting_walls_match_oracle(emu_lib, precision):
    """Conducting lower walls along x and z (the heat source of the CUDA
    variant), plain walls elsewhere: bit for bit the oracle, over steps that
    include a captured graph."""
    import deepgreen
    interior = (6, 5, 7)
    w, ds = cases.smooth_random(interior, seed=31, amp_v=0.1, p_base=100.0)
    bcs = [3, 2, 2, 1, 3, 2]
    prm = dict(gamma=5.0 / 3.0, cfl=0.2, g=0.0)
    sim = deepgreen.Simulation(*interior, ds, bcs, precision=precision, tile_y=4, **prm)
    with deepgreen.Solver(sim, lib=emu_lib) as s:
        s.set_heat(300.0, 4.0, 287.0)
        s.set_primitives(w)
        s.run(17)
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    o = Oracle(interior, ds, bcs, precision="f64" if precision == "fp64" else "f32", t_out=300.0, v_out=4.0,
               rs=287.0, **prm)
    o.set_primitives(w)
    o.compute(1e30, 17)
    np.testing.assert_array_equal(U, np.stack([np.array(o.interior_view(a)) for a in o.cons]))


@pytest.mark.parametrize("precision, tol", [("fp32", 2e-5), ("fp64", 1e-12)])
def test_emulated_fast_conducting_walls_close_to_oracle(emu_lib, precision, tol):
    """Fast math with conducting walls: the heat source's wall temperature
    comes from the fast Hancock state, whose pressure word holds
    2 p / (gamma - 1).  The gas is hot enough (p / (R rho) near half of
    T_out) that a wrong scale of that word would flip the sign of the heat
    flux; the energy stays within rounding of the oracle's."""
    import deepgreen
    interior = (6, 5, 7)
    w, ds = cases.smooth_random(interior, seed=33, amp_v=0.1, p_base=4.0e4)
    bcs = [3, 2, 2, 1, 3, 2]
    prm = dict(gamma=5.0 / 3.0, cfl=0.2, g=0.0)
    sim = deepgreen.Simulation(*interior, ds, bcs, precision=precision, math="fast", tile_y=4, **prm)
    with deepgreen.Solver(sim, lib=emu_lib) as s:
        s.set_heat(300.0, 4.0, 287.0)
        s.set_primitives(w)
        s.run(6)
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    o = Oracle(interior, ds, bcs, precision="f64" if precision == "fp64" else "f32", t_out=300.0, v_out=4.0,
               rs=287.0, **prm)
    o.set_primitives(w)
    o.compute(1e30, 6)
    Uo = np.stack([np.array(o.interior_view(a)) for a in o.cons]).astype(np.float64)
    p_over_rho = np.asarray(w[4], np.float64)[2:-2, 2:-2, 2:-2] / np.asarray(w[0], np.float64)[2:-2, 2:-2, 2:-2]
    assert 0.3 < float(np.median(p_over_rho)) / 287.0 / 300.0 < 0.7
    rel = np.abs(U - Uo).reshape(5, -1).sum(1) / np.abs(Uo).reshape(5, -1).sum(1)
    assert np.all(rel <= tol), rel
    # the heat source is visible at this tolerance: the same run with plain
    # walls differs by over 10x tol in energy, and a heat flux of the wrong
    # sign would be off by twice that
    o2 = Oracle(interior, ds, [2 if b == 3 else b for b in bcs], precision="f64" if precision == "fp64" else "f32",
                **prm)
    o2.set_primitives(w)
    o2.compute(1e30, 6)
    U2 = np.stack([np.array(o2.interior_view(a)) for a in o2.cons]).astype(np.float64)
    assert np.abs(U2 - Uo)[4].sum() / np.abs(Uo[4]).sum() > 10 * tol


def test_emulated_noalloc_step_matches_dg_step(emu_lib):
    """dg_step_noalloc (static scratch array, caller-owned buffers) gives the
    state and dt of dg_step bit for bit, on two calls in a row, the second
    on a smaller grid."""
    import ctypes

    import deepgreen
    for interior, bcs, precision, limiter in (((7, 6, 9), [0, 0, 1, 1, 2, 2], "fp32", "none"),
                                              ((5, 5, 6), [1, 1, 0, 0, 2, 2], "fp64", "minmod")):
        w, ds = cases.smooth_random(interior, seed=4)
        sim = deepgreen.Simulation(*interior, ds, bcs, precision=precision, limiter=limiter, gamma=1.4, cfl=0.4,
                                   g=0.0, tile_y=4)
        with deepgreen.Solver(sim, lib=emu_lib) as ref, deepgreen.Solver(sim, lib=emu_lib) as a, \
                deepgreen.Solver(sim, lib=emu_lib) as b:
            for s in (ref, a, b):
                s.set_primitives(w)
            dt_ref = ref.compute_step()
            U_ref = ref.get_state()
            cfg = sim.config(emu_lib)
            ins = (ctypes.c_void_p * 5)(*a.device_pointers())
            outs = (ctypes.c_void_p * 5)(*b.device_pointers())
            dt = ctypes.c_double()
            assert emu_lib.dg_step_noalloc(ctypes.byref(cfg), ins, outs, None, ctypes.byref(dt)) == 0
            assert dt.value == dt_ref
            U = b.get_state()
            for f in range(5):
                np.testing.assert_array_equal(b.interior(U[f]), ref.interior(U_ref[f]))


@pytest.mark.parametrize("precision, limiter, bcs, gravity", [
    ("fp32", "none", [0, 0, 1, 1, 2, 2], True),
    ("fp64", "minmod", [2, 1, 0, 0, 1, 2], False),
])
def test_emulated_step_dump_matches_oracle_arrays(emu_lib, precision, limiter, bcs, gravity):
    """dg_step_dump hands back the oracle's primitives (start of the step),
    gradients and fluxes -- ghost cells included -- after the same steps."""
    import deepgreen
    interior = (6, 5, 7)
    w, ds = cases.smooth_random(interior, seed=8, amp_v=0.1, p_base=10.0)
    prm = dict(gamma=5.0 / 3.0, cfl=0.2, g=-0.5)
    sim = deepgreen.Simulation(*interior, ds, bcs, precision=precision, limiter=limiter, gravity=gravity, tile_y=4,
                               **prm)
    o = Oracle(interior, ds, bcs, precision="f64" if precision == "fp64" else "f32", limiter=limiter,
               gravity=gravity, gravity_k_shift=0, **prm)
    o.set_primitives(w)
    o.calc_conserved()
    with deepgreen.Solver(sim, lib=emu_lib) as s:
        s.set_primitives(w)
        for step in range(3):
            dt, prim, grad, flux = s.compute_step_dump()
            dto = o.steps(1)[0]
            assert dt == float(dto)
            for q in range(5):
                np.testing.assert_array_equal(prim[q], np.asarray(o.prim[q]))
                np.testing.assert_array_equal(grad[q], np.asarray(o.grad[q]))
                np.testing.assert_array_equal(flux[q], np.asarray(o.flux[q]))
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    np.testing.assert_array_equal(U, np.stack([np.array(o.interior_view(a)) for a in o.cons]))


@pytest.mark.parametrize("limiter, bcs", [("none", [1, 1, 0, 0, 2, 2]), ("minmod", [0, 0, 2, 1, 0, 0])])
def test_emulated_fp64_fast_close_to_oracle(emu_lib, limiter, bcs):
    """FP64 fast math (reassociated, refined reciprocals) stays within 1e-13
    relative L1 per field of the FP64 oracle over a few steps."""
    interior = (7, 6, 9)
    w, ds = cases.smooth_random(interior, seed=12)
    U, dts, Uo, dto = run_both(emu_lib, interior, w, ds, bcs, 4, "fp64", limiter=limiter, math="fast")
    rel = np.abs(U - Uo).reshape(5, -1).sum(1) / np.abs(Uo).reshape(5, -1).sum(1)
    assert np.all(rel <= 1e-13), rel
    np.testing.assert_allclose(dts, dto, rtol=1e-14)
    assert not np.array_equal(U, Uo)   # the fast arithmetic is not the reference's order


@pytest.mark.parametrize("precision, tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_emulated_fast_math_on_sod(emu_lib, precision, tol):
    """Fast math on a Sod tube (outflow): the unlimited scheme's face states
    at the jump have negative density and pressure (p / rho > 0), where the
    fast sound speed must still be sqrt(gamma p / rho); dt sequence and state
    stay within rounding of the oracle over 20 steps."""
    interior = (12, 12, 12)
    w, ds = cases.sod(interior, axis=0)
    U, dts, Uo, dto = run_both(emu_lib, interior, w, ds, [cases.OUTFLOW] * 6, 20, precision, math="fast")
    np.testing.assert_allclose(dts, dto.astype(np.float64), rtol=tol)
    ref = np.abs(Uo).reshape(5, -1).sum(1)
    rel = np.abs(U - Uo).reshape(5, -1).sum(1) / np.where(ref > 0, ref, 1.0)
    assert np.all(rel <= tol), rel
