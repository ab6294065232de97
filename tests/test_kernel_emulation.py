"""The HIP step kernels compiled for the CPU (tests/cpu_emu: each workgroup
runs as host threads with real barriers; LDS is poisoned with NaN at kernel
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


def test_emulated_periodic_gravity(emu_lib):
    """Periodic z with gravity: the lower boundary z-flux is the copy of the
    face (n-1, n), whose gravity height belongs to z = n-1."""
    interior = (5, 6, 8)
    w, ds = cases.smooth_random(interior, seed=6, amp_v=0.1, p_base=100.0)
    U, dts, Uo, dto = run_both(emu_lib, interior, w, ds, [cases.PERIODIC] * 6, 2, "fp64", params="native")
    np.testing.assert_array_equal(dts, dto.astype(np.float64))
    np.testing.assert_array_equal(U, Uo)
