"""GPU parity: the HIP step library (through its C ABI) against the golden
vectors written by the reference's Fortran physics and against the oracle.

Exact math is checked bit for bit: FP64 against the real(8) reference build,
FP32 against the real(4) build (whose arithmetic equals the Rust f32
solver's).  Cases the Fortran cannot represent (the Rust gravity height,
upper walls, limiters, partial tiles) are checked bit for bit against the
oracle, itself pinned by tests/test_oracle_*.py.
"""
import os
import sys

import numpy as np
import pytest

import cases
from dg_oracle import Oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lib_run(c, precision, math="exact", limiter="none", tile_y=0, chunk_x=0, per_step=True, params=None):
    from deepgreen import Simulation, Solver
    prm = dict(cases.PARAMS[params or c["params"]])
    prm.pop("gravity_k_shift")
    sim = Simulation(*c["interior"], c["ds"], c["bcs"], precision=precision, math=math, limiter=limiter,
                     tile_y=tile_y, chunk_x=chunk_x, **prm)
    _say(f"  lib {c['interior']} {precision}/{math}/{limiter} ty={tile_y} lx={chunk_x} bcs={c['bcs']}")
    with Solver(sim) as s:
        s.set_primitives(c["w"])
        if per_step:
            dts = np.array([s.compute_step() for _ in range(c["nsteps"])])
        else:
            t, n = s.run(c["nsteps"])
            assert n == c["nsteps"]
            dts = None
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
        sums = s.conservation_sums()
    return U, dts, sums


def oracle_run(c, precision, limiter="none", params=None):
    prm = dict(cases.PARAMS[params or c["params"]])
    prm["gravity_k_shift"] = 0  # the library follows the Rust height (physics.rs:368)
    o = Oracle(c["interior"], c["ds"], c["bcs"], precision="f64" if precision == "fp64" else "f32",
               limiter=limiter, nthreads=4, **prm)
    o.set_primitives(c["w"])
    n, _, dts = o.compute(1e30, c["nsteps"])
    return np.stack([np.array(o.interior_view(a)) for a in o.cons]), dts


CFG_CASES = [n for n in cases.CASES if cases.CASES[n][5] == "cfg"]


@pytest.mark.parametrize("name", CFG_CASES)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_matches_reference_fortran(golden_dir, name, precision):
    _say(f"case {name} {precision}")
    c = cases.build_case(name)
    g = np.load(os.path.join(golden_dir, f"{name}.npz"))
    key = "dp" if precision == "fp64" else "sp"
    U, dts, _ = lib_run(c, precision)
    np.testing.assert_array_equal(dts, g[f"{key}_dt"].astype(np.float64))
    if name == "sod64":
        mid = [s // 2 for s in U.shape[1:]]
        np.testing.assert_array_equal(U[:, :, mid[1], mid[2]], g[f"{key}_line"])
        # the whole field: the tube along x stays exactly uniform in y and z,
        # and the per-field checksums (sum, sum |.|, min, max) of all 64^3
        # cells equal those of the reference's field
        np.testing.assert_array_equal(U, np.broadcast_to(U[:, :, :1, :1], U.shape))
        summary = np.stack([[u.astype(np.float64).sum(), np.abs(u.astype(np.float64)).sum(), u.min(), u.max()]
                            for u in U])
        np.testing.assert_array_equal(summary, g[f"{key}_summary"])
    else:
        np.testing.assert_array_equal(U, g[f"{key}_U"])


@pytest.mark.parametrize("name", ["smooth_lowerwalls_native", "smooth_periodic_native"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gravity_and_walls_match_oracle(name, precision):
    _say(f"case {name} {precision}")
    c = cases.build_case(name)
    U, dts, _ = lib_run(c, precision)
    Uo, dto = oracle_run(c, precision)
    np.testing.assert_array_equal(dts, dto.astype(np.float64))
    np.testing.assert_array_equal(U, Uo)


def _walls_case():
    w, ds = cases.smooth_random((10, 9, 11), seed=21, amp_v=0.1, p_base=100.0)
    return dict(w=w, ds=ds, interior=(10, 9, 11), bcs=[cases.WALL] * 6, nsteps=12, params="native")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_all_walls_match_oracle(precision):
    c = _walls_case()
    U, dts, _ = lib_run(c, precision)
    Uo, dto = oracle_run(c, precision)
    np.testing.assert_array_equal(dts, dto.astype(np.float64))
    np.testing.assert_array_equal(U, Uo)


@pytest.mark.parametrize("limiter", ["minmod", "vanleer"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_limited_blast_matches_oracle(limiter, precision):
    w, ds = cases.blast((20, 18, 22), radius=0.2, p_in=10.0, p_out=0.1)
    c = dict(w=w, ds=ds, interior=(20, 18, 22), bcs=[cases.PERIODIC] * 2 + [cases.OUTFLOW] * 2 + [cases.WALL] * 2,
             nsteps=15, params="cfg")
    U, dts, _ = lib_run(c, precision, limiter=limiter)
    Uo, dto = oracle_run(c, precision, limiter=limiter)
    np.testing.assert_array_equal(dts, dto.astype(np.float64))
    np.testing.assert_array_equal(U, Uo)


@pytest.mark.parametrize("tile_y,chunk_x", [(4, 3), (8, 5), (16, 64), (8, 1)])
def test_tiling_does_not_change_bits(tile_y, chunk_x):
    """Partial tiles in y and z, x chunks of every length: same bits."""
    w, ds = cases.smooth_random((21, 19, 70), seed=8)
    c = dict(w=w, ds=ds, interior=(21, 19, 70), bcs=[cases.PERIODIC] * 2 + [cases.OUTFLOW] * 4, nsteps=6,
             params="cfg")
    U, dts, _ = lib_run(c, "fp32", tile_y=tile_y, chunk_x=chunk_x)
    Uo, dto = oracle_run(c, "fp32")
    np.testing.assert_array_equal(U, Uo)
    np.testing.assert_array_equal(dts, dto.astype(np.float64))


def test_graph_run_equals_single_steps():
    c = cases.build_case("smooth_periodic")
    c["nsteps"] = 37  # two graph replays of 16 plus 5 single launches
    U1, _, s1 = lib_run(c, "fp32", per_step=True)
    U2, _, s2 = lib_run(c, "fp32", per_step=False)
    np.testing.assert_array_equal(U1, U2)
    np.testing.assert_array_equal(s1, s2)


def test_graph_runs_across_buffer_parity():
    """run(17) leaves the state in the other buffer (one single launch after a
    graph replay); the next run(16) must replay the graph captured for that
    parity.  Equals 33 single steps."""
    from deepgreen import Simulation, Solver
    c = cases.build_case("smooth_mixed")
    c["nsteps"] = 33
    U1, _, _ = lib_run(c, "fp32", per_step=True)
    prm = {k: v for k, v in cases.PARAMS[c["params"]].items() if k != "gravity_k_shift"}
    with Solver(Simulation(*c["interior"], c["ds"], c["bcs"], precision="fp32", **prm)) as s:
        s.set_primitives(c["w"])
        s.run(17)
        t, n = s.run(16)
        assert n == 33
        U2 = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    np.testing.assert_array_equal(U1, U2)


def test_conservation_sums_and_determinism():
    c = cases.build_case("blast16")
    U1, _, s1 = lib_run(c, "fp64", per_step=False)
    U2, _, s2 = lib_run(c, "fp64", per_step=False)
    np.testing.assert_array_equal(U1, U2)
    np.testing.assert_array_equal(s1, s2)
    np.testing.assert_allclose(s1, U1.reshape(5, -1).sum(1), rtol=1e-13, atol=1e-13)


def test_compute_with_total_time_matches_reference_loop():
    """main.rs:51-67: stop when tnow > total_time (f32 clock for FP32)."""
    from deepgreen import Simulation, Solver
    c = cases.build_case("smooth_mixed")
    prm = dict(cases.PARAMS["cfg"])
    prm.pop("gravity_k_shift")
    o = Oracle(c["interior"], c["ds"], c["bcs"], precision="f32", **cases.PARAMS["cfg"])
    o.set_primitives(c["w"])
    n_o, t_o, _ = o.compute(0.1, 1000)
    sim = Simulation(*c["interior"], c["ds"], c["bcs"], precision="fp32", **prm)
    with Solver(sim) as s:
        s.set_primitives(c["w"])
        t, n = s.compute(0.1, 1000)
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    assert n == n_o
    assert np.float32(t) == np.float32(t_o)
    np.testing.assert_array_equal(U, np.stack([np.array(o.interior_view(a)) for a in o.cons]))


def test_nan_wave_speed_falls_back_to_cfl():
    from deepgreen import Simulation, Solver
    c = cases.build_case("smooth_periodic")
    c["w"][4][5, 5, 5] = -5.0  # negative pressure in one interior cell
    prm = dict(cases.PARAMS["cfg"])
    prm.pop("gravity_k_shift")
    sim = Simulation(*c["interior"], c["ds"], c["bcs"], precision="fp32", **prm)
    with Solver(sim) as s:
        s.set_primitives(c["w"])
        assert s.compute_step() == float(np.float32(0.4) * np.float32(1.0))


def test_get_state_ghosts_follow_boundary_maps():
    from deepgreen import Simulation, Solver, PRIMITIVE
    c = cases.build_case("smooth_lowerwalls_native")
    prm = dict(cases.PARAMS["native"])
    prm.pop("gravity_k_shift")
    sim = Simulation(*c["interior"], c["ds"], c["bcs"], precision="fp64", **prm)
    o = Oracle(c["interior"], c["ds"], c["bcs"], precision="f64", **cases.PARAMS["native"])
    o.set_primitives(c["w"])
    o.calc_conserved()
    o.calc_primitives()
    o.update_primitive_ghost_cells()
    with Solver(sim) as s:
        s.set_primitives(c["w"])
        W = s.get_state(PRIMITIVE)
    # faces and edges: every cell whose ghost status is along at most two axes
    for a, b in zip(W, o.prim):
        np.testing.assert_array_equal(a, b)


def test_fp32_bulk_and_fast_errors_against_fp64():
    """FP32-bulk and FP32-fast against the FP64 oracle (limited blast): the
    errors are reported and bounded."""
    w, ds = cases.blast((32, 32, 32), radius=0.2, p_in=10.0, p_out=0.1)
    c = dict(w=w, ds=ds, interior=(32, 32, 32), bcs=[cases.PERIODIC] * 6, nsteps=30, params="cfg")
    Uo, _ = oracle_run(c, "fp64", limiter="minmod")
    for prec, math in (("fp32_bulk", "exact"), ("fp32", "exact"), ("fp32", "fast"), ("fp32_bulk", "fast")):
        U, _, _ = lib_run(c, prec, math=math, limiter="minmod", per_step=False)
        l1 = np.abs(U - Uo).reshape(5, -1).sum(1) / np.abs(Uo).reshape(5, -1).sum(1)
        linf = np.abs(U - Uo).reshape(5, -1).max(1) / np.abs(Uo).reshape(5, -1).max(1)
        _say(f"  {prec}/{math}: rel L1 {l1}, rel Linf {linf}")
        assert np.all(l1 < 1e-4) and np.all(linf < 1e-2)


def test_blast_fp64_minmod_64cubed_matches_oracle():
    """The blast configuration (rho 1, p 0.1, p 10 inside r < 0.1, periodic,
    gamma 1.4, CFL 0.4, minmod) at 64^3 for 100 steps: FP64 on the GPU equals
    the FP64 oracle bit for bit (relative L1 0 <= 1e-12)."""
    from deepgreen import PERIODIC, Simulation, Solver
    from deepgreen.setups import blast
    n = 64
    w, ds = blast(n)
    _say("  blast 64^3 fp64 minmod, 100 steps")
    sim = Simulation(n, n, n, ds, [PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0, precision="fp64", limiter="minmod")
    with Solver(sim) as s:
        s.set_primitives(w)
        t, steps = s.run(100)
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    o = Oracle((n, n, n), ds, [PERIODIC] * 6, precision="f64", limiter="minmod", nthreads=16, gamma=1.4, cfl=0.4,
               g=0.0)
    o.set_primitives(w)
    o.compute(1e30, 100)
    Uo = np.stack([np.array(o.interior_view(a)) for a in o.cons])
    rel_l1 = np.abs(U - Uo).reshape(5, -1).sum(1) / np.abs(Uo).reshape(5, -1).sum(1)
    assert np.all(rel_l1 <= 1e-12), rel_l1
    np.testing.assert_array_equal(U, Uo)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_conducting_walls_match_oracle(precision):
    """Conducting lower walls along x and y (heat source), a wall, an outflow
    face: bit for bit the oracle over 40 steps (graph replays included)."""
    from deepgreen import Simulation, Solver
    interior = (20, 18, 70)
    w, ds = cases.smooth_random(interior, seed=32, amp_v=0.1, p_base=100.0)
    bcs = [3, 2, 3, 1, 2, 2]
    prm = dict(gamma=5.0 / 3.0, cfl=0.2, g=0.0)
    _say(f"  conducting walls {precision}")
    with Solver(Simulation(*interior, ds, bcs, precision=precision, **prm)) as s:
        s.set_heat(300.0, 4.0, 287.0)
        s.set_primitives(w)
        s.run(40)
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    o = Oracle(interior, ds, bcs, precision="f64" if precision == "fp64" else "f32", nthreads=8, t_out=300.0,
               v_out=4.0, rs=287.0, **prm)
    o.set_primitives(w)
    o.compute(1e30, 40)
    np.testing.assert_array_equal(U, np.stack([np.array(o.interior_view(a)) for a in o.cons]))


@pytest.mark.parametrize("precision, tol", [("fp32", 2e-5), ("fp64", 1e-12)])
def test_conducting_walls_fast_math_close_to_oracle(precision, tol):
    """Fast math with conducting walls (the wall temperature from the fast
    Hancock state, whose pressure word holds 2 p / (gamma - 1)), in a gas hot
    enough (p / (R rho) near half of T_out) that a wrong scale would flip the
    heat flux: relative L1 per field within rounding of the oracle after 20
    steps."""
    from deepgreen import Simulation, Solver
    interior = (20, 18, 70)
    w, ds = cases.smooth_random(interior, seed=34, amp_v=0.1, p_base=4.0e4)
    bcs = [3, 2, 3, 1, 2, 2]
    prm = dict(gamma=5.0 / 3.0, cfl=0.2, g=0.0)
    _say(f"  conducting walls, fast math {precision}")
    with Solver(Simulation(*interior, ds, bcs, precision=precision, math="fast", **prm)) as s:
        s.set_heat(300.0, 4.0, 287.0)
        s.set_primitives(w)
        s.run(20)
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    o = Oracle(interior, ds, bcs, precision="f64" if precision == "fp64" else "f32", nthreads=8, t_out=300.0,
               v_out=4.0, rs=287.0, **prm)
    o.set_primitives(w)
    o.compute(1e30, 20)
    Uo = np.stack([np.array(o.interior_view(a)) for a in o.cons])
    rel = _rel_l1(U, Uo)
    _say(f"  rel L1 {rel}")
    assert np.all(rel <= tol), rel


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_conducting_wall_heat_from_extrapolated_state(axis):
    """A gas whose extrapolated state differs from its start-of-step state at
    the wall (uniform rho, p; tangential velocity growing linearly along
    another axis): one FP64 step with a conducting lower wall is bit for bit
    the oracle's, whose heat source reads the extrapolated ghost state
    (pinned by a closed form in test_oracle_invariants.py)."""
    from deepgreen import Simulation, Solver
    n, ds = (8, 8, 8), 0.125
    along = (axis + 1) % 3
    shape = tuple(k + 4 for k in n)
    vel = 1.0 * ((np.arange(shape[along]) - 2 + 0.5) * ds - 0.5)
    sl = [None, None, None]
    sl[along] = slice(None)
    w = [np.ones(shape), np.zeros(shape), np.zeros(shape), np.zeros(shape), np.ones(shape)]
    w[1 + along] = np.broadcast_to(vel[tuple(sl)], shape).copy()
    bcs = [1] * 6
    bcs[2 * axis], bcs[2 * axis + 1] = 3, 2
    prm = dict(gamma=1.4, cfl=0.2, g=0.0)
    _say(f"  conducting wall along {axis}, extrapolated state")
    with Solver(Simulation(*n, ds, bcs, precision="fp64", **prm)) as s:
        s.set_heat(300.0, 4.0, 0.004)
        s.set_primitives(w)
        dt = s.compute_step()
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    o = Oracle(n, ds, bcs, precision="f64", t_out=300.0, v_out=4.0, rs=0.004, **prm)
    o.set_primitives(w)
    o.calc_conserved()
    dto = o.steps(1)[0]
    assert dt == dto
    np.testing.assert_array_equal(U, np.stack([np.array(o.interior_view(a)) for a in o.cons]))


@pytest.mark.parametrize("precision, math, limiter, bcs, gravity", [
    ("fp32", "exact", "none", [0, 0, 1, 1, 2, 2], True),
    ("fp64", "exact", "minmod", [2, 1, 0, 0, 1, 2], False),
    ("fp32", "fast", "vanleer", [1, 2, 2, 1, 0, 0], True),
])
def test_step_dump_matches_oracle_arrays(precision, math, limiter, bcs, gravity):
    """dg_step_dump: the primitives of the step's start, the gradients and the
    fluxes the reference grid holds after compute_step (ghost cells
    included) equal the oracle's bit for bit, over several steps and two
    tiles per axis; the dump does not change the step (fast math: the dumped
    arrays are the exact arithmetic's on the fast run's state)."""
    from deepgreen import Simulation, Solver
    interior = (20, 18, 70)
    w, ds = cases.smooth_random(interior, seed=9, amp_v=0.1, p_base=10.0)
    prm = dict(gamma=5.0 / 3.0, cfl=0.2, g=-0.5)
    sim = Simulation(*interior, ds, bcs, precision=precision, math=math, limiter=limiter, gravity=gravity, **prm)
    _say(f"  step dump {precision}/{math}/{limiter} bcs={bcs}")
    oprec = "f64" if precision == "fp64" else "f32"
    with Solver(sim) as s, Solver(sim) as plain:
        s.set_primitives(w)
        plain.set_primitives(w)
        for step in range(3):
            U_start = s.get_state()
            o = Oracle(interior, ds, bcs, precision=oprec, limiter=limiter, gravity=gravity, gravity_k_shift=0,
                       **prm)
            o.set_conserved(U_start)
            dt, prim, grad, flux = s.compute_step_dump()
            assert dt == plain.compute_step()
            o.steps(1)
            for q in range(5):
                np.testing.assert_array_equal(prim[q], np.asarray(o.prim[q]))
                np.testing.assert_array_equal(grad[q], np.asarray(o.grad[q]))
                if math == "exact":
                    np.testing.assert_array_equal(flux[q], np.asarray(o.flux[q]))
                else:   # the fast run's dt (fast wave speeds) enters the extrapolated state
                    fo = np.asarray(o.flux[q])
                    np.testing.assert_allclose(flux[q], fo, rtol=1e-5, atol=1e-5 * np.abs(fo).max())
            for a, b in zip(s.get_state(), plain.get_state()):
                np.testing.assert_array_equal(a, b)


def _rel_l1(U, Uo):
    d = np.abs(U - Uo).reshape(5, -1).sum(1)
    ref = np.abs(Uo).reshape(5, -1).sum(1)
    return np.where(ref > 0, d / np.where(ref > 0, ref, 1.0), d)


@pytest.mark.parametrize("case", ["sod64", "blast64", "walls48"])
def test_fp64_fast_within_1e12_of_oracle(case):
    """FP64 fast math (the scheme reassociated around FMAs, refined hardware
    reciprocals and reciprocal square roots) against the FP64 oracle: relative
    L1 per conserved field <= 1e-12 after 100 steps of the Sod tube (64^3,
    outflow, the benchmark list's CPU case), of the blast (64^3, periodic,
    minmod) and of a smooth random state in a box with walls, outflow and
    periodic faces, gravity and van Leer slopes (48 x 40 x 70)."""
    from deepgreen import Simulation, Solver
    prm = dict(gamma=1.4, cfl=0.4, g=0.0)
    gravity = False
    if case == "sod64":
        c = cases.build_case("sod64")
        w, ds, n, bcs, limiter = c["w"], c["ds"], c["interior"], c["bcs"], "none"
    elif case == "blast64":
        n = (64, 64, 64)
        w, ds = cases.blast(n, radius=0.1 * 64 / 64, p_in=10.0, p_out=0.1)
        bcs, limiter = [cases.PERIODIC] * 6, "minmod"
    else:
        n = (48, 40, 70)
        w, ds = cases.smooth_random(n, seed=11, amp_v=0.2, p_base=5.0)
        w = [np.asarray(a, np.float64) for a in w]
        bcs, limiter, gravity = [2, 1, 0, 0, 1, 2], "vanleer", True
        prm = dict(gamma=5.0 / 3.0, cfl=0.3, g=-0.5)
    _say(f"  fp64 fast {case}")
    with Solver(Simulation(*n, ds, bcs, precision="fp64", math="fast", limiter=limiter, gravity=gravity,
                           **prm)) as s:
        s.set_primitives(w)
        t, steps = s.run(100)
        assert steps == 100
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    o = Oracle(n, ds, bcs, precision="f64", limiter=limiter, nthreads=8, gravity=gravity, gravity_k_shift=0, **prm)
    o.set_primitives(w)
    _, to, _ = o.compute(1e30, 100)
    Uo = np.stack([np.array(o.interior_view(a)) for a in o.cons])
    rel = _rel_l1(U, Uo)
    _say(f"  rel L1 {rel}  t {t} vs {to}")
    assert np.all(rel <= 1e-12), rel
    assert abs(t - to) <= 1e-12 * to


def test_khi512_translation_equivariance_and_conservation():
    """The headline configuration at its size (KHI 512^3 periodic, FP32 fast
    math, the benchmark's kernel and tiling), through properties that hold
    at any size: each cell's update depends only on its neighbours and the
    global dt (a minimum, independent of order), so a periodically shifted
    initial state gives the shifted state and the same dt sequence bit for
    bit after 4 steps; the fluxes telescope, so the totals of mass, momentum
    and energy change only by the FP32 rounding of the updates."""
    from deepgreen import PERIODIC, Simulation, Solver
    from deepgreen.setups import khi
    n = 512
    w, ds = khi(n, np.float32)
    shift = (37, 91, 13)

    def run(fields):
        sim = Simulation(n, n, n, ds, [PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0, precision="fp32", math="fast")
        with Solver(sim) as s:
            s.set_primitives(fields)
            s0 = s.conservation_sums()
            dts = [s.compute_step() for _ in range(4)]
            s1 = s.conservation_sums()
            U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
        return U, dts, s0, s1

    _say("  KHI 512^3: reference run")
    U1, dt1, a0, a1 = run(w)
    shifted = []
    for f in w:
        g = np.zeros_like(f)
        g[2:-2, 2:-2, 2:-2] = np.roll(f[2:-2, 2:-2, 2:-2], shift, axis=(0, 1, 2))
        shifted.append(g)
    del w
    _say("  KHI 512^3: shifted run")
    U2, dt2, b0, b1 = run(shifted)
    del shifted
    assert dt1 == dt2
    np.testing.assert_array_equal(U2, np.roll(U1, shift, axis=(1, 2, 3)))
    scale = np.array([a0[0], a0[0], a0[0], a0[0], a0[4]])
    assert np.all(np.abs(a1 - a0) <= 1e-6 * scale), (a0, a1)
    assert np.all(np.isfinite(U1))


@pytest.mark.parametrize("precision, math", [("fp64", "exact"), ("fp32_bulk", "exact"), ("fp32", "fast")])
def test_blast256_mirror_symmetry(precision, math):
    """The blast configuration at its size (256^3 periodic, minmod, 20
    steps): the initial state is exactly symmetric under the reflection of
    each axis about the centre (density and energy even, the reflected
    momentum component odd).  The scheme's slopes and Rusanov fluxes are
    sign-symmetric, and the exact-math update takes the x flux difference
    first (physics.rs:539-593: (((Fx+ - Fx-) + Fy+) - Fy-) + ...), so the x
    mirror holds bit for bit; the y and z differences are added to a running
    sum, whose rounding is not mirror-symmetric, so those mirrors (and, in
    fast math, whose update is reassociated into fused multiply-adds, all
    three) hold to a few ulps of the state (FP64) or a few FP32 roundings
    amplified by the limiter.  A mistake in an axis's slot rotation, a face
    index or a sign breaks them by far more."""
    from deepgreen import PERIODIC, Simulation, Solver
    from deepgreen.setups import blast
    n = 256
    w, ds = blast(n, np.float64 if precision == "fp64" else np.float32)
    sim = Simulation(n, n, n, ds, [PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0, precision=precision, math=math,
                     limiter="minmod")
    with Solver(sim) as s:
        s.set_primitives(w)
        t, steps = s.run(20)
        U = np.stack([np.array(s.interior(a)) for a in s.get_state()])
    assert steps == 20 and np.isfinite(U).all()
    assert U[4].max() > 1.001 * U[4].min()   # the blast has spread: a nontrivial state
    tol = 1e-13 if precision == "fp64" else 1e-6   # relative L1; relative Linf within 10x
    for axis in range(3):
        R = np.flip(U, axis=axis + 1)
        for q in range(5):
            sign = -1 if q == axis + 1 else 1
            if axis == 0 and math == "exact":
                np.testing.assert_array_equal(U[q], sign * R[q], err_msg=f"field {q}, x mirror")
            else:
                d = np.abs(U[q].astype(np.float64) - sign * R[q].astype(np.float64))
                ref = np.abs(U[q].astype(np.float64))
                l1, linf = d.sum() / max(ref.sum(), 1e-300), d.max() / max(ref.max(), 1e-300)
                _say(f"  field {q}, axis {axis}: rel L1 {l1:.2e}, rel Linf {linf:.2e}")
                assert l1 <= tol and linf <= 10 * tol, (q, axis, l1, linf)
