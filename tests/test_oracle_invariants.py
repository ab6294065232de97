"""Closed forms and invariants that pin the oracle where the Fortran golden
vectors cannot: the Rust-only behaviours (dt NaN fallback physics.rs:191-194,
upper-wall gradient ghosts, gravity's 0-based height), the limiter extension,
and the physics itself (a single Rusanov face evaluated in exact rational
arithmetic, uniform-state fixed points, symmetries, conservation)."""
from fractions import Fraction

import numpy as np
import pytest

import cases
from dg_oracle import OUTFLOW, PERIODIC, WALL, Oracle


def make(interior, bcs, prec="f64", **kw):
    ds = kw.pop("ds", 1.0 / max(interior))
    return Oracle(interior, ds, bcs, precision=prec, **kw), ds


def uniform_w(interior, rho=1.3, v=(0.2, -0.1, 0.3), p=0.9):
    shape = tuple(n + 4 for n in interior)
    return [np.full(shape, rho), np.full(shape, v[0]), np.full(shape, v[1]), np.full(shape, v[2]), np.full(shape, p)]


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("bc", [PERIODIC, OUTFLOW, WALL])
def test_uniform_state_is_fixed_point(prec, bc):
    # walls reflect the normal velocity, so only a resting state is uniform there
    v = (0.0, 0.0, 0.0) if bc == WALL else (0.2, -0.1, 0.3)
    o, ds = make((6, 5, 7), [bc] * 6, prec)
    o.set_primitives(uniform_w((6, 5, 7), v=v))
    o.calc_conserved()
    u0 = [np.array(o.interior_view(a)) for a in o.cons]
    o.steps(3)
    for a, b in zip(u0, o.cons):
        np.testing.assert_array_equal(a, o.interior_view(b))


@pytest.mark.parametrize("prec,dtype", [("f32", np.float32), ("f64", np.float64)])
def test_timestep_closed_form(prec, dtype):
    # gamma*p/rho = 4 and |v| = 5 exactly: dt = cfl * ds / 7
    o, ds = make((4, 4, 4), [PERIODIC] * 6, prec, gamma=2.0, cfl=0.5, ds=0.125)
    o.set_primitives(uniform_w((4, 4, 4), rho=1.0, v=(3.0, 0.0, 4.0), p=2.0))
    o.calc_conserved()
    o.calc_primitives()
    o.update_primitive_ghost_cells()
    assert o.calc_timestep() == float(dtype(0.5) * (dtype(0.125) / dtype(7.0)))


@pytest.mark.parametrize("prec,dtype", [("f32", np.float32), ("f64", np.float64)])
def test_timestep_nan_falls_back_to_cfl(prec, dtype):
    o, ds = make((4, 4, 4), [PERIODIC] * 6, prec, cfl=0.4)
    w = uniform_w((4, 4, 4))
    w[4][3, 3, 3] = -1.0  # sqrt of a negative -> NaN in one cell
    o.set_primitives(w)
    o.update_primitive_ghost_cells()
    assert o.calc_timestep() == float(dtype(0.4) * dtype(1.0))
    o2, _ = make((4, 4, 4), [PERIODIC] * 6, prec, cfl=0.4, dt_nan_fallback=False)
    o2.set_primitives(w)
    o2.update_primitive_ghost_cells()
    assert np.isnan(o2.calc_timestep())


def _exact_rusanov(l, r, gamma):
    """Rusanov flux of the averaged state, in exact rationals except for the
    two square roots (done in float64, then made exact)."""
    g = Fraction(gamma)

    def cons(s):
        rho, u, v, w, p = s
        e = rho * (u * u + v * v + w * w) / 2 + p / (g - 1)
        return [rho, rho * u, rho * v, rho * w, e]

    L, Rr = [Fraction(x) for x in l], [Fraction(x) for x in r]
    Ul, Ur = cons(L), cons(Rr)
    Us = [(a + b) / 2 for a, b in zip(Ul, Ur)]
    rho, mn, mt1, mt2, e = Us
    p = (g - 1) * (e - (mn * mn + mt1 * mt1 + mt2 * mt2) / (2 * rho))
    un = mn / rho
    F = [mn, mn * un + p, mt1 * un, mt2 * un, (e + p) * un]
    cl = Fraction(np.sqrt(float(g * L[4] / L[0]))) + abs(L[1])
    cr = Fraction(np.sqrt(float(g * Rr[4] / Rr[0]))) + abs(Rr[1])
    c = max(cl, cr)
    return [f - c / 2 * (a - b) for f, a, b in zip(F, Ul, Ur)]


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_single_face_rusanov_matches_exact(axis):
    """Set xtr and gradients by hand so the face states are known exactly,
    then compare the oracle's face flux with the textbook flux."""
    interior = (4, 4, 4)
    o, ds = make(interior, [PERIODIC] * 6, "f64", ds=0.5, gamma=1.4)
    # cell c = (3,3,3) is "r" (its +axis face state), c+e is "l"
    c = (3, 3, 3)
    ce = list(c)
    ce[axis] += 1
    ce = tuple(ce)
    sr = (1.25, 0.5, -0.25, 0.125, 2.0)    # rho, v_normal, v_t1, v_t2, p at face
    sl = (0.75, -0.375, 0.25, 0.5, 1.5)
    # rotated slots (physics.rs:637-677): normal velocity -> its own axis
    vel_axes = {0: (1, 2, 3), 1: (2, 1, 3), 2: (3, 1, 2)}[axis]
    fields = [0] + list(vel_axes) + [4]
    for slot, f in enumerate(fields):
        o.xtr[f][c] = sr[slot]
        o.xtr[f][ce] = sl[slot]
        o.grad[f][c + (axis,)] = 0.0
        o.grad[f][ce + (axis,)] = 0.0
    o.calc_fluxes()
    want = _exact_rusanov(sl, sr, 1.4)  # note: the "l" argument is cell c+e
    # the library stores F = ds * (Rusanov), slot order rotated back
    got = [o.flux[f][c + (axis,)] / ds for f in fields]
    # want uses l = c+e as the *left* argument of the textbook formula's
    # (U_l - U_r) term, exactly as physics.rs:396-400
    for g_, w_ in zip(got, want):
        assert g_ == pytest.approx(float(w_), rel=1e-14, abs=1e-14)


def _run(o, w, nsteps):
    o.set_primitives(w)
    o.compute(1e30, nsteps)
    return [np.array(o.interior_view(a)) for a in o.cons]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_axis_permutation_symmetry(prec):
    """Rotating the problem x->y->z permutes the result: pins the momentum-slot
    rotation of the three flux calls.  Not bit for bit, because the reference
    sums the axes in a fixed x, y, z order (physics.rs:254-259, 548-554); a
    slot mix-up would show as an O(1) difference."""
    w, ds = cases.smooth_random((8, 6, 5), seed=2)
    ref = _run(Oracle((8, 6, 5), ds, [PERIODIC, PERIODIC, OUTFLOW, OUTFLOW, PERIODIC, PERIODIC],
                      precision=prec), w, 6)
    perm = (2, 0, 1)  # new axes (x', y', z') = old (z, x, y)
    wp = [np.transpose(a, perm) for a in w]
    wp = [wp[0], wp[3], wp[1], wp[2], wp[4]]  # vx' = vz, vy' = vx, vz' = vy
    out = _run(Oracle((5, 8, 6), ds, [PERIODIC, PERIODIC, PERIODIC, PERIODIC, OUTFLOW, OUTFLOW],
                      precision=prec), wp, 6)
    back = [np.transpose(a, (1, 2, 0)) for a in out]
    back = [back[0], back[2], back[3], back[1], back[4]]
    tol = 2e-5 if prec == "f32" else 1e-12
    for a, b in zip(ref, back):
        np.testing.assert_allclose(b, a, rtol=tol, atol=tol * np.abs(a).max())


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("bc", [PERIODIC, WALL])
def test_mirror_symmetry(prec, bc):
    """Mirroring x (and negating vx) mirrors the result.  With walls this pins
    the upper-wall ghosts (gradient computed on the ghost, boundaries.rs
    mirror map) against the lower-wall ones, which the Fortran vectors pin."""
    interior = (9, 6, 5)
    w, ds = cases.smooth_random(interior, seed=4, amp_v=0.1)
    bcs = [bc, bc, PERIODIC, PERIODIC, PERIODIC, PERIODIC]
    ref = _run(Oracle(interior, ds, bcs, precision=prec), w, 6)
    wm = [a[::-1].copy() for a in w]
    wm[1] = -wm[1]
    out = _run(Oracle(interior, ds, bcs, precision=prec), wm, 6)
    out = [a[::-1] for a in out]
    out[1] = -out[1]
    for a, b in zip(ref, out):
        np.testing.assert_allclose(a, b, rtol=0, atol=0)


def test_periodic_conservation():
    w, ds = cases.smooth_random((10, 8, 6), seed=9)
    o = Oracle((10, 8, 6), ds, [PERIODIC] * 6, precision="f64")
    o.set_primitives(w)
    o.calc_conserved()
    s0 = np.array([np.array(o.interior_view(a)).sum() for a in o.cons])
    o.steps(10)
    s1 = np.array([np.array(o.interior_view(a)).sum() for a in o.cons])
    np.testing.assert_allclose(s1, s0, rtol=1e-13, atol=1e-15)


def test_outflow_lower_face_copies_first_interior_face():
    """boundaries.rs:107-111 on fluxes: F[1] <- F[2], so the first interior
    cell has zero divergence along x while the last one does not."""
    interior = (8, 3, 3)
    w, ds = cases.smooth_random(interior, seed=1)
    for a in w:  # make the state depend on x only
        a[...] = a[:, :1, :1]
    o = Oracle(interior, ds, [OUTFLOW, OUTFLOW, PERIODIC, PERIODIC, PERIODIC, PERIODIC], precision="f64")
    o.set_primitives(w)
    o.calc_conserved()
    u0 = [np.array(o.interior_view(a)) for a in o.cons]
    o.steps(1)
    u1 = [np.array(o.interior_view(a)) for a in o.cons]
    for a, b in zip(u0, u1):
        np.testing.assert_array_equal(a[0], b[0])
    assert any(not np.array_equal(a[-1], b[-1]) for a, b in zip(u0, u1))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_conserved_primitive_round_trip(prec):
    interior = (5, 4, 3)
    w, ds = cases.smooth_random(interior, seed=12)
    o = Oracle(interior, ds, [PERIODIC] * 6, precision=prec)
    o.set_primitives(w)
    w0 = [np.array(a) for a in o.prim]
    o.calc_conserved()
    o.calc_primitives()
    tol = 4e-7 if prec == "f32" else 1e-15
    for a, b in zip(w0, o.prim):
        np.testing.assert_allclose(b, a, rtol=tol, atol=tol)


def test_gravity_height_rust_vs_fortran_convention():
    """physics.rs:368 uses the 0-based k, physics.f90:138 the 1-based one:
    the two heights differ by one ds, i.e. shifting k by one reproduces the
    Rust height on a grid one cell taller in z.  Checked on the flux of a
    resting isothermal column, where F_mz = p* + rho* g h exactly."""
    interior = (3, 3, 6)
    for shift in (0, 1):
        o = Oracle(interior, 0.25, [PERIODIC] * 6, precision="f64", g=-2.0, gravity=True,
                   gravity_k_shift=shift, gamma=1.4)
        o.set_primitives(uniform_w(interior, rho=1.0, v=(0, 0, 0), p=5.0))
        o.calc_conserved()
        o.calc_primitives()
        o.update_primitive_ghost_cells()
        o.calc_gradients()
        o.update_gradients_ghost_cells()
        o.extrapolate(0.0)
        o.update_primitive_xtr_ghost_cells()
        o.calc_fluxes()
        nzg = interior[2] + 4
        for k in range(1, nzg - 1):
            h = (nzg - 4) * 0.25 - (k + shift - 2) * 0.25
            assert o.flux[3][2, 2, k, 2] == pytest.approx((5.0 + 1.0 * -2.0 * h) * 0.25, rel=1e-14)


@pytest.mark.parametrize("lim", ["minmod", "vanleer"])
def test_limited_slopes_reduce_to_central_on_linear_data(lim):
    interior = (8, 5, 4)
    o = Oracle(interior, 0.125, [OUTFLOW] * 6, precision="f64", limiter=lim)
    x = np.arange(interior[0] + 4)[:, None, None] * np.ones((1, interior[1] + 4, interior[2] + 4))
    w = uniform_w(interior)
    w[0] = 1.0 + 0.01 * x
    o.set_primitives(w)
    o.calc_gradients()
    inner = o.grad[0][2:-2, 2:-2, 2:-2, 0]
    np.testing.assert_allclose(inner, 0.01 / 0.125, rtol=1e-12)


@pytest.mark.parametrize("lim", ["minmod", "vanleer"])
def test_limited_scheme_blast_stays_positive(lim):
    """The benchmark blast (pressure ratio 100) breaks the unlimited reference
    scheme within two steps; with a limiter it stays positive."""
    n = 24
    w, ds = cases.blast((n, n, n), radius=0.1, p_in=10.0, p_out=0.1)
    o = Oracle((n, n, n), ds, [PERIODIC] * 6, precision="f64", limiter=lim, nthreads=4)
    U = _run(o, w, 40)
    assert all(np.isfinite(u).all() for u in U)
    assert U[0].min() > 0
    o2 = Oracle((n, n, n), ds, [PERIODIC] * 6, precision="f64", nthreads=4)
    U2 = _run(o2, w, 5)
    assert not all(np.isfinite(u).all() for u in U2)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_conducting_wall_heat_closed_form(axis):
    """A uniform gas at rest in a box whose lower wall along `axis` is
    conducting: the flow stays at rest (a fixed point of the walled scheme)
    and, after one step, only the first interior layer along that axis has
    gained energy, by dt ds^2 Q / V with Q = (T_out - p/(RS rho)) ds^2 /
    (0.008/0.2 + 1/hc), hc = 10 - v + 10 sqrt(v) (double-precision closed form
    of the CUDA variant's source, src/cu/include/kernels.h:561-661)."""
    n, ds = (6, 5, 7), 0.1
    rho, p, t_out, v_out, rs = 1.293, 101325.0, 300.0, 4.0, 287.0
    bcs = [2] * 6
    bcs[2 * axis] = 3
    o = Oracle(n, ds, bcs, precision="f64", gamma=5.0 / 3.0, cfl=0.2, g=0.0, t_out=t_out, v_out=v_out, rs=rs)
    shape = tuple(k + 4 for k in n)
    o.set_primitives([np.full(shape, rho), np.zeros(shape), np.zeros(shape), np.zeros(shape), np.full(shape, p)])
    o.calc_conserved()
    E0 = np.array(o.interior_view(o.cons[4])).copy()
    U0 = [np.array(o.interior_view(a)).copy() for a in o.cons[:4]]
    dt = o.steps(1)[0]
    E1 = np.array(o.interior_view(o.cons[4]))
    hc = 10.0 - v_out + 10.0 * np.sqrt(v_out)
    Q = (t_out - p / (rs * rho)) * ds * ds / (0.008 / 0.2 + 1.0 / hc)
    expect = E0.copy()
    sl = [slice(None)] * 3
    sl[axis] = 0
    expect[tuple(sl)] += dt * ds * ds * Q / ds ** 3
    np.testing.assert_allclose(E1, expect, rtol=1e-12, atol=0)
    assert np.abs(E1 - E0).max() > 0
    for a, b in zip(U0, o.cons[:4]):
        np.testing.assert_allclose(np.array(o.interior_view(b)), a, rtol=0, atol=1e-10 * rho * ds ** 3)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_conducting_wall_heat_reads_the_extrapolated_state(axis):
    """The heat source takes t_in from the half-step extrapolated (xtr) state of
    the wall ghost, as the CUDA flux kernel reads rho_xtr / p_xtr
    (physics.hpp:437-442).  Closed form: uniform rho, p and a velocity
    tangential to the wall that grows linearly along another axis (divergence
    alpha), so that away from the other boundaries the Hancock state of the
    first layer is rho (1 - alpha dt/2), p (1 - gamma alpha dt/2) exactly.
    The energy of a run with the conducting wall minus that of a run with a
    plain wall is dt ds^2 Q / V in the first layer and 0 elsewhere, with Q
    from that t_in; the start-of-step t_in would give a different increment."""
    n, ds = [8, 8, 8], 0.125
    t_out, v_out, rs, gamma = 300.0, 4.0, 0.004, 1.4
    along = (axis + 1) % 3              # the velocity's direction and the axis it varies along
    shape = tuple(k + 4 for k in n)
    coord = (np.arange(shape[along]) - 2 + 0.5) * ds
    alpha = 1.0
    vel = alpha * (coord - 0.5)
    sl = [None, None, None]
    sl[along] = slice(None)
    w = [np.ones(shape), np.zeros(shape), np.zeros(shape), np.zeros(shape), np.ones(shape)]
    w[1 + along] = np.broadcast_to(vel[tuple(sl)], shape).copy()
    runs = {}
    for bc_lo in (2, 3):
        bcs = [1] * 6
        bcs[2 * axis], bcs[2 * axis + 1] = bc_lo, 2
        o = Oracle(tuple(n), ds, bcs, precision="f64", gamma=gamma, cfl=0.2, g=0.0, t_out=t_out, v_out=v_out,
                   rs=rs)
        o.set_primitives(w)
        o.calc_conserved()
        dt = o.steps(1)[0]
        runs[bc_lo] = (dt, np.array(o.interior_view(o.cons[4])).copy())
    (dt, e_wall), (dt2, e_heat) = runs[2], runs[3]
    assert dt == dt2
    diff = e_heat - e_wall
    layer = [slice(None)] * 3
    layer[axis] = 0
    assert np.all(np.delete(diff, 0, axis=axis) == 0)
    hc = 10.0 - v_out + 10.0 * np.sqrt(v_out)

    def incr(t_in):
        return dt * ds * ((t_out - t_in) * (ds * ds / (0.008 / 0.2 + 1.0 / hc)) / ds ** 3 * ds)

    t_xtr = (1.0 - gamma * alpha * dt / 2) / (rs * (1.0 - alpha * dt / 2))
    inner = [slice(None)] * 3
    inner[along] = slice(1, -1)          # away from the outflow faces along `along`
    got = diff[tuple(layer)][tuple(inner[:axis] + inner[axis + 1:])]
    np.testing.assert_allclose(got, incr(t_xtr), rtol=1e-11)
    t_start = 1.0 / rs                   # p / (rs rho) at the start of the step
    assert abs(incr(t_start) - incr(t_xtr)) > 1e-3 * abs(incr(t_xtr))
