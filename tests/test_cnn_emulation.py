"""The flux-correction network's HIP kernels compiled for the CPU
(tests/cpu_emu: MFMA operations emulated with the documented lane layouts)
against the torch-CPU float32 reference with the kernels' bfloat16 rounding
points (tests/cnn_reference.py): forward, parameter gradients, the
divergence loss and its gradient, and Adam."""
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import cnn_reference as ref  # noqa: E402

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
EMU = os.path.join(ROOT, "tests", "cpu_emu")
EMU_LIB = os.path.join(EMU, "libdeepgreen_emu.so")


@pytest.fixture(scope="module")
def emu_lib():
    subprocess.run(["make", "-s", "-C", EMU, "EXTRA=-DDG_EMU_POISON"], check=True)
    import deepgreen
    return deepgreen.load_library(EMU_LIB)


def model(lib, seed=1):
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=seed, lib=lib)
    p = m.get_params()
    # a larger output layer than the default (0.1x) init, so the test sees it
    blocks = ref.split(p)
    blocks["W3"] = blocks["W3"] * 10.0
    blocks["b1"] = np.linspace(-0.2, 0.2, 32, dtype=np.float32)
    blocks["b3"] = np.linspace(-0.1, 0.1, 15, dtype=np.float32)
    from deepgreen.learned import join_params
    m.set_params(join_params(blocks))
    return m


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-30))


def check_grads(g, gr, tol=3e-2):
    """Per parameter block: relative error below tol, measured against the
    block's norm or 1e-2 of the whole gradient's norm, whichever is larger (the
    output bias of the divergence loss has a zero gradient in exact
    arithmetic; bf16 rounding of dG leaves a small residue there)."""
    total = np.linalg.norm(gr)
    o = 0
    for name, shape in ref.LAYOUT:
        n = int(np.prod(shape))
        err = np.linalg.norm(g[o:o + n] - gr[o:o + n])
        assert err < tol * max(np.linalg.norm(gr[o:o + n]), 1e-2 * total), (name, err, np.linalg.norm(gr[o:o + n]))
        o += n


def test_forward_matches_torch(emu_lib):
    m = model(emu_lib)
    x = np.random.default_rng(0).normal(size=(2, 9, 8, 10, 5)).astype(np.float32)
    y = m.forward(x)
    yr = ref.Net(m.get_params()).forward(x).detach().numpy()
    assert y.shape == yr.shape == (2, 3, 2, 4, 15)
    assert rel(y, yr) < 1e-2


def test_backward_matches_torch(emu_lib):
    m = model(emu_lib, seed=2)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 9, 10, 8, 5)).astype(np.float32)
    dy = rng.normal(size=(1, 3, 4, 2, 15)).astype(np.float32)
    g = m.backward(x, dy)
    net = ref.Net(m.get_params())
    (net.forward(x) * torch.as_tensor(dy)).sum().backward()
    check_grads(g, net.grads_flat())


def test_train_batch_matches_torch(emu_lib):
    m = model(emu_lib, seed=3)
    rng = np.random.default_rng(2)
    L = 3
    x = rng.normal(size=(2, L + 7, L + 7, L + 7, 5)).astype(np.float32)
    t = rng.normal(size=(2, L, L, L, 5)).astype(np.float32)
    loss, g = m.train_batch(x, t)
    net = ref.Net(m.get_params())
    lr_ = ref.divergence_loss(net.forward(x), t)
    lr_.backward()
    assert abs(loss - lr_.item()) < 1e-2 * lr_.item()
    check_grads(g, net.grads_flat())


def test_adam_closed_form(emu_lib):
    """Constant gradient g: the bias-corrected moments are g and g^2, so each
    step moves a parameter by lr * g / (|g| + eps); after k steps
    p_k = p_0 - k lr g / (|g| + eps)."""
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=4, lib=emu_lib)
    p0 = m.get_params()
    g = np.random.default_rng(3).normal(size=p0.size).astype(np.float32)
    lr, eps, k = 1e-3, 1e-8, 5
    for _ in range(k):
        m.adam_step(g, lr=lr, eps=eps)
    expect = p0 - k * lr * g / (np.abs(g) + eps)
    np.testing.assert_allclose(m.get_params(), expect, rtol=0, atol=1e-6)


def periodic_solver(lib, n, seed, ds=None, math="fast"):
    import deepgreen
    sim = deepgreen.Simulation(n, n, n, ds or 1.0 / n, [deepgreen.PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0,
                               precision="fp32", math=math, limiter="minmod", tile_y=4)
    s = deepgreen.Solver(sim, lib=lib)
    rng = np.random.default_rng(seed)
    shape = sim.shape
    w = [1.0 + 0.3 * rng.random(shape), 0.3 * rng.normal(size=shape), 0.3 * rng.normal(size=shape),
         0.3 * rng.normal(size=shape), 1.0 + 0.5 * rng.random(shape)]
    s.set_primitives([a.astype(np.float32) for a in w])
    return s


def interior(U):
    return np.stack([a[2:-2, 2:-2, 2:-2] for a in U])


def test_learned_zero_correction_is_plain_step(emu_lib):
    """With the output layer zeroed the learned step is the plain step, bit for bit."""
    from deepgreen.learned import FluxCNN, join_params
    m = FluxCNN(seed=5, lib=emu_lib)
    b = ref.split(m.get_params())
    b["W3"][:] = 0.0
    b["b3"][:] = 0.0
    m.set_params(join_params(b))
    a, p = periodic_solver(emu_lib, 6, 1), periodic_solver(emu_lib, 6, 1)
    for _ in range(2):
        assert m.step(a) == p.compute_step()
    np.testing.assert_array_equal(interior(a.get_state()), interior(p.get_state()))
    assert a.next_dt() == p.next_dt()
    a.close(), p.close()


def correction_reference(Up, G, dt, ds, scale, gamma):
    """numpy: masked telescoping correction of a plain update Up (interior
    [5, n, n, n], extensive) by G [n, n, n, 16] (periodic)."""
    V, gm1 = ds ** 3, gamma - 1.0

    def delta(mask_hi=None, mask_lo=None):
        d = np.zeros_like(Up, dtype=np.float64)
        for a in range(3):
            Gd = G[..., 5 * a:5 * a + 5].astype(np.float64)          # face c + e_a/2
            Glo = np.roll(Gd, 1, axis=a)                              # face c - e_a/2
            mh = 1.0 if mask_hi is None else mask_hi[a][..., None]
            ml = 1.0 if mask_lo is None else mask_lo[a][..., None]
            d += np.moveaxis(mh * Gd - ml * Glo, -1, 0)
        return -dt * ds * ds * np.asarray(scale, np.float64)[:, None, None, None] * d

    def invalid(U):
        rho = U[0] / V
        p = (U[4] / V - 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / (rho * V * V)) * gm1
        return ~((rho > 0) & (p > 0))

    def masks(flag):
        return ([~(flag | np.roll(flag, -1, axis=a)) for a in range(3)],
                [~(flag | np.roll(flag, 1, axis=a)) for a in range(3)])

    flag = invalid(Up + delta())
    if flag.any():
        # two passes: cells invalid once the flagged cells' faces keep the plain flux
        for _ in range(2):
            flag = flag | invalid(Up + delta(*masks(flag)))
    return Up + delta(*masks(flag)), flag


def test_learned_correction_fallback_and_conservation(emu_lib):
    """A host-given correction with one huge face value: the cell that loses
    mass through that face would get a negative density, is flagged and keeps
    the plain flux on all its faces (its neighbour then gets the plain flux
    on the shared face); every other face is corrected; the totals are
    unchanged up to rounding."""
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=6, lib=emu_lib)
    n, ds = 6, 1.0 / 6
    a, p = periodic_solver(emu_lib, n, 2), periodic_solver(emu_lib, n, 2)
    sums0 = a.conservation_sums()
    rng = np.random.default_rng(7)
    G = (0.05 * rng.normal(size=(n, n, n, 16))).astype(np.float32)
    G[..., 15] = 0.0
    G[2, 3, 1, 0] = 1e4       # mass flux through the +x face of cell (2, 3, 1)
    dt = m.step_given(a, G)
    assert dt == p.compute_step()
    Up = interior(p.get_state()).astype(np.float64)
    scale = m.get_normalization()[10:15]
    expect, flag = correction_reference(Up, G, dt, ds, scale, 1.4)
    assert flag.sum() == 1 and flag[2, 3, 1]
    assert m.flagged() == 1
    got = interior(a.get_state())
    np.testing.assert_allclose(got, expect, rtol=2e-6, atol=2e-6 * np.abs(expect).max())
    np.testing.assert_array_equal(got[:, 2, 3, 1], Up[:, 2, 3, 1].astype(np.float32))
    sums1 = a.conservation_sums()
    np.testing.assert_allclose(sums1, sums0, rtol=0, atol=1e-5 * np.abs(Up).sum(axis=(1, 2, 3)).max())
    a.close(), p.close()


def test_learned_correction_without_flags(emu_lib):
    """A small host-given correction that flags no cell: every face is
    corrected (the apply pass uses the increments the flag pass kept), twice
    in a row so the second step reads increments of its own step."""
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=6, lib=emu_lib)
    n, ds = 6, 1.0 / 6
    a, p = periodic_solver(emu_lib, n, 2), periodic_solver(emu_lib, n, 2)
    sums0 = a.conservation_sums()
    rng = np.random.default_rng(11)
    scale = m.get_normalization()[10:15]
    for k in range(2):
        p.set_conserved(a.get_state())
        G = (0.05 * rng.normal(size=(n, n, n, 16))).astype(np.float32)
        G[..., 15] = 0.0
        dt = m.step_given(a, G)
        assert dt == p.compute_step()
        Up = interior(p.get_state()).astype(np.float64)
        expect, flag = correction_reference(Up, G, dt, ds, scale, 1.4)
        assert not flag.any() and m.flagged() == 0
        got = interior(a.get_state())
        assert not np.array_equal(got, Up.astype(np.float32))
        np.testing.assert_allclose(got, expect, rtol=2e-6, atol=2e-6 * np.abs(expect).max())
    np.testing.assert_allclose(a.conservation_sums(), sums0, rtol=0, atol=1e-5 * np.abs(Up).sum(axis=(1, 2, 3)).max())
    a.close(), p.close()


def test_learned_correction_fallback_reaches_the_neighbours(emu_lib):
    """Two huge mass fluxes in a row along x: cell (2, 3, 1) empties through
    its +x face and is flagged; cell (3, 3, 1) passes the mass on through its
    own +x face (net zero with every face corrected), but once the face it
    shares with the flagged cell keeps the plain flux it would empty too, so
    the masked pass flags it as well.  The step is then the plain step on
    every cell those faces touch, and the state stays valid."""
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=6, lib=emu_lib)
    n, ds = 6, 1.0 / 6
    a, p = periodic_solver(emu_lib, n, 2), periodic_solver(emu_lib, n, 2)
    sums0 = a.conservation_sums()
    G = np.zeros((n, n, n, 16), np.float32)
    G[2, 3, 1, 0] = 1e4
    G[3, 3, 1, 0] = 1e4
    dt = m.step_given(a, G)
    assert dt == p.compute_step()
    Up = interior(p.get_state()).astype(np.float64)
    scale = m.get_normalization()[10:15]
    expect, flag = correction_reference(Up, G, dt, ds, scale, 1.4)
    assert flag.sum() == 2 and flag[2, 3, 1] and flag[3, 3, 1]
    assert m.flagged() == 2
    got = interior(a.get_state())
    # every nonzero face belongs to a flagged cell: the learned step is the plain step
    np.testing.assert_array_equal(got, Up.astype(np.float32))
    np.testing.assert_allclose(got, expect, rtol=0, atol=0)
    assert (got[0] > 0).all() and np.isfinite(a.next_dt())
    np.testing.assert_allclose(a.conservation_sums(), sums0, rtol=0, atol=1e-5 * np.abs(Up).sum(axis=(1, 2, 3)).max())
    a.close(), p.close()


def to_bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32)
    u = (u + (((u >> 16) & 1) + 0x7FFF)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def test_learned_step_applies_the_bf16_correction(emu_lib):
    """A learned step applies the output layer rounded to bfloat16 (nearest,
    ties to even): bit for bit a host-given correction made of
    dg_learned_correction's fp32 output rounded so, from the same state."""
    m = model(emu_lib, seed=10)
    a, b = periodic_solver(emu_lib, 6, 4), periodic_solver(emu_lib, 6, 4)
    G = m.correction(b)
    Gh = to_bf16(G)
    assert not np.array_equal(Gh, G) and np.abs(Gh - G).max() <= 2.0 ** -8 * np.abs(G).max()
    assert to_bf16(np.float32([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8]))[0] == 1.0   # ties to even
    dt_b = m.step_given(b, Gh)
    dt_a = m.step(a)
    assert dt_a == dt_b
    np.testing.assert_array_equal(interior(a.get_state()), interior(b.get_state()))
    assert a.next_dt() == b.next_dt()
    a.close(), b.close()


def test_learned_inputs_are_normalized_primitives(emu_lib):
    """dg_learned_correction: the network applied to the normalized
    primitives of the periodic state (halo of 3 wrapped cells)."""
    from deepgreen.learned import FluxCNN
    m = model(emu_lib, seed=8)
    norm = np.array([1.1, 0.0, 0.01, -0.02, 1.2, 0.2, 0.3, 0.3, 0.35, 0.15] + [1.0] * 5, np.float32)
    m.set_normalization(norm)
    s = periodic_solver(emu_lib, 5, 3)
    G = m.correction(s)
    W = interior(s.get_state(kind=0)).astype(np.float32)       # primitives
    x = (np.moveaxis(W, 0, -1) - norm[:5]) / norm[5:10]
    x = np.pad(x, ((3, 3), (3, 3), (3, 3), (0, 0)), mode="wrap")[None]
    Gr = ref.Net(m.get_params()).forward(x).detach().numpy()[0]
    assert rel(G[..., :15], Gr) < 1e-2
    s.close()


def test_learned_sample_and_training_smoke(emu_lib):
    """Samples from a 12^3 / 6^3 pair, normalization, a few Adam steps on
    3^3 crops (the loss is finite), then learned steps conserve the totals."""
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=9, lib=emu_lib)
    fine = periodic_solver(emu_lib, 12, 4)
    coarse = periodic_solver(emu_lib, 6, 5, ds=2.0 / 12)
    for _ in range(2):
        m.add_sample(fine, coarse)
        fine.run(2)
    assert m.samples == 2
    m.fit_normalization()
    norm = m.get_normalization()
    assert np.all(norm[5:] > 0)
    losses = m.train(3, batch=2, crop=3, lr=1e-3, seed=1)
    assert np.all(np.isfinite(losses)) and losses[0] > 0
    coarse.restrict_from(fine)
    s0 = coarse.conservation_sums()
    m.run(coarse, 3)
    s1 = coarse.conservation_sums()
    np.testing.assert_allclose(s1[0], s0[0], rtol=1e-5)
    fine.close(), coarse.close()


def test_learned_run_after_solver_replaced(emu_lib):
    """A model used on one grid and then on a new grid of the same size (the
    first one destroyed, so its memory may be reused at the same addresses)
    rebuilds its input tables and captured steps: the second grid's run
    equals that of a fresh model."""
    from deepgreen.learned import FluxCNN
    m = model(emu_lib, seed=11)
    p = m.get_params()
    a = periodic_solver(emu_lib, 5, 1)
    m.run(a, 16)
    a.close()
    b = periodic_solver(emu_lib, 5, 2)
    m.run(b, 16)
    got = interior(b.get_state())
    b.close()
    m2 = FluxCNN(seed=12, lib=emu_lib)
    m2.set_params(p)
    m2.set_normalization(m.get_normalization())
    c = periodic_solver(emu_lib, 5, 2)
    m2.run(c, 16)
    np.testing.assert_array_equal(got, interior(c.get_state()))
    c.close()


def test_learned_run_equals_single_learned_steps(emu_lib):
    """dg_learned_run (captured graphs of 16 steps plus single steps) equals
    the same number of dg_learned_step calls, bit for bit, with a non-zero
    correction."""
    m = model(emu_lib, seed=13)
    a, b = periodic_solver(emu_lib, 5, 3), periodic_solver(emu_lib, 5, 3)
    m.run(a, 17)
    for _ in range(17):
        m.step(b)
    np.testing.assert_array_equal(interior(a.get_state()), interior(b.get_state()))
    assert a.run(0) == b.run(0)
    a.close(), b.close()


def test_learned_run_after_workspace_growth(emu_lib):
    """A learned run, then a training batch whose crops need larger
    workspaces than the run's (the slots the captured steps use are freed and
    reallocated), then a learned run again: equal to the same runs of a model
    that never trained in between (the captured steps are rebuilt, not replayed
    on freed buffers)."""
    from deepgreen.learned import FluxCNN
    m = model(emu_lib, seed=15)
    p, norm = m.get_params(), m.get_normalization()
    a = periodic_solver(emu_lib, 5, 6)
    m.run(a, 16)
    rng = np.random.default_rng(0)
    L = 8
    x = rng.normal(size=(2, L + 7, L + 7, L + 7, 5)).astype(np.float32)
    t = rng.normal(size=(2, L, L, L, 5)).astype(np.float32)
    m.train_batch(x, t)                      # grows the input, activation and output slots
    m.run(a, 16)
    got = interior(a.get_state())
    a.close()
    m2 = FluxCNN(seed=16, lib=emu_lib)
    m2.set_params(p)
    m2.set_normalization(norm)
    b = periodic_solver(emu_lib, 5, 6)
    m2.run(b, 32)
    np.testing.assert_array_equal(got, interior(b.get_state()))
    b.close()


def test_trajectory_samples_follow_the_coarse_grids_own_steps(emu_lib):
    """add_sample(keep_coarse=True): the coarse grid is not reset to the
    restricted fine state; it takes its own plain step with the common dt
    (the smaller of its CFL dt and twice the fine one), and store=False keeps
    no sample."""
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=10, lib=emu_lib)
    fine = periodic_solver(emu_lib, 12, 6)
    coarse = periodic_solver(emu_lib, 6, 7, ds=2.0 / 12)
    twin = periodic_solver(emu_lib, 6, 7, ds=2.0 / 12)
    coarse.restrict_from(fine)
    fine.run(3)                                   # the fine grid moves on; the coarse one keeps its state
    for store in (False, True, True):
        U0 = coarse.get_state()
        dt = min(coarse.next_dt(), 2.0 * fine.next_dt())
        twin.set_conserved(U0)
        twin.step_dt(dt)
        n0 = m.samples
        m.add_sample(fine, coarse, keep_coarse=True, store=store)
        assert m.samples == n0 + int(store)
        np.testing.assert_array_equal(interior(coarse.get_state()), interior(twin.get_state()))
    fine.close(), coarse.close(), twin.close()


def test_sample_and_step_dt_advance_after_a_finished_run(emu_lib):
    """A finished dg_run (a clock query is a run of zero steps) leaves a
    context's scalars inactive; training samples and dg_step_dt must still
    advance their grids (an inactive step carries the state unchanged)."""
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=11, lib=emu_lib)
    fine = periodic_solver(emu_lib, 12, 8)
    coarse = periodic_solver(emu_lib, 6, 9, ds=2.0 / 12)
    fine.run(2)
    _, n0 = fine.run(0)
    U0 = interior(fine.get_state())
    m.add_sample(fine, coarse)
    _, n1 = fine.run(0)
    assert n1 == n0 + 2
    assert not np.array_equal(interior(fine.get_state()), U0)
    coarse.run(0)
    W0 = interior(coarse.get_state())
    coarse.step_dt(1e-4)
    assert coarse.run(0)[1] >= 1
    assert not np.array_equal(interior(coarse.get_state()), W0)
    fine.close(), coarse.close()
