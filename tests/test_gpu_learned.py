"""GPU: the learned flux correction's MFMA kernels and the learned step.

Forward values and gradients against the torch-CPU float32 reference with
the kernels' bfloat16 rounding points (tests/cnn_reference.py); a
directional finite-difference check of the gradient on the GPU network
itself; Adam against its closed form; the learned step against the plain
step (zero correction: bit for bit), against a numpy restatement of the
masked telescoping correction (host-given corrections, with the positivity
fallback), and for conservation.
"""
import sys

import numpy as np
import pytest

torch = pytest.importorskip("torch")
import cnn_reference as ref  # noqa: E402
from test_cnn_emulation import check_grads, correction_reference, rel  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def model(seed=1, scale3=10.0):
    from deepgreen.learned import FluxCNN, join_params
    m = FluxCNN(seed=seed)
    b = ref.split(m.get_params())
    b["W3"] = b["W3"] * scale3
    b["b1"] = np.linspace(-0.2, 0.2, 32, dtype=np.float32)
    b["b3"] = np.linspace(-0.1, 0.1, 15, dtype=np.float32)
    m.set_params(join_params(b))
    return m


def solver(n, seed, ds=None):
    import deepgreen
    sim = deepgreen.Simulation(n, n, n, ds or 1.0 / n, [deepgreen.PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0,
                               precision="fp32", math="fast", limiter="minmod")
    s = deepgreen.Solver(sim)
    rng = np.random.default_rng(seed)
    shape = sim.shape
    w = [1.0 + 0.3 * rng.random(shape), 0.3 * rng.normal(size=shape), 0.3 * rng.normal(size=shape),
         0.3 * rng.normal(size=shape), 1.0 + 0.5 * rng.random(shape)]
    s.set_primitives([a.astype(np.float32) for a in w])
    return s


def interior(U):
    return np.stack([a[2:-2, 2:-2, 2:-2] for a in U])


def test_forward_matches_torch():
    m = model()
    x = np.random.default_rng(0).normal(size=(2, 20, 18, 24, 5)).astype(np.float32)
    _say("  cnn forward 2x20x18x24")
    y = m.forward(x)
    yr = ref.Net(m.get_params()).forward(x).detach().numpy()
    assert rel(y, yr) < 1e-2


def test_backward_matches_torch():
    m = model(seed=2)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 16, 17, 15, 5)).astype(np.float32)
    dy = rng.normal(size=(2, 10, 11, 9, 15)).astype(np.float32)
    _say("  cnn backward 2x16x17x15")
    g = m.backward(x, dy)
    net = ref.Net(m.get_params())
    (net.forward(x) * torch.as_tensor(dy)).sum().backward()
    check_grads(g, net.grads_flat())


def test_train_batch_matches_torch():
    m = model(seed=3)
    rng = np.random.default_rng(2)
    L = 8
    x = rng.normal(size=(3, L + 7, L + 7, L + 7, 5)).astype(np.float32)
    t = rng.normal(size=(3, L, L, L, 5)).astype(np.float32)
    _say("  cnn train batch 3 x 8^3")
    loss, g = m.train_batch(x, t)
    net = ref.Net(m.get_params())
    lt = ref.divergence_loss(net.forward(x), t)
    lt.backward()
    assert abs(loss - lt.item()) < 1e-2 * lt.item()
    check_grads(g, net.grads_flat())


def test_gradient_directional_finite_difference():
    """d/dh L(theta + h v) at h = 0 along v = g / |g| equals |g| (central
    differences of the GPU network's own loss; h moves the weights by many
    bf16 steps so that rounding does not dominate)."""
    from deepgreen.learned import FluxCNN
    m = model(seed=4)
    rng = np.random.default_rng(3)
    L = 6
    x = rng.normal(size=(2, L + 7, L + 7, L + 7, 5)).astype(np.float32)
    t = rng.normal(size=(2, L, L, L, 5)).astype(np.float32)
    theta = m.get_params()
    _, g = m.train_batch(x, t)
    v = g / np.linalg.norm(g)
    h = 2e-2 * np.linalg.norm(theta) / np.sqrt(theta.size)
    m.set_params(theta + h * v)
    lp, _ = m.train_batch(x, t)
    m.set_params(theta - h * v)
    lm, _ = m.train_batch(x, t)
    fd = (lp - lm) / (2 * h)
    assert abs(fd - np.linalg.norm(g)) < 5e-2 * np.linalg.norm(g), (fd, np.linalg.norm(g))


def test_adam_closed_form():
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=5)
    p0 = m.get_params()
    g = np.random.default_rng(4).normal(size=p0.size).astype(np.float32)
    lr, eps, k = 1e-3, 1e-8, 10
    for _ in range(k):
        m.adam_step(g, lr=lr, eps=eps)
    np.testing.assert_allclose(m.get_params(), p0 - k * lr * g / (np.abs(g) + eps), rtol=0, atol=2e-6)


def test_zero_correction_is_plain_step():
    from deepgreen.learned import FluxCNN, join_params
    m = FluxCNN(seed=6)
    b = ref.split(m.get_params())
    b["W3"][:] = 0.0
    b["b3"][:] = 0.0
    m.set_params(join_params(b))
    a, p = solver(48, 1), solver(48, 1)
    _say("  learned step, zero correction, 48^3")
    for _ in range(3):
        assert m.step(a) == p.compute_step()
    np.testing.assert_array_equal(interior(a.get_state()), interior(p.get_state()))
    t, n = m.run(a, 20)
    t2, n2 = p.run(20)
    assert (t, n) == (t2, n2)
    np.testing.assert_array_equal(interior(a.get_state()), interior(p.get_state()))


def test_correction_fallback_and_conservation():
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=7)
    n = 24
    a, p = solver(n, 2), solver(n, 2)
    sums0 = a.conservation_sums()
    rng = np.random.default_rng(7)
    G = (0.05 * rng.normal(size=(n, n, n, 16))).astype(np.float32)
    G[..., 15] = 0.0
    G[2, 3, 1, 0] = 1e4
    G[10, 20, 5, 7] = -1e4   # y-momentum... component 2 of the y face: mass of the y face is channel 5
    G[10, 20, 5, 5] = -1e4   # mass into (10,20,5) from (10,21,5): the upper neighbour loses mass
    _say("  learned step, host-given correction, 24^3")
    dt = m.step_given(a, G)
    assert dt == p.compute_step()
    Up = interior(p.get_state()).astype(np.float64)
    expect, flag = correction_reference(Up, G, dt, 1.0 / n, m.get_normalization()[10:15], 1.4)
    assert flag[2, 3, 1] and flag[10, 21, 5]
    assert m.flagged() == int(flag.sum())
    got = interior(a.get_state())
    np.testing.assert_allclose(got, expect, rtol=2e-6, atol=2e-6 * np.abs(expect).max())
    sums1 = a.conservation_sums()
    np.testing.assert_allclose(sums1, sums0, rtol=0, atol=1e-5 * np.abs(Up).sum(axis=(1, 2, 3)).max())


def test_learned_step_applies_the_bf16_correction():
    """A learned step applies the output layer rounded to bfloat16 (nearest,
    ties to even): bit for bit a host-given correction made of
    dg_learned_correction's fp32 output rounded so, from the same state."""
    from test_cnn_emulation import to_bf16
    m = model(seed=10)
    a, b = solver(24, 4), solver(24, 4)
    G = m.correction(b)
    Gh = to_bf16(G)
    assert not np.array_equal(Gh, G)
    dt_b = m.step_given(b, Gh)
    dt_a = m.step(a)
    assert dt_a == dt_b
    np.testing.assert_array_equal(interior(a.get_state()), interior(b.get_state()))
    assert a.next_dt() == b.next_dt()
    a.close(), b.close()


def test_learned_inputs_are_normalized_primitives():
    m = model(seed=8)
    norm = np.array([1.1, 0.0, 0.01, -0.02, 1.2, 0.2, 0.3, 0.3, 0.35, 0.15] + [1.0] * 5, np.float32)
    m.set_normalization(norm)
    s = solver(20, 3)
    G = m.correction(s)
    W = interior(s.get_state(kind=0)).astype(np.float32)
    x = (np.moveaxis(W, 0, -1) - norm[:5]) / norm[5:10]
    x = np.pad(x, ((3, 3), (3, 3), (3, 3), (0, 0)), mode="wrap")[None]
    Gr = ref.Net(m.get_params()).forward(x).detach().numpy()[0]
    assert rel(G[..., :15], Gr) < 1e-2


def test_samples_training_and_learned_run():
    """Samples from a 64^3 / 32^3 pair; 60 Adam steps on 16^3 crops lower the
    loss; learned runs conserve mass."""
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=9)
    fine, coarse = solver(64, 4), solver(32, 5, ds=2.0 / 64)
    _say("  samples 64^3 -> 32^3")
    for _ in range(6):
        m.add_sample(fine, coarse)
        fine.run(3)
    m.fit_normalization()
    losses = m.train(60, batch=4, crop=16, lr=1e-3, seed=1)
    _say(f"  losses {losses[:3]} ... {losses[-3:]}")
    assert np.all(np.isfinite(losses))
    assert losses[-10:].mean() < losses[:10].mean()
    coarse.restrict_from(fine)
    s0 = coarse.conservation_sums()
    _, n0 = coarse.run(0)
    t, n = m.run(coarse, 40)
    assert n - n0 == 40
    s1 = coarse.conservation_sums()
    np.testing.assert_allclose(s1[0], s0[0], rtol=2e-6)
    assert np.isfinite(s1).all()


def test_learned_workload_128():
    """The learned configuration at its size: the 128^3 periodic grid of the
    learned workload's initial condition (8 density modes, Gaussian pressure
    blasts), FP32 fast math, minmod, with a fixed-seed network (a 10x output
    layer, so the correction is not negligible).  Over 24 learned steps the
    total mass is conserved to rounding (the correction telescopes), the
    state stays finite, and the per-step fallback count (cells keeping the
    plain flux) is reported; with a zero output layer the learned run equals
    the plain run bit for bit."""
    from deepgreen import Solver
    from deepgreen.learned import FluxCNN, _sim, join_params
    from deepgreen.setups import learned_ic
    n = 128
    w, ds = learned_ic(n, 7)
    _say(f"  learned {n}^3 workload IC ready")
    m = model(seed=3)
    m.set_normalization(np.array([1.0, 0, 0, 0, 2.0, 0.2, 0.1, 0.1, 0.1, 2.0] + [0.1, 0.03, 0.03, 0.03, 0.3],
                                 np.float32))
    with Solver(_sim(n, ds)) as s, Solver(_sim(n, ds)) as plain, Solver(_sim(n, ds)) as lz:
        for x in (s, plain, lz):
            x.set_primitives(w)
        M0 = s.conservation_sums()[0]
        flagged = []
        for _ in range(24):
            m.step(s)
            flagged.append(m.flagged())
        sums = s.conservation_sums()
        _say(f"  mass drift {(sums[0] - M0) / M0:.3e}, fallback cells per step {flagged}")
        assert np.isfinite(sums).all()
        assert all(isinstance(f, int) and 0 <= f < n ** 3 for f in flagged)
        assert abs(sums[0] - M0) <= 1e-8 * M0
        U = interior(s.get_state())
        assert np.isfinite(U).all()
        # zero correction: the learned run is the plain run
        b = ref.split(m.get_params())
        b["W3"] = np.zeros_like(b["W3"])
        b["b3"] = np.zeros_like(b["b3"])
        mz = FluxCNN(seed=3)
        mz.set_params(join_params(b))
        mz.set_normalization(m.get_normalization())
        mz.run(lz, 24)
        plain.run(24)
        np.testing.assert_array_equal(interior(lz.get_state()), interior(plain.get_state()))
        assert lz.run(0) == plain.run(0)
