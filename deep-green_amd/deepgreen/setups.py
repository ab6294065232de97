"""Initial conditions of the benchmark and ensemble workloads.

Every function returns primitive fields (rho, vx, vy, vz, p) as five ghosted
C-order arrays of shape (n+4,)*3 -- the host layout of the reference
(grid.rs:79-109) -- with cell-centred coordinates on the unit cube,
x_i = (i - 2 + 1/2) ds.  Ghost cells hold values too (the library reads only
the interior).
"""
from __future__ import annotations

import numpy as np

ENSEMBLE_SIZES = (32, 48, 64, 96, 128)


def centres(n):
    ds = 1.0 / n
    return (np.arange(n + 4) - 2 + 0.5) * ds, ds


def khi(n, dtype=np.float32, seed=1):
    """Kelvin-Helmholtz shear layer along y: |y - 1/2| < 1/4 holds rho 2,
    u +0.5, elsewhere rho 1, u -0.5; p 2.5, w 0; v = 0.01 sin(4 pi x) plus
    seeded uniform noise in [-1e-3, 1e-3]."""
    c, ds = centres(n)
    shape = (n + 4,) * 3
    inner = (np.abs(c - 0.5) < 0.25)[None, :, None]
    rho = np.empty(shape, dtype)
    rho[...] = np.where(inner, 2.0, 1.0)
    u = np.empty(shape, dtype)
    u[...] = np.where(inner, 0.5, -0.5)
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1e-3, 1e-3, size=shape).astype(dtype)
    v += (0.01 * np.sin(4 * np.pi * c)).astype(dtype)[:, None, None]
    w = np.zeros(shape, dtype)
    p = np.full(shape, 2.5, dtype)
    return [rho, u, v, w, p], ds


def blast(n, dtype=np.float64, p_in=10.0, radius=0.1):
    """Spherical blast: rho 1, p 0.1 and zero velocity, p = p_in inside the
    sphere of the given radius about the centre of the cube."""
    c, ds = centres(n)
    r2 = (c[:, None, None] - 0.5) ** 2 + (c[None, :, None] - 0.5) ** 2 + (c[None, None, :] - 0.5) ** 2
    shape = (n + 4,) * 3
    rho = np.ones(shape, dtype)
    p = np.where(r2 < radius * radius, p_in, 0.1).astype(dtype)
    z = np.zeros(shape, dtype)
    return [rho, z, z.copy(), z.copy(), p], ds


def sod(n, dtype=np.float64):
    """Sod shock tube along x: (rho, p) = (1, 1) for x < 1/2, (0.125, 0.1)
    beyond, at rest."""
    c, ds = centres(n)
    shape = (n + 4,) * 3
    left = (c < 0.5)[:, None, None]
    rho = np.empty(shape, dtype)
    rho[...] = np.where(left, 1.0, 0.125)
    p = np.empty(shape, dtype)
    p[...] = np.where(left, 1.0, 0.1)
    z = np.zeros(shape, dtype)
    return [rho, z, z.copy(), z.copy(), p], ds


def planar_riemann(n, rng, dtype=np.float32):
    """Random planar Riemann problem: a unit normal uniform on the sphere, the
    plane through the centre of the cube; on each side rho, p ~ U[0.1, 1] and
    each velocity component ~ U[-1, 1].  Returns (fields, ds, description)."""
    c, ds = centres(n)
    normal = rng.normal(size=3)
    normal /= np.linalg.norm(normal)
    left = dict(rho=rng.uniform(0.1, 1.0), p=rng.uniform(0.1, 1.0), v=rng.uniform(-1.0, 1.0, size=3))
    right = dict(rho=rng.uniform(0.1, 1.0), p=rng.uniform(0.1, 1.0), v=rng.uniform(-1.0, 1.0, size=3))
    s = ((c[:, None, None] - 0.5) * normal[0] + (c[None, :, None] - 0.5) * normal[1]
         + (c[None, None, :] - 0.5) * normal[2])
    side = s < 0.0
    out = []
    for q in ("rho", "v0", "v1", "v2", "p"):
        if q.startswith("v"):
            a, b = left["v"][int(q[1])], right["v"][int(q[1])]
        else:
            a, b = left[q], right[q]
        out.append(np.where(side, a, b).astype(dtype))
    return out, ds, dict(normal=normal, left=left, right=right)


def ensemble_members(count=256, seed=0, sizes=ENSEMBLE_SIZES):
    """Edge lengths of the ensemble drawn by seed from `sizes`, and one
    independent generator per member (seeded by (seed, member))."""
    rng = np.random.default_rng(seed)
    edges = [int(x) for x in rng.choice(np.asarray(sizes), size=count)]
    return [(n, np.random.default_rng([seed, i])) for i, n in enumerate(edges)]


def ensemble_workload(count=256, seed=0, precision="fp32", math="fast", limiter="minmod", choice=None):
    """The batched-ensemble workload: `count` cubic members with edges drawn by
    seed, each a random planar Riemann problem on the unit cube with outflow
    boundaries, gamma 1.4, CFL 0.4.  `choice` (an autotuner entry) sets the
    tiling.  Returns (Simulations, state generator): the generator yields
    (index, primitive fields) one member at a time so the host never holds
    every state at once."""
    from . import OUTFLOW, Simulation
    members = ensemble_members(count, seed)
    dtype = np.float64 if precision == "fp64" else np.float32
    sims = []
    for n, _ in members:
        ty = int(choice["tile_y"]) if choice else 0
        lx = int(choice["chunk_x"].get(str(n), 0)) if choice else 0
        sims.append(Simulation(n, n, n, 1.0 / n, [OUTFLOW] * 6, gamma=1.4, cfl=0.4, g=0.0, precision=precision,
                               math=math, limiter=limiter, tile_y=ty, chunk_x=lx))

    def states():
        for i, (n, rng) in enumerate(members):
            yield i, planar_riemann(n, np.random.default_rng(rng.bit_generator.seed_seq), dtype)[0]

    return sims, states


def learned_ic_params(seed):
    """Random initial condition of the learned-correction workload: density 1
    plus 8 Fourier modes (integer wave vectors with components in [-4, 4],
    amplitudes in [-0.2, 0.2], random phases), pressure 1 plus 1-3 Gaussian
    blasts of peak 5-20 (width 0.05, periodic distance), at rest."""
    rng = np.random.default_rng(seed)
    modes = []
    while len(modes) < 8:
        k = rng.integers(-4, 5, size=3)
        if not k.any():
            continue
        modes.append((k.astype(np.float64), rng.uniform(-0.2, 0.2), rng.uniform(0.0, 2 * np.pi)))
    blasts = [(rng.uniform(0.0, 1.0, size=3), rng.uniform(5.0, 20.0)) for _ in range(int(rng.integers(1, 4)))]
    return modes, blasts, 0.05


def learned_ic(n, seed, dtype=np.float32, rho_min=0.2):
    """Primitive fields (ghosted, (n+4)^3) of learned_ic_params(seed) at the
    cell centres of an n^3 periodic unit cube.  Where the 8 modes would take
    the density below rho_min, the mode sum is scaled down uniformly so that
    its minimum is rho_min.  Modes and blasts are evaluated as outer products
    of per-axis factors (sin(a+b+c) = Im(e^ia e^ib e^ic); a Gaussian in the
    periodic distance factorizes per axis), a slab of planes at a time."""
    modes, blasts, sigma = learned_ic_params(seed)
    c, ds = centres(n)
    m = n + 4
    shape = (m,) * 3
    tw = 2 * np.pi
    ex = [np.exp(1j * (tw * k[0] * c + ph)) * a for k, a, ph in modes]
    ey = [np.exp(1j * tw * k[1] * c) for k, _, _ in modes]
    ez = [np.exp(1j * tw * k[2] * c) for k, _, _ in modes]

    def gauss(ctr):
        d = np.abs(c - ctr)
        d = np.minimum(d, 1.0 - d)
        return np.exp(-d * d / (2 * sigma * sigma))

    g = [(gauss(ctr[0]), gauss(ctr[1]), gauss(ctr[2]), peak) for ctr, peak in blasts]
    S = np.empty(shape, np.float64)
    p = np.ones(shape, np.float64)
    slab = 16
    for i0 in range(0, m, slab):
        i1 = min(i0 + slab, m)
        acc = np.zeros((i1 - i0, m, m))
        for q in range(len(modes)):
            acc += (ex[q][i0:i1, None, None] * (ey[q][:, None] * ez[q][None, :])[None]).imag
        S[i0:i1] = acc
        for gx, gy, gz, peak in g:
            p[i0:i1] += (peak - 1.0) * gx[i0:i1, None, None] * (gy[:, None] * gz[None, :])[None]
    smin = S.min()
    scale = 1.0 if 1.0 + smin >= rho_min else (1.0 - rho_min) / (-smin)
    rho = (1.0 + scale * S).astype(dtype)
    del S
    z = np.zeros(shape, dtype)
    return [rho, z, z.copy(), z.copy(), p.astype(dtype)], ds
