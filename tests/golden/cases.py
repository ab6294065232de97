"""Golden-vector cases: initial states and parameters.

Shared by make_golden.py (which runs the reference's Fortran physics on them)
and by the tests (which run the oracle and the GPU library on them).  Each
case builds ghosted primitive arrays (interior + 4 per axis, C order) with
the reference's convention of filling every cell, ghosts included
(init.rs:323-906 fill loops run over 0..nx).
"""
from __future__ import annotations

import numpy as np

PERIODIC, OUTFLOW, WALL = 0, 1, 2

# parameter sets compiled into the reference harness (oracle/Makefile)
PARAMS = {
    "cfg": dict(gamma=1.4, cfl=0.4, g=0.0, gravity=False, gravity_k_shift=0),
    # reference constants (global.f90:5); the Fortran z-flux always applies
    # gravity with a 1-based k in its height (physics.f90:135-141)
    "native": dict(gamma=5.0 / 3.0, cfl=0.2, g=-9.81, gravity=True, gravity_k_shift=1),
}


def ghosted_coords(interior, ds):
    """Ghost-inclusive coordinates x = i*ds, as init.rs:423-425, 743-748."""
    return [np.arange(n + 4) * ds for n in interior]


def sod(interior, axis=0):
    """Sod shock tube along `axis`, split at half the ghosted length
    (init.rs:735-755 convention).  Returns (w, ds)."""
    ds = 1.0 / interior[axis]
    shape = tuple(n + 4 for n in interior)
    c = np.arange(shape[axis]) * ds
    left = c < shape[axis] * ds / 2.0
    sl = [None, None, None]
    sl[axis] = slice(None)
    left = np.broadcast_to(left[tuple(sl)], shape)
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros(shape)
    return [rho, z.copy(), z.copy(), z.copy(), p], ds


def smooth_random(interior, seed, amp_v=0.3, p_base=1.0):
    """Smooth periodic state from a few seeded Fourier modes: exercises every
    term of the step (all velocity components, gradients along every axis)."""
    rng = np.random.default_rng(seed)
    ds = 1.0 / max(interior)
    x, y, z = np.meshgrid(*ghosted_coords(interior, ds), indexing="ij")
    L = [n * ds for n in interior]

    def field(base, amp):
        f = np.full(x.shape, base, dtype=np.float64)
        for _ in range(3):
            kx, ky, kz = rng.integers(1, 3, size=3)
            ph = rng.uniform(0, 2 * np.pi, size=3)
            a = rng.uniform(-amp, amp)
            f += a * (np.sin(2 * np.pi * kx * x / L[0] + ph[0]) * np.cos(2 * np.pi * ky * y / L[1] + ph[1])
                      + 0.5 * np.sin(2 * np.pi * kz * z / L[2] + ph[2]))
        return f

    w = [field(1.0, 0.15), field(0.0, amp_v), field(0.0, amp_v), field(0.0, amp_v), field(p_base, 0.15 * p_base)]
    return w, ds


def blast(interior, radius=0.1, p_in=10.0, p_out=0.1):
    """Pressure sphere at the domain centre (cell centres, interior coords)."""
    ds = 1.0 / interior[0]
    c = [(np.arange(n + 4) - 2 + 0.5) * ds for n in interior]
    x, y, z = np.meshgrid(*c, indexing="ij")
    r2 = (x - 0.5 * interior[0] * ds) ** 2 + (y - 0.5 * interior[1] * ds) ** 2 + (z - 0.5 * interior[2] * ds) ** 2
    p = np.where(r2 < radius * radius, p_in, p_out)
    rho = np.ones_like(p)
    z0 = np.zeros_like(p)
    return [rho, z0.copy(), z0.copy(), z0.copy(), p], ds


# name -> (builder, kwargs, interior, bcs, nsteps, params)
CASES = {
    "sod64": (sod, dict(axis=0), (64, 64, 64), [OUTFLOW] * 6, 100, "cfg"),
    "sod_y": (sod, dict(axis=1), (6, 24, 5), [OUTFLOW] * 6, 30, "cfg"),
    "sod_z": (sod, dict(axis=2), (5, 6, 24), [OUTFLOW] * 6, 30, "cfg"),
    "smooth_periodic": (smooth_random, dict(seed=7), (12, 10, 8), [PERIODIC] * 6, 25, "cfg"),
    # periodic faces come in pairs: the Fortran applies periodic maps only
    # when both faces of an axis are periodic (physics.f90:272-294)
    "smooth_mixed": (smooth_random, dict(seed=11), (10, 9, 12),
                     [OUTFLOW, OUTFLOW, PERIODIC, PERIODIC, OUTFLOW, OUTFLOW], 25, "cfg"),
    # pressure ratio 2: the unlimited reference scheme goes non-finite on
    # the ratio-100 blast of the benchmark config within two steps
    "blast16": (blast, dict(radius=0.25, p_in=2.0, p_out=1.0), (16, 16, 16), [PERIODIC] * 6, 20, "cfg"),
    # lower walls only: the reference Fortran leaves the gradient at the upper
    # wall ghost uncomputed (physics.f90:248-250), unlike the Rust
    # p ~ 100 keeps g = -9.81 (on a unit box) from driving p negative
    "smooth_lowerwalls_native": (smooth_random, dict(seed=3, amp_v=0.1, p_base=100.0), (9, 8, 10),
                                 [WALL, OUTFLOW, PERIODIC, PERIODIC, WALL, OUTFLOW], 20, "native"),
    "smooth_periodic_native": (smooth_random, dict(seed=5, p_base=100.0), (8, 10, 12), [PERIODIC] * 6, 20, "native"),
}


def build_case(name):
    fn, kw, interior, bcs, nsteps, params = CASES[name]
    w, ds = fn(interior, **kw)
    return dict(w=w, ds=ds, interior=interior, bcs=list(bcs), nsteps=nsteps, params=params)
