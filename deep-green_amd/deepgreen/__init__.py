"""deepgreen -- Python host for the MI355X compressible-Euler step library.

Mirrors the reference Rust host's interface (src/rs/src/grid.rs, physics.rs,
main.rs): a `Simulation` describes the grid (sizes without ghosts, ds, six
boundary types, gravity) and a `Solver` owns the device state and offers
`compute_step` (physics.rs:595) and `compute` (main.rs:38-68).  All work is
done by the HIP library libdeepgreen.so through its C ABI
(include/deepgreen.h); there is no CPU fallback: importing raises if the
library is missing.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libdeepgreen.so")

PERIODIC, OUTFLOW, WALL = 0, 1, 2          # grid.rs:24-28
CONDUCTING_WALL = 3                        # src/cu/include/constants.h:44
FP64, FP32, FP32_BULK = 0, 1, 2
EXACT, FAST = 0, 1
LIMITERS = {"none": 0, "minmod": 1, "vanleer": 2}
PRIMITIVE, CONSERVED = 0, 1
PRECISIONS = {"fp64": FP64, "fp32": FP32, "fp32_bulk": FP32_BULK}
MATHS = {"exact": EXACT, "fast": FAST}

STATUS = {0: "ok", -1: "invalid argument", -2: "device error", -3: "allocation failed",
          -4: "non-finite dt", -5: "no state"}

# every entry point declared in include/deepgreen.h
EXPORTS = ("dg_config_default", "dg_create", "dg_destroy", "dg_last_error", "dg_set_state",
           "dg_get_state", "dg_step", "dg_step_dump", "dg_step_noalloc", "dg_run", "dg_reset_clock", "dg_conservation_sums",
           "dg_next_dt", "dg_device_state", "dg_time_steps", "dg_kernel_info", "dg_device_count", "dg_set_heat",
           "dg_device_info",
           "dg_ensemble_create", "dg_ensemble_destroy", "dg_ensemble_last_error", "dg_ensemble_set_state",
           "dg_ensemble_get_state", "dg_ensemble_run", "dg_ensemble_clock", "dg_ensemble_conservation_sums",
           "dg_ensemble_time_steps", "dg_ensemble_retile", "dg_ensemble_info",
           "dg_step_dt", "dg_restrict",
           "dg_cnn_create", "dg_cnn_destroy", "dg_cnn_last_error", "dg_cnn_param_count", "dg_cnn_get_params",
           "dg_cnn_set_params", "dg_cnn_get_normalization", "dg_cnn_set_normalization", "dg_cnn_forward",
           "dg_cnn_backward", "dg_cnn_train_batch", "dg_cnn_adam_step", "dg_cnn_add_sample", "dg_cnn_add_sample_mode", "dg_cnn_sample_count", "dg_cnn_set_input_noise",
           "dg_cnn_fit_normalization", "dg_cnn_train", "dg_learned_step", "dg_learned_run", "dg_learned_flagged",
           "dg_learned_correction", "dg_learned_step_given", "dg_learned_time")


class DgConfig(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64 * 3), ("nghosts", ctypes.c_int32), ("bc", ctypes.c_int32 * 6),
        ("ds", ctypes.c_double), ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
        ("g", ctypes.c_double), ("gravity", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("math", ctypes.c_int32), ("limiter", ctypes.c_int32), ("dt_nan_fallback", ctypes.c_int32),
        ("device", ctypes.c_int32), ("tile_y", ctypes.c_int32), ("chunk_x", ctypes.c_int32),
    ]


class DeepGreenError(RuntimeError):
    pass


_lib = None
_other = {}


def load_library(path=None):
    """Load libdeepgreen.so from the package directory (built by
    `make -C deep-green_amd` or __graft_entry__.build()).  `path` loads
    another build of the same C ABI (the CPU-emulation build of the tests)
    and returns its handle without making it the default: it is used only by
    a Solver that is given it explicitly."""
    global _lib
    if path is None and _lib is not None:
        return _lib
    if path is not None and path in _other:
        return _other[path]
    target = path or LIB_PATH
    if not os.path.exists(target):
        raise DeepGreenError(f"{target} is missing: build it with `make -C deep-green_amd`")
    L = ctypes.CDLL(target)
    vp, i32, i64, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
    P = ctypes.POINTER
    L.dg_config_default.argtypes = [P(DgConfig)]
    L.dg_config_default.restype = None
    L.dg_create.argtypes = [P(DgConfig), P(vp)]
    L.dg_destroy.argtypes = [vp]
    L.dg_destroy.restype = None
    L.dg_last_error.argtypes = [vp]
    L.dg_last_error.restype = ctypes.c_char_p
    L.dg_set_state.argtypes = [vp, P(vp), i32]
    L.dg_get_state.argtypes = [vp, P(vp), i32]
    L.dg_step.argtypes = [vp, P(d)]
    L.dg_step_noalloc.argtypes = [P(DgConfig), P(vp), P(vp), vp, P(d)]
    L.dg_step_dump.argtypes = [vp, P(d), P(vp), P(vp), P(vp)]
    L.dg_run.argtypes = [vp, i64, d, P(d), P(i64)]
    L.dg_reset_clock.argtypes = [vp]
    L.dg_conservation_sums.argtypes = [vp, P(d)]
    L.dg_next_dt.argtypes = [vp, P(d)]
    L.dg_device_state.argtypes = [vp, P(vp)]
    L.dg_time_steps.argtypes = [vp, i64, P(d), P(d), P(d)]
    L.dg_kernel_info.argtypes = [vp, P(i32), P(i32), P(i64), P(i32), P(i32)]
    L.dg_device_count.argtypes = [P(i32)]
    L.dg_set_heat.argtypes = [vp, d, d, d]
    L.dg_ensemble_create.argtypes = [P(DgConfig), i32, i32, P(vp)]
    L.dg_ensemble_destroy.argtypes = [vp]
    L.dg_ensemble_destroy.restype = None
    L.dg_ensemble_last_error.argtypes = [vp]
    L.dg_ensemble_last_error.restype = ctypes.c_char_p
    L.dg_ensemble_set_state.argtypes = [vp, i32, P(vp), i32]
    L.dg_ensemble_get_state.argtypes = [vp, i32, P(vp), i32]
    L.dg_ensemble_run.argtypes = [vp, i64, d]
    L.dg_ensemble_clock.argtypes = [vp, i32, P(d), P(i64), P(d)]
    L.dg_ensemble_conservation_sums.argtypes = [vp, i32, P(d)]
    L.dg_ensemble_time_steps.argtypes = [vp, i64, P(d), P(d)]
    L.dg_ensemble_retile.argtypes = [vp, i32, P(ctypes.c_int32), i32, i32]
    L.dg_ensemble_info.argtypes = [vp, P(i64), P(i32), P(i32), P(i32)]
    L.dg_device_info.argtypes = [i32, ctypes.c_char_p, i32, P(i32)]
    fp = P(ctypes.c_float)
    L.dg_step_dt.argtypes = [vp, d]
    L.dg_restrict.argtypes = [vp, vp]
    L.dg_cnn_create.argtypes = [i32, ctypes.c_uint64, P(vp)]
    L.dg_cnn_destroy.argtypes = [vp]
    L.dg_cnn_destroy.restype = None
    L.dg_cnn_last_error.argtypes = [vp]
    L.dg_cnn_last_error.restype = ctypes.c_char_p
    L.dg_cnn_param_count.argtypes = []
    L.dg_cnn_param_count.restype = ctypes.c_int64
    L.dg_cnn_get_params.argtypes = [vp, fp]
    L.dg_cnn_set_params.argtypes = [vp, fp]
    L.dg_cnn_get_normalization.argtypes = [vp, fp]
    L.dg_cnn_set_normalization.argtypes = [vp, fp]
    L.dg_cnn_forward.argtypes = [vp, i32, P(ctypes.c_int32), fp, fp]
    L.dg_cnn_backward.argtypes = [vp, i32, P(ctypes.c_int32), fp, fp, fp]
    L.dg_cnn_train_batch.argtypes = [vp, i32, i32, fp, fp, P(d), fp]
    L.dg_cnn_adam_step.argtypes = [vp, fp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.dg_cnn_add_sample.argtypes = [vp, vp, vp]
    L.dg_cnn_add_sample_mode.argtypes = [vp, vp, vp, i32, i32]
    L.dg_cnn_sample_count.argtypes = [vp]
    L.dg_cnn_set_input_noise.argtypes = [vp, ctypes.c_float]
    L.dg_cnn_fit_normalization.argtypes = [vp]
    L.dg_cnn_train.argtypes = [vp, i32, i32, i32, ctypes.c_float, ctypes.c_uint64, fp]
    L.dg_learned_step.argtypes = [vp, vp, d, P(d)]
    L.dg_learned_run.argtypes = [vp, vp, i64, d, P(d), P(i64)]
    L.dg_learned_flagged.argtypes = [vp, P(i64)]
    L.dg_learned_correction.argtypes = [vp, vp, fp]
    L.dg_learned_step_given.argtypes = [vp, vp, fp, P(d)]
    L.dg_learned_time.argtypes = [vp, vp, i32, P(d)]
    for name in EXPORTS:
        if name not in ("dg_config_default", "dg_destroy", "dg_last_error", "dg_ensemble_destroy",
                        "dg_ensemble_last_error", "dg_cnn_destroy", "dg_cnn_last_error", "dg_cnn_param_count"):
            getattr(L, name).restype = ctypes.c_int
    if path is None:
        _lib = L
    else:
        _other[path] = L
    return L


class Simulation:
    """Simulation (grid.rs:31-76): interior sizes, ds, boundary types in the
    order x-, x+, y-, y+, z-, z+, gravity flag, and the reference constants
    (globals.rs) as parameters."""

    def __init__(self, nx, ny, nz, ds, bcs, gravity=False, gamma=5.0 / 3.0, cfl=0.2, g=-0.1,
                 precision="fp32", math="exact", limiter="none", dt_nan_fallback=True,
                 device=0, tile_y=0, chunk_x=0):
        self.interior = (int(nx), int(ny), int(nz))
        self.nghosts = 2
        self.shape = tuple(n + 2 * self.nghosts for n in self.interior)  # grid.rs:64-66
        self.ds = float(ds)
        self.bcs = [int(b) for b in bcs]
        if len(self.bcs) != 6:
            raise ValueError("six boundary types expected")
        self.gravity = bool(gravity)
        self.gamma, self.cfl, self.g = float(gamma), float(cfl), float(g)
        self.precision = precision
        self.math = math
        self.limiter = limiter
        self.dt_nan_fallback = bool(dt_nan_fallback)
        self.device = int(device)
        self.tile_y = int(tile_y)
        self.chunk_x = int(chunk_x)

    @property
    def dtype(self):
        return np.float64 if self.precision == "fp64" else np.float32

    def config(self, lib=None):
        c = DgConfig()
        (lib or load_library()).dg_config_default(ctypes.byref(c))
        c.n[:] = list(self.interior)
        c.nghosts = 2
        c.bc[:] = self.bcs
        c.ds, c.gamma, c.cfl, c.g = self.ds, self.gamma, self.cfl, self.g
        c.gravity = int(self.gravity)
        c.precision = PRECISIONS[self.precision]
        c.math = MATHS[self.math]
        c.limiter = LIMITERS[self.limiter]
        c.dt_nan_fallback = int(self.dt_nan_fallback)
        c.device = self.device
        c.tile_y = self.tile_y
        c.chunk_x = self.chunk_x
        return c


class Solver:
    """Device-resident grid of one Simulation (SimulationGrid, grid.rs:111-118,
    without the 40 scratch arrays the fused kernel does not need)."""

    def __init__(self, sim: Simulation, lib=None):
        self.sim = sim
        self._L = lib or load_library()
        self._h = ctypes.c_void_p()
        cfg = sim.config(self._L)
        rc = self._L.dg_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if rc != 0:
            raise DeepGreenError(f"dg_create failed: {STATUS.get(rc, rc)}")

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.dg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.dg_last_error(self._h)
            raise DeepGreenError(f"{what}: {STATUS.get(rc, rc)} ({msg.decode() if msg else ''})")

    def _ptrs(self, arrays, write=False):
        dt = self.sim.dtype
        out = []
        for a in arrays:
            if a.shape != self.sim.shape or a.dtype != dt or not a.flags.c_contiguous:
                raise ValueError(f"expected C-contiguous {dt.__name__} arrays of shape {self.sim.shape}")
            if write and not a.flags.writeable:
                raise ValueError("output arrays must be writeable")
            out.append(a.ctypes.data)
        return (ctypes.c_void_p * 5)(*out)

    def set_state(self, fields, kind=PRIMITIVE):
        arrs = [np.ascontiguousarray(f, dtype=self.sim.dtype) for f in fields]
        self._check(self._L.dg_set_state(self._h, self._ptrs(arrs), int(kind)), "dg_set_state")

    def set_primitives(self, w):
        self.set_state(w, PRIMITIVE)

    def set_conserved(self, u):
        self.set_state(u, CONSERVED)

    def get_state(self, kind=CONSERVED):
        arrs = [np.empty(self.sim.shape, dtype=self.sim.dtype) for _ in range(5)]
        self._check(self._L.dg_get_state(self._h, self._ptrs(arrs, True), int(kind)), "dg_get_state")
        return arrs

    def compute_step(self):
        """physics::compute_step: one step; returns the dt it used."""
        dt = ctypes.c_double()
        self._check(self._L.dg_step(self._h, ctypes.byref(dt)), "dg_step")
        return dt.value

    def compute_step_dump(self):
        """One step that also returns the reference grid's other dumped
        arrays (dg_step_dump): (dt, primitives of the step's start [5 arrays
        of the ghosted shape], gradients and fluxes [5 arrays of the ghosted
        shape + (3,)])."""
        dt = ctypes.c_double()
        sh = self.sim.shape
        prim = [np.empty(sh, dtype=self.sim.dtype) for _ in range(5)]
        grad = [np.empty(sh + (3,), dtype=self.sim.dtype) for _ in range(5)]
        flux = [np.empty(sh + (3,), dtype=self.sim.dtype) for _ in range(5)]
        ptr = lambda arrs: (ctypes.c_void_p * 5)(*[a.ctypes.data for a in arrs])  # noqa: E731
        self._check(self._L.dg_step_dump(self._h, ctypes.byref(dt), ptr(prim), ptr(grad), ptr(flux)),
                    "dg_step_dump")
        return dt.value, prim, grad, flux

    def set_heat(self, t_out_kelvin=273.0, v_out=0.0, rs=287.0):
        """Constants of the conducting-wall heat source (dg_set_heat)."""
        self._check(self._L.dg_set_heat(self._h, float(t_out_kelvin), float(v_out), float(rs)), "dg_set_heat")

    def step_dt(self, dt):
        """One step with the given dt (the next step uses the CFL dt again)."""
        self._check(self._L.dg_step_dt(self._h, float(dt)), "dg_step_dt")

    def restrict_from(self, fine):
        """This (coarse) grid's state := the 2x restriction of `fine`'s state."""
        self._check(self._L.dg_restrict(fine._h, self._h), "dg_restrict")

    @property
    def handle(self):
        return self._h

    def run(self, max_steps, t_end=math.inf):
        t = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self._L.dg_run(self._h, int(max_steps), float(t_end), ctypes.byref(t), ctypes.byref(n)),
                    "dg_run")
        return t.value, n.value

    def compute(self, total_time, max_steps):
        """main.rs compute() loop without output (the state must be set)."""
        self._check(self._L.dg_reset_clock(self._h), "dg_reset_clock")
        return self.run(max_steps, total_time)

    def reset_clock(self):
        self._check(self._L.dg_reset_clock(self._h), "dg_reset_clock")

    def conservation_sums(self):
        out = (ctypes.c_double * 5)()
        self._check(self._L.dg_conservation_sums(self._h, out), "dg_conservation_sums")
        return np.array(out[:])

    def next_dt(self):
        dt = ctypes.c_double()
        self._check(self._L.dg_next_dt(self._h, ctypes.byref(dt)), "dg_next_dt")
        return dt.value

    def device_pointers(self):
        out = (ctypes.c_void_p * 5)()
        self._check(self._L.dg_device_state(self._h, out), "dg_device_state")
        return [int(p) for p in out]

    def time_steps(self, nsteps):
        """(ms total, mean ms of the fused step kernel, mean ms of the
        reduction kernel) over nsteps steps, from HIP events on the
        library's stream."""
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self._L.dg_time_steps(self._h, int(nsteps), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
                    "dg_time_steps")
        return a.value, b.value, c.value

    def kernel_info(self):
        ty, lx, nb, lds, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.dg_kernel_info(self._h, ctypes.byref(ty), ctypes.byref(lx), ctypes.byref(nb),
                                           ctypes.byref(lds), ctypes.byref(nt)), "dg_kernel_info")
        return dict(tile_y=ty.value, chunk_x=lx.value, workgroups=nb.value, lds_bytes=lds.value,
                    threads=nt.value)

    def interior(self, a):
        return a[2:-2, 2:-2, 2:-2]


def _field_ptrs(sim, arrays, write=False):
    dt = sim.dtype
    out = []
    for a in arrays:
        if a.shape != sim.shape or a.dtype != dt or not a.flags.c_contiguous:
            raise ValueError(f"expected C-contiguous {dt.__name__} arrays of shape {sim.shape}")
        if write and not a.flags.writeable:
            raise ValueError("output arrays must be writeable")
        out.append(a.ctypes.data)
    return (ctypes.c_void_p * 5)(*out)


class Ensemble:
    """Many independent grids advanced together by one persistent step
    launch per step (dg_ensemble_*).  `sims` share precision, math, limiter,
    device and tile_y; everything else is per instance.  Each instance
    follows exactly the trajectory a Solver of the same Simulation follows."""

    def __init__(self, sims, grid=0, lib=None):
        self.sims = list(sims)
        self._L = lib or load_library()
        self._h = ctypes.c_void_p()
        cfgs = (DgConfig * len(self.sims))()
        for i, s in enumerate(self.sims):
            cfgs[i] = s.config(self._L)
        rc = self._L.dg_ensemble_create(cfgs, len(self.sims), int(grid), ctypes.byref(self._h))
        if rc != 0:
            raise DeepGreenError(f"dg_ensemble_create failed: {STATUS.get(rc, rc)}")

    def __len__(self):
        return len(self.sims)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.dg_ensemble_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            msg = self._L.dg_ensemble_last_error(self._h)
            raise DeepGreenError(f"{what}: {STATUS.get(rc, rc)} ({msg.decode() if msg else ''})")

    def set_state(self, i, fields, kind=PRIMITIVE):
        sim = self.sims[i]
        arrs = [np.ascontiguousarray(f, dtype=sim.dtype) for f in fields]
        self._check(self._L.dg_ensemble_set_state(self._h, int(i), _field_ptrs(sim, arrs), int(kind)),
                    f"dg_ensemble_set_state({i})")

    def get_state(self, i, kind=CONSERVED):
        sim = self.sims[i]
        arrs = [np.empty(sim.shape, dtype=sim.dtype) for _ in range(5)]
        self._check(self._L.dg_ensemble_get_state(self._h, int(i), _field_ptrs(sim, arrs, True), int(kind)),
                    f"dg_ensemble_get_state({i})")
        return arrs

    def run(self, max_steps, t_end=math.inf):
        self._check(self._L.dg_ensemble_run(self._h, int(max_steps), float(t_end)), "dg_ensemble_run")

    def clock(self, i):
        """(t, steps, next dt) of instance i."""
        t, n, dt = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        self._check(self._L.dg_ensemble_clock(self._h, int(i), ctypes.byref(t), ctypes.byref(n), ctypes.byref(dt)),
                    "dg_ensemble_clock")
        return t.value, n.value, dt.value

    def conservation_sums(self, i):
        out = (ctypes.c_double * 5)()
        self._check(self._L.dg_ensemble_conservation_sums(self._h, int(i), out), "dg_ensemble_conservation_sums")
        return np.array(out[:])

    def time_steps(self, nsteps):
        """(mean ms of the persistent step kernel, mean ms of the dt kernel)
        over nsteps steps, from HIP events on the ensemble's stream."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self._L.dg_ensemble_time_steps(self._h, int(nsteps), ctypes.byref(a), ctypes.byref(b)),
                    "dg_ensemble_time_steps")
        return a.value, b.value

    def retile(self, tile_y=0, chunk_x=None, grid=0, waves=0):
        """New tiling (results unchanged): tile_y for all (0 keeps it),
        chunk_x per instance (None keeps them, 0 automatic), grid persistent
        workgroups (0: device maximum), waves per SIMD of the FP32 fast
        kernel (3 or 4; 0 keeps it)."""
        arr = None
        if chunk_x is not None:
            arr = (ctypes.c_int32 * len(self.sims))(*[int(c) for c in chunk_x])
        self._check(self._L.dg_ensemble_retile(self._h, int(tile_y), arr, int(grid), int(waves)),
                    "dg_ensemble_retile")

    def info(self):
        cells, items, grid, ty = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.dg_ensemble_info(self._h, ctypes.byref(cells), ctypes.byref(items), ctypes.byref(grid),
                                             ctypes.byref(ty)), "dg_ensemble_info")
        return dict(cells=cells.value, items=items.value, grid=grid.value, tile_y=ty.value)


def device_count():
    n = ctypes.c_int(0)
    load_library().dg_device_count(ctypes.byref(n))
    return n.value


def device_info(device=0, lib=None):
    """(architecture name such as "gfx950", compute units) of a HIP device."""
    buf = ctypes.create_string_buffer(64)
    cus = ctypes.c_int(0)
    rc = (lib or load_library()).dg_device_info(int(device), buf, 64, ctypes.byref(cus))
    if rc != 0:
        raise DeepGreenError(f"dg_device_info: {STATUS.get(rc, rc)}")
    return buf.value.decode(), cus.value
