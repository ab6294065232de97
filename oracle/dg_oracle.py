"""ctypes view of the CPU oracle (oracle/libdg_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker, never as the path under
test.  The oracle restates the reference Rust step (src/rs/src/physics.rs,
boundaries.rs, grid.rs, main.rs); see dg_oracle.c for what pins it.

Also wraps the reference-Fortran harness binaries in oracle/_ref/ (built by
`make -C oracle ref` where the reference sources exist).
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdg_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

PERIODIC, OUTFLOW, WALL = 0, 1, 2  # grid.rs:24-28
LIMITERS = {"none": 0, "minmod": 1, "vanleer": 2}


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("nghosts", ctypes.c_int32), ("bcs", ctypes.c_int32 * 6),
        ("ds", ctypes.c_double), ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
        ("g", ctypes.c_double), ("gravity", ctypes.c_int32),
        ("gravity_k_shift", ctypes.c_int32), ("dt_nan_fallback", ctypes.c_int32),
        ("nthreads", ctypes.c_int32), ("ref_threading", ctypes.c_int32),
        ("limiter", ctypes.c_int32), ("heat_mask", ctypes.c_int32),
        ("t_out", ctypes.c_double), ("v_out", ctypes.c_double), ("rs", ctypes.c_double),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        assert _lib.dgo_params_size() == ctypes.sizeof(_Params)
        for sfx in ("f32", "f64"):
            getattr(_lib, f"dgo_new_{sfx}").restype = ctypes.c_void_p
            getattr(_lib, f"dgo_new_{sfx}").argtypes = [ctypes.POINTER(_Params)]
            getattr(_lib, f"dgo_free_{sfx}").argtypes = [ctypes.c_void_p]
            getattr(_lib, f"dgo_field_{sfx}").restype = ctypes.c_void_p
            getattr(_lib, f"dgo_field_{sfx}").argtypes = [ctypes.c_void_p, ctypes.c_int]
            rt = ctypes.c_float if sfx == "f32" else ctypes.c_double
            getattr(_lib, f"dgo_calc_timestep_{sfx}").restype = rt
            getattr(_lib, f"dgo_calc_timestep_{sfx}").argtypes = [ctypes.c_void_p]
            getattr(_lib, f"dgo_extrapolate_{sfx}").argtypes = [ctypes.c_void_p, rt]
            getattr(_lib, f"dgo_add_fluxes_{sfx}").argtypes = [ctypes.c_void_p, rt]
            getattr(_lib, f"dgo_compute_step_{sfx}").argtypes = [ctypes.c_void_p, ctypes.POINTER(rt)]
            getattr(_lib, f"dgo_compute_{sfx}").restype = ctypes.c_int64
            getattr(_lib, f"dgo_compute_{sfx}").argtypes = [
                ctypes.c_void_p, ctypes.c_double, ctypes.c_int64,
                ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
            getattr(_lib, f"dgo_steps_{sfx}").restype = ctypes.c_int64
            getattr(_lib, f"dgo_steps_{sfx}").argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
            getattr(_lib, f"dgo_limited_{sfx}").restype = rt
            getattr(_lib, f"dgo_limited_{sfx}").argtypes = [rt, rt, ctypes.c_int]
            for nm in ("calc_primitives", "calc_conserved", "calc_gradients", "calc_fluxes",
                       "update_primitive_ghost_cells", "update_primitive_xtr_ghost_cells",
                       "update_gradients_ghost_cells", "update_fluxes_ghost_cells"):
                getattr(_lib, f"dgo_{nm}_{sfx}").argtypes = [ctypes.c_void_p]
    return _lib


class Oracle:
    """One reference-layout grid (grid.rs SimulationGrid) driven by the oracle.

    `interior` is the interior cell count per axis; arrays are ghosted
    (interior + 4), C order, as in the reference.
    """

    def __init__(self, interior, ds, bcs, gamma=1.4, cfl=0.4, g=0.0, gravity=False,
                 precision="f64", gravity_k_shift=0, dt_nan_fallback=True,
                 nthreads=1, ref_threading=False, limiter="none", t_out=273.0, v_out=0.0, rs=287.0):
        self.sfx = {"f32": "f32", "f64": "f64", "float32": "f32", "float64": "f64"}[precision]
        self.dtype = np.float32 if self.sfx == "f32" else np.float64
        self.interior = tuple(int(v) for v in interior)
        self.shape = tuple(v + 4 for v in self.interior)
        p = _Params()
        p.nx, p.ny, p.nz = self.shape
        p.nghosts = 2
        # a conducting wall (3) is a wall for the flow; lower ones carry the
        # heat source (kernels.h:561-661)
        for i, b in enumerate(bcs):
            p.bcs[i] = 2 if int(b) == 3 else int(b)
        p.heat_mask = sum(1 << (i // 2) for i, b in enumerate(bcs) if int(b) == 3 and i % 2 == 0)
        p.t_out, p.v_out, p.rs = float(t_out), float(v_out), float(rs)
        p.ds, p.gamma, p.cfl, p.g = float(ds), float(gamma), float(cfl), float(g)
        p.gravity = int(bool(gravity))
        p.gravity_k_shift = int(gravity_k_shift)
        p.dt_nan_fallback = int(bool(dt_nan_fallback))
        p.nthreads = int(nthreads)
        p.ref_threading = int(bool(ref_threading))
        p.limiter = LIMITERS[limiter]
        self._p = p
        self._L = lib()
        self._h = getattr(self._L, f"dgo_new_{self.sfx}")(ctypes.byref(p))
        if not self._h:
            raise MemoryError("oracle allocation failed")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            getattr(self._L, f"dgo_free_{self.sfx}")(h)
            self._h = None

    def _view(self, which, four=False):
        ptr = getattr(self._L, f"dgo_field_{self.sfx}")(self._h, which)
        shape = self.shape + ((3,) if four else ())
        n = int(np.prod(shape))
        ct = ctypes.c_float if self.sfx == "f32" else ctypes.c_double
        buf = (ct * n).from_address(ptr)
        return np.frombuffer(buf, dtype=self.dtype).reshape(shape)

    @property
    def prim(self):
        return [self._view(i) for i in range(5)]

    @property
    def xtr(self):
        return [self._view(5 + i) for i in range(5)]

    @property
    def cons(self):
        return [self._view(10 + i) for i in range(5)]

    @property
    def grad(self):
        return [self._view(15 + i, True) for i in range(5)]

    @property
    def flux(self):
        return [self._view(20 + i, True) for i in range(5)]

    def _call(self, name, *args):
        return getattr(self._L, f"dgo_{name}_{self.sfx}")(self._h, *args)

    # reference-named passes -------------------------------------------
    def set_primitives(self, w):
        for dst, src in zip(self.prim, w):
            dst[...] = np.asarray(src, dtype=self.dtype)

    def set_conserved(self, u):
        for dst, src in zip(self.cons, u):
            dst[...] = np.asarray(src, dtype=self.dtype)

    def calc_conserved(self):
        self._call("calc_conserved")

    def calc_primitives(self):
        self._call("calc_primitives")

    def update_primitive_ghost_cells(self):
        self._call("update_primitive_ghost_cells")

    def calc_timestep(self):
        return float(self._call("calc_timestep"))

    def calc_gradients(self):
        self._call("calc_gradients")

    def update_gradients_ghost_cells(self):
        self._call("update_gradients_ghost_cells")

    def extrapolate(self, dt):
        self._call("extrapolate", dt)

    def update_primitive_xtr_ghost_cells(self):
        self._call("update_primitive_xtr_ghost_cells")

    def calc_fluxes(self):
        self._call("calc_fluxes")

    def update_fluxes_ghost_cells(self):
        self._call("update_fluxes_ghost_cells")

    def add_fluxes(self, dt):
        self._call("add_fluxes", dt)

    def compute_step(self):
        ct = ctypes.c_float if self.sfx == "f32" else ctypes.c_double
        dt = ct(0.5)
        self._call("compute_step", ctypes.byref(dt))
        return float(dt.value)

    def steps(self, n):
        """n compute_step calls on the current conserved state; returns dts."""
        dts = np.zeros(max(int(n), 1), dtype=self.dtype)
        self._call("steps", int(n), dts.ctypes.data)
        return dts[: int(n)]

    def compute(self, total_time, max_steps):
        """main.rs compute(): calc_conserved, then the stepping loop."""
        dts = np.zeros(max(int(max_steps), 1), dtype=self.dtype)
        t = ctypes.c_double(0.0)
        n = self._call("compute", float(total_time), int(max_steps), ctypes.byref(t), dts.ctypes.data)
        return int(n), float(t.value), dts[:n]

    def interior_view(self, a):
        return a[2:-2, 2:-2, 2:-2]


# ------------------------------------------------------------ Fortran ref --
def ref_binary(precision="dp", params="cfg"):
    path = os.path.join(REF_DIR, f"dg_ref_{precision}_{params}")
    return path if os.path.exists(path) else None


def run_reference_fortran(w, ds, bcs, nsteps, precision="dp", params="cfg", nthreads=1):
    """Run the reference Fortran physics (via oracle/_ref harness) from the
    ghosted primitives `w` (5 C-order arrays); returns (U[5] C-order float64,
    dts, wall_seconds)."""
    exe = ref_binary(precision, params)
    if exe is None:
        raise FileNotFoundError("oracle/_ref binaries are not built")
    shape = np.asarray(w[0]).shape
    interior = [s - 4 for s in shape]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<4i", *interior, int(nsteps)))
            f.write(struct.pack("<6i", *[int(b) for b in bcs]))
            f.write(struct.pack("<i", int(nthreads)))
            f.write(struct.pack("<d", float(ds)))
            for a in w:
                f.write(np.asarray(a, dtype=np.float64).ravel(order="F").tobytes())
        subprocess.run([exe, fin, fout], check=True)
        raw = np.fromfile(fout, dtype=np.float64)
    nd = max(int(nsteps), 1)
    dts, wall = raw[:nd][: int(nsteps)], float(raw[nd])
    n3 = int(np.prod(shape))
    fields = [raw[nd + 1 + i * n3: nd + 1 + (i + 1) * n3].reshape(shape, order="F") for i in range(5)]
    return [np.ascontiguousarray(f) for f in fields], dts, wall


def limited(a, b, limiter, precision="f64"):
    """The oracle's limited slope of each (backward, forward) difference pair
    (a[i], b[i]), in the given precision."""
    L = lib()
    fn = getattr(L, f"dgo_limited_{precision}")
    dt = np.float32 if precision == "f32" else np.float64
    lim = LIMITERS[limiter]
    return np.array([fn(float(x), float(y), lim) for x, y in zip(np.asarray(a, dt), np.asarray(b, dt))], dtype=dt)


def run_reference_limiters(a, b, precision="dp"):
    """The reference Fortran module's minmod, maxmod and vanleerlimiter
    (physics.f90:205-238) of each pair (a[i], b[i]), via oracle/_ref/dg_lim_*;
    returns three float64 arrays."""
    exe = os.path.join(REF_DIR, f"dg_lim_{precision}")
    if not os.path.exists(exe):
        raise FileNotFoundError("oracle/_ref binaries are not built")
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<i", a.size))
            f.write(a.tobytes())
            f.write(b.tobytes())
        subprocess.run([exe, fin, fout], check=True)
        raw = np.fromfile(fout, dtype=np.float64)
    n = a.size
    return raw[:n], raw[n:2 * n], raw[2 * n:3 * n]
