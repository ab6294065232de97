"""Learned conservative flux correction (dg_cnn_* / dg_learned_* of the C ABI).

`FluxCNN` owns the network (3D CNN 5 -> 32 -> 32 -> 15, 3x3x3 kernels, GELU;
bf16 MFMA kernels with fp32 accumulation, fp32 master weights and Adam on the
device), its normalization and its training samples.  Samples come from the
solver: a fine grid restricted 2x to the coarse grid, one coarse step against
two fine steps.  `step` / `run` advance a periodic FP32 coarse Solver with the
correction added to its face fluxes (mass, momentum and energy totals are
unchanged up to rounding; cells whose corrected density or pressure would not
be positive keep the plain flux on all their faces).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import STATUS, DeepGreenError, load_library

# parameter blocks (name, shape) in the order of the flat parameter vector;
# weights are [tap][cin][cout] with tap = (dx*3 + dy)*3 + dz
LAYOUT = (("W1", (27, 5, 32)), ("b1", (32,)), ("W2", (27, 32, 32)), ("b2", (32,)),
          ("W3", (27, 32, 15)), ("b3", (15,)))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def split_params(flat):
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape)
        o += n
    return out


def join_params(blocks):
    return np.concatenate([np.asarray(blocks[name], np.float32).reshape(-1) for name, _ in LAYOUT])


class FluxCNN:
    def __init__(self, seed=0, device=0, lib=None):
        self._L = lib or load_library()
        self._h = ctypes.c_void_p()
        rc = self._L.dg_cnn_create(int(device), int(seed), ctypes.byref(self._h))
        if rc != 0:
            raise DeepGreenError(f"dg_cnn_create failed: {STATUS.get(rc, rc)}")
        self.nparams = int(self._L.dg_cnn_param_count())

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._L.dg_cnn_destroy(self._h)
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
            msg = self._L.dg_cnn_last_error(self._h)
            raise DeepGreenError(f"{what}: {STATUS.get(rc, rc)} ({msg.decode() if msg else ''})")

    # parameters -----------------------------------------------------------
    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        self._check(self._L.dg_cnn_get_params(self._h, _fp(p)), "dg_cnn_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        self._check(self._L.dg_cnn_set_params(self._h, _fp(p)), "dg_cnn_set_params")

    def get_normalization(self):
        n = np.empty(15, np.float32)
        self._check(self._L.dg_cnn_get_normalization(self._h, _fp(n)), "dg_cnn_get_normalization")
        return n

    def set_normalization(self, norm):
        n = np.ascontiguousarray(norm, np.float32)
        self._check(self._L.dg_cnn_set_normalization(self._h, _fp(n)), "dg_cnn_set_normalization")

    # network on host arrays (tests) -----------------------------------------
    def forward(self, x):
        """x [B, D0, D1, D2, 5] (normalized) -> y [B, D0-6, D1-6, D2-6, 15]."""
        x = np.ascontiguousarray(x, np.float32)
        B, dims = x.shape[0], (ctypes.c_int32 * 3)(*x.shape[1:4])
        y = np.empty((B,) + tuple(d - 6 for d in x.shape[1:4]) + (15,), np.float32)
        self._check(self._L.dg_cnn_forward(self._h, B, dims, _fp(x), _fp(y)), "dg_cnn_forward")
        return y

    def backward(self, x, dy):
        """Parameter gradients of sum(forward(x) * dy)."""
        x = np.ascontiguousarray(x, np.float32)
        dy = np.ascontiguousarray(dy, np.float32)
        B, dims = x.shape[0], (ctypes.c_int32 * 3)(*x.shape[1:4])
        g = np.empty(self.nparams, np.float32)
        self._check(self._L.dg_cnn_backward(self._h, B, dims, _fp(x), _fp(dy), _fp(g)), "dg_cnn_backward")
        return g

    def train_batch(self, x, target):
        """(loss, gradients) of one batch: x [B, L+7, L+7, L+7, 5], target [B, L, L, L, 5]."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        B, L = x.shape[0], t.shape[1]
        loss = ctypes.c_double()
        g = np.empty(self.nparams, np.float32)
        self._check(self._L.dg_cnn_train_batch(self._h, B, L, _fp(x), _fp(t), ctypes.byref(loss), _fp(g)),
                    "dg_cnn_train_batch")
        return loss.value, g

    def adam_step(self, grads=None, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        gp = None
        if grads is not None:
            grads = np.ascontiguousarray(grads, np.float32)
            gp = _fp(grads)
        self._check(self._L.dg_cnn_adam_step(self._h, gp, float(lr), float(beta1), float(beta2), float(eps)),
                    "dg_cnn_adam_step")

    # training from the solver -----------------------------------------------
    def add_sample(self, fine, coarse):
        self._check(self._L.dg_cnn_add_sample(self._h, fine.handle, coarse.handle), "dg_cnn_add_sample")

    @property
    def samples(self):
        return int(self._L.dg_cnn_sample_count(self._h))

    def fit_normalization(self):
        self._check(self._L.dg_cnn_fit_normalization(self._h), "dg_cnn_fit_normalization")

    def train(self, steps, batch=8, crop=32, lr=1e-3, seed=0):
        losses = np.empty(int(steps), np.float32)
        self._check(self._L.dg_cnn_train(self._h, int(steps), int(batch), int(crop), float(lr), int(seed),
                                         _fp(losses)), "dg_cnn_train")
        return losses

    # learned solver -----------------------------------------------------------
    def step(self, solver):
        dt = ctypes.c_double()
        self._check(self._L.dg_learned_step(solver.handle, self._h, ctypes.byref(dt)), "dg_learned_step")
        return dt.value

    def run(self, solver, max_steps, t_end=math.inf):
        t, n = ctypes.c_double(), ctypes.c_int64()
        self._check(self._L.dg_learned_run(solver.handle, self._h, int(max_steps), float(t_end), ctypes.byref(t),
                                           ctypes.byref(n)), "dg_learned_run")
        return t.value, n.value

    def step_given(self, solver, G):
        """A learned step with the host's correction G [n0, n1, n2, 16]."""
        G = np.ascontiguousarray(G, np.float32)
        assert G.shape == tuple(solver.sim.interior) + (16,)
        dt = ctypes.c_double()
        self._check(self._L.dg_learned_step_given(self._h, solver.handle, _fp(G), ctypes.byref(dt)),
                    "dg_learned_step_given")
        return dt.value

    def flagged(self):
        n = ctypes.c_int64()
        self._check(self._L.dg_learned_flagged(self._h, ctypes.byref(n)), "dg_learned_flagged")
        return n.value

    def correction(self, solver):
        """Network output for the solver's current state: [n0, n1, n2, 16]."""
        n = solver.sim.interior
        G = np.empty(n + (16,), np.float32)
        self._check(self._L.dg_learned_correction(self._h, solver.handle, _fp(G)), "dg_learned_correction")
        return G
