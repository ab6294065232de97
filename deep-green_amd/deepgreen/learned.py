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
    def add_sample(self, fine, coarse, keep_coarse=False, store=True):
        """One training sample (dg_cnn_add_sample_mode): U^n is the restriction
        of `fine` or, with keep_coarse, the coarse grid's own state; the fine
        grid advances two steps and the coarse grid one (plain) step with a
        common dt; store=False only advances them."""
        self._check(self._L.dg_cnn_add_sample_mode(self._h, fine.handle, coarse.handle, int(bool(keep_coarse)),
                                                   int(bool(store))), "dg_cnn_add_sample_mode")

    @property
    def samples(self):
        return int(self._L.dg_cnn_sample_count(self._h))

    def fit_normalization(self):
        self._check(self._L.dg_cnn_fit_normalization(self._h), "dg_cnn_fit_normalization")

    def set_input_noise(self, std):
        """Gaussian noise (normalized units) added to the inputs while training."""
        self._check(self._L.dg_cnn_set_input_noise(self._h, float(std)), "dg_cnn_set_input_noise")

    def train(self, steps, batch=8, crop=32, lr=1e-3, seed=0, chunk=250, log=None):
        """Adam on random crops of the samples, `chunk` steps per library call
        (crops of chunk k drawn from seed + k); `log` gets a progress line after
        every chunk.  Returns the per-step losses."""
        out = []
        done, k = 0, 0
        while done < int(steps):
            n = min(int(chunk), int(steps) - done)
            losses = np.empty(n, np.float32)
            self._check(self._L.dg_cnn_train(self._h, n, int(batch), int(crop), float(lr), int(seed) + k,
                                             _fp(losses)), "dg_cnn_train")
            out.append(losses)
            done += n
            k += 1
            if log:
                log(f"[learned] Adam step {done}/{int(steps)}: loss {float(losses[-min(n, 50):].mean()):.4g}")
        return np.concatenate(out)

    # learned solver -----------------------------------------------------------
    def step(self, solver, dt=0.0):
        """One learned step (CFL dt, or the given dt when dt > 0); returns the dt used."""
        used = ctypes.c_double()
        self._check(self._L.dg_learned_step(solver.handle, self._h, float(dt), ctypes.byref(used)),
                    "dg_learned_step")
        return float(dt) if dt > 0 else used.value

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

    def time_forward(self, solver, nrep=10):
        """Mean ms of [input preparation, layer 1, layer 2, layer 3, total]
        of the network's forward on the solver's grid (HIP events)."""
        ms = (ctypes.c_double * 5)()
        self._check(self._L.dg_learned_time(self._h, solver.handle, int(nrep), ms), "dg_learned_time")
        return list(ms)

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


# ------------------------------------------------------------ workflow --
def _sim(n, ds):
    from . import PERIODIC, Simulation
    return Simulation(n, n, n, ds, [PERIODIC] * 6, gamma=1.4, cfl=0.4, g=0.0, precision="fp32", math="fast",
                      limiter="minmod")


def clock(solver):
    return solver.run(0)


def run_to(solver, T, model=None, guard=4):
    """Advance a periodic FP32 Solver (plain, or learned with `model`) to t = T
    exactly: captured-graph runs until a few steps before T, then single
    steps, the last one with dt = T - t.  Returns the number of steps."""
    big = 1 << 40
    t, n0 = clock(solver)
    dt = solver.next_dt()
    if t + guard * dt < T:
        if model is None:
            solver.run(big, T - guard * dt)
        else:
            model.run(solver, big, T - guard * dt)
    while True:
        t, n = clock(solver)
        dt = solver.next_dt()
        if t + dt >= T:
            last = T - t
            if last > 0:
                if model is None:
                    solver.step_dt(last)
                else:
                    model.step(solver, dt=last)
            return clock(solver)[1] - n0
        if model is None:
            solver.compute_step()
        else:
            model.step(solver)


def make_training_set(model, seeds, n_fine=256, t_end=0.03, every=8, log=None, trajectory=False):
    """Samples from plain runs of learned_ic(seed) on n_fine^3 against the 2x
    coarser grid up to t_end.  trajectory=False: every `every` fine steps the
    coarse grid is reset to the restricted fine state and one sample taken
    (the one-step error of the coarse step).  trajectory=True: the coarse grid
    starts from the restricted initial state and runs its own plain
    trajectory in step with the fine one (one coarse step per two fine steps,
    common dt); every `every // 2` coarse steps a sample of its state is kept,
    with the target that returns it onto the restricted fine trajectory."""
    from . import Solver
    from .setups import learned_ic
    for seed in seeds:
        w, ds = learned_ic(n_fine, seed)
        fine = Solver(_sim(n_fine, ds))
        coarse = Solver(_sim(n_fine // 2, 2 * ds))
        fine.set_primitives(w)
        del w
        if log:
            log(f"[learned] seed {seed}: {n_fine}^3 initial condition ready")
        t, k = 0.0, 0
        if trajectory:
            coarse.restrict_from(fine)
        while t < t_end:
            if trajectory:
                model.add_sample(fine, coarse, keep_coarse=True, store=k % max(every // 2, 1) == 0)
                k += 1
            else:
                model.add_sample(fine, coarse)
                fine.run(max(every - 2, 0))
            t = clock(fine)[0]
            if log and model.samples % 4 == 0:
                log(f"[learned] seed {seed}: {model.samples} samples, t = {t:.4f}")
        if log:
            log(f"[learned] seed {seed}: {model.samples} samples, t = {t:.4f}")
        fine.close()
        coarse.close()


def density(solver):
    U = solver.get_state()
    V = solver.sim.ds ** 3
    return np.asarray(U[0][2:-2, 2:-2, 2:-2], np.float64) / V


def evaluate(model, seed, n_truth=512, T=0.03, warmup_steps=0, log=None, ic=None):
    """L1 density errors at t = T on the coarsest grid (n_truth / 4) of the
    plain n_truth/2 and n_truth/4 solvers and of the learned n_truth/4
    solver, against the plain n_truth solution restricted to that grid (all
    start from restrictions of the n_truth initial condition), and their
    wall times (the state resident on the device, graph-captured runs).
    warmup_steps: plain steps the learned solver takes before the correction
    starts (the initial state, at rest with sharp blasts, lies outside the
    states the network is trained on)."""
    import time

    from . import Solver
    from .setups import learned_ic
    n2, n4 = n_truth // 2, n_truth // 4
    w, ds = ic if ic is not None else learned_ic(n_truth, seed)
    if log:
        log(f"[learned] evaluation IC {n_truth}^3 (seed {seed}) ready")
    truth = Solver(_sim(n_truth, ds))
    truth.set_primitives(w)
    del w
    s2, p4, l4 = Solver(_sim(n2, 2 * ds)), Solver(_sim(n4, 4 * ds)), Solver(_sim(n4, 4 * ds))
    s2.restrict_from(truth)
    p4.restrict_from(s2)
    l4.restrict_from(s2)
    out = {}
    for name, s, mdl in (("plain_truth", truth, None), ("plain_half", s2, None), ("plain_quarter", p4, None),
                         ("learned_quarter", l4, model)):
        s.next_dt()
        t0 = time.perf_counter()
        if mdl is not None and warmup_steps > 0:
            s.run(warmup_steps)
        steps = run_to(s, T, mdl) + (warmup_steps if mdl is not None else 0)
        s.next_dt()
        out[name] = {"steps": int(steps), "wall_s": time.perf_counter() - t0, "n": s.sim.interior[0]}
        if log:
            log(f"[learned] {name} {s.sim.interior[0]}^3: {steps} steps in {out[name]['wall_s']:.3f} s")
    # restrict everything to the coarsest grid
    r2, r4 = Solver(_sim(n2, 2 * ds)), Solver(_sim(n4, 4 * ds))
    r2.restrict_from(truth)
    r4.restrict_from(r2)
    ref4 = density(r4)
    q4 = Solver(_sim(n4, 4 * ds))
    q4.restrict_from(s2)
    errs = {"plain_half": float(np.abs(density(q4) - ref4).mean()),
            "plain_quarter": float(np.abs(density(p4) - ref4).mean()),
            "learned_quarter": float(np.abs(density(l4) - ref4).mean())}
    mass = (l4.conservation_sums()[0], p4.conservation_sums()[0])
    for s in (truth, s2, p4, l4, r2, r4, q4):
        s.close()
    return {"runs": out, "l1_rho": errs, "mass_learned_vs_plain": mass, "T": T, "seed": seed}
