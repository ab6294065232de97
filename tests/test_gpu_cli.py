"""GPU: dg_cli, the reference's command-line host (main.rs) on the library.

The loop and the dump cadence follow main.rs:38-68 (a dump after the first
step and whenever wtime > tout); each dump holds the reference's HDF5
datasets (io.rs) as .npy files.  For every demo, the library's step between
two dumps and the dumped arrays are checked bit for bit against the oracle
started from the first dump (FP32 exact math, the reference's constants,
walls and gravity); a restarted run is checked against an uninterrupted one.
The demos' initial conditions themselves (init.rs) are parity unpinned: the
reference's Rust cannot run here."""
import os
import subprocess

import numpy as np
import pytest

from dg_oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
CLI = os.path.join(ROOT, "deep-green_amd", "dg_cli")
FIELDS = ("mass", "momentum_x", "momentum_y", "momentum_z", "energy")


def run_cli(args, cwd):
    r = subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def load(dirpath, group, names):
    return [np.load(os.path.join(dirpath, f"{group}.{n}.npy")) for n in names]


PRIMS = ("rho", "vx", "vy", "vz", "pressure")

# the demos of main.rs:79-108 as init.rs sets them up: interior size, ds,
# boundaries (x-, x+, y-, y+, z-, z+), gravity
DEMOS = {
    "KELVIN_HELMHOLTZ": ((256, 2, 128), 1.0 / 256, [0, 0, 0, 0, 0, 0], False),
    "RAYLEIGH_TAYLOR": ((100, 2, 300), 2.0 * 0.0025, [0, 0, 0, 0, 2, 2], True),
    "RAYLEIGH_BERNARD": ((64, 64, 64), 1.0 / 64, [2, 2, 2, 2, 2, 2], True),
    "EXPLOSION": ((128, 2, 128), 1.0 / 128, [2, 2, 0, 0, 2, 2], False),
    "RELAXATION": ((30, 30, 30), 0.1, [2, 2, 2, 2, 2, 1], True),
    "SHOCK_TUBE": ((5000, 2, 2), 0.0002, [2, 2, 0, 0, 0, 0], False),
    "THERMAL_RISING_BUBBLE": ((150, 2, 512), 1.0, [2, 2, 0, 0, 2, 1], True),
    "THERMAL_RISING_BUBBLES": ((128, 2, 500), 2.0, [2, 2, 0, 0, 2, 1], True),
}


def demo_oracle(demo):
    n, ds, bcs, gravity = DEMOS[demo]
    return Oracle(n, np.float32(ds), bcs, precision="f32", gamma=5.0 / 3.0, cfl=0.2, g=-0.1, gravity=gravity,
                  gravity_k_shift=0)


@pytest.mark.parametrize("demo", sorted(DEMOS))
def test_cli_demo_dumps_follow_oracle(tmp_path, demo):
    """Every demo at its reference size: the step between two dumps equals the
    oracle's step bit for bit, and the dump holds what the reference grid
    holds after compute_step (io.rs:5-196): the conserved state after the
    step, the primitives of the step's start and the step's gradients and
    fluxes, ghost cells included (those of the conserved state follow the
    boundary maps)."""
    n = DEMOS[demo][0]
    out = run_cli(["--demo", demo, "--tout", "0", "--total-time", "1e9", "--max-steps", "2", "--out", str(tmp_path)],
                  str(tmp_path))
    assert "Computing step 1" in out
    dumps = sorted(d for d in os.listdir(tmp_path) if d.startswith("state."))
    assert dumps == ["state.00000", "state.00001"]
    U0 = load(tmp_path / dumps[0], "conserved", FIELDS)
    assert U0[0].shape == tuple(k + 4 for k in n) and U0[0].dtype == np.float64
    assert all(np.isfinite(u).all() for u in U0)
    o = demo_oracle(demo)
    o.set_conserved([u.astype(np.float32) for u in U0])
    o.steps(1)
    d1 = tmp_path / dumps[1]
    U1 = load(d1, "conserved", FIELDS)
    W1, G1, F1 = load(d1, "primitives", PRIMS), load(d1, "gradients", PRIMS), load(d1, "fluxes", FIELDS)
    assert G1[0].shape == U0[0].shape + (3,)
    for q in range(5):
        np.testing.assert_array_equal(np.array(o.interior_view(o.cons[q])), U1[q][2:-2, 2:-2, 2:-2].astype(np.float32))
        np.testing.assert_array_equal(np.asarray(o.prim[q]), W1[q].astype(np.float32))
        np.testing.assert_array_equal(np.asarray(o.grad[q]), G1[q].astype(np.float32))
        np.testing.assert_array_equal(np.asarray(o.flux[q]), F1[q].astype(np.float32))


def test_cli_restart_continues_the_run_exactly(tmp_path):
    """--restart from a dump continues the loop (state, clock, dump numbering)
    and ends where the uninterrupted run ends, bit for bit."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    probe = run_cli(["--demo", "RAYLEIGH_TAYLOR", "--tout", "1", "--total-time", "1e9", "--max-steps", "2",
                     "--no-dump"], str(a))
    dt = float(probe.split("Computing step 1 and time ")[1].split(" with dt ")[1].split()[0].rstrip("."))
    # a dump about every third step
    args = ["--demo", "RAYLEIGH_TAYLOR", "--tout", repr(2.5 * dt), "--total-time", "1e9", "--max-steps", "11",
            "--quiet"]
    out_a = run_cli(args + ["--out", str(a)], str(a))
    dumps = sorted(d for d in os.listdir(a) if d.startswith("state."))
    assert len(dumps) >= 3, dumps
    out_b = run_cli(args + ["--out", str(b), "--restart", str(a / dumps[1])], str(b))
    assert "Restarting simulation RAYLEIGH_TAYLOR" in out_b
    assert out_a.split("Simulation done")[1].split("s!")[1] == out_b.split("Simulation done")[1].split("s!")[1]
    assert sorted(d for d in os.listdir(b) if d.startswith("state.")) == dumps[2:]
    last = dumps[-1]
    for group, names in (("conserved", FIELDS), ("primitives", PRIMS), ("gradients", PRIMS), ("fluxes", FIELDS)):
        for x, y in zip(load(a / last, group, names), load(b / last, group, names)):
            np.testing.assert_array_equal(x, y)
    assert open(a / last / "meta.txt").read() == open(b / last / "meta.txt").read()


@pytest.mark.parametrize("flag, value", [("--precision", "f64"), ("--math", "quick"), ("--limiter", "superbee")])
def test_cli_rejects_unknown_option_values(tmp_path, flag, value):
    r = subprocess.run([CLI, "--demo", "SHOCK_TUBE", "--tout", "1", "--total-time", "1", "--max-steps", "1", flag,
                        value], capture_output=True, text=True, timeout=60, cwd=str(tmp_path))
    assert r.returncode == 2 and "unknown" in r.stderr


def test_cli_unknown_demo_lists_demos(tmp_path):
    out = subprocess.run([CLI, "--demo", "NOPE", "--tout", "1", "--total-time", "1", "--max-steps", "1"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "KELVIN_HELMHOLTZ" in out.stderr


def test_cli_benchmark_config_sod64_follows_oracle(tmp_path):
    """--config sod64 (gamma 1.4, CFL 0.4, outflow, FP64): one library step
    between two dumps equals the FP64 oracle's step bit for bit."""
    out = run_cli(["--config", "sod64", "--steps", "2", "--tout", "0", "--out", str(tmp_path), "--quiet"],
                  str(tmp_path))
    assert "2 steps" in out
    dumps = sorted(d for d in os.listdir(tmp_path) if d.startswith("state."))
    assert dumps == ["state.00000", "state.00001"]
    U0 = load(tmp_path / dumps[0], "conserved", FIELDS)
    U1 = load(tmp_path / dumps[1], "conserved", FIELDS)
    assert U0[0].shape == (68, 68, 68)
    o = Oracle((64, 64, 64), 1.0 / np.float32(64), [1] * 6, precision="f64", gamma=1.4, cfl=0.4, g=0.0,
               gravity=False, gravity_k_shift=0)
    o.set_conserved(U0)
    o.steps(1)
    for q in range(5):
        np.testing.assert_array_equal(np.array(o.interior_view(o.cons[q])), U1[q][2:-2, 2:-2, 2:-2])
    # the tube: left state rho 1, right 0.125, far from the interface untouched after two steps
    W1 = load(tmp_path / dumps[1], "primitives", ("rho", "vx", "vy", "vz", "pressure"))
    assert W1[0][2, 10, 10] == 1.0 and abs(W1[0][-3, 10, 10] - 0.125) < 1e-7


def test_cli_benchmark_config_without_out_writes_nothing(tmp_path):
    out = run_cli(["--config", "blast256", "--steps", "1", "--precision", "fp32", "--quiet"], str(tmp_path))
    assert "1 steps" in out and not os.listdir(tmp_path)


def test_cli_benchmark_config_khi512_runs(tmp_path):
    """The 512^3 benchmark configuration end to end from the command line
    (fast math, as bench.py times it)."""
    out = run_cli(["--config", "khi512", "--steps", "3", "--math", "fast", "--quiet"], str(tmp_path))
    assert "Running simulation khi512: 512x512x512 cells" in out and "3 steps" in out


def test_cli_config_ensemble_runs(tmp_path):
    """--config ensemble: a seeded batch of planar Riemann grids (edges from
    {32, 48, 64, 96, 128}, outflow, minmod, FP32 fast math) advanced in one
    batched run; every member's conserved totals stay finite."""
    out = run_cli(["--config", "ensemble", "--steps", "3", "--members", "16", "--seed", "5"], str(tmp_path))
    assert "Running ensemble of 16 grids" in out
    assert "(3 steps," in out and "state finite" in out


def test_cli_config_learned128_trains_and_runs(tmp_path):
    """--config learned128: samples from one 256^3 plain run, a short
    training, then learned steps on the 128^3 grid: mass conserved to
    rounding (the correction telescopes), state finite."""
    out = run_cli(["--config", "learned128", "--steps", "4", "--train-ics", "1", "--train-steps", "20"],
                  str(tmp_path))
    assert "Training run 1 of 1" in out and "Trained on" in out
    line = [ln for ln in out.splitlines() if ln.startswith("Learned run done")][0]
    assert "(4 steps" in line and "state finite" in line
    drift = float(line.split("mass drift ")[1].split(",")[0])
    assert abs(drift) < 1e-6


def test_cli_config_learned128_reads_parameters(tmp_path):
    """--params: the network's parameters and normalization as FluxCNN
    writes them (float32), so a network trained from Python runs from the
    command line; a file of the wrong size is rejected."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "deep-green_amd"))
    from deepgreen.learned import FluxCNN
    m = FluxCNN(seed=4)
    blob = np.concatenate([m.get_params(), m.get_normalization()]).astype(np.float32)
    m.close()
    path = tmp_path / "net.f32"
    blob.tofile(path)
    out = run_cli(["--config", "learned128", "--steps", "2", "--params", str(path)], str(tmp_path))
    assert "Network parameters from" in out and "state finite" in out
    blob[:-1].tofile(path)
    r = subprocess.run([CLI, "--config", "learned128", "--steps", "2", "--params", str(path)], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "cannot read" in r.stderr
