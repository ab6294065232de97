"""Write the golden vectors in tests/golden/ by running the reference's own
Fortran physics module (src/f90/single/physics.f90, compiled unmodified by
`make -C oracle ref` together with the oracle/ref_f90 harness).

Run in the development container only (it needs /root/reference):

    make -C oracle ref && python tests/golden/make_golden.py

For every case in cases.py and both builds of the reference (real(8) "dp"
and real(4) "sp"), the fixture stores the per-step dt and the conserved
fields on the interior after the last step; the 64^3 Sod case stores one
x-line plus per-field checksums instead of the full fields.  limiters.npz
holds the module's limiter functions on a set of difference pairs
(`python tests/golden/make_golden.py limiters` writes only that file).
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import cases  # noqa: E402
import dg_oracle  # noqa: E402


def summarize(u):
    return np.array([u.sum(), np.abs(u).sum(), u.min(), u.max()])


def main():
    for name in cases.CASES:
        c = cases.build_case(name)
        out = {}
        for prec in ("dp", "sp"):
            U, dts, _ = dg_oracle.run_reference_fortran(
                c["w"], c["ds"], c["bcs"], c["nsteps"], precision=prec, params=c["params"], nthreads=4)
            Ui = np.stack([u[2:-2, 2:-2, 2:-2] for u in U])
            if prec == "sp":
                Ui = Ui.astype(np.float32)
                dts = dts.astype(np.float32)
            out[f"{prec}_dt"] = dts
            if name == "sod64":
                mid = [s // 2 for s in Ui.shape[1:]]
                out[f"{prec}_line"] = Ui[:, :, mid[1], mid[2]]
                out[f"{prec}_summary"] = np.stack([summarize(u.astype(np.float64)) for u in Ui])
            else:
                out[f"{prec}_U"] = Ui
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {os.path.getsize(path)} bytes")


def limiter_pairs():
    """(backward, forward) difference pairs for the limiter fixture: hand-made
    cases (unequal magnitudes of both signs, sign changes, zeros, ties) and
    seeded random pairs over several magnitudes."""
    hand = [(1, 3), (3, 1), (-3, -1), (-1, -3), (1, -1), (-2, 5), (0, 4), (4, 0), (0, 0),
            (2, 2), (-4, -4), (0.5, 0.25), (1e-3, 7e-3), (-1e3, -2.5e2), (1e-30, 2e-30)]
    rng = np.random.default_rng(20240607)
    mag = 10.0 ** rng.uniform(-6, 3, size=(2, 200))
    sgn = np.where(rng.random((2, 200)) < 0.3, -1.0, 1.0)
    a = np.concatenate([[h[0] for h in hand], sgn[0] * mag[0]])
    b = np.concatenate([[h[1] for h in hand], sgn[1] * mag[1]])
    return a, b


def write_limiters():
    """tests/golden/limiters.npz: the reference module's minmod, maxmod and
    vanleerlimiter (physics.f90:205-238) of every pair, in both real kinds."""
    a, b = limiter_pairs()
    out = {"a": a, "b": b}
    for prec in ("dp", "sp"):
        mn, mx, vl = dg_oracle.run_reference_limiters(a, b, precision=prec)
        out[f"{prec}_minmod"], out[f"{prec}_maxmod"], out[f"{prec}_vanleerlimiter"] = mn, mx, vl
    path = os.path.join(HERE, "limiters.npz")
    np.savez_compressed(path, **out)
    print(f"limiters: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    if sys.argv[1:] == ["limiters"]:
        write_limiters()
    else:
        main()
        write_limiters()
