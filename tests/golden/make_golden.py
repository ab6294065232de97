"""Write the golden vectors in tests/golden/ by running the reference's own
Fortran physics module (src/f90/single/physics.f90, compiled unmodified by
`make -C oracle ref` together with the oracle/ref_f90 harness).

Run in the development container only (it needs /root/reference):

    make -C oracle ref && python tests/golden/make_golden.py

For every case in cases.py and both builds of the reference (real(8) "dp"
and real(4) "sp"), the fixture stores the per-step dt and the conserved
fields on the interior after the last step; the 64^3 Sod case stores one
x-line plus per-field checksums instead of the full fields.
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


if __name__ == "__main__":
    main()
