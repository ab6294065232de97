"""Pin the CPU oracle to golden vectors written by the reference's own Fortran
physics module (tests/golden/make_golden.py, oracle/Makefile `ref`).

The real(8) build ("dp") pins the oracle's double instantiation and the
real(4) build ("sp"), whose expression order equals the Rust f32 solver's,
pins the float instantiation.  Both must agree bit for bit.
"""
import os

import numpy as np
import pytest

import cases
from dg_oracle import Oracle

CASE_NAMES = list(cases.CASES)


def run_oracle(name, precision, nthreads=4):
    c = cases.build_case(name)
    prm = cases.PARAMS[c["params"]]
    o = Oracle(c["interior"], c["ds"], c["bcs"], precision=precision, nthreads=nthreads, **prm)
    o.set_primitives(c["w"])
    n, _, dts = o.compute(1e30, c["nsteps"])
    assert n == c["nsteps"]
    U = np.stack([np.array(o.interior_view(a)) for a in o.cons])
    return U, dts


@pytest.mark.parametrize("name", CASE_NAMES)
@pytest.mark.parametrize("prec", ["dp", "sp"])
def test_oracle_matches_reference_fortran(golden_dir, name, prec):
    g = np.load(os.path.join(golden_dir, f"{name}.npz"))
    U, dts = run_oracle(name, "f64" if prec == "dp" else "f32")
    np.testing.assert_array_equal(dts, g[f"{prec}_dt"])
    if name == "sod64":
        mid = [s // 2 for s in U.shape[1:]]
        np.testing.assert_array_equal(U[:, :, mid[1], mid[2]], g[f"{prec}_line"])
        summ = np.stack([[u.astype(np.float64).sum(), np.abs(u.astype(np.float64)).sum(), u.min(), u.max()]
                         for u in U])
        np.testing.assert_array_equal(summ, g[f"{prec}_summary"])
        # Sod along x stays exactly independent of y and z
        assert np.array_equal(U, np.broadcast_to(U[:, :, :1, :1], U.shape))
    else:
        np.testing.assert_array_equal(U, g[f"{prec}_U"])


def test_golden_cases_are_finite(golden_dir):
    for name in CASE_NAMES:
        g = np.load(os.path.join(golden_dir, f"{name}.npz"))
        for k in g.files:
            assert np.isfinite(g[k]).all(), (name, k)
