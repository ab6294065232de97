"""Pins the oracle's limited slopes (oracle/dg_oracle_impl.h, `limited`).

The reference's step uses unlimited central slopes; its Fortran physics
module carries minmod, maxmod and a van Leer limiter as functions its step
never calls (src/f90/single/physics.f90:205-238).  The library offers
textbook minmod and van Leer (harmonic-mean) slopes as an extension, because
the blast workload breaks the unlimited scheme.  They are pinned here by:

* the reference module's own functions, compiled where they lie
  (oracle/_ref/dg_lim_*, values in tests/golden/limiters.npz):
  - minmod equals the module's `minmod` on every pair with |a| != |b|; on a
    tie the module returns 0 and the oracle returns a (= b), the textbook
    value that keeps linear data second order;
  - van Leer equals the module's `vanleerlimiter(a, b)` (a dimensionless
    ratio (sign a + sign b) min(|a|,|b|) / (|a| + |b| + 1e-30)) times
    max(|a|, |b|), which is 2ab/(a+b) for pairs of one sign;
* closed forms on a piecewise-linear profile with kinks, through the whole
  calc_gradients pass, along each axis;
* checks that the fixture separates the plausible mistakes (maxmod for
  minmod, an arithmetic mean for van Leer, a swapped comparison).
"""
import os

import numpy as np
import pytest

import dg_oracle
from dg_oracle import OUTFLOW, Oracle

PREC = {"f64": ("dp", np.float64), "f32": ("sp", np.float32)}


@pytest.fixture(scope="module")
def fixture(golden_dir):
    return np.load(os.path.join(golden_dir, "limiters.npz"))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_minmod_matches_reference_module(fixture, prec):
    tag, dt = PREC[prec]
    a, b = fixture["a"].astype(dt), fixture["b"].astype(dt)
    ours = dg_oracle.limited(a, b, "minmod", prec)
    ref = fixture[f"{tag}_minmod"].astype(dt)
    tie = np.abs(a) == np.abs(b)
    assert (~tie).sum() > 200
    np.testing.assert_array_equal(ours[~tie], ref[~tie])
    # ties: the module gives 0, the textbook minmod the common value
    assert np.all(ref[tie] == 0)
    np.testing.assert_array_equal(ours[tie], np.where(a[tie] * b[tie] > 0, a[tie], 0))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_vanleer_matches_reference_module(fixture, prec):
    tag, dt = PREC[prec]
    a, b = fixture["a"].astype(dt), fixture["b"].astype(dt)
    ours = dg_oracle.limited(a, b, "vanleer", prec).astype(np.float64)
    phi = fixture[f"{tag}_vanleerlimiter"]
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    same = a64 * b64 > 0
    big = np.abs(a64) + np.abs(b64) > 1e-20        # the module's 1e-30 regularizer is negligible
    expect = phi * np.maximum(np.abs(a64), np.abs(b64))
    rtol = 1e-14 if prec == "f64" else 1e-6
    sel = same & big
    assert sel.sum() > 100
    np.testing.assert_allclose(ours[sel], expect[sel], rtol=rtol, atol=0)
    # a sign change or a zero difference gives a zero slope in both
    assert np.all(ours[~same] == 0)
    assert np.all(phi[~same & big] == 0)


def test_fixture_separates_plausible_mistakes(fixture):
    """The pins above fail for maxmod in place of minmod (the module's own
    maxmod), for an arithmetic mean in place of van Leer, and for a minmod
    whose comparison is swapped."""
    a, b = fixture["a"], fixture["b"]
    tie = np.abs(a) == np.abs(b)
    ref_mm = fixture["dp_minmod"]
    assert np.any(fixture["dp_maxmod"][~tie] != ref_mm[~tie])
    swapped = np.where(a * b > 0, np.where(np.abs(a) >= np.abs(b), a, b), 0.0)
    assert np.any(swapped[~tie] != ref_mm[~tie])
    same = (a * b > 0) & (np.abs(a) + np.abs(b) > 1e-20)
    vl = fixture["dp_vanleerlimiter"] * np.maximum(np.abs(a), np.abs(b))
    mean = 0.5 * (a + b)
    assert np.any(np.abs(mean[same] - vl[same]) > 1e-6 * np.abs(vl[same]))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_closed_forms(prec):
    a = np.array([1.0, 3.0, -3.0, -1.0, 1.0, -2.0, 0.0, 2.0])
    b = np.array([3.0, 1.0, -1.0, -4.0, -1.0, 5.0, 4.0, 2.0])
    mm = dg_oracle.limited(a, b, "minmod", prec)
    vl = dg_oracle.limited(a, b, "vanleer", prec)
    np.testing.assert_array_equal(mm, [1.0, 1.0, -1.0, -1.0, 0.0, 0.0, 0.0, 2.0])
    dt = PREC[prec][1]
    np.testing.assert_array_equal(vl, np.array([1.5, 1.5, -1.5, -1.6, 0.0, 0.0, 0.0, 2.0], dtype=dt))


# piecewise-linear profile along one axis (12 ghosted cells, 8 interior) and
# its limited slopes at ghosted indices 1..10, worked out by hand with ds 0.5:
# differences (a, b) = (1,3) (3,1) (1,-2) (-2,0) (0,4) (4,-1) (-1,-4) (-4,-4)
# (-4,-1) (-1,0)
PROFILE = np.array([0.0, 1.0, 4.0, 5.0, 3.0, 3.0, 7.0, 6.0, 2.0, -2.0, -3.0, -3.0])
SLOPES = {
    "minmod": [2.0, 2.0, 0.0, 0.0, 0.0, 0.0, -2.0, -8.0, -2.0, 0.0],
    "vanleer": [3.0, 3.0, 0.0, 0.0, 0.0, 0.0, -3.2, -8.0, -3.2, 0.0],
}


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("lim", ["minmod", "vanleer"])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_kinked_profile_through_calc_gradients(prec, lim, axis):
    interior = [5, 6, 7]
    interior[axis] = 8
    shape = tuple(n + 4 for n in interior)
    o = Oracle(tuple(interior), 0.5, [OUTFLOW] * 6, precision=prec, limiter=lim)
    sl = [None, None, None]
    sl[axis] = slice(None)
    prof = np.broadcast_to(PROFILE[tuple(sl)], shape)
    # every field carries the profile (offsets keep rho and p positive)
    w = [10.0 + prof, prof, -prof, 2.0 * prof, 20.0 + prof]
    o.set_primitives(w)
    o.calc_gradients()
    dt = PREC[prec][1]
    expect = np.array(SLOPES[lim], dtype=dt)
    scale = [1.0, 1.0, -1.0, 2.0, 1.0]
    for f in range(5):
        g = np.moveaxis(np.asarray(o.grad[f])[..., axis], axis, 0)   # profile axis first
        line = g[1:-1, 2, 2]
        np.testing.assert_array_equal(line, (expect * dt(scale[f])) if scale[f] != -1.0 else -expect)
        for other in range(3):
            if other != axis:   # constant along the other axes: zero slopes
                go = np.asarray(o.grad[f])[..., other]
                inner = go[tuple(slice(1, -1) for _ in range(3))]
                assert np.all(inner == 0)


def test_minmod_is_the_median_with_zero(fixture):
    """The fast FP32 step evaluates minmod as the median of (a, b, 0) (one
    hardware median instruction): equal to the oracle's minmod on every
    fixture pair, ties, zeros and sign changes included, except where the
    oracle's sign test a * b > 0 underflows in FP32 (|a b| below the smallest
    subnormal: the oracle gives 0, the median the smaller of two tiny
    same-sign differences); and different from the median with the wrong
    third operand."""
    a, b = fixture["a"].astype(np.float32), fixture["b"].astype(np.float32)
    med = np.median(np.stack([a, b, np.zeros_like(a)]), axis=0).astype(np.float32)
    ours = dg_oracle.limited(a, b, "minmod", "f32")
    under = (a * b == 0) & (a != 0) & (b != 0)
    assert 0 < under.sum() < 5
    np.testing.assert_array_equal(med[~under], ours[~under])
    assert np.all(ours[under] == 0) and np.all(np.abs(med[under]) < 1e-19)
    assert np.any(np.abs(a) == np.abs(b)) and np.any(a * b < 0) and np.any(a == 0)
    wrong = np.median(np.stack([a, b, np.ones_like(a)]), axis=0)
    assert np.any(wrong != ours)
