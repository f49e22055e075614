"""Closed-form pins of the two slicers' cut rules.

pEVSL_SpslicerKpm (SRC/spslicer.c:62-156) and pEVSL_SpslicerLan
(SRC/landos.c:331-355) turn an integrated DOS into slice boundaries.  Here the
integral is known exactly, so every cut follows from the published rule
alone:

* KPM: with mu = (1/pi, 0, ..., 0) the integral at the grid points is
  Y(x) = (acos x_0 - acos x) / pi exactly (intChx, SRC/spslicer.c:12-38).
  Cut l is the grid point nearest to the running target (the first point with
  Y >= target or the one before it, SRC/spslicer.c:109-118), and the next
  target spreads what is left of the integral evenly over the remaining
  slices (:120).
* Lanczos DOS: a constant ydos on an integer grid has the Simpson integral
  Y_k = c k exactly; a cut falls at the first point whose integral since the
  last cut reaches want = Y_total / n_int, and the last cut resets the origin
  (SRC/landos.c:343-352).

Each test also evaluates the plausible mistakes (no best-of-two choice, a
fixed target, '>' for '>=', no reset of the origin) on the same input and
checks that they give other cuts, so that a slicer carrying one of them fails.
The oracle runs on the CPU; the library's SpslicerLan is host code and runs
here too; the library's SpslicerKpm (integral on the GPU) is checked in
tests/test_gpu_c2c3.py against the same rule.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402


def ref_linspace(a, b, n):
    # SRC/vector.c:15-25: end points exact, a + i h inside
    h = (b - a) / (n - 1)
    x = a + np.arange(n) * h
    x[0], x[-1] = a, b
    return x


def kpm_cuts_rule(intv, n_int, npts, best_of_two=True, running=True, scale=1.0):
    """Cuts of the greedy walk from the closed-form integral of
    mu = (scale/pi, 0, ..., 0); the grid index of each cut is the first
    point whose integral reaches the target.  Returns (rc, sli)."""
    a, b, lmin, lmax = intv
    ctr, wid = (lmax + lmin) / 2, (lmax - lmin) / 2
    aL = max((max(a, lmin) - ctr) / wid, -1.0)
    bL = min((min(b, lmax) - ctr) / wid, 1.0)
    npts = max(npts, 2 * n_int + 1)
    x = ref_linspace(aL, bL, npts)
    th = np.arccos(x)
    Y = (scale / np.pi) * (th[0] - th)
    sli = np.zeros(n_int + 1)
    sli[0] = a
    target = Y[-1] / n_int
    ls = 0
    while True:
        ls += 1
        if ls >= n_int:
            break
        reach = np.nonzero(Y >= target)[0]
        if len(reach) == 0:
            break
        i = int(reach[0])
        if best_of_two and i > 0 and (target - Y[i - 1]) < (Y[i] - target):
            i -= 1
        sli[ls] = ctr + wid * x[i]
        target = (Y[i] + (Y[-1] - Y[i]) / (n_int - ls)) if running else (ls + 1) * Y[-1] / n_int
    sli[n_int] = b
    rc = 0 if ls == n_int else 1
    if np.any(np.diff(sli) < 0):
        rc += 2
    return rc, sli


MU0 = np.zeros(31)
MU0[0] = 1.0 / np.pi

# coarse grids make the nearest-point choice and the running target matter
KPM_CASES = [([0.6, 1.2, 0.0, 12.0], 5, 11), ([-3.0, 7.5, -4.0, 8.0], 7, 15),
             ([1.0, 11.0, 0.0, 12.0], 4, 9), ([0.6, 1.2, 0.0, 12.0], 20, 4000)]


@pytest.mark.parametrize("intv,n_int,npts", KPM_CASES)
def test_oracle_spslicer_kpm_closed_form(intv, n_int, npts):
    rc, sli = orc.spslicer_kpm(MU0, intv, n_int, npts)
    rc_ref, sli_ref = kpm_cuts_rule(intv, n_int, npts)
    assert rc == rc_ref == 0
    np.testing.assert_allclose(sli, sli_ref, rtol=0, atol=1e-13 * (intv[3] - intv[2]))


def test_kpm_rule_mutants_are_detected():
    """On the coarse cases the rule's two ingredients each change a cut."""
    diff_b2, diff_fixed = 0, 0
    for intv, n_int, npts in KPM_CASES[:3]:
        _, good = kpm_cuts_rule(intv, n_int, npts)
        _, nob2 = kpm_cuts_rule(intv, n_int, npts, best_of_two=False)
        _, fixed = kpm_cuts_rule(intv, n_int, npts, running=False)
        diff_b2 += int(np.max(np.abs(good - nob2)) > 1e-9)
        diff_fixed += int(np.max(np.abs(good - fixed)) > 1e-9)
    assert diff_b2 >= 1 and diff_fixed >= 1


def test_oracle_spslicer_kpm_nonpositive_integral():
    """mu = 0 (and a negative density) give an integral <= 0: every target is
    reached at the first grid point, whose cut is taken without reading
    before the grid.  The cuts then sit at the rounded interval start
    ctr + wid x_0; the return code flags a cut rounded below intv[0] (+2,
    SRC/spslicer.c:130-135) exactly as the rule does."""
    for intv in ([0.6, 1.2, 0.0, 12.0], [1.0, 5.0, 0.0, 12.0]):
        for scale in (0.0, -1.0):
            rc, sli = orc.spslicer_kpm(scale * MU0[:11], intv, 4, 9)
            rc_ref, sli_ref = kpm_cuts_rule(intv, 4, 9, scale=scale)
            assert rc == rc_ref
            np.testing.assert_array_equal(sli, sli_ref)
            np.testing.assert_allclose(sli[1:-1], intv[0], atol=1e-14)


def lan_cuts_rule(y_int, x, n_int, geq=True, reset=True):
    """Cuts of SpslicerLan from an exact integral y_int at the points x."""
    want = (y_int[-1] - y_int[0]) / n_int
    sli = np.zeros(n_int + 1)
    sli[0] = x[0]
    t, ls = y_int[0], 0
    for k in range(1, len(x)):
        d = y_int[k] - t
        if (d >= want) if geq else (d > want):
            ls += 1
            if ls <= n_int:
                sli[ls] = x[k]
            if reset:
                t = y_int[k]
    sli[n_int] = x[-1]
    return sli


LAN_CASES = [(13, 4, 1.0), (11, 3, 0.5), (25, 6, 2.0)]  # (npts, n_int, ydos)


@pytest.mark.parametrize("npts,n_int,c", LAN_CASES)
def test_oracle_spslicer_lan_closed_form(npts, n_int, c):
    x = np.arange(npts, dtype=np.float64)
    yd = np.full(npts, c)
    # Simpson of a constant on an integer grid is exact: Y_k = c k
    np.testing.assert_array_equal(orc.simpson(x, yd.copy()), c * x)
    sli = orc.spslicer_lan(x, yd.copy(), n_int)
    np.testing.assert_array_equal(sli, lan_cuts_rule(c * x, x, n_int))


@pytest.mark.parametrize("npts,n_int,c", LAN_CASES)
def test_library_spslicer_lan_closed_form(lib, npts, n_int, c):
    from pevsl_amd import driver
    x = np.arange(npts, dtype=np.float64)
    sli = driver.spslicer_lan(lib, x, np.full(npts, c), n_int)
    np.testing.assert_array_equal(sli, lan_cuts_rule(c * x, x, n_int))


def test_lan_rule_mutants_are_detected():
    """13 points, 4 slices: want = 3 exactly, cuts at 3, 6, 9 ('>' cuts at
    4, 8; no reset cuts at every point from 3 on).  11 points, 3 slices:
    want = 10/3 is not a grid step, so an origin advanced by want instead of
    reset to the cut point (t += want) moves the second cut from 8 to 7."""
    x = np.arange(13.0)
    good = lan_cuts_rule(x, x, 4)
    np.testing.assert_array_equal(good, [0, 3, 6, 9, 12])
    assert not np.array_equal(lan_cuts_rule(x, x, 4, geq=False), good)
    assert not np.array_equal(lan_cuts_rule(x, x, 4, reset=False), good)
    x = np.arange(11.0)
    good = lan_cuts_rule(x, x, 3)
    np.testing.assert_array_equal(good, [0, 4, 8, 10])
    want, t, cuts = 10.0 / 3, 0.0, []
    for k in range(1, 11):
        if x[k] - t >= want:
            cuts.append(k)
            t += want
    assert cuts[:2] == [4, 7]
