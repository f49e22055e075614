"""CPU checks of the generalized-problem oracle (oracle/gen_oracle.py) and of
the library's least-squares polynomial fit (host code; no GPU)."""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import gen_oracle as go
from pevsl_amd import capi


def test_lspol_rec_matches_closed_form():
    """1/(cc + dd t) = (1/dd) 1/(r + t), r = cc/dd > 1, has the Chebyshev
    series (2/dd/sqrt(r^2-1)) sum' (-rho)^k T_k(t), rho = r - sqrt(r^2-1)."""
    for a, b in ((2.0, 6.0), (0.5, 3.0), (1.0, 1.5)):
        mu, cc, dd = go.lspol_coeffs(lambda x: 1.0 / x, a, b, 60, 1e-12)
        r = cc / dd
        rho = r - math.sqrt(r * r - 1)
        exact = 2.0 / dd / math.sqrt(r * r - 1) * (-rho) ** np.arange(len(mu))
        exact[0] /= 2.0
        assert np.max(np.abs(mu - exact)) < 1e-13 * np.max(np.abs(exact))


def test_lspol_sqrt_matches_interpolant():
    a, b = 2.0, 6.0
    mu, cc, dd = go.lspol_coeffs(lambda x: 1.0 / math.sqrt(x), a, b, 50, 1e-6)
    ref = np.polynomial.chebyshev.chebinterpolate(lambda t: (cc + dd * t) ** -0.5, 200)
    assert np.max(np.abs(mu - ref[:len(mu)])) < 1e-13


@pytest.mark.parametrize("max_deg,tol", [(50, 1e-6), (50, 1e-9), (8, 1e-14), (30, 1e-3)])
def test_lspol_degree_rule(max_deg, tol):
    """The degree is the first k whose sum of relative errors at 50 max_deg
    equispaced points falls below tol (or max_deg - 1), recomputed with
    numpy's Chebyshev evaluation (tolerances well above the rounding level
    of a 50 max_deg-term error sum, where two evaluation orders could
    disagree on the step)."""
    f = lambda x: 1.0 / math.sqrt(x)
    mu, cc, dd = go.lspol_coeffs(f, 2.0, 6.0, max_deg, tol)
    xs = np.linspace(-1.0, 1.0, 50 * max_deg)
    fx = (cc + dd * xs) ** -0.5
    err = lambda m: np.sum(np.abs((np.polynomial.chebyshev.chebval(xs, m) - fx) / fx))
    deg = len(mu) - 1
    if deg < max_deg - 1:
        assert err(mu) < tol * (1 + 1e-6)
        for k in range(deg):
            assert err(mu[:k + 1]) >= tol * (1 - 1e-6)
    else:
        assert deg == max_deg - 1


def _lspol_lib(lib, which, max_deg, tol, a, b, n=7):
    B = capi.Parcsr()
    B.ncol_global = B.ncol_local = B.nrow_global = B.nrow_local = n
    vd = C.c_void_p()
    setup = lib.pEVSL_SetupLSPolRec if which == "rec" else lib.pEVSL_SetupLSPolSqrt
    setup(max_deg, tol, a, b, C.byref(B), C.byref(vd))
    deg = C.c_int()
    mu = np.zeros(max_deg + 1)
    cc, dd = C.c_double(), C.c_double()
    assert lib.pEVSL_LSPolInfo(vd, C.byref(deg), mu.ctypes.data_as(capi.c_double_p),
                               C.byref(cc), C.byref(dd)) == 0
    lib.pEVSL_LSPolFree(vd)
    return mu[:deg.value + 1], cc.value, dd.value


@pytest.mark.parametrize("which,max_deg,tol,a,b", [
    ("rec", 50, 1e-12, 2.0, 6.0), ("rec", 20, 1e-6, 0.3, 2.0),
    ("sqrt", 50, 1e-6, 2.0, 6.0), ("sqrt", 40, 1e-10, 1.0, 9.0)])
def test_library_lspol_fit_matches_oracle(lib, which, max_deg, tol, a, b):
    """pEVSL_SetupLSPolRec / Sqrt (host fit) against the oracle: same degree,
    same coefficients to rounding."""
    mu, cc, dd = _lspol_lib(lib, which, max_deg, tol, a, b)
    f = (lambda x: 1.0 / x) if which == "rec" else (lambda x: 1.0 / math.sqrt(x))
    mo, cco, ddo = go.lspol_coeffs(f, a, b, max_deg, tol)
    assert len(mu) == len(mo)
    assert (cc, dd) == (cco, ddo)
    assert np.max(np.abs(mu - mo)) <= 1e-14 * np.max(np.abs(mo))


def _pencil(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    A = A + A.T
    Q = rng.standard_normal((n, n))
    B = Q @ Q.T / n + np.eye(n)
    return A, B


def test_cheb_av_gen_matches_spectral_form():
    """p(A B^{-1}) v = B X p(Lambda) X^T v with X^T B X = I."""
    A, B = _pencil(40, 3)
    lam, X = sla.eigh(A, B)
    cc, dd = 0.5 * (lam[-1] + lam[0]), 0.5 * (lam[-1] - lam[0]) * 1.01
    mu = np.random.default_rng(4).standard_normal(25)
    v = np.random.default_rng(5).standard_normal(40)
    y = go.cheb_av_gen(A, lambda r: np.linalg.solve(B, r), mu, cc, dd, v)
    p = np.polynomial.chebyshev.chebval((lam - cc) / dd, mu)
    ref = B @ (X @ (p * (X.T @ v)))
    assert np.linalg.norm(y - ref) < 1e-11 * np.linalg.norm(ref)


def test_kpm_gen_dots_match_spectral_form():
    """<v_{k+1}, B w> = sum_j (x_j^T B w)^2 T_{k+1}((lambda_j - ctr)/wid)."""
    A, B = _pencil(30, 7)
    lam, X = sla.eigh(A, B)
    ctr, wid = 0.5 * (lam[-1] + lam[0]), 0.5 * (lam[-1] - lam[0]) * 1.05
    ev, U = np.linalg.eigh(B)
    Bmh = U @ np.diag(ev ** -0.5) @ U.T
    vecs = np.random.default_rng(8).uniform(-1, 1, (3, 30))
    dots = go.kpm_gen_dots(A, B, lambda r: np.linalg.solve(B, r), lambda r: Bmh @ r,
                           vecs, 12, ctr, wid)
    for m, v in enumerate(vecs):
        w = Bmh @ (v / np.linalg.norm(v))
        c = (X.T @ (B @ w)) ** 2
        x = (lam - ctr) / wid
        for k in range(12):
            tk = np.cos((k + 1) * np.arccos(np.clip(x, -1, 1)))
            assert abs(dots[m, k] - c @ tk) < 1e-11 * max(1.0, np.sum(c))


@pytest.mark.parametrize("deg", [1, 3, 6])
def test_chebiter_solv2_residual_polynomial(deg):
    """After k steps from x_0 = 0, Chebyshev iteration leaves the residual
    r_k = T_k((theta - B)/delta) / T_k(theta/delta) b (Saad 2003, sec. 12.3)."""
    A, _ = _pencil(25, 11)
    ev, U = np.linalg.eigh(A @ A.T + np.eye(25))
    lb, ub = ev[0] * 0.9, ev[-1] * 1.1
    B = U @ np.diag(ev) @ U.T
    b = np.random.default_rng(12).standard_normal(25)
    x = go.chebiter_solv2(B, b, lb, ub, deg)
    theta, delta = 0.5 * (ub + lb), 0.5 * (ub - lb)
    tk = np.polynomial.chebyshev.chebval((theta - ev) / delta, [0] * deg + [1])
    tk0 = np.polynomial.chebyshev.chebval(theta / delta, [0] * deg + [1])
    r = U @ ((tk / tk0) * (U.T @ b))
    assert np.linalg.norm((b - B @ x) - r) < 1e-12 * np.linalg.norm(b)


def test_chebiter_both_forms_converge_to_the_solve():
    A, _ = _pencil(30, 13)
    ev, U = np.linalg.eigh(A @ A.T / 30 + np.eye(30))
    B = U @ np.diag(ev) @ U.T
    b = np.random.default_rng(14).standard_normal(30)
    ex = np.linalg.solve(B, b)
    for f in (go.chebiter_solv1, go.chebiter_solv2):
        x = f(B, b, ev[0], ev[-1], 200)
        assert np.linalg.norm(x - ex) < 1e-12 * np.linalg.norm(ex)
    assert go.chebiter_solv1(B, b, ev[0], ev[-1], 0) is None


def test_landos_gen_full_lanczos_is_the_spectral_measure():
    """With msteps = n the Lanczos quadrature of the pencil is exact: nodes =
    eigenvalues of (A, B), weights = (x_j^T B v)^2 for the B-normalised start
    v = L^{-T} r; the DOS is then the sum of the weighted Gaussians."""
    A, B = _pencil(20, 21)
    lam, X = sla.eigh(A, B)
    ev, U = np.linalg.eigh(B)
    Bmh = U @ np.diag(ev ** -0.5) @ U.T
    r = np.random.default_rng(22).uniform(-1, 1, (2, 20))
    intv = [lam[0], lam[-1], lam[0] - 0.1, lam[-1] + 0.1]
    xd, yd = go.landos_gen(A, B, lambda q: np.linalg.solve(B, q), lambda q: Bmh @ q, r, 20,
                           300, intv)
    M = 20
    sigma = (intv[3] - intv[2]) / (M - 1) / math.sqrt(8 * math.log(1.25))
    ref = np.zeros_like(xd)
    for q in r:
        v = Bmh @ q
        v = v / math.sqrt(v @ B @ v)
        w = (X.T @ (B @ v)) ** 2
        assert abs(w.sum() - 1) < 1e-12
        for lj, wj in zip(lam, w):
            sel = np.abs(xd - lj) < sigma * math.sqrt(-2 * math.log(1e-8))
            ref[sel] += wj * np.exp(-((xd[sel] - lj) ** 2) / (2 * sigma * sigma))
    ref /= 2 * math.sqrt(2 * sigma * sigma * math.pi)
    assert np.max(np.abs(yd - ref)) < 1e-10 * np.max(ref)
