"""Dense numpy restatement (FP64) of the generalized-problem pieces of the
path: the least-squares Chebyshev fit of the B-solves, the generalized
Chebyshev filter and the generalized KPM moments.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check results.  The product
(pevsl_amd) never imports it.

How each function is pinned (tests/test_gen_oracle.py):
  lspol_coeffs   closed-form Chebyshev series of 1/(cc + dd t); a degree-200
                 Chebyshev interpolant of t -> (cc + dd t)^(-1/2); the
                 degree rule recomputed with numpy.polynomial.chebyshev
  cheb_av_gen    the spectral form p(A B^{-1}) = B X p(Lambda) X^T of the
                 pencil (X B-orthonormal eigenvectors, scipy.linalg.eigh)
  kpm_gen_dots   the spectral form w^T B X T_k(Lambda) X^T B w
"""
import math

import numpy as np


def lspol_coeffs(f, a, b, max_deg, tol):
    """Chebyshev coefficients of f on [a, b] by Gauss-Chebyshev quadrature
    on 4 max_deg nodes; the degree grows until the sum of the relative
    errors at 50 max_deg equispaced points of [-1, 1] is below tol
    (SRC/lspolapprox.c:63-146).  Returns (mu[0..deg], cc, dd)."""
    cc, dd = (a + b) / 2.0, (b - a) / 2.0
    npts = 4 * max_deg
    th = (2 * np.arange(npts) + 1) * (math.pi / (2 * npts))
    g = np.array([f(cc + dd * math.cos(t)) for t in th])
    m = 50 * max_deg
    h = 2.0 / (m - 1) if m > 1 else 0.0  # linspace of SRC/vector.c:15-25
    xs = np.array([-1.0 + i * h for i in range(m)])
    xs[0], xs[-1] = -1.0, 1.0
    fx = np.array([f(cc + dd * x) for x in xs])
    mu = np.zeros(max(max_deg, 1))
    k = 0
    while k < max_deg:
        mu[k] = (1.0 if k == 0 else 2.0) / npts * float(np.sum(np.cos(k * th) * g))
        err = float(np.sum(np.abs((cheb_series(mu[:k + 1], xs) - fx) / fx)))
        if err < tol:
            k += 1
            break
        k += 1
    deg = k - 1
    return mu[:deg + 1].copy(), cc, dd


def cheb_series(mu, x):
    """sum_j mu_j T_j(x) by the three-term recurrence (chebxPltd,
    SRC/chebpol.c:100-131)."""
    x = np.asarray(x, dtype=np.float64)
    tkm1, tk = np.zeros_like(x), np.ones_like(x)
    y = mu[0] * np.ones_like(x)
    for k in range(len(mu) - 1):
        tkp1 = (1.0 if k == 0 else 2.0) * x * tk - tkm1
        tkm1, tk = tk, tkp1
        y = y + mu[k + 1] * tk
    return y


def lspol_apply(B, mu, cc, dd, b):
    """sum_k mu_k T_k((B - cc)/dd) b (pnav, SRC/lspolapprox.c:245-288)."""
    vk, vkm1 = b.copy(), np.zeros_like(b)
    y = mu[0] * vk
    for k in range(1, len(mu)):
        t = (1.0 if k == 1 else 2.0) / dd
        vkp1 = (B @ vk - cc * vk) * t - (vkm1 if k > 1 else 0.0)
        y = y + mu[k] * vkp1
        vkm1, vk = vk, vkp1
    return y


def cheb_av_gen(A, solveB, mu, cc, dd, v):
    """Generalized ChebAv (SRC/chebpol.c:509-574 with ifGenEv): per degree
    w2 = B \\ v_k, v_{k+1} = t (A w2 - cc v_k) - v_{k-1}, y += mu_k v_{k+1}."""
    vk, vkm1 = v.copy(), np.zeros_like(v)
    y = mu[0] * vk
    for k in range(1, len(mu)):
        t = (1.0 if k == 1 else 2.0) / dd
        vkp1 = (A @ solveB(vk) - cc * vk) * t - (vkm1 if k > 1 else 0.0)
        y = y + mu[k] * vkp1
        vkm1, vk = vk, vkp1
    return y


def kpm_gen_dots(A, B, solveB, solveLT, vecs, Mdeg, ctr, wid):
    """dots[m, k] of the generalized KPM (SRC/kpmdos.c:105-156): for each
    random vector (rows of vecs), v = unit(v), w = L^{-T} v, v = B w,
    v_0 = w, v_{k+1} = scal (B^{-1} A v_k - ctr v_k) - v_{k-1},
    dots[m, k] = <v_{k+1}, v>."""
    out = np.zeros((len(vecs), Mdeg))
    for m, v in enumerate(vecs):
        v = v / np.linalg.norm(v)
        w = solveLT(v)
        v = B @ w
        vk, vkm1 = w.copy(), np.zeros_like(w)
        for k in range(Mdeg):
            scal = (1.0 if k == 0 else 2.0) / wid
            vkp1 = (solveB(A @ vk) - ctr * vk) * scal - vkm1
            vkm1, vk = vk, vkp1
            out[m, k] = vk @ v
    return out


def chebiter_solv1(B, b, lb, ub, deg):
    """Chebyshev iteration, Templates form (EXTERNAL/ITERSOL/chebiter.c:
    121-200): x_1 = (2/d) b, then p = r + bet p, x += alp p with
    bet = (c alp / 2)^2, alp = 1/(d - bet), d and c the centre and half
    width of [lb, ub].  Returns None for deg < 1 (x left untouched)."""
    if deg < 1:
        return None
    d, c = (ub + lb) * 0.5, (ub - lb) * 0.5
    alp = 2.0 / d
    p = b.copy()
    x = alp * p
    r = b - B @ x
    for _ in range(1, deg):
        t = c * alp * 0.5
        bet = t * t
        alp = 1.0 / (d - bet)
        p = bet * p + r
        x = x + alp * p
        r = b - B @ x
    return x


def chebiter_solv2(B, b, lb, ub, deg):
    """Chebyshev iteration of Saad, Iterative Methods for Sparse Linear
    Systems (2nd ed.), Algorithm 12.1 (EXTERNAL/ITERSOL/chebiter.c:204-270),
    zero initial guess."""
    theta, delta = (ub + lb) * 0.5, (ub - lb) * 0.5
    sigma = theta / delta
    rho = 1.0 / sigma
    x = np.zeros_like(b)
    r = b.copy()
    d = r / theta
    for _ in range(deg):
        x = x + d
        r = r - B @ d
        rho1 = 1.0 / (2.0 * sigma - rho)
        d = rho1 * rho * d + (2.0 * rho1 / delta) * r
        rho = rho1
    return x


def landos_gen(A, B, solveB, solveLT, vecs, msteps, npts, intv, sigma_mult=1.0):
    """Generalized Lanczos DOS (SRC/landos.c:40-314 with ifGenEv): per start
    vector r (rows of vecs), v = L^{-T} r normalised in the B-norm, msteps
    Lanczos steps with Z = B V, CGS_DGKS2 (two passes) and B-solves; the DOS
    is the sum of Gaussians of width sigma at the Ritz values weighted by the
    squared first eigenvector components.  Returns (xdos, ydos) before the
    Simpson integration; a breakdown before the last step raises (the
    library draws a fresh random vector there)."""
    n = A.shape[0]
    maxit = min(n, msteps)
    lm, lM = intv[2], intv[3]
    aa, bb = max(intv[0], intv[2]), min(intv[1], intv[3])
    M = min(msteps, 30)
    H = (lM - lm) / (M - 1)
    sigma = H / math.sqrt(8 * math.log(1.25)) * sigma_mult
    sigma2 = 2 * sigma * sigma
    width = sigma * math.sqrt(-2.0 * math.log(1e-08))
    h = (bb - aa) / (npts - 1) if npts > 1 else 0.0  # linspace of SRC/vector.c:15-25
    xdos = np.array([aa + i * h for i in range(npts)])
    xdos[0], xdos[-1] = aa, bb
    y = np.zeros(npts)
    for r in vecs:
        v = solveLT(np.asarray(r, dtype=np.float64))
        z = B @ v
        t = 1.0 / math.sqrt(v @ z)
        V, Z = [v * t], [z * t]
        dT, eT = np.zeros(maxit), np.zeros(maxit)
        beta, wn, nwn = 0.0, 0.0, 0
        for k in range(maxit):
            znew = A @ V[k]
            if k > 0:
                znew = znew - beta * Z[k - 1]
            alpha = V[k] @ znew
            dT[k] = alpha
            wn += abs(alpha)
            znew = znew - alpha * Z[k]
            Vm, Zm = np.array(V).T, np.array(Z).T
            for _ in range(2):
                znew = znew - Zm @ (Vm.T @ znew)
            vnew = solveB(znew)
            beta = math.sqrt(vnew @ znew)
            wn += 2.0 * beta
            nwn += 3
            if beta * nwn < 1e-14 * wn:
                if k < maxit - 1:
                    raise RuntimeError("breakdown at step %d" % k)
                beta = 0.0
                vnew, znew = np.zeros(n), np.zeros(n)
            else:
                vnew, znew = vnew / beta, znew / beta
            eT[k] = beta
            V.append(vnew)
            Z.append(znew)
        th, S = np.linalg.eigh(np.diag(dT) + np.diag(eT[:-1], 1) + np.diag(eT[:-1], -1))
        g2 = S[0, :] ** 2
        for i in range(maxit):
            sel = np.abs(xdos - th[i]) < width
            y[sel] += g2[i] * np.exp(-((xdos[sel] - th[i]) ** 2) / sigma2)
    y *= 1.0 / (len(vecs) * math.sqrt(sigma2 * math.pi))
    return xdos, y
