"""Pins the learned slicer's CPU oracle (oracle/slicer_oracle.py) against
mathematics and independent library routines, and checks the library's host
parts of the slicer (bin masses, cuts, features, judge) against it.  No GPU.
"""
import ctypes as C
import math

import numpy as np
import pytest
import scipy.integrate as si

from oracle import slicer_oracle as so

SMALL = (5, 7, 2, 6)  # nin, nhidden, nlayers, nout
NSOFT = 4


def _problem(seed=0, batch=3, shape=SMALL, nsoft=NSOFT):
    rng = np.random.default_rng(seed)
    nin, nh, nl, nout = shape
    nw = sum(o * i + o for (o, i) in so.layer_shapes(nin, nh, nl, nout))
    w = rng.normal(0, 0.5, nw)
    X = rng.normal(0, 1, (batch, nin))
    Q = rng.normal(0, 0.3, (batch, nout))
    T = np.zeros((batch, nout))
    t = rng.random((batch, nsoft))
    T[:, :nsoft] = t / t.sum(1, keepdims=True)
    T[:, nsoft:] = rng.normal(0, 1, (batch, nout - nsoft))
    return w, X, Q, T


def test_mlp_gradient_matches_finite_differences():
    w, X, Q, T = _problem()
    loss, g = so.mlp_loss_grad(w, SMALL, NSOFT, X, Q, T)
    h = 1e-6
    fd = np.empty_like(w)
    for i in range(len(w)):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd[i] = (so.mlp_loss_grad(wp, SMALL, NSOFT, X, Q, T)[0] -
                 so.mlp_loss_grad(wm, SMALL, NSOFT, X, Q, T)[0]) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))
    # a dropped GELU derivative term would fail this
    assert np.max(np.abs(g)) > 1e-3


def test_mlp_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    w, X, Q, T = _problem(seed=1, batch=4)
    loss, g = so.mlp_loss_grad(w, SMALL, NSOFT, X, Q, T)
    wt = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    nin, nh, nl, nout = SMALL
    a = torch.tensor(X)
    off = 0
    shapes = so.layer_shapes(nin, nh, nl, nout)
    for l, (o, i) in enumerate(shapes):
        W = wt[off:off + o * i].reshape(o, i)
        off += o * i
        b = wt[off:off + o]
        off += o
        a = a @ W.T + b
        if l < len(shapes) - 1:
            a = torch.nn.functional.gelu(a)
    y = a + torch.tensor(Q)
    Tt = torch.tensor(T)
    xent = -(Tt[:, :NSOFT] * torch.log_softmax(y[:, :NSOFT], 1)).sum(1)
    reg = ((y[:, NSOFT:] - Tt[:, NSOFT:]) ** 2).sum(1)
    lt = (xent + reg).mean()
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12 * max(1, abs(loss))
    assert np.max(np.abs(wt.grad.numpy() - g)) < 1e-12 * max(1.0, np.max(np.abs(g)))


def test_adam_matches_torch_and_first_step_closed_form():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    w0 = rng.normal(size=20)
    g1, g2 = rng.normal(size=20), rng.normal(size=20)
    lr = 3e-3
    # first step from zero moments: w - lr * g / (|g| + eps)
    w1, m, v = so.adam_step(w0, g1, np.zeros(20), np.zeros(20), 1, lr)
    assert np.allclose(w1, w0 - lr * g1 / (np.abs(g1) + 1e-8), rtol=0, atol=1e-15)
    w2, m, v = so.adam_step(w1, g2, m, v, 2, lr)
    p = torch.nn.Parameter(torch.tensor(w0))
    opt = torch.optim.Adam([p], lr=lr, betas=(0.9, 0.999), eps=1e-8)
    for g in (g1, g2):
        p.grad = torch.tensor(g)
        opt.step()
    assert np.max(np.abs(p.detach().numpy() - w2)) < 1e-14


def _moments_of(ev, lmin, lmax, M, jackson=True):
    c, h = (lmax + lmin) / 2, (lmax - lmin) / 2
    th = np.arccos(np.clip((ev - c) / h, -1, 1))
    k = np.arange(M + 1)
    mu = np.cos(np.outer(k, th)).mean(1) * 2 / np.pi
    if jackson:
        tj = np.pi / (M + 2)
        jac = (1 / (M + 2)) * np.sin((k + 1) * tj) / np.sin(tj) + \
            (1 - (k + 1) / (M + 2)) * np.cos(k * tj)
        mu = mu * jac
    mu[0] = 1 / np.pi
    return mu


def test_kpm_cdf_matches_quadrature_of_the_density():
    rng = np.random.default_rng(3)
    ev = rng.uniform(1.0, 9.0, 400)
    mu = _moments_of(ev, 0.0, 12.0, 30)

    def dens_no_edge(t):  # sum mu_k T_k(t) / sqrt(1 - t); weight (1+t)^-1/2
        return np.polynomial.chebyshev.chebval(t, mu) / math.sqrt(1 - t)

    for x in (-0.9, -0.3, 0.0, 0.45, 0.8):
        q, _ = si.quad(dens_no_edge, -1.0, x, weight="alg", wvar=(-0.5, 0.0), limit=200)
        assert abs(float(so.kpm_cdf(mu, x)) - q) < 1e-10
    # total mass 1 and mu_0 convention of the reference (mu_0 = 1/pi)
    assert abs(float(so.kpm_cdf(mu, 1.0)) - 1.0) < 1e-14


def test_kpm_cdf_matches_oracle_intchx():
    from oracle import oracle as orc
    rng = np.random.default_rng(4)
    mu = _moments_of(rng.uniform(0.5, 11, 300), 0, 12, 60)
    x = np.linspace(-0.95, 0.9, 9)
    y = orc.intchx(mu, x)
    F = so.kpm_cdf(mu, x)
    assert np.max(np.abs((F - F[0]) - y)) < 1e-12


def test_bin_masses_sum_and_per_bin_quadrature():
    mu = _moments_of(np.random.default_rng(5).uniform(2, 10, 500), 0, 12, 40)
    intv = [3.0, 5.0, 0.0, 12.0]
    m, fr = so.bin_masses(mu, intv, 16)
    assert abs(m.sum() - fr) < 1e-14
    for i in (0, 7, 15):
        lo, hi = 3.0 + 2.0 * i / 16, 3.0 + 2.0 * (i + 1) / 16
        q, _ = si.quad(lambda t: np.polynomial.chebyshev.chebval((t - 6) / 6, mu) /
                       math.sqrt(1 - ((t - 6) / 6) ** 2) / 6, lo, hi)
        assert abs(m[i] - q) < 1e-12


def test_cuts_from_bins_on_known_quantiles():
    # uniform density -> uniform cuts
    s = so.cuts_from_bins(np.ones(64), 1.0, 3.0, 8)
    assert np.allclose(s, np.linspace(1, 3, 9), atol=1e-14)
    # all mass in bin 5 of 10 over [0, 10): cuts inside [5, 6], evenly spaced
    m = np.zeros(10)
    m[5] = 2.0
    s = so.cuts_from_bins(m, 0.0, 10.0, 4)
    assert np.allclose(s[1:-1], [5.25, 5.5, 5.75], atol=1e-14)
    # two equal bins at the ends: the middle cut sits at the end of bin 0
    m = np.zeros(8)
    m[0] = m[7] = 1.0
    s = so.cuts_from_bins(m, 0.0, 8.0, 2)
    assert abs(s[1] - 1.0) < 1e-14
    # brute force: a piecewise-constant density, quantiles of its CDF
    rng = np.random.default_rng(6)
    m = rng.random(50)
    s = so.cuts_from_bins(m, -1.0, 4.0, 7)
    grid = np.linspace(-1, 4, 500001)
    cdf = np.interp(grid, np.linspace(-1, 4, 51), np.concatenate([[0], np.cumsum(m)]))
    for j in range(1, 7):
        q = grid[np.searchsorted(cdf, m.sum() * j / 7)]
        assert abs(q - s[j]) < 2e-5


def test_judge_counts_of_exact_moments():
    # with many moments of a discrete spectrum, the judge's counts approach
    # the true counts (Jackson kernel resolution ~ pi/M)
    ev = np.sort(np.random.default_rng(7).uniform(0.0, 12.0, 20000))
    mu = _moments_of(ev, 0, 12, 2000)
    sli = np.array([2.0, 3.0, 4.5, 6.0])
    c, imb = so.judge(mu, [2.0, 6.0, 0.0, 12.0], len(ev), sli)
    exact = np.histogram(ev, bins=sli)[0]
    assert np.max(np.abs(c - exact)) < 0.02 * exact.max()
    assert abs(imb - (c.max() / c.mean() - 1)) < 1e-15


# ------------------------------------------------ library host functions
@pytest.fixture(scope="module")
def hostlib():
    from pevsl_amd import capi
    return capi.load()


def test_library_bin_masses_match_oracle(hostlib):
    from pevsl_amd import slicer
    rng = np.random.default_rng(8)
    mu = _moments_of(rng.uniform(0.3, 11, 1000), 0, 12, 300)
    for iv in ([0.6, 1.2, 0, 12], [4.0, 7.5, 0, 12], [-1.0, 0.4, 0, 12], [11.0, 13.0, 0, 12]):
        m, fr = slicer.bin_masses(hostlib, mu, iv, 512)
        mo, fro = so.bin_masses(mu, iv, 512)
        assert np.max(np.abs(m - mo)) < 1e-13
        assert abs(fr - fro) < 1e-13


def test_library_cuts_match_oracle(hostlib):
    from pevsl_amd import slicer
    rng = np.random.default_rng(9)
    for trial in range(20):
        nb = int(rng.integers(1, 600))
        m = rng.random(nb) * (rng.random(nb) < 0.7)
        a, b = sorted(rng.uniform(-3, 9, 2))
        k = int(rng.integers(1, 25))
        s = slicer.cuts_from_bins(hostlib, m, a, b, k)
        so_ = so.cuts_from_bins(m, a, b, k)
        assert np.max(np.abs(s - so_)) < 1e-12 * max(1, abs(b))
        assert np.all(np.diff(s) >= 0)


def test_library_features_match_oracle(hostlib):
    from pevsl_amd import slicer
    rng = np.random.default_rng(10)
    mu = _moments_of(rng.uniform(0.0, 12, 5000), 0, 13.0, 20)
    for iv in ([0.6, 1.2, 0, 13.0], [5.0, 8.0, 0, 13.0], [12.0, 14.0, 0, 13.0]):
        f, q = slicer.features_from_mu(hostlib, mu, 125000, iv)
        fo, qo = so.features(mu, 125000, iv)
        assert f.shape == (so.NIN,) and q.shape == (so.NBINS + 1,)
        assert np.max(np.abs(f - fo)) < 1e-5
        assert np.max(np.abs(q - qo)) < 1e-4 * max(1, np.max(np.abs(qo)))


def test_library_judge_matches_oracle(hostlib):
    from pevsl_amd import slicer
    mu = _moments_of(np.random.default_rng(11).uniform(0, 12, 3000), 0, 12, 300)
    iv = [2.0, 6.0, 0, 12]
    sli = np.array([2.0, 2.7, 3.9, 5.0, 6.0])
    c, imb = slicer.judge(hostlib, mu, iv, 1e6, sli)
    co, imbo = so.judge(mu, iv, 1e6, sli)
    assert np.max(np.abs(c - co)) < 1e-7
    assert abs(imb - imbo) < 1e-12
