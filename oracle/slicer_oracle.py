"""CPU restatement (numpy, FP64) of the learned slicer: the network's forward
and backward passes, its loss head, the Adam update, the KPM bin masses, the
equal-share cuts of a binned DOS, the feature map and the full-KPM judge.

TEST INFRASTRUCTURE ONLY: imported by tests/ to check results.  The product
(pevsl_amd) never imports it.

The slicer has no counterpart in the reference; the DOS convention it builds
on is pEVSL_Kpmdos's (SRC/kpmdos.c:104-157: mu_0 = 1/pi, fraction of the
spectrum below x = mu_0 (pi - acos x) - sum_k mu_k sin(k acos x)/k, the
closed form of SRC/kpmdos.c:127,155,182 and intChx, SRC/spslicer.c:12-38).

How each function is pinned (tests/test_slicer_oracle.py):
  mlp_forward / mlp_loss_grad   central finite differences of the loss, and
                                torch autograd in float64
  adam_step                     torch.optim.Adam (two steps), and the
                                closed-form first step lr*g/(|g|+eps)
  kpm_cdf / bin_masses          adaptive quadrature of the KPM density
                                sum_k mu_k T_k(x)/sqrt(1-x^2); the oracle C
                                intchx (pinned against the README run)
  cuts_from_bins                brute force on distributions with known
                                quantiles (uniform, one bin, two bins)
  features                      the cheap-KPM CDF features are those of
                                kpm_cdf; moment, interval and Fourier
                                features are direct formulas of the inputs
"""
import math

import numpy as np

NBINS = 512
MDEG = 20
NFREQ = 24
NIN = 42 + 4 * NFREQ


# ---------------------------------------------------------------- network
def layer_shapes(nin, nh, nl, nout):
    dims = [nin] + [nh] * nl + [nout]
    return [(dims[i + 1], dims[i]) for i in range(nl + 1)]


def unpack(w, nin, nh, nl, nout):
    """Packed parameters -> [(W_l, b_l)], W_l of shape (out, in)."""
    out, off = [], 0
    for (o, i) in layer_shapes(nin, nh, nl, nout):
        W = w[off:off + o * i].reshape(o, i)
        off += o * i
        b = w[off:off + o]
        off += o
        out.append((W, b))
    assert off == len(w)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for (W, b) in layers])


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(
        2.0 * math.pi)


def mlp_forward(w, shape, X, Q=None):
    """Raw network output z, plus the cache of the backward pass."""
    layers = unpack(np.asarray(w, np.float64), *shape)
    a = np.asarray(X, np.float64)
    acts, pres = [a], []
    for l, (W, b) in enumerate(layers):
        z = a @ W.T + b
        pres.append(z)
        if l < len(layers) - 1:
            a = gelu(z)
            acts.append(a)
        else:
            a = z
    return a, (layers, acts, pres)


def head(z, Q, T, nsoft):
    """y = z + Q; softmax over the first nsoft; loss (mean) and d loss/d z."""
    y = z + (0.0 if Q is None else np.asarray(Q, np.float64))
    batch = y.shape[0]
    Y = y.copy()
    dz = np.zeros_like(y)
    loss = np.zeros(batch)
    if nsoft:
        s = y[:, :nsoft]
        lse = np.log(np.exp(s - s.max(1, keepdims=True)).sum(1)) + s.max(1)
        lp = s - lse[:, None]
        p = np.exp(lp)
        Y[:, :nsoft] = p
        if T is not None:
            t = np.asarray(T, np.float64)[:, :nsoft]
            loss -= np.where(t > 0, t * lp, 0.0).sum(1)
            dz[:, :nsoft] = (p - t) / batch
    if T is not None and nsoft < y.shape[1]:
        e = y[:, nsoft:] - np.asarray(T, np.float64)[:, nsoft:]
        loss += (e * e).sum(1)
        dz[:, nsoft:] = 2.0 * e / batch
    return Y, loss.mean(), dz


def mlp_loss_grad(w, shape, nsoft, X, Q, T):
    """Mean loss and its gradient with respect to the packed parameters."""
    z, (layers, acts, pres) = mlp_forward(w, shape, X, Q)
    _, loss, d = head(z, Q, T, nsoft)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (d.T @ acts[l], d.sum(0))
        if l:
            d = (d @ W) * gelu_grad(pres[l - 1])
    return loss, pack(grads)


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); step counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------------------- DOS pieces
def kpm_cdf(mu, x):
    """Fraction of the spectrum below x (frame coordinates) from KPM moments."""
    th = np.arccos(np.clip(np.asarray(x, np.float64), -1.0, 1.0))
    k = np.arange(1, len(mu))
    return mu[0] * (np.pi - th) - (np.asarray(mu[1:])[:, None] *
                                   np.sin(np.outer(k, np.atleast_1d(th))) /
                                   k[:, None]).sum(0).reshape(np.shape(th))


def frame_x(intv, t):
    c, h = (intv[3] + intv[2]) / 2.0, (intv[3] - intv[2]) / 2.0
    return np.clip((np.asarray(t, np.float64) - c) / h, -1.0, 1.0)


def bin_masses(mu, intv, nbins):
    a, b = max(intv[0], intv[2]), min(intv[1], intv[3])
    edges = a + (b - a) * np.arange(nbins + 1) / nbins
    edges[-1] = b
    F = kpm_cdf(mu, frame_x(intv, edges))
    return np.diff(F), F[-1] - F[0]


def cuts_from_bins(masses, a, b, n_int):
    """Equal-share cuts of a binned density, linear inside a bin."""
    m = np.maximum(np.asarray(masses, np.float64), 0.0)
    nb = len(m)
    w = (b - a) / nb
    tot = m.sum()
    sli = np.empty(n_int + 1)
    sli[0], sli[-1] = a, b
    if not tot > 0:
        sli[1:-1] = a + (b - a) * np.arange(1, n_int) / n_int
        return sli
    below, k = 0.0, 0
    for j in range(1, n_int):
        target = tot * j / n_int
        while k < nb - 1 and below + m[k] < target:
            below += m[k]
            k += 1
        fr = min(1.0, max(0.0, (target - below) / m[k])) if m[k] > 0 else 0.0
        sli[j] = max(sli[j - 1], a + w * (k + fr))
    return sli


def features(mu, n, intv):
    """Inputs and prior of the slicer network from cheap-KPM moments."""
    mu = np.asarray(mu, np.float64)[:MDEG + 1]
    a, b = max(intv[0], intv[2]), min(intv[1], intv[3])
    xa, xb = float(frame_x(intv, a)), float(frame_x(intv, b))
    f = list(np.pi * mu[1:MDEG + 1])
    f += [xa, xb, math.acos(xa) / np.pi, math.acos(xb) / np.pi,
          10.0 * (math.acos(xa) - math.acos(xb)) / np.pi, math.log10(n) / 6.0]
    fa, fb = float(kpm_cdf(mu, xa)), float(kpm_cdf(mu, xb))
    fr = max(fb - fa, 1e-12)
    f.append(math.log(fr) / 10.0)
    t = a + (b - a) * np.arange(1, 16) / 16.0
    f += list((kpm_cdf(mu, frame_x(intv, t)) - fa) / fr - np.arange(1, 16) / 16.0)
    ta, tb = math.acos(xa), math.acos(xb)
    for j in range(1, NFREQ + 1):
        f += [math.cos(j * ta), math.sin(j * ta), math.cos(j * tb), math.sin(j * tb)]
    m, fr2 = bin_masses(mu, intv, NBINS)
    tot = np.maximum(m, 0).sum()
    q = np.log(np.maximum(m, 1e-9 * max(tot, 1e-300)) / max(tot, 1e-300))
    q = np.concatenate([q, [math.log(max(fr2, 1e-12))]])
    return np.array(f), q


def judge(mu, intv, n, sli):
    """Full-KPM counts per slice and the imbalance max/mean - 1."""
    F = kpm_cdf(mu, frame_x(intv, np.asarray(sli)))
    c = n * np.diff(F)
    return c, c.max() / c.mean() - 1.0
