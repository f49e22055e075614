"""Pins the CPU oracle (oracle/) against the reference's own known answers
and against mathematics that a plausible mistake would violate.  CPU only."""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from oracle import oracle as orc
from pevsl_amd import driver

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_laplacian_matches_closed_form_spectrum():
    # dense spectrum of the generated matrix == ExactEigLap3 closed form
    # (TESTS/COMMON/lapl.c:139-169) on a small non-cubic grid
    A = orc.Csr(lap=(4, 3, 5))
    n = A.n
    M = sp.csr_matrix((A.a, A.ja, A.ia), shape=(n, n)).toarray()
    assert np.allclose(M, M.T)
    ev = np.linalg.eigvalsh(M)
    ex = orc.exact_eig_lap3(4, 3, 5, -1.0, 13.0)
    assert len(ex) == n
    np.testing.assert_allclose(ev, ex, rtol=0, atol=1e-12)


def test_laplacian_2d_case_and_row_order():
    # nz == 1 -> 5-point stencil with diagonal 4 (TESTS/COMMON/lapl.c:20-21);
    # row entries ordered -z,-y,-x,diag,+x,+y,+z (columns increasing)
    A = orc.Csr(lap=(5, 4, 1))
    assert A.n == 20 and A.ia[-1] == 5 * 20 - 2 * 5 - 2 * 4
    d = [A.a[A.ia[r]:A.ia[r + 1]][A.ja[A.ia[r]:A.ia[r + 1]] == r][0] for r in range(20)]
    assert set(d) == {4.0}
    for r in range(A.n):
        cols = A.ja[A.ia[r]:A.ia[r + 1]]
        assert np.all(np.diff(cols) > 0)
    ev = np.linalg.eigvalsh(sp.csr_matrix((A.a, A.ja, A.ia), shape=(20, 20)).toarray())
    np.testing.assert_allclose(ev, orc.exact_eig_lap3(5, 4, 1, -1, 9), atol=1e-12)


def test_product_laplacian_generator_is_identical():
    A = orc.Csr(lap=(7, 6, 5))
    ia, ja, a = driver.laplacian3d(7, 6, 5)
    assert np.array_equal(A.ia, ia) and np.array_equal(A.ja, ja)
    assert np.array_equal(A.a, a)


def test_csrmv_sin_recipe():
    # TESTS0/Matvec/test1.c:70-94: x_i = sin(i), compare with a sequential
    # CSR product; one sequential sum per row -> bitwise equal to scipy's
    A = orc.Csr(lap=(20, 20, 20))
    x = np.sin(np.arange(A.n, dtype=np.float64))
    y = A.matvec(x)
    ys = sp.csr_matrix((A.a, A.ja, A.ia), shape=(A.n, A.n)) @ x
    assert np.max(np.abs(y - ys)) <= 1e-13
    # integer stencil: interior row = 6 x_i - sum of 6 neighbours
    i = 20 * 20 * 10 + 20 * 10 + 10
    nb = [i - 400, i - 20, i - 1, i + 1, i + 20, i + 400]
    assert abs(y[i] - (6 * x[i] - sum(x[j] for j in nb))) < 1e-14


def test_tridiag_eig_vs_lapack():
    rng = np.random.default_rng(1)
    for n in (1, 2, 7, 60):
        d = rng.uniform(-1, 1, n)
        e = rng.uniform(-1, 1, max(n - 1, 0))
        w, z = orc.tridiag_eig(d, e, vectors=True)
        w2 = sla.eigh_tridiagonal(d, e, eigvals_only=True)
        np.testing.assert_allclose(w, w2, rtol=0, atol=1e-13)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        assert np.max(np.abs(T @ z - z * w)) < 1e-12
        assert np.max(np.abs(z.T @ z - np.eye(n))) < 1e-12
        assert np.array_equal(orc.tridiag_eig(d, e), w)


def test_tridiag_eig_multiple_eigenvalues():
    # block-diagonal T with repeated eigenvalues (Lanczos on the Laplacian
    # produces near-multiple Ritz values)
    d = np.array([2.0, 2.0, 2.0, 5.0, 5.0])
    e = np.array([0.0, 0.0, 1e-300, 0.0])
    w = orc.tridiag_eig(d, e)
    np.testing.assert_allclose(w, [2, 2, 2, 5, 5], atol=1e-15)


def test_dampcf_against_textbook_kernels():
    # Jackson kernel, Weisse et al. Rev. Mod. Phys. 78 (2006) eq. (71), N=m+1:
    # g_k = [(N-k+1) cos(pi k/(N+1)) + sin(pi k/(N+1)) cot(pi/(N+1))]/(N+1)
    for m in (1, 5, 300):
        jac = orc.dampcf(m, 1)
        N = m + 1
        k = np.arange(1, m + 1)
        g = ((N - k + 1) * np.cos(np.pi * k / (N + 1)) +
             np.sin(np.pi * k / (N + 1)) / np.tan(np.pi / (N + 1))) / (N + 1)
        assert jac[0] == 0.5
        np.testing.assert_allclose(jac[1:], g, rtol=1e-12, atol=1e-14)
        # Lanczos sigma factors sinc(k pi/(m+1))
        sig = orc.dampcf(m, 2)
        np.testing.assert_allclose(sig[1:], np.sinc(k / (m + 1.0)), rtol=1e-13)
        assert np.all(orc.dampcf(m, 0)[1:] == 1.0)


def test_chebxPltd_is_cosine_series():
    rng = np.random.default_rng(2)
    mu = rng.normal(size=41)
    x = np.linspace(-1, 1, 33)
    y = orc.chebxPltd(mu, x)
    ref = np.polynomial.chebyshev.chebval(x, mu)
    np.testing.assert_allclose(y, ref, atol=1e-12)
    th = np.arccos(x)
    ref2 = (mu[:, None] * np.cos(np.arange(41)[:, None] * th[None, :])).sum(0)
    np.testing.assert_allclose(y, ref2, atol=1e-12)


def test_findpol_readme_degrees():
    # README.md:111-115: slices of the published run -> polydeg 61..110
    g = json.load(open(os.path.join(GOLD, "readme_c1.json")))
    sli = g["sli"]
    for s in g["slices"]:
        pol = orc.Pol([sli[s["slice"]], sli[s["slice"] + 1], 0.0, 12.0])
        assert pol.rc == 0 and pol.type == 0
        assert pol.deg == s["deg"], (s, pol.deg)


def test_findpol_invariants():
    for intv in ([0.6, 0.75, 0, 12], [3.0, 3.2, 0, 12], [0.0, 0.3, 0, 12],
                 [11.5, 12.0, 0, 12], [0.6, 1.2, 0, 12]):
        pol = orc.Pol(intv)
        assert pol.rc == 0
        cc, dd = pol.cc, pol.dd
        assert cc == 6.0 and dd == 6.0
        a = (max(intv[0], intv[2]) - cc) / dd
        b = (min(intv[1], intv[3]) - cc) / dd
        if pol.type == 0:
            # p(gam) = 1 after scaling, p(a) = p(b) (balanced), bar = p(edge)
            assert abs(orc.chebxPltd(pol.mu, [pol.gam])[0] - 1.0) < 1e-12
            pa, pb = orc.chebxPltd(pol.mu, [a, b])
            assert abs(pa - pb) < 1e-9
            assert abs(min(pa, pb) - pol.bar) < 1e-12
            assert pol.bar <= 0.8 + 1e-12
            # p below bar outside the slice, p peaks inside it
            out = np.concatenate([np.linspace(-1, a - 0.05, 50), np.linspace(b + 0.05, 1, 50)])
            assert np.max(orc.chebxPltd(pol.mu, out)) < pol.bar
        else:
            assert pol.type == (1 if intv[0] <= 0 else 2)
            assert abs(pol.gam) == 1.0
            # value at the inner edge is the threshold bar, below thresh_ext
            edge = b if pol.type == 1 else a
            assert abs(orc.chebxPltd(pol.mu, [edge])[0] - pol.bar) < 1e-10
            assert pol.bar < 0.2


def test_findpol_invalid_interval():
    assert orc.Pol([1.0, 0.5, 0, 12]).rc == -1
    assert orc.Pol([13.0, 14.0, 0, 12]).rc == -1


def test_cheb_av_diagonal_matrix_closed_form():
    lam = np.linspace(0.1, 11.9, 50)
    n = len(lam)
    A = orc.Csr(np.arange(n + 1), np.arange(n), lam)
    pol = orc.Pol([3.0, 3.4, 0.0, 12.0])
    v = np.random.default_rng(3).normal(size=n)
    y = orc.cheb_av(A, pol, v)
    x = (lam - 6.0) / 6.0
    p = np.cos(np.arange(pol.deg + 1)[:, None] * np.arccos(x)[None, :])
    ref = (pol.mu[:, None] * p).sum(0) * v
    np.testing.assert_allclose(y, ref, atol=1e-12)


def test_cgs_dgks_orthogonalizes():
    rng = np.random.default_rng(4)
    n, k = 300, 40
    Q, _ = np.linalg.qr(rng.normal(size=(n, k)))
    v = rng.normal(size=n) + Q @ rng.normal(size=k) * 100
    w, nrm = orc.cgs_dgks(Q, v)
    assert np.max(np.abs(Q.T @ w)) < 1e-13 * np.linalg.norm(w)
    assert abs(nrm - np.linalg.norm(w)) < 1e-13 * nrm
    # the component in span(Q) is removed, the rest kept
    ref = v - Q @ (Q.T @ v)
    np.testing.assert_allclose(w, ref, atol=1e-11)


def test_libc_rand_stream():
    # (rand() - RAND_MAX/2)/(RAND_MAX/2) in [-1, 1], reproducible after srand
    orc.srand(0)
    a = orc.rand_vec(1000)
    orc.srand(0)
    b = orc.rand_vec(1000)
    assert np.array_equal(a, b) and np.all(np.abs(a) <= 1.0)
    assert abs(np.mean(a)) < 0.1 and 0.25 < np.var(a) < 0.42


def test_kpm_moments_against_spectral_sum():
    # diagonal A: v^T T_k(A') v = sum_i v_i^2 cos(k acos(l_i')) -- evaluated
    # here by the closed form instead of the three-term recurrence
    lam = np.linspace(0.05, 11.95, 200)
    n = len(lam)
    A = orc.Csr(np.arange(n + 1), np.arange(n), lam)
    Mdeg, nvec = 40, 5
    orc.srand(7)
    rc, mu, ecnt = orc.kpmdos(A, Mdeg, 1, nvec, [2.0, 4.0, 0.0, 12.0])
    assert rc == 0
    orc.srand(7)
    V = np.array([orc.rand_vec(n) for _ in range(nvec)])
    V /= np.linalg.norm(V, axis=1)[:, None]
    th = np.arccos((lam - 6.0) / 6.0)
    T = np.cos(np.arange(Mdeg + 1)[:, None] * th[None, :])
    jac = orc.dampcf(Mdeg, 1)
    jac[0] = 1.0
    fac = np.where(np.arange(Mdeg + 1) == 0, 1.0, 2.0) * jac
    ref = fac * (T @ (V ** 2).sum(0)) / (nvec * np.pi)
    np.testing.assert_allclose(mu, ref, rtol=1e-10, atol=1e-14)
    assert abs(mu[0] - 1.0 / np.pi) < 1e-15


@pytest.mark.parametrize("a,b,damping", [(2.0, 4.0, 1), (0.0, 12.0, 1), (5.5, 5.6, 1),
                                         (-1.0, 3.0, 1), (2.0, 4.0, 0), (2.0, 4.0, 2)])
def test_kpm_count_is_the_integral_of_its_dos(a, b, damping):
    """The count the KPM returns (accumulated from sine terms during the
    recurrence, SRC/kpmdos.c:127,155,182) equals n times the integral of the
    DOS its own moments define over [a, b] clipped to the frame, computed by
    the independently pinned intchx quadrature.  A dropped term, a wrong sign
    or a wrong damping factor in either formula fails this."""
    lam = np.linspace(0.05, 11.95, 200)
    n = len(lam)
    A = orc.Csr(np.arange(n + 1), np.arange(n), lam)
    orc.srand(7)
    rc, mu, ecnt = orc.kpmdos(A, 40, damping, 5, [a, b, 0.0, 12.0])
    assert rc == 0
    xa, xb = (max(a, 0.0) - 6.0) / 6.0, (min(b, 12.0) - 6.0) / 6.0
    y = orc.intchx(mu, np.array([xa, xb]))
    assert abs(ecnt - n * y[1]) <= 1e-8 * max(1.0, abs(ecnt))


def test_kpm_count_c1_within_stochastic_error():
    A = orc.Csr(lap=(20, 20, 20))
    orc.srand(0)
    rc, mu, ecnt = orc.kpmdos(A, 300, 1, 60, [0.6, 1.2, 0.0, 12.0])
    assert rc == 0
    exact = len(orc.exact_eig_lap3(20, 20, 20, 0.6, 1.2))
    assert exact == 118
    assert abs(ecnt - exact) < 15


def test_intchx_matches_quadrature_of_dos():
    rng = np.random.default_rng(5)
    mu = rng.normal(size=11) * np.exp(-np.arange(11) / 3.0)
    xi = np.linspace(-0.9, 0.7, 9)
    y = orc.intchx(mu, xi)
    # integral of sum mu_k T_k(t)/sqrt(1-t^2) from xi[0]: t = cos(th)
    for j in range(len(xi)):
        ths = np.linspace(np.arccos(xi[j]), np.arccos(xi[0]), 20001)
        f = (mu[:, None] * np.cos(np.arange(11)[:, None] * ths[None, :])).sum(0)
        ref = np.trapezoid(f, ths)
        assert abs(y[j] - ref) < 1e-7


def test_spslicer_equal_shares():
    A = orc.Csr(lap=(20, 20, 20))
    orc.srand(0)
    rc, mu, ecnt = orc.kpmdos(A, 300, 1, 60, [0.6, 1.2, 0.0, 12.0])
    npts = int(10 * ecnt)
    rc, sli = orc.spslicer_kpm(mu, [0.6, 1.2, 0.0, 12.0], 5, npts)
    assert rc == 0
    assert sli[0] == 0.6 and sli[-1] == 1.2 and np.all(np.diff(sli) > 0)
    xs = (sli - 6.0) / 6.0
    Y = orc.intchx(mu, xs)
    share = np.diff(Y)
    step = (1.2 - 0.6) / 6.0 / (npts - 1)
    dmax = np.max(np.abs(orc.intchx(mu, np.linspace(xs[0], xs[-1], npts))[1:] -
                         orc.intchx(mu, np.linspace(xs[0], xs[-1], npts))[:-1]))
    assert np.max(np.abs(share - Y[-1] / 5)) <= 5 * dmax + 1e-12
    assert step > 0


def test_spslicer_trivial_and_invalid():
    mu = np.zeros(11)
    rc, sli = orc.spslicer_kpm(mu, [0.6, 1.2, 0.0, 12.0], 1, 100)
    assert rc == 0 and list(sli) == [0.6, 1.2]
    rc, _ = orc.spslicer_kpm(mu, [1.2, 0.6, 0.0, 12.0], 3, 100)
    assert rc == -1


def test_simpson_exact_on_cubics():
    x = np.linspace(-1.0, 2.0, 31)
    f = 3 * x ** 3 - x ** 2 + 2 * x - 5
    F = lambda t: 0.75 * t ** 4 - t ** 3 / 3 + t ** 2 - 5 * t  # noqa: E731
    y = orc.simpson(x, f)
    # even-indexed points are Simpson sums: exact for cubics
    np.testing.assert_allclose(y[::2], F(x[::2]) - F(x[0]), atol=1e-11)


def test_cheblan_readme_slices_counts():
    # README.md:111-115: on the published slices the counts are 18/27/27/19/27
    g = json.load(open(os.path.join(GOLD, "readme_c1.json")))
    A = orc.Csr(lap=(20, 20, 20))
    orc.srand(0)
    vinit = orc.rand_vec(A.n)
    sli = g["sli"]
    for s in g["slices"][:2]:
        iv = [sli[s["slice"]], sli[s["slice"] + 1], 0.0, 12.0]
        pol = orc.Pol(iv)
        r = orc.cheblan_nr(A, iv, 300, 1e-8, vinit, pol)
        assert r["rc"] == 0
        assert len(r["lam"]) == s["found"] == s["exact"]
        ex = orc.exact_eig_lap3(20, 20, 20, iv[0], iv[1])
        np.testing.assert_allclose(np.sort(r["lam"]), ex, rtol=1e-12)
        assert np.all(r["res"] < 1e-8)
        assert r["its"] % 20 == 11 and r["its"] <= 300  # tests at k = 30, 50, ...
        # the README prints 210/270 (2 ranks per slice, another start vector);
        # with this start vector the trace test first holds at k = 150
        assert r["its"] == 151


def test_cheblan_vectors_are_eigenvectors():
    A = orc.Csr(lap=(10, 10, 10))
    orc.srand(0)
    vinit = orc.rand_vec(A.n)
    iv = [0.6, 1.2, 0.0, 12.0]
    r = orc.cheblan_nr(A, iv, 300, 1e-8, vinit, orc.Pol(iv), want_vectors=True)
    Y = r["vecs"]
    for lam, y, res in zip(r["lam"], Y, r["res"]):
        assert abs(np.linalg.norm(y) - 1) < 1e-12
        assert abs(np.linalg.norm(A.matvec(y) - lam * y) - res) < 1e-12
    assert np.max(np.abs(Y @ Y.T - np.eye(len(Y)))) < 1e-8


def test_cheblan_invalid_interval_and_bad_gamma():
    A = orc.Csr(lap=(6, 6, 6))
    v = np.ones(A.n)
    pol = orc.Pol([0.6, 1.2, 0, 12])
    r = orc.cheblan_nr(A, [1.2, 0.6, 0, 12], 50, 1e-8, v, pol)
    assert r["rc"] == 0 and len(r["lam"]) == 0


def test_driver_c1_counts_match_exact():
    d = orc.run_driver(["-nx", 20, "-ny", 20, "-nz", 20, "-a", 0.6, "-b", 1.2,
                        "-nslices", 5], timeout=120)
    assert abs(d["ecount"] - 118) < 15
    for s in d["slices"]:
        assert s["nev"] == s["nev_exact"]
        ex = orc.exact_eig_lap3(20, 20, 20, *s["intv"])
        np.testing.assert_allclose(s["lam"], ex, rtol=1e-12)


def test_landos_count_and_slices():
    A = orc.Csr(lap=(10, 10, 10))
    orc.srand(0)
    rc, xd, yd, neig = orc.landos(A, 20, 40, 200, [0.6, 1.2, 0.0, 12.0])
    assert rc == 0
    exact = len(orc.exact_eig_lap3(10, 10, 10, 0.6, 1.2))
    assert abs(neig - exact) < 0.5 * exact
    sli = orc.spslicer_lan(xd, yd, 4)
    assert sli[0] == 0.6 and sli[-1] == 1.2 and np.all(np.diff(sli) > 0)


def test_landos_full_lanczos_is_the_spectral_measure():
    """With msteps = n the Lanczos quadrature is exact: the DOS curve of
    orc_landos is the sum of Gaussians at the eigenvalues weighted by the
    squared components (u_j . v)^2 of the normalised start vector (the libc
    stream after srand(0)).  One start vector, so the breakdown draw at the
    last step cannot shift the stream."""
    ia, ja, a = driver.laplacian3d(9, 1, 1)
    ia2, ja2, a2 = driver.laplacian3d(4, 5, 1)
    # a 20 x 20 matrix with distinct eigenvalues: the 9-chain and the 4x5 grid
    import scipy.sparse as sps
    M = sps.block_diag([sps.csr_matrix((a, ja, ia)), sps.csr_matrix((a2, ja2, ia2)) * 0.9],
                       format="csr")
    M = (M + sps.diags(np.linspace(0.0, 0.3, M.shape[0]))).tocsr()
    M.sort_indices()
    A = orc.Csr(M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data)
    n = M.shape[0]
    lam, U = np.linalg.eigh(M.toarray())
    intv = [lam[0], lam[-1], lam[0] - 0.2, lam[-1] + 0.2]
    orc.srand(0)
    rc, xd, yd, neig = orc.landos(A, 1, n, 400, intv)
    assert rc == 0
    orc.srand(0)
    v = orc.rand_vec(n)
    v = v / np.linalg.norm(v)
    w = (U.T @ v) ** 2
    Mq = min(n, 30)
    sigma = (intv[3] - intv[2]) / (Mq - 1) / math.sqrt(8 * math.log(1.25))
    width = sigma * math.sqrt(-2 * math.log(1e-8))
    ref = np.zeros_like(xd)
    for lj, wj in zip(lam, w):
        sel = np.abs(xd - lj) < width
        ref[sel] += wj * np.exp(-((xd[sel] - lj) ** 2) / (2 * sigma * sigma))
    ref /= math.sqrt(2 * sigma * sigma * math.pi)
    assert np.max(np.abs(yd - ref)) < 1e-10 * np.max(ref)


def test_driver_output_independent_of_thread_count():
    """The oracle's threaded loops (CSR product, vector updates, CGS) keep
    every element's sequence of operations, so a run on 41^3 (n = 68,921,
    above the threading threshold) prints bitwise the same moments, slices,
    eigenvalues and residuals on 1 and 4 threads."""
    args = ["-nx", 41, "-ny", 41, "-nz", 41, "-a", 0.6, "-b", 0.7, "-nslices", 2, "-Mdeg", 30,
            "-nvec", 4, "-slice", 0, "-maxsteps", 80, "-reorthk", 64]
    r1 = orc.run_driver(args, timeout=300, threads=1)
    r4 = orc.run_driver(args, timeout=300, threads=4)
    assert r1["threads"] == 1 and r4["threads"] == 4

    def strip(d):
        if isinstance(d, dict):
            return {k: strip(v) for k, v in d.items() if not k.startswith("t_") and k != "threads"}
        if isinstance(d, list):
            return [strip(x) for x in d]
        return d

    assert strip(r1) == strip(r4)
    assert r1["slices"][0]["nev"] > 0
