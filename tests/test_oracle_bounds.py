"""Pins the oracle's dense symmetric eigensolver and its thick-restart
Lanczos bounds (SRC/lantrbnd.c) against LAPACK (numpy) and the closed-form
Laplacian spectrum.  No GPU."""
import numpy as np
import pytest

from oracle import oracle as orc
from pevsl_amd import driver


@pytest.mark.parametrize("n", [1, 2, 5, 17, 51])
def test_sym_eig_matches_lapack(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n))
    A = A + A.T
    rc, w, Z = orc.sym_eig(A)
    assert rc == 0
    wr = np.linalg.eigvalsh(A)
    assert np.max(np.abs(w - wr)) < 1e-12 * max(1.0, np.max(np.abs(wr)))
    assert np.max(np.abs(A @ Z - Z * w)) < 1e-11 * max(1.0, np.max(np.abs(wr)))
    assert np.max(np.abs(Z.T @ Z - np.eye(n))) < 1e-12


def test_sym_eig_arrowhead_with_clusters():
    # the thick-restart T: diagonal Ritz values, spike row/column, tridiagonal tail
    n = 40
    rng = np.random.default_rng(3)
    T = np.diag(rng.uniform(0, 1, n))
    T[0, 0] = T[1, 1] = 0.5  # a double eigenvalue in the leading block
    s = rng.normal(size=3) * 1e-9
    T[2, :3] = T[:3, 2] = s
    for i in range(2, n - 1):
        T[i, i + 1] = T[i + 1, i] = rng.uniform(0.1, 1)
    rc, w, Z = orc.sym_eig(T)
    assert np.max(np.abs(w - np.linalg.eigvalsh(T))) < 1e-13
    assert np.max(np.abs(T @ Z - Z * w)) < 1e-12


@pytest.mark.parametrize("dims,bndtype", [((30, 30, 1), 1), ((12, 12, 12), 1), ((15, 10, 8), 2)])
def test_lan_trbounds_encloses_closed_form_extremes(dims, bndtype):
    nx, ny, nz = dims
    ia, ja, a = driver.laplacian3d(nx, ny, nz)
    A = orc.Csr(ia, ja, a)
    orc.srand(0)
    v = orc.rand_vec(nx * ny * nz)
    rc, lo, hi, its = orc.lan_trbounds(A, 50, 200, 1e-8, v, bndtype)
    assert rc == 0 and its <= 200
    ev = driver.exact_eig_lap3(nx, ny, nz, -1.0, 13.0)
    scale = abs(lo) + abs(hi)
    # the stopping rule: t1 + t2 < tol (|lmin| + |lmax|)
    assert abs(lo - ev[0]) < 1e-6 * scale
    assert abs(hi - ev[-1]) < 1e-6 * scale
    assert lo <= ev[0] + 1e-12 and hi >= ev[-1] - 1e-12
