"""CPU checks of the synthetic workload generators the benchmarks and GPU
tests use: the irregular power-law matrix (the C4 recipe) and the 3-D
Laplacian (TESTS/COMMON/lapl.c:4-70)."""
import numpy as np
import scipy.sparse as sp

from oracle import oracle as orc
from pevsl_amd import driver


def test_powerlaw_sym_recipe():
    """Seeded, symmetric, sorted int32 CSR; about 2 * mean * n off-diagonal
    entries (A0 + A0^T, few duplicates at random columns) plus the diagonal;
    a heavy row-length tail up to about 2 * max_len; diagonal shift
    0.1 * sum_j |A_ij| of the symmetrized off-diagonal part."""
    n, mean, max_len = 20000, 16.0, 2000
    ia, ja, a = driver.powerlaw_sym(n, mean=mean, max_len=max_len, seed=7)
    ia2, ja2, a2 = driver.powerlaw_sym(n, mean=mean, max_len=max_len, seed=7)
    assert np.array_equal(ia, ia2) and np.array_equal(ja, ja2) and np.array_equal(a, a2)
    assert ia.dtype == np.int32 and ja.dtype == np.int32 and a.dtype == np.float64
    A = sp.csr_matrix((a, ja, ia), shape=(n, n))
    assert abs(A - A.T).max() == 0.0
    for i in range(0, n, 997):
        row = ja[ia[i]:ia[i + 1]]
        assert np.all(np.diff(row) > 0)
    nnz = ia[-1]
    assert 0.9 * (2 * mean * n + n) < nnz < 1.05 * (2 * mean * n + n)
    lens = np.diff(ia)
    assert lens.max() > 20 * mean and lens.max() <= 2 * max_len + 1
    off = A - sp.diags(A.diagonal())
    d_off = np.asarray(abs(off).sum(axis=1)).ravel()
    # diagonal = (A0 + A0^T)_ii + 0.1 * sum_j |A_ij| (the sum includes the
    # symmetrized diagonal entry itself), so it is at least 0.1 * d_off - 2
    assert np.all(A.diagonal() >= 0.1 * d_off - 2.0)


def test_laplacian3d_matches_oracle_and_stencil():
    """driver.laplacian3d equals the oracle's LocalLapGen restatement, and
    its spectrum is the closed form of ExactEigLap3."""
    ia, ja, a = driver.laplacian3d(6, 5, 4)
    o = orc.Csr(lap=(6, 5, 4))
    assert np.array_equal(ia, o.ia) and np.array_equal(ja, o.ja) and np.array_equal(a, o.a)
    n = len(ia) - 1
    ev = np.linalg.eigvalsh(sp.csr_matrix((a, ja, ia), shape=(n, n)).toarray())
    ex = driver.exact_eig_lap3(6, 5, 4, -1.0, 13.0)
    np.testing.assert_allclose(np.sort(ev), np.sort(ex), atol=1e-12)
