"""Writes mm_stiff1_eigs.json: the exact spectrum facts of the Matrix Market
fixture mm_stiff1.mtx (the reference's MATRICES/A.mtx, 'stiff1', copied as
test data) that the MM-flow tests check against: the extreme eigenvalues and
every eigenvalue in [0.10, 0.14] (the intervals of the reference's
TESTS/MM/matfile: [0.10, 0.12] in 3 slices, [0.12, 0.14] in 2), from a dense
LAPACK eigensolve (numpy.linalg.eigvalsh) of the symmetric expansion.

usage: python tests/golden/make_mm_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from pevsl_amd import driver  # noqa: E402

ia, ja, a = driver.read_matrix_market(os.path.join(HERE, "mm_stiff1.mtx"), 1)
n = len(ia) - 1
A = np.zeros((n, n))
rows = np.repeat(np.arange(n), np.diff(ia))
A[rows, ja] = a
assert np.array_equal(A, A.T)
ev = np.linalg.eigvalsh(A)
sel = ev[(ev >= 0.10) & (ev <= 0.14)]
json.dump({"source": "MATRICES/A.mtx of the reference (stiff1), MM1, symmetric",
           "n": n, "nnz": int(ia[-1]), "lambda_min": float(ev[0]), "lambda_max": float(ev[-1]),
           "eigs_0.10_0.14": [float(x) for x in sel]},
          open(os.path.join(HERE, "mm_stiff1_eigs.json"), "w"), indent=1)
print(n, ia[-1], ev[0], ev[-1], len(sel))
