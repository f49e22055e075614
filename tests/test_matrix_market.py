"""Matrix Market reader (driver layer, the rules of the reference's
TESTS/COMMON/parmm.c:11-162) on hand-written files and on the reference's
stiff1 matrix kept as a fixture.  No GPU."""
import json
import os

import numpy as np
import pytest

from pevsl_amd import driver

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _write(tmp_path, text):
    p = tmp_path / "m.mtx"
    p.write_text(text)
    return str(p)


def test_symmetric_file_is_mirrored_and_mm0_mm1_indexing(tmp_path):
    text = ("%%MatrixMarket matrix coordinate real symmetric\n% comment\n3 3 4\n"
            "1 1 2.0\n2 1 -1.0\n3 2 -0.5\n3 3 4.0\n")
    ia, ja, a = driver.read_matrix_market(_write(tmp_path, text), 1)
    dense = np.zeros((3, 3))
    dense[np.repeat(np.arange(3), np.diff(ia)), ja] = a
    assert np.array_equal(dense, [[2, -1, 0], [-1, 0, -0.5], [0, -0.5, 4]])
    # MM0: the same entries with 0-based indices
    text0 = text.replace("1 1 2.0", "0 0 2.0").replace("2 1 -1.0", "1 0 -1.0") \
        .replace("3 2 -0.5", "2 1 -0.5").replace("3 3 4.0", "2 2 4.0")
    ia0, ja0, a0 = driver.read_matrix_market(_write(tmp_path, text0), 0)
    assert np.array_equal(ia, ia0) and np.array_equal(ja, ja0) and np.array_equal(a, a0)
    # any positive idxin means 1-based (parmm.c: idxin = idxin > 0)
    ia2, ja2, a2 = driver.read_matrix_market(_write(tmp_path, text), 2)
    assert np.array_equal(ia, ia2) and np.array_equal(ja, ja2) and np.array_equal(a, a2)


def test_general_file_is_not_mirrored(tmp_path):
    text = "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 2 3.0\n2 1 5.0\n"
    ia, ja, a = driver.read_matrix_market(_write(tmp_path, text), 1)
    assert list(ia) == [0, 1, 2] and list(ja) == [1, 0] and list(a) == [3.0, 5.0]


def test_rejects_dense_and_complex(tmp_path):
    with pytest.raises(ValueError):
        driver.read_matrix_market(_write(tmp_path, "%%MatrixMarket matrix array real general\n1 1\n1.0\n"))
    with pytest.raises(ValueError):
        driver.read_matrix_market(_write(
            tmp_path, "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1.0 0.0\n"))


def test_stiff1_fixture():
    ia, ja, a = driver.read_matrix_market(os.path.join(GOLDEN, "mm_stiff1.mtx"), 1)
    g = json.load(open(os.path.join(GOLDEN, "mm_stiff1_eigs.json")))
    n = len(ia) - 1
    assert n == g["n"] == 5795
    # symmetric expansion: 2 * 71180 stored entries minus the 5795 diagonal ones
    assert ia[-1] == g["nnz"] == 2 * 71180 - 5795
    rows = np.repeat(np.arange(n), np.diff(ia))
    key = rows.astype(np.int64) * n + ja
    keyt = ja.astype(np.int64) * n + rows
    o, ot = np.argsort(key), np.argsort(keyt)
    assert np.array_equal(key[o], keyt[ot]) and np.array_equal(a[o], a[ot])
    assert len(g["eigs_0.10_0.14"]) == 312


def test_index_outside_declared_size_is_an_error(tmp_path):
    text = "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n3 1 2.0\n"
    import pytest
    with pytest.raises(ValueError):
        driver.read_matrix_market(_write(tmp_path, text), 1)
