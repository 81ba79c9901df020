"""Matrix Market reading (io/matrix_market.hpp) on the reference's own test
files (libs/numeric/mtl/test/matrix_market/*.mtx, copied as fixtures): the
Python reader and the C++ drop-in reader must build the same CSR, and both
must show the properties the files declare (symmetric / skew-symmetric /
pattern / dense column-major layout)."""
import os
import subprocess

import numpy as np
import pytest

from mtl4_amd import io as mio

HERE = os.path.dirname(os.path.abspath(__file__))
MM = os.path.join(HERE, "golden", "matrix_market")
CPP = os.path.join(HERE, "cpp", "mm_read")
FILES = ["bcsstk01", "bcspwr02", "jgl009", "plskz362", "dense_read"]


def dense(nr, nc, rp, ci, va):
    D = np.zeros((nr, nc))
    for r in range(nr):
        D[r, ci[rp[r]:rp[r + 1]]] = va[rp[r]:rp[r + 1]]
    return D


def pattern(nr, nc, rp, ci):
    P = np.zeros((nr, nc), bool)
    for r in range(nr):
        P[r, ci[rp[r]:rp[r + 1]]] = True
    return P


def file_entries(name):
    """Count of entry lines of a coordinate file (after the size line)."""
    lines = [l for l in open(os.path.join(MM, name + ".mtx")) if l.strip() and not l.startswith("%")]
    return len(lines) - 1, lines[0].split()


def test_bcsstk01_real_symmetric():
    nr, nc, rp, ci, va = mio.read_matrix_market(os.path.join(MM, "bcsstk01.mtx"))
    assert (nr, nc) == (48, 48)
    n_file, size_line = file_entries("bcsstk01")
    assert int(size_line[2]) == n_file == 224
    D = dense(nr, nc, rp, ci, va)
    assert np.array_equal(D, D.T)
    assert np.all(np.diag(D) > 0)                   # stiffness matrix: positive diagonal
    assert len(ci) == 2 * 224 - 48                  # lower triangle mirrored, diagonal once
    assert D[0, 0] == 2.8322685185200e+06 and D[4, 0] == D[0, 4] == 1.0e+06
    assert np.all(np.linalg.eigvalsh(D) > 0)        # SPD: a CG case
    assert np.all(np.diff(rp) > 0) and all(np.all(np.diff(ci[rp[r]:rp[r + 1]]) > 0) for r in range(nr))


def test_plskz362_skew_symmetric():
    nr, nc, rp, ci, va = mio.read_matrix_market(os.path.join(MM, "plskz362.mtx"))
    assert (nr, nc) == (362, 362)
    D = dense(nr, nc, rp, ci, va)
    assert np.array_equal(D, -D.T)
    assert not np.any(np.diag(D))
    assert len(ci) == 2 * 880


def test_pattern_files_store_zeros():
    nr, nc, rp, ci, va = mio.read_matrix_market(os.path.join(MM, "jgl009.mtx"))
    assert (nr, nc, len(ci)) == (9, 9, 50) and not np.any(va)
    nr, nc, rp, ci, va = mio.read_matrix_market(os.path.join(MM, "bcspwr02.mtx"))
    P = pattern(nr, nc, rp, ci)
    assert (nr, nc) == (49, 49) and np.array_equal(P, P.T) and not np.any(va)
    ndiag = int(np.trace(P))
    assert len(ci) == 2 * 108 - ndiag


def test_dense_array_is_column_major():
    nr, nc, rp, ci, va = mio.read_matrix_market(os.path.join(MM, "dense_read.mtx"))
    assert np.array_equal(dense(nr, nc, rp, ci, va), np.array([[1., 2, 3, 4], [5, 6, 7, 8]]))


def test_duplicates_overwrite_and_complex_rejected(tmp_path):
    p = tmp_path / "dup.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n% c\n2 2 3\n1 1 1.5\n2 1 2\n1 1 -3\n")
    nr, nc, rp, ci, va = mio.read_matrix_market(str(p))
    assert np.array_equal(dense(nr, nc, rp, ci, va), np.array([[-3.0, 0], [2, 0]]))
    q = tmp_path / "c.mtx"
    q.write_text("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1.0 2.0\n")
    with pytest.raises(mio.runtime_error):
        mio.read_matrix_market(str(q))
    with pytest.raises(mio.file_not_found):
        mio.read_matrix_market(str(tmp_path / "missing.mtx"))


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    nr, nc, rp, ci, va = mio.read_matrix_market(os.path.join(MM, "bcsstk01.mtx"))
    va = va * rng.uniform(0.5, 2.0, len(va))        # values with full mantissas
    p = str(tmp_path / "w.mtx")
    mio.write_matrix_market(p, nr, nc, rp, ci, va)
    nr2, nc2, rp2, ci2, va2 = mio.read_matrix_market(p)
    assert (nr2, nc2) == (nr, nc) and np.array_equal(rp2, rp) and np.array_equal(ci2, ci)
    assert np.array_equal(va2, va)


def cpp_read(path, out=None):
    args = [CPP, path] + ([out] if out else [])
    p = subprocess.run(args, capture_output=True, text=True)
    return p.returncode, p.stdout


@pytest.mark.skipif(not os.path.exists(CPP), reason="C++ test programs not built")
@pytest.mark.parametrize("name", FILES)
def test_cpp_reader_matches(name, tmp_path):
    path = os.path.join(MM, name + ".mtx")
    rc, out = cpp_read(path, str(tmp_path / "back.mtx"))
    assert rc == 0, out
    lines = out.strip().split("\n")
    nr, nc, nnz = map(int, lines[0].split())
    nr_p, nc_p, rp, ci, va = mio.read_matrix_market(path)
    assert (nr, nc, nnz) == (nr_p, nc_p, len(ci))
    rows = np.array([int(l.split()[0]) for l in lines[1:]], np.int64)
    cols = np.array([int(l.split()[1]) for l in lines[1:]], np.int64)
    vals = np.array([float(l.split()[2]) for l in lines[1:]])
    assert np.array_equal(cols, ci) and np.array_equal(vals, va)
    assert np.array_equal(np.bincount(rows, minlength=nr), np.diff(rp))
    # what the C++ writer wrote reads back to the same matrix
    nr_w, nc_w, rp_w, ci_w, va_w = mio.read_matrix_market(str(tmp_path / "back.mtx"))
    assert np.array_equal(rp_w, rp) and np.array_equal(ci_w, ci) and np.array_equal(va_w, va)


@pytest.mark.skipif(not os.path.exists(CPP), reason="C++ test programs not built")
def test_cpp_reader_errors(tmp_path):
    q = tmp_path / "c.mtx"
    q.write_text("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1.0 2.0\n")
    assert cpp_read(str(q))[0] == 3
    assert cpp_read(str(tmp_path / "missing.mtx"))[0] == 4
