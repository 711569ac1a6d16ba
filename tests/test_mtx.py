"""MatrixMarket system I/O against the reference's own data files (tests/golden:
simple.mtx / simple_rhs.mtx of src/movesemantic/tests, mesh1em6.mtx of the
plugin examples) and an independent reader (scipy.io.mmread)."""
import pathlib

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

import alien_amd as aa

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def csr_of(path):
    A = sp.csr_matrix(sio.mmread(str(path)))
    A.sum_duplicates()
    A.sort_indices()
    return A


@pytest.mark.parametrize("name", ["simple.mtx", "mesh1em6.mtx"])
def test_read_matches_scipy(name):
    ro, cols, vals = aa.mtx_read_csr(GOLD / name)
    A = csr_of(GOLD / name)
    np.testing.assert_array_equal(ro, A.indptr)
    np.testing.assert_array_equal(cols, A.indices)
    np.testing.assert_array_equal(vals, A.data)
    assert np.all(np.diff(cols[ro[0]:ro[1]]) > 0)


def test_symmetric_file_is_mirrored():
    ro, cols, vals = aa.mtx_read_csr(GOLD / "mesh1em6.mtx")  # 'coordinate real symmetric', lower triangle stored
    A = sp.csr_matrix((vals, cols, ro))
    assert (A - A.T).nnz == 0 and A.shape == (48, 48) and A.nnz == 2 * 177 - 48


def test_read_vector():
    """The fixture mixes number formats (8.1, 9., 10, 11.23, 1.6e1, 25.666)."""
    v = aa.mtx_read_vector(GOLD / "simple_rhs.mtx")
    ref = np.arange(1.0, 26.0)
    ref[[7, 10, 24]] = [8.1, 11.23, 25.666]
    np.testing.assert_array_equal(v, ref)
    np.testing.assert_array_equal(v, np.asarray(sio.mmread(str(GOLD / "simple_rhs.mtx"))).ravel())


def test_write_read_round_trip_exact(tmp_path):
    ro, cols, vals = aa.gen_csr(aa.GEN_LOGNORMAL7, 5, params=[2.0, 1, 1, 1], seed=3)
    aa.mtx_write_csr(tmp_path / "a.mtx", ro, cols, vals, "log-normal 5^3")
    r2, c2, v2 = aa.mtx_read_csr(tmp_path / "a.mtx")
    np.testing.assert_array_equal(r2, ro)
    np.testing.assert_array_equal(c2, cols)
    np.testing.assert_array_equal(v2, vals)  # 17 significant digits: exact
    b = aa.uniform01(7, len(ro) - 1)
    aa.mtx_write_vector(tmp_path / "b.mtx", b, "rhs")
    np.testing.assert_array_equal(aa.mtx_read_vector(tmp_path / "b.mtx"), b)
    A = csr_of(tmp_path / "a.mtx")  # and an independent reader agrees
    np.testing.assert_array_equal(A.data, vals)


def test_repeated_entries_are_summed(tmp_path):
    f = tmp_path / "dup.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1.5\n2 2 2.0\n1 1 0.25\n3 2 -1\n")
    ro, cols, vals = aa.mtx_read_csr(f)
    np.testing.assert_array_equal(ro, [0, 1, 2, 3])
    np.testing.assert_array_equal(cols, [0, 1, 1])
    np.testing.assert_array_equal(vals, [1.75, 2.0, -1.0])


def test_bad_files_raise(tmp_path):
    with pytest.raises(aa.AlienHipError):
        aa.mtx_read_csr(tmp_path / "missing.mtx")
    f = tmp_path / "arr.mtx"
    f.write_text("%%MatrixMarket matrix array real general\n2 1\n1\n2\n")
    with pytest.raises(aa.AlienHipError):
        aa.mtx_read_csr(f)  # a matrix must be 'coordinate'
    with pytest.raises(aa.AlienHipError):
        aa.mtx_read_vector(GOLD / "simple.mtx")  # a vector must be 'array'
