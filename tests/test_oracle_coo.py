"""Pins of the COO product restatements -- CPU only.

seg_matmul: the worked examples of the reference's docstring
(sparse/_jax_prim.py:113-140, int32 values and operands).  coomv and the
two-dimensional seg_matmul: scipy.sparse's COO products (duplicates summed,
any entry order), both orientations, scalar and per-entry values; coomv of
a CSR's COO form equals the CSR oracle.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import coo_ref, orc
from tests.helpers import random_csr

_EX = {'data': np.int32([2, 4, 1, 3, 2]), 'index': (np.int32([0, 0, 1, 2, 2]), np.int32([1, 3, 0, 1, 3])),
       'shape': (3, 4)}


def test_seg_matmul_docstring_examples():
    left = coo_ref.seg_matmul(_EX, np.arange(4, dtype=np.int32))
    assert left.dtype == np.int32 and left.tolist() == [14, 0, 9]
    right = coo_ref.seg_matmul(np.arange(3, dtype=np.int32), _EX)
    assert right.dtype == np.int32 and right.tolist() == [1, 6, 0, 4]


def _coo(nse, shape, seed, dup=True):
    rng = np.random.default_rng(seed)
    row = rng.integers(0, shape[0], nse).astype(np.int32)
    col = rng.integers(0, shape[1], nse).astype(np.int32)
    if dup and nse > 4:
        row[1], col[1] = row[0], col[0]       # a duplicate entry
    return row, col


@pytest.mark.parametrize("shape", [(50, 70), (70, 50), (1, 9), (9, 1)])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("homo", [False, True])
def test_coomv_matches_scipy(shape, transpose, homo):
    nse = 3 * (shape[0] + shape[1])
    row, col = _coo(nse, shape, sum(shape))
    rng = np.random.default_rng(1)
    data = np.float64([0.7]) if homo else rng.standard_normal(nse)
    A = sp.coo_matrix((np.broadcast_to(data, (nse,)), (row, col)), shape=shape).tocsr()
    v = rng.standard_normal(shape[0] if transpose else shape[1])
    ref = A.T @ v if transpose else A @ v
    np.testing.assert_allclose(coo_ref.coomv(data, row, col, v, shape, transpose), ref, rtol=1e-12, atol=1e-12)


def test_coomv_int32_wraps():
    row, col = np.int32([0, 0, 1]), np.int32([0, 1, 1])
    data = np.int32([2 ** 30, 2 ** 30, -3])
    v = np.int32([2, 1])
    out = coo_ref.coomv(data, row, col, v, (2, 2))
    # 2**31 + 2**30 wraps to -2**30 in int32
    assert out.dtype == np.int32 and out.tolist() == [-(2 ** 30), -3]


@pytest.mark.parametrize("left", [True, False])
@pytest.mark.parametrize("k", [1, 5, 40])
def test_seg_matmul_matches_scipy(left, k):
    shape = (30, 45)
    nse = 200
    row, col = _coo(nse, shape, k)
    rng = np.random.default_rng(k)
    vals = rng.standard_normal(nse)
    A = sp.coo_matrix((vals, (row, col)), shape=shape).tocsr()
    sparse = {'data': vals, 'index': (row, col), 'shape': shape}
    if left:
        D = rng.standard_normal((shape[1], k))
        np.testing.assert_allclose(coo_ref.seg_matmul(sparse, D), A @ D, rtol=1e-12, atol=1e-12)
    else:
        D = rng.standard_normal((k, shape[0]))
        np.testing.assert_allclose(coo_ref.seg_matmul(D, sparse), D @ A, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("transpose", [False, True])
def test_coomv_of_csr_matches_csr_oracle(transpose):
    m, n = 80, 60
    indices, indptr = random_csr(m, n, 0.2, seed=3, dtype=np.int32)
    row = np.repeat(np.arange(m, dtype=np.int32), np.diff(indptr))
    rng = np.random.default_rng(4)
    data = rng.standard_normal(indices.size).astype(np.float32)
    v = rng.standard_normal(m if transpose else n).astype(np.float32)
    ref = orc.csrmv(data, indices, indptr, v, (m, n), transpose)
    np.testing.assert_allclose(coo_ref.coomv(data, row, indices, v, (m, n), transpose), ref, rtol=1e-5, atol=1e-5)
