"""Pins of the block-sparse restatements -- CPU only.

bsrmv (both orientations, block sizes 1..8, empty block rows) against
scipy.sparse.bsr_matrix products, which use the same row-major block
layout; the reference's Numba CPU loop as written agrees where every block
sits in the block column of its own index; the value cotangent against a
finite-difference-free identity (the product is linear in the values, so
<grad, dvalues> = <ct, bsrmv(dvalues)>); bcsrmm against A @ B with B
assembled element by element and against a scipy CSR product.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import bsr_ref


def random_bsr(mb, nb, bs, density, seed, empty_row=True):
    rng = np.random.default_rng(seed)
    mask = rng.random((mb, nb)) < density
    if empty_row and mb > 1:
        mask[mb // 2] = False
    indptr = np.zeros(mb + 1, np.int32)
    indptr[1:] = np.cumsum(mask.sum(1))
    indices = np.nonzero(mask)[1].astype(np.int32)
    data = rng.standard_normal((indices.size, bs, bs))
    return data, indices, indptr


@pytest.mark.parametrize("bs", [1, 2, 3, 4, 8])
@pytest.mark.parametrize("transpose", [False, True])
def test_bsrmv_matches_scipy(bs, transpose):
    mb, nb = 7, 5
    data, indices, indptr = random_bsr(mb, nb, bs, 0.4, bs)
    A = sp.bsr_matrix((data, indices, indptr), shape=(mb * bs, nb * bs))
    rng = np.random.default_rng(3)
    x = rng.standard_normal((mb if transpose else nb) * bs)
    ref = A.T @ x if transpose else A @ x
    out = bsrmv(data, indices, indptr, x, bs, (mb, nb), transpose)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)


def bsrmv(*a):
    return bsr_ref.bsrmv(*a)


def test_bsrmv_skips_out_of_range_columns():
    data, indices, indptr = random_bsr(3, 3, 2, 0.9, 0, empty_row=False)
    x = np.arange(6, dtype=np.float64)
    base = bsr_ref.bsrmv(data, indices, indptr, x, 2, (3, 3))
    bad = indices.copy()
    bad[0] = 7
    keep = np.ones(indices.size, bool)
    keep[0] = False
    ptr2 = indptr.copy()
    ptr2[1:] -= 1
    ref = bsr_ref.bsrmv(data[keep], indices[keep], ptr2, x, 2, (3, 3))
    np.testing.assert_allclose(bsr_ref.bsrmv(data, bad, indptr, x, 2, (3, 3)), ref)
    assert not np.allclose(base, ref)


def test_numba_literal_agrees_where_block_index_equals_column():
    # one block per row on the diagonal: block j sits in block column j
    mb, bs = 5, 4
    rng = np.random.default_rng(2)
    data = rng.standard_normal((mb * bs, bs))
    indices = np.arange(mb, dtype=np.int32)
    indptr = np.arange(mb + 1, dtype=np.int32)
    x = rng.standard_normal(mb * bs)
    np.testing.assert_allclose(bsr_ref.bsrmv_numba_literal(data, indices, indptr, x, bs, (mb, mb)),
                               bsr_ref.bsrmv(data, indices, indptr, x, bs, (mb, mb)), rtol=1e-12)


@pytest.mark.parametrize("transpose", [False, True])
def test_value_grad_is_the_adjoint(transpose):
    mb, nb, bs = 4, 6, 3
    data, indices, indptr = random_bsr(mb, nb, bs, 0.5, 5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((mb if transpose else nb) * bs)
    ct = rng.standard_normal((nb if transpose else mb) * bs)
    g = bsr_ref.bsr_value_grad(indices, indptr, x, ct, bs, (mb, nb), transpose)
    dv = rng.standard_normal(data.shape)
    lhs = float((g * dv).sum())
    rhs = float(ct @ bsr_ref.bsrmv(dv, indices, indptr, x, bs, (mb, nb), transpose))
    assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(rhs))


def random_bcsr(k, n, density, seed):
    rng = np.random.default_rng(seed)
    kb, nbk = k // 32, n // 32
    mask = rng.random((nbk, kb)) < density
    ptr = np.zeros(nbk + 1, np.int32)
    ptr[1:] = np.cumsum(mask.sum(1))
    ki = np.nonzero(mask)[1].astype(np.int32)
    nblk = ki.size
    bi = rng.permutation(nblk).astype(np.int32)          # data blocks in any order
    data = rng.standard_normal((nblk, 32, 32))
    return data, np.stack([ki, bi], 1), ptr


def test_bcsrmm_matches_dense_assembly():
    k, n, m = 96, 128, 5
    data, idx, ptr = random_bcsr(k, n, 0.5, 1)
    A = np.random.default_rng(2).standard_normal((m, k))
    B = bsr_ref.bcsr_dense(data, idx, ptr, (k, n))
    np.testing.assert_allclose(bsr_ref.bcsrmm(A, data, idx, ptr, (k, n)), A @ B, rtol=1e-12, atol=1e-12)
    # the same B through scipy: a column-block CSC of transposed blocks is the
    # BSR of B^T with block (ni, ki) = data[bi]
    nbk, kb = n // 32, k // 32
    BT = sp.bsr_matrix((data[idx[:, 1]], idx[:, 0], ptr), shape=(n, k))
    np.testing.assert_allclose(B, BT.T.toarray(), rtol=0, atol=0)
    assert nbk * 32 == n and kb * 32 == k
