"""GPU: block-sparse products (``cusparse_bcsr_matvec``, ``bcsrmm``) against
the numpy oracle (oracle/bsr_ref.py).

Floating point: every output is compared with the FP64 oracle within
``n * u`` times the sum of the magnitudes of its terms, n the largest number
of terms of one output and u the unit roundoff of the dtype -- the
worst-case bound of a sum in any order (the transposed matvec adds block
partials with atomics, so its order is run-dependent).  Block sizes 4, 8,
16, 32 and 64 run the wave-per-block-row kernel, the others the generic
one; bcsrmm f32 runs on MFMA."""
import numpy as np
import pytest
import torch

import brainpy_amd.math as bm
from oracle import bsr_ref
from tests.test_oracle_bsr import random_bcsr, random_bsr

pytestmark = pytest.mark.gpu

U = {np.float32: 2.0 ** -24, np.float64: 2.0 ** -53}


def _check(got, ref, mag, n_terms, dtype):
    tol = (n_terms + 2) * U[dtype] * mag
    err = np.abs(np.asarray(got, np.float64) - ref)
    assert np.all(err <= tol + 1e-300), float(np.max(err - tol))


def _max_terms(indices, indptr, bs, nb, transpose):
    if transpose:
        counts = np.bincount(np.asarray(indices, np.int64), minlength=nb)
        return int(counts.max(initial=0)) * bs
    return int(np.diff(indptr).max(initial=0)) * bs


def _bsrmv(data, indices, indptr, x, bs, shape, transpose, index_dtype=np.int32):
    return bm.sparse.cusparse_bcsr_matvec(data, indices.astype(index_dtype), indptr.astype(index_dtype), x,
                                          blocksize=bs, nnzb=len(indices), shape=shape, transpose=transpose)


@pytest.mark.parametrize("bs", [1, 2, 3, 4, 5, 8, 16, 32, 64])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bsrmv(bs, transpose, dtype):
    mb, nb = 37, 29
    data, indices, indptr = random_bsr(mb, nb, bs, 0.3, bs)
    data = data.astype(dtype)
    rng = np.random.default_rng(bs + 100)
    x = rng.standard_normal((mb if transpose else nb) * bs).astype(dtype)
    got = _bsrmv(data.reshape(-1, bs), indices, indptr, x, bs, (mb, nb), transpose).cpu().numpy()
    assert got.dtype == dtype and got.shape == ((nb if transpose else mb) * bs,)
    ref = bsr_ref.bsrmv(data, indices, indptr, x, bs, (mb, nb), transpose)
    mag = bsr_ref.bsrmv(np.abs(data), indices, indptr, np.abs(x), bs, (mb, nb), transpose)
    _check(got, ref, mag, _max_terms(indices, indptr, bs, nb, transpose), dtype)


@pytest.mark.parametrize("bs", [4, 16, 32])
@pytest.mark.parametrize("transpose", [False, True])
def test_bsrmv_long_rows_int64_and_3d_data(bs, transpose):
    # every block row holds all 203 blocks: the unrolled block loop's tail
    mb, nb = 3, 203
    data, indices, indptr = random_bsr(mb, nb, bs, 1.1, 7, empty_row=False)
    data = data.astype(np.float32)
    x = np.random.default_rng(8).standard_normal((mb if transpose else nb) * bs).astype(np.float32)
    got = _bsrmv(data, indices, indptr, x, bs, (mb, nb), transpose, np.int64).cpu().numpy()
    ref = bsr_ref.bsrmv(data, indices, indptr, x, bs, (mb, nb), transpose)
    mag = bsr_ref.bsrmv(np.abs(data), indices, indptr, np.abs(x), bs, (mb, nb), transpose)
    _check(got, ref, mag, _max_terms(indices, indptr, bs, nb, transpose), np.float32)


@pytest.mark.parametrize("transpose", [False, True])
def test_bsrmv_unaligned_data_takes_the_generic_kernel(transpose):
    mb, nb, bs = 9, 11, 8
    data, indices, indptr = random_bsr(mb, nb, bs, 0.5, 9)
    flat = np.concatenate([[0.0], data.reshape(-1)]).astype(np.float32)
    buf = torch.tensor(flat, device="cuda")
    view = buf[1:]                                         # 4-byte offset: not 16-byte aligned
    assert view.data_ptr() % 16 != 0
    x = np.random.default_rng(10).standard_normal((mb if transpose else nb) * bs).astype(np.float32)
    got = _bsrmv(view, indices, indptr, x, bs, (mb, nb), transpose).cpu().numpy()
    ref = bsr_ref.bsrmv(data, indices, indptr, x, bs, (mb, nb), transpose)
    mag = bsr_ref.bsrmv(np.abs(data), indices, indptr, np.abs(x), bs, (mb, nb), transpose)
    _check(got, ref, mag, _max_terms(indices, indptr, bs, nb, transpose), np.float32)


def test_bsrmv_edges():
    bs = 4
    # no blocks at all
    out = bm.sparse.cusparse_bcsr_matvec(np.zeros(0, np.float32), np.zeros(0, np.int32), np.zeros(4, np.int32),
                                         np.ones(8, np.float32), blocksize=bs, nnzb=0, shape=(3, 2))
    assert out.shape == (12,) and float(out.abs().sum()) == 0.0
    outT = bm.sparse.cusparse_bcsr_matvec(np.zeros(0, np.float32), np.zeros(0, np.int32), np.zeros(4, np.int32),
                                          np.ones(12, np.float32), blocksize=bs, nnzb=0, shape=(3, 2),
                                          transpose=True)
    assert outT.shape == (8,) and float(outT.abs().sum()) == 0.0
    # a block column outside the shape is skipped
    data, indices, indptr = random_bsr(4, 4, bs, 0.9, 11, empty_row=False)
    data = data.astype(np.float32)
    bad = indices.copy()
    bad[2] = 99
    x = np.random.default_rng(12).standard_normal(16).astype(np.float32)
    got = _bsrmv(data, bad, indptr, x, bs, (4, 4), False).cpu().numpy()
    ref = bsr_ref.bsrmv(data, bad, indptr, x, bs, (4, 4), False)
    mag = bsr_ref.bsrmv(np.abs(data), bad, indptr, np.abs(x), bs, (4, 4), False)
    _check(got, ref, mag, 4 * bs, np.float32)


def test_bsrmv_errors():
    d, i, p = random_bsr(3, 3, 2, 0.9, 13, empty_row=False)
    d = d.astype(np.float32)
    x = np.ones(6, np.float32)
    kw = dict(blocksize=2, nnzb=len(i), shape=(3, 3))
    with pytest.raises(ValueError, match="Only support methods"):
        bm.sparse.cusparse_bcsr_matvec(d, i, p, x, method='cusparse', **kw)
    with pytest.raises(TypeError, match="Only support float32 and float64"):
        bm.sparse.cusparse_bcsr_matvec(d.astype(np.int32), i, p, x.astype(np.int32), **kw)
    with pytest.raises(TypeError, match="should be the same"):
        bm.sparse.cusparse_bcsr_matvec(d, i, p, x.astype(np.float64), **kw)
    with pytest.raises(ValueError, match="elements"):
        bm.sparse.cusparse_bcsr_matvec(d[:-1], i, p, x, **kw)
    with pytest.raises(ValueError, match="Shape mismatch"):
        bm.sparse.cusparse_bcsr_matvec(d, i, p, np.ones(5, np.float32), **kw)


@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("bs", [3, 8])
def test_bsrmv_grads_match_dense_autograd(transpose, bs):
    mb, nb = 6, 5
    data, indices, indptr = random_bsr(mb, nb, bs, 0.5, 20 + bs)
    rng = np.random.default_rng(21)
    data = data.astype(np.float32)
    x = rng.standard_normal((mb if transpose else nb) * bs).astype(np.float32)
    ct = rng.standard_normal((nb if transpose else mb) * bs).astype(np.float32)
    d_t = torch.tensor(data.reshape(-1, bs), device="cuda", requires_grad=True)
    x_t = torch.tensor(x, device="cuda", requires_grad=True)
    y = bm.sparse.cusparse_bcsr_matvec(d_t, indices, indptr, x_t, blocksize=bs, nnzb=len(indices),
                                       shape=(mb, nb), transpose=transpose)
    y.backward(torch.tensor(ct, device="cuda"))
    d64 = torch.tensor(data, dtype=torch.float64, requires_grad=True)
    x64 = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    A = torch.zeros(mb * bs, nb * bs, dtype=torch.float64)
    rows = np.repeat(np.arange(mb), np.diff(indptr))
    for j, (i, c) in enumerate(zip(rows, indices)):
        A = A.index_put((torch.arange(i * bs, (i + 1) * bs)[:, None], torch.arange(c * bs, (c + 1) * bs)[None, :]),
                        d64[j], accumulate=True)
    y64 = x64 @ A if transpose else A @ x64
    y64.backward(torch.tensor(ct, dtype=torch.float64))
    np.testing.assert_allclose(y.detach().cpu().numpy(), y64.detach().numpy(), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(x_t.grad.cpu().numpy(), x64.grad.numpy(), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(d_t.grad.cpu().numpy().reshape(-1, bs, bs), d64.grad.numpy(), rtol=1e-5, atol=1e-5)
    g = bsr_ref.bsr_value_grad(indices, indptr, x, ct, bs, (mb, nb), transpose)
    np.testing.assert_allclose(d_t.grad.cpu().numpy().reshape(-1, bs, bs), g, rtol=1e-6, atol=1e-6)


def test_bsrmv_vmap():
    mb, nb, bs = 10, 12, 16
    data, indices, indptr = random_bsr(mb, nb, bs, 0.4, 30)
    data = torch.tensor(data.astype(np.float32), device="cuda")
    xs = torch.tensor(np.random.default_rng(31).standard_normal((3, nb * bs)).astype(np.float32), device="cuda")
    f = lambda x: bm.sparse.cusparse_bcsr_matvec(data, indices, indptr, x, blocksize=bs, nnzb=len(indices),
                                                 shape=(mb, nb))
    Y = torch.func.vmap(f)(xs)
    for b in range(3):
        np.testing.assert_array_equal(Y[b].cpu().numpy(), f(xs[b]).cpu().numpy())


@pytest.mark.parametrize("m", [1, 31, 32, 100, 300])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bcsrmm(m, dtype):
    k, n = 160, 224
    data, idx, ptr = random_bcsr(k, n, 0.4, m)
    data = data.astype(dtype)
    A = np.random.default_rng(m + 1).standard_normal((m, k)).astype(dtype)
    got = bm.sparse.bcsrmm(A, data, idx, ptr, shape=(k, n), block_size=(32, 32)).cpu().numpy()
    assert got.shape == (m, n) and got.dtype == dtype
    ref = bsr_ref.bcsrmm(A, data, idx, ptr, (k, n))
    mag = bsr_ref.bcsrmm(np.abs(A), np.abs(data), idx, ptr, (k, n))
    _check(got, ref, mag, int(np.diff(ptr).max(initial=0)) * 32, dtype)


def test_bcsrmm_int64_unaligned_empty_and_skipped_blocks():
    k, n, m = 128, 160, 70
    data, idx, ptr = random_bcsr(k, n, 0.5, 40)
    ptr = ptr.copy()
    # column block 2 empty; one entry with a k block outside the shape
    cnt = np.diff(ptr)
    cnt[2] = 0
    keep = np.ones(idx.shape[0], bool)
    keep[ptr[2]:ptr[3]] = False
    idx = idx[keep]
    data = data[idx[:, 1]].astype(np.float32)             # one data block per entry, renumbered
    idx = np.stack([idx[:, 0], np.arange(idx.shape[0])], 1).astype(np.int64)
    ptr = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    idx[0, 0] = 99
    A = np.random.default_rng(41).standard_normal((m, k)).astype(np.float32)
    ref = bsr_ref.bcsrmm(A, data, idx, ptr, (k, n))
    mag = bsr_ref.bcsrmm(np.abs(A), np.abs(data), idx, ptr, (k, n))
    terms = int(np.diff(ptr).max()) * 32
    got = bm.sparse.bcsrmm(A, data, idx, ptr, shape=(k, n), block_size=(32, 32)).cpu().numpy()
    _check(got, ref, mag, terms, np.float32)
    assert np.all(got[:, 64:96] == 0)
    # A at a 4-byte offset: the FMA kernel
    buf = torch.tensor(np.concatenate([[0.0], A.reshape(-1)]).astype(np.float32), device="cuda")
    A_view = buf[1:].view(m, k)
    assert A_view.data_ptr() % 16 != 0
    got2 = bm.sparse.bcsrmm(A_view, data, idx, ptr, shape=(k, n), block_size=(32, 32)).cpu().numpy()
    _check(got2, ref, mag, terms, np.float32)


def test_bcsrmm_errors():
    k, n = 64, 64
    data, idx, ptr = random_bcsr(k, n, 0.5, 50)
    data = data.astype(np.float32)
    A = np.ones((4, k), np.float32)
    with pytest.raises(AssertionError, match="block size"):
        bm.sparse.bcsrmm(A, data, idx, ptr, shape=(k, n), block_size=(16, 16))
    with pytest.raises(ValueError):
        bm.sparse.bcsrmm(A, data, idx, ptr, shape=(k, n), block_size=(32, 32), method='other')
    with pytest.raises(NotImplementedError):
        bm.sparse.bcsrmm(A, data, idx, ptr, shape=(k, n), block_size=(32, 32), transpose=True)
    with pytest.raises(AssertionError):
        bm.sparse.bcsrmm(np.ones((4, 32), np.float32), data, idx, ptr, shape=(k, n), block_size=(32, 32))
    with pytest.raises(NotImplementedError, match="differentiation"):
        bm.sparse.bcsrmm(torch.ones((4, k), device="cuda", requires_grad=True), torch.tensor(data, device="cuda"),
                         idx, ptr, shape=(k, n), block_size=(32, 32))
