"""Sparse CSR matvec with a dense vector (``bm.sparse.csrmv``,
``bm.sparse.csrmv_taichi``), COO products (``bm.sparse.coomv``,
``bm.sparse.seg_matmul``), block-sparse products (``cusparse_bcsr_matvec``
and ``bcsrmm`` of ``brainpy._src.math.sparse``) and the CSR/COO/dense
converters on the gfx950 library.

References: brainpy/_src/math/sparse/_csr_mv.py:29-123,
sparse/_csr_mv_taichi.py:188-264, sparse/_coo_mv.py:17-196,
sparse/_jax_prim.py:15-166, sparse/_utils.py:20-120, sparse/_bsr_mv.py:24-210,
sparse/_bsr_mm.py:210-307.
"""
from typing import Tuple

import numpy as np
import torch

from .. import _lib
from ._common import acc_dtype, batch_rows, data_tensor, index_tensor, ndim, needs_ad
from .event import _PLANS

__all__ = ["csrmv", "csrmv_taichi", "coomv", "seg_matmul", "csr_to_dense", "csr_to_coo", "coo_to_csr",
           "cusparse_bcsr_matvec", "bcsrmm"]

_METHODS = ("cusparse", "scalar", "vector", "adaptive")


def _as_array(x):
    """Arrays pass through; python scalars and sequences take JAX's default
    dtypes without x64 (float -> float32, int -> int32)."""
    if isinstance(x, (torch.Tensor, np.ndarray)):
        return x
    a = np.asarray(x)
    if a.dtype == np.float64:
        a = a.astype(np.float32)
    elif a.dtype == np.int64:
        a = a.astype(np.int32)
    return a


def _launch(d, idx, ptr, v, shape, transpose):
    # raw pointers go to the kernel: views (an expanded cotangent, a row of
    # a batch) are made dense first
    d, idx, ptr, v = d.contiguous(), idx.contiguous(), ptr.contiguous(), v.contiguous()
    m, n = int(shape[0]), int(shape[1])
    if v.shape[0] != (m if transpose else n):
        raise ValueError(f'Shape mismatch, vector ({v.shape[0]},) with matrix {shape}, transpose={transpose}.')
    if ptr.shape[0] != m + 1:
        raise ValueError(f'indptr should have shape ({m + 1},), but we got {tuple(ptr.shape)}.')
    if v.dtype == torch.float32:
        # repeated calls on the same arrays run on the cached tiled plan
        # (shared with bm.event.csrmv): the transpose sums its products in
        # LDS with no global atomic per synapse, the gather reads each column
        # tile of v from LDS instead of one random gather per synapse (up to
        # 64 column tiles: its per-(tile, row) partials grow with the tiles)
        plan = _PLANS.lookup(d, idx, ptr, m, n, v, None)
        if plan is not None:
            if transpose:
                return plan.matvec(v)
            if plan.plan.tiles <= 64:
                return plan._gather(v, False, _lib.RULE_POSITIVE)
    out = torch.empty(n if transpose else m, dtype=acc_dtype(d.dtype), device=v.device)
    rc = _lib.lib().bpx_csrmv(_lib.stream_ptr(), _lib.ptr(d), _lib.bpx_dtype(d), d.shape[0], _lib.ptr(idx),
                              _lib.ptr(ptr), _lib.bpx_dtype(idx), idx.shape[0], _lib.ptr(v), _lib.bpx_dtype(v), m, n,
                              int(bool(transpose)), _lib.ptr(out))
    _lib.check(rc, "bpx_csrmv")
    return out


class _CSRMV(torch.autograd.Function):
    """Differentiation rules of sparse/_csr_mv.py:211-244: tangents pass
    through the same operator; the cotangent of a scalar weight is
    <ct, csrmv(1, x)>, of heterogeneous weights x[row] * ct[col] (transpose)
    or x[col] * ct[row]; the vector's is csrmv(data, ct) with `transpose`
    flipped."""

    @staticmethod
    def forward(d, v, idx, ptr, shape, transpose):
        return _launch(d, idx, ptr, v, shape, transpose)

    @staticmethod
    def setup_context(ctx, inputs, output):
        d, v, idx, ptr, shape, transpose = inputs
        ctx.save_for_backward(d, v, idx, ptr)
        ctx.shape, ctx.transpose = shape, transpose

    @staticmethod
    def backward(ctx, ct):
        d, v, idx, ptr = ctx.saved_tensors
        gd = gv = None
        if ctx.needs_input_grad[0]:
            if d.shape[0] == 1:
                ones = torch.ones(1, dtype=d.dtype, device=d.device)
                base = _CSRMV.apply(ones, v, idx, ptr, ctx.shape, ctx.transpose)
                gd = torch.dot(ct.to(base.dtype), base).reshape(1)
            else:
                row, col = csr_to_coo(idx, ptr)
                x = v.to(ct.dtype)
                gd = x[row.long()] * ct[col.long()] if ctx.transpose else x[col.long()] * ct[row.long()]
            gd = gd.to(d.dtype)
        if ctx.needs_input_grad[1]:
            gv = _CSRMV.apply(d, ct.to(v.dtype), idx, ptr, ctx.shape, not ctx.transpose)
        return gd, gv, None, None, None, None

    @staticmethod
    def vmap(info, in_dims, d, v, idx, ptr, shape, transpose):
        return batch_rows(info, in_dims, (d, v, idx, ptr, shape, transpose), _CSRMV.apply)


def _run(data, indices, indptr, vector, shape, transpose):
    d = data_tensor(data)
    if d.dtype == torch.float16:
        d = d.to(torch.float32)
    idx = index_tensor(indices)
    ptr = index_tensor(indptr)
    if ptr.dtype != idx.dtype:
        ptr = ptr.to(idx.dtype)
    v = _lib.to_device(vector)
    acc = acc_dtype(d.dtype)
    if v.dtype != acc:
        v = v.to(acc)
    if needs_ad(d, v):
        return _CSRMV.apply(d, v, idx, ptr, (int(shape[0]), int(shape[1])), bool(transpose))
    return _launch(d, idx, ptr, v, shape, transpose)


def csrmv(data, indices, indptr, vector, *, shape: Tuple[int, int], transpose: bool = False,
          method: str = 'cusparse') -> torch.Tensor:
    """Product of a CSR sparse matrix and a dense vector.

    ``method`` is accepted for signature compatibility; every method runs the
    same gfx950 kernels (row-group gather, or coalesced scatter with FP32
    atomics for ``transpose=True``, which the reference's scalar/vector
    methods reject).  Bool vectors are cast to the data dtype; data must be
    float32/float64 (or bfloat16, accumulated in f32) and the vector of the
    same float dtype.  Reference: sparse/_csr_mv.py:29-123.
    """
    data = _as_array(data)
    if ndim(data) == 0:
        data = data.reshape(1)
    vector = _as_array(vector)
    if method not in _METHODS:
        raise ValueError(f'Only support methods: cusparse, scalar, vector, and adaptive. But we got {method}.')
    ddt = data.dtype
    if isinstance(vector, torch.Tensor) and vector.dtype == torch.bool:
        vector = vector.to(ddt if isinstance(ddt, torch.dtype) else torch.from_numpy(np.zeros(0, ddt)).dtype)
    elif isinstance(vector, np.ndarray) and vector.dtype == np.bool_:
        vector = vector.astype(ddt if not isinstance(ddt, torch.dtype) else np.float32)
    d_t = data_tensor(data)
    v_t = _lib.to_device(vector)
    if d_t.dtype not in (torch.float32, torch.float64, torch.bfloat16):
        raise TypeError(f'Only support float32 and float64. But we got {d_t.dtype}.')
    if acc_dtype(d_t.dtype) != v_t.dtype:
        raise TypeError('The types of data and vector should be the same. '
                        f'But we got {d_t.dtype} != {v_t.dtype}.')
    return _run(d_t, indices, indptr, v_t, shape, transpose)


def csrmv_taichi(data, indices, indptr, vector, *, shape: Tuple[int, int], transpose: bool = False):
    """Taichi twin of :func:`csrmv`; returns a 1-tuple.
    Reference: sparse/_csr_mv_taichi.py:188-264."""
    d_t = data_tensor(_as_array(data))
    v_t = _lib.to_device(_as_array(vector))
    if v_t.dtype == torch.bool:
        v_t = v_t.to(d_t.dtype)
    if d_t.dtype not in (torch.float16, torch.float32, torch.float64):
        raise TypeError('Only support float16, float32 or float64 type. '
                        f'But we got {d_t.dtype}.')
    if d_t.dtype != v_t.dtype:
        raise TypeError('The types of data and vector should be the same. '
                        f'But we got {d_t.dtype} != {v_t.dtype}.')
    out_dtype = d_t.dtype
    out = _run(d_t, indices, indptr, v_t, shape, transpose)
    return (out.to(out_dtype),)


def csr_to_coo(indices, indptr):
    """``(row, col)`` of a CSR structure (sparse/_utils.py:45-52)."""
    ptr = index_tensor(indptr)
    idx = index_tensor(indices)
    counts = ptr[1:] - ptr[:-1]
    rows = torch.repeat_interleave(torch.arange(counts.shape[0], device=ptr.device, dtype=idx.dtype), counts)
    return rows, idx


def coo_to_csr(pre_ids, post_ids, *, num_row: int):
    """``(indices, indptr)`` of COO pairs sorted by pre id (sparse/_utils.py:55-66)."""
    pre = index_tensor(pre_ids)
    post = index_tensor(post_ids)
    order = torch.argsort(pre, stable=True)
    post = post[order]
    counts = torch.bincount(pre, minlength=num_row)
    indptr = torch.zeros(num_row + 1, dtype=post.dtype, device=post.device)
    indptr[1:] = torch.cumsum(counts, 0)
    return post, indptr


def csr_to_dense(data, indices, indptr, *, shape: Tuple[int, int]):
    """Dense ``(shape)`` matrix of a CSR (sparse/_utils.py:69-80)."""
    d = data_tensor(data)
    rows, cols = csr_to_coo(indices, indptr)
    mat = torch.zeros(shape, dtype=d.dtype, device=d.device)
    vals = d.expand(cols.shape[0]) if d.shape[0] == 1 else d
    mat.index_put_((rows.long(), cols.long()), vals, accumulate=True)
    return mat


# ---------------------------------------------------------------------------
# COO products
# ---------------------------------------------------------------------------
_COO_DTYPES = (torch.float32, torch.float64, torch.int32)


def _coomv_launch(d, row, col, v, shape, transpose):
    d, row, col, v = d.contiguous(), row.contiguous(), col.contiguous(), v.contiguous()
    m, n = int(shape[0]), int(shape[1])
    if d.dtype == torch.float32 and v.dtype == torch.float32 and row.shape[0] > 0:
        # repeated calls on the same arrays run on a tiled plan of the matrix
        # ordered by row (shared cache with csrmv): the transpose sums its
        # products in LDS, the gather stages column tiles of v in LDS
        plan = _PLANS.lookup(d, col, row, m, n, v, None, coo=True)
        if plan is not None:
            if transpose:
                return plan.matvec(v)
            if plan.plan.tiles <= 64:
                return plan._gather(v, False, _lib.RULE_POSITIVE)
    return _coomv_stateless(d, row, col, v, shape, transpose)


def _coomv_stateless(d, row, col, v, shape, transpose):
    d, row, col, v = d.contiguous(), row.contiguous(), col.contiguous(), v.contiguous()
    m, n = int(shape[0]), int(shape[1])
    out = torch.empty(n if transpose else m, dtype=d.dtype, device=v.device)
    rc = _lib.lib().bpx_coomv(_lib.stream_ptr(), _lib.ptr(d), _lib.bpx_dtype(d), d.shape[0], _lib.ptr(row),
                              _lib.ptr(col), _lib.bpx_dtype(row), row.shape[0], _lib.ptr(v), _lib.bpx_dtype(v), m, n,
                              int(bool(transpose)), _lib.ptr(out))
    _lib.check(rc, "bpx_coomv")
    return out


def _coo_sddmm(row, col, a, a_keys, a_stride, a_col_stride, b, b_keys, b_stride, b_col_stride, ncols):
    """out[e] = sum_c a[row[e]*a_stride + c*a_col_stride] * b[col[e]*b_stride + c*b_col_stride]."""
    row, col, a, b = row.contiguous(), col.contiguous(), a.contiguous(), b.contiguous()
    out = torch.empty(row.shape[0], dtype=a.dtype, device=a.device)
    rc = _lib.lib().bpx_coo_sddmm(_lib.stream_ptr(), _lib.ptr(row), _lib.ptr(col), _lib.bpx_dtype(row),
                                  row.shape[0], _lib.ptr(a), a_keys, a_stride, a_col_stride, _lib.ptr(b), b_keys,
                                  b_stride, b_col_stride, ncols, _lib.bpx_dtype(a), _lib.ptr(out))
    _lib.check(rc, "bpx_coo_sddmm")
    return out


class _COOMV(torch.autograd.Function):
    """Differentiation rules of sparse/_coo_mv.py:155-186: tangents pass
    through the same operator; the vector's cotangent is coomv(data, ct) with
    `transpose` flipped; the values' cotangent is ct[row] * v[col], and for
    ``transpose=True`` ct[col] * v[row] (the derivative of the transposed
    product; the reference's rule returns ct[row] * v[col] for both, which
    only matches the forward map when transpose is False).  A scalar value's
    cotangent is the sum of those products."""

    @staticmethod
    def forward(d, v, row, col, shape, transpose):
        return _coomv_launch(d, row, col, v, shape, transpose)

    @staticmethod
    def setup_context(ctx, inputs, output):
        d, v, row, col, shape, transpose = inputs
        ctx.save_for_backward(d, v, row, col)
        ctx.shape, ctx.transpose = shape, transpose

    @staticmethod
    def backward(ctx, ct):
        d, v, row, col = ctx.saved_tensors
        gd = gv = None
        m, n = ctx.shape
        ct = ct.to(d.dtype)
        if ctx.needs_input_grad[0]:
            okey, gkey = (col, row) if ctx.transpose else (row, col)
            out_len, in_len = (n, m) if ctx.transpose else (m, n)
            per = _coo_sddmm(okey, gkey, ct, out_len, 1, 0, v.to(d.dtype), in_len, 1, 0, 1)
            gd = per.sum().reshape(1) if d.shape[0] == 1 else per
        if ctx.needs_input_grad[1]:
            gv = _COOMV.apply(d, ct, row, col, ctx.shape, not ctx.transpose).to(v.dtype)
        return gd, gv, None, None, None, None

    @staticmethod
    def vmap(info, in_dims, d, v, row, col, shape, transpose):
        return batch_rows(info, in_dims, (d, v, row, col, shape, transpose), _COOMV.apply)


def _coo_prepare(data, row, col):
    d = _lib.to_device(_as_array(data))
    if d.ndim == 0:
        d = d.reshape(1)
    return d, index_tensor(_as_array(row)), index_tensor(_as_array(col))


def coomv(data, row, col, vector, *, shape: Tuple[int, int], rows_sorted: bool = False, cols_sorted: bool = False,
          transpose: bool = False, method: str = 'cusparse') -> torch.Tensor:
    """Product of a COO sparse matrix and a dense vector.

    ``out = A @ vector`` (``A.T @ vector`` with ``transpose``), where A has
    the ``nse`` entries ``(row[k], col[k], data[k])`` in any order; ``data``
    of one element is broadcast.  ``rows_sorted`` / ``cols_sorted`` only
    choose a cuSPARSE call in the reference and do not change the result:
    the gfx950 kernel sums runs of equal output rows inside each 256-element
    tile (a row-sorted matrix costs one atomic per row and tile).  Data and
    vector share one dtype (float32, float64 or int32); indices int32 or
    int64 (unsigned become int64).  Reference: sparse/_coo_mv.py:17-89 and
    :99-105 (semantics), :108-117 (checks).
    """
    if method != 'cusparse':
        raise ValueError
    d, r, c = _coo_prepare(data, row, col)
    v = _lib.to_device(_as_array(vector))
    m, n = int(shape[0]), int(shape[1])
    if r.shape != c.shape or (d.shape[0] != 1 and d.shape != r.shape):
        raise AssertionError(f'data, row and col must have one shape (data may have one element), got '
                             f'{tuple(d.shape)}, {tuple(r.shape)}, {tuple(c.shape)}.')
    if r.dtype != c.dtype:
        raise AssertionError(f'row and col must have one dtype, got {r.dtype} and {c.dtype}.')
    if d.dtype != v.dtype:
        raise AssertionError(f'data and vector must have one dtype, got {d.dtype} and {v.dtype}.')
    if d.dtype not in _COO_DTYPES:
        raise TypeError(f'coomv supports float32, float64 and int32 values, got {d.dtype}.')
    if v.ndim != 1 or v.shape[0] != (m if transpose else n):
        raise AssertionError(f'vector of shape {tuple(v.shape)} does not fit matrix {tuple(shape)} '
                             f'(transpose={transpose}).')
    if needs_ad(d, v):
        return _COOMV.apply(d, v, r, c, (m, n), bool(transpose))
    return _coomv_launch(d, r, c, v, (m, n), transpose)


def _matmat_launch(d, row, col, dense, shape, left):
    d, row, col, dense = d.contiguous(), row.contiguous(), col.contiguous(), dense.contiguous()
    m, n = int(shape[0]), int(shape[1])
    k = int(dense.shape[1] if left else dense.shape[0])
    out = torch.empty((m, k) if left else (k, n), dtype=d.dtype, device=dense.device)
    rc = _lib.lib().bpx_coo_matmat(_lib.stream_ptr(), _lib.ptr(d), _lib.bpx_dtype(d), d.shape[0], _lib.ptr(row),
                                   _lib.ptr(col), _lib.bpx_dtype(row), row.shape[0], _lib.ptr(dense),
                                   _lib.bpx_dtype(dense), m, n, k, int(bool(left)), _lib.ptr(out))
    _lib.check(rc, "bpx_coo_matmat")
    return out


class _SegMatmul(torch.autograd.Function):
    """Gradients of the JAX expression of sparse/_jax_prim.py:15-94 (a
    gather, a product, a segment sum) with a two-dimensional dense operand:
    left, Y = S D: dD = S^T dY, dvalues[e] = <dY[row[e], :], D[col[e], :]>;
    right, Y = D S: dD = dY S^T, dvalues[e] = <dY[:, col[e]], D[:, row[e]]>.
    A scalar value's cotangent is the sum over entries."""

    @staticmethod
    def forward(d, dense, row, col, shape, left):
        return _matmat_launch(d, row, col, dense, shape, left)

    @staticmethod
    def setup_context(ctx, inputs, output):
        d, dense, row, col, shape, left = inputs
        ctx.save_for_backward(d, dense, row, col)
        ctx.shape, ctx.left = shape, left

    @staticmethod
    def backward(ctx, ct):
        d, dense, row, col = ctx.saved_tensors
        m, n = ctx.shape
        ct = ct.to(d.dtype).contiguous()
        D = dense.to(d.dtype)
        gd = gD = None
        if ctx.left:
            k = D.shape[1]
            if ctx.needs_input_grad[0]:
                per = _coo_sddmm(row, col, ct, m, k, 1, D, n, k, 1, k)
                gd = per.sum().reshape(1) if d.shape[0] == 1 else per
            if ctx.needs_input_grad[1]:
                gD = _SegMatmul.apply(d, ct, col, row, (n, m), True)
        else:
            k = D.shape[0]
            if ctx.needs_input_grad[0]:
                per = _coo_sddmm(col, row, ct, n, 1, n, D, m, 1, m, k)
                gd = per.sum().reshape(1) if d.shape[0] == 1 else per
            if ctx.needs_input_grad[1]:
                gD = _SegMatmul.apply(d, ct, col, row, (n, m), False)
        if gD is not None:
            gD = gD.to(dense.dtype)
        return gd, gD, None, None, None, None

    @staticmethod
    def vmap(info, in_dims, d, dense, row, col, shape, left):
        return batch_rows(info, in_dims, (d, dense, row, col, shape, left), _SegMatmul.apply)


def _is_dense(x):
    return isinstance(x, (torch.Tensor, np.ndarray))


def _seg_matmul(sparse, dense, left):
    if dense.ndim not in (1, 2):
        raise AssertionError('Dense matrix must be a one- or two-dimensional matrix.')
    shape = sparse['shape']
    if len(shape) != 2:
        raise ValueError(f'Sparse matrix must be a two-dimensional matrix. But we got {shape}')
    rows, cols = sparse['index']
    d, r, c = _coo_prepare(sparse['data'], rows, cols)
    if r.dtype != c.dtype:
        r, c = r.to(torch.int64), c.to(torch.int64)
    D = _lib.to_device(_as_array(dense))
    dt = torch.promote_types(d.dtype, D.dtype)
    if dt not in _COO_DTYPES:
        raise TypeError(f'seg_matmul supports float32, float64 and int32 operands, got {d.dtype} and {D.dtype}.')
    d, D = d.to(dt), D.to(dt)
    m, n = int(shape[0]), int(shape[1])
    if d.shape[0] != 1 and d.shape[0] != r.shape[0]:
        raise ValueError(f'values of shape {tuple(d.shape)} do not match {r.shape[0]} indices.')
    inner = D.shape[0] if (left or D.ndim == 1) else D.shape[1]
    if inner != (n if left else m):
        raise ValueError(f'dense operand of shape {tuple(D.shape)} does not fit sparse matrix {tuple(shape)}.')
    if D.ndim == 1:
        return coomv(d, r, c, D, shape=(m, n), transpose=not left)
    if needs_ad(d, D):
        return _SegMatmul.apply(d, D, r, c, (m, n), left)
    return _matmat_launch(d, r, c, D, (m, n), left)


def seg_matmul(A, B):
    """Sparse-dense matrix product ``A @ B`` where exactly one of A, B is a
    COO sparse matrix given as ``{'data': values, 'index': (rows, cols),
    'shape': (N, M)}``.

    Left sparse: ``(N, M) @ (M,)`` or ``(M, K)``; right sparse: ``(N,)`` or
    ``(K, N) @ (N, M)``.  Values and dense operand are promoted to one dtype
    (float32, float64 or int32).  One-dimensional dense operands run the
    coomv kernel; two-dimensional ones the COO matrix-matrix kernels.
    Reference: sparse/_jax_prim.py:97-166 (dispatch and errors), :15-94
    (semantics).
    """
    if isinstance(A, dict):
        if not _is_dense(B):
            raise ValueError('A and B cannot be both sparse. \n'
                             f'A:\n{A}\n'
                             f'B:\n{B}')
        return _seg_matmul(A, B, True)
    if not isinstance(B, dict):
        raise ValueError('A and B cannot be both dense. \n'
                         f'A:\n{A}\n'
                         f'B:\n{B}')
    return _seg_matmul(B, A, False)


# ---------------------------------------------------------------------------
# block-sparse (BSR) products
# ---------------------------------------------------------------------------
_BSR_METHODS = ("scalar", "vector", "adaptive")


def _bsr_launch(d, idx, ptr, v, bs, nnzb, shape, transpose):
    d, idx, ptr, v = d.contiguous(), idx.contiguous(), ptr.contiguous(), v.contiguous()
    mb, nb = int(shape[0]), int(shape[1])
    out = torch.empty((nb if transpose else mb) * bs, dtype=d.dtype, device=v.device)
    rc = _lib.lib().bpx_bsrmv(_lib.stream_ptr(), _lib.ptr(d), _lib.bpx_dtype(d), _lib.ptr(idx), _lib.ptr(ptr),
                              _lib.bpx_dtype(idx), nnzb, bs, _lib.ptr(v), mb, nb, int(bool(transpose)), _lib.ptr(out))
    _lib.check(rc, "bpx_bsrmv")
    return out


def _bsr_sddmm(idx, ptr, bs, nnzb, rowv, colv, shape):
    """out[j][p][q] = rowv[i*bs + p] * colv[indices[j]*bs + q], i = block row of block j."""
    idx, ptr, rowv, colv = idx.contiguous(), ptr.contiguous(), rowv.contiguous(), colv.contiguous()
    out = torch.empty(nnzb * bs * bs, dtype=rowv.dtype, device=rowv.device)
    rc = _lib.lib().bpx_bsr_sddmm(_lib.stream_ptr(), _lib.ptr(idx), _lib.ptr(ptr), _lib.bpx_dtype(idx), nnzb, bs,
                                  _lib.ptr(rowv), _lib.ptr(colv), _lib.bpx_dtype(rowv), int(shape[0]), int(shape[1]),
                                  _lib.ptr(out))
    _lib.check(rc, "bpx_bsr_sddmm")
    return out


class _BSRMV(torch.autograd.Function):
    """Differentiation rules of sparse/_bsr_mv.py:170-199: tangents of the
    values pass through the same operator; the vector's cotangent is the
    product with `transpose` flipped; the values' cotangent of block j in
    block row i is ct[i*bs + p] * vector[indices[j]*bs + q] (transposed:
    vector[i*bs + p] * ct[indices[j]*bs + q]), as the reference's rule
    writes it."""

    @staticmethod
    def forward(d, v, idx, ptr, bs, nnzb, shape, transpose):
        return _bsr_launch(d.reshape(-1), idx, ptr, v, bs, nnzb, shape, transpose)

    @staticmethod
    def setup_context(ctx, inputs, output):
        d, v, idx, ptr, bs, nnzb, shape, transpose = inputs
        ctx.save_for_backward(d, v, idx, ptr)
        ctx.bs, ctx.nnzb, ctx.shape, ctx.transpose = bs, nnzb, shape, transpose

    @staticmethod
    def backward(ctx, ct):
        d, v, idx, ptr = ctx.saved_tensors
        gd = gv = None
        ct = ct.to(d.dtype)
        if ctx.needs_input_grad[0]:
            rowv, colv = (v.to(d.dtype), ct) if ctx.transpose else (ct, v.to(d.dtype))
            gd = _bsr_sddmm(idx, ptr, ctx.bs, ctx.nnzb, rowv, colv, ctx.shape).reshape(d.shape)
        if ctx.needs_input_grad[1]:
            gv = _BSRMV.apply(d, ct, idx, ptr, ctx.bs, ctx.nnzb, ctx.shape, not ctx.transpose).to(v.dtype)
        return gd, gv, None, None, None, None, None, None

    @staticmethod
    def vmap(info, in_dims, d, v, idx, ptr, bs, nnzb, shape, transpose):
        return batch_rows(info, in_dims, (d, v, idx, ptr, bs, nnzb, shape, transpose), _BSRMV.apply)


def cusparse_bcsr_matvec(data, indices, indptr, vector, *, blocksize: int, nnzb: int, shape: Tuple[int, int],
                         method: str = 'vector', transpose: bool = False) -> torch.Tensor:
    """Product of a block-sparse (BSR) matrix and a dense vector.

    The matrix has ``shape = (mb, nb)`` blocks of ``blocksize x blocksize``;
    ``data`` holds the ``nnzb`` blocks row-major (any shape with
    ``nnzb * blocksize**2`` elements: ``(nnzb * bs, bs)`` as the reference's
    benchmark builds it, or ``(nnzb, bs, bs)``), ``indices`` the block
    columns and ``indptr`` the ``mb + 1`` block-row pointers.  Returns
    ``A @ vector`` (``mb * bs`` entries), or ``A.T @ vector`` (``nb * bs``
    entries) with ``transpose`` -- the form the reference's transpose rule
    calls; its own lowerings take no transpose.  The GPU lowering of the
    reference is cuSPARSE bsrmv, whose block-j semantics this follows; its
    Numba CPU loop slices block ``indices[j]`` instead of block ``j``
    (sparse/_bsr_mv.py:28-33), which agrees only where the two are equal.
    ``method`` is checked and otherwise ignored.  Data and vector share one
    dtype, float32 or float64.  Reference: sparse/_bsr_mv.py:60-91.
    """
    data = _as_array(data)
    if ndim(data) == 0:
        data = data.reshape(1)
    vector = _as_array(vector)
    if method not in _BSR_METHODS:
        raise ValueError('Only support methods: scalar, vector, and adaptive. '
                         f'But we got {method}.')
    d = _lib.to_device(data)
    if d.ndim == 0:
        d = d.reshape(1)
    v = _lib.to_device(vector)
    if d.dtype not in (torch.float32, torch.float64):
        raise TypeError(f'Only support float32 and float64. But we got {d.dtype}.')
    if d.dtype != v.dtype:
        raise TypeError('The types of data and vector should be the same. '
                        f'But we got {d.dtype} != {v.dtype}.')
    bs, nnzb = int(blocksize), int(nnzb)
    mb, nb = int(shape[0]), int(shape[1])
    if bs < 1:
        raise ValueError(f'blocksize must be positive, got {blocksize}.')
    idx = index_tensor(_as_array(indices))
    ptr = index_tensor(_as_array(indptr))
    if ptr.dtype != idx.dtype:
        ptr = ptr.to(idx.dtype)
    if d.numel() != nnzb * bs * bs:
        raise ValueError(f'data has {d.numel()} elements, {nnzb} blocks of {bs} x {bs} need {nnzb * bs * bs}.')
    if idx.ndim != 1 or idx.shape[0] != nnzb:
        raise ValueError(f'indices should have shape ({nnzb},), but we got {tuple(idx.shape)}.')
    if ptr.ndim != 1 or ptr.shape[0] != mb + 1:
        raise ValueError(f'indptr should have shape ({mb + 1},), but we got {tuple(ptr.shape)}.')
    if v.ndim != 1 or v.shape[0] != (mb if transpose else nb) * bs:
        raise ValueError(f'Shape mismatch, vector {tuple(v.shape)} with a matrix of {shape} blocks of {bs} x {bs}, '
                         f'transpose={transpose}.')
    if needs_ad(d, v):
        return _BSRMV.apply(d, v, idx, ptr, bs, nnzb, (mb, nb), bool(transpose))
    return _bsr_launch(d.reshape(-1), idx, ptr, v, bs, nnzb, (mb, nb), transpose)


def bcsrmm(A_data, B_data, B_indices, B_ptr, *, shape: Tuple[int, int], block_size: Tuple[int, int],
           transpose: bool = False, method: str = 'cutlass') -> torch.Tensor:
    """Dense-times-block-sparse product ``C = A @ B``.

    ``A`` is dense ``(m, k)``; ``B`` of ``shape = (k, n)`` is stored by
    32-column blocks: ``B_ptr`` (``n // 32 + 1``) points into ``B_indices``,
    whose rows are ``(k block, data block)`` pairs, and ``B_data`` holds the
    data blocks as ``(nblk, 32, 32)`` (or ``(nblk * 32, 32)``) with
    ``B[ki*32 + b, ni*32 + a] = B_data[bi, a, b]`` -- the layout of the
    reference's loop (sparse/_bsr_mm.py:292-307, reading the k slice as
    ``ki*32 : ki*32 + 32``).  Returns ``C`` of shape ``(m, n)``.  float32 runs
    on MFMA, float64 on FMA.  ``block_size`` must be ``(32, 32)`` and
    ``method`` ``'cutlass'``, as in the reference; ``transpose=True`` and
    differentiation are not provided (the reference's primitive rejects the
    transpose keyword and its differentiation rules are stubs,
    sparse/_bsr_mm.py:246-259, 377-391).  Reference: sparse/_bsr_mm.py:210-259.
    """
    A = _lib.to_device(_as_array(A_data))
    Bd = _lib.to_device(_as_array(B_data))
    Bi = index_tensor(_as_array(B_indices))
    Bp = index_tensor(_as_array(B_ptr))
    if Bp.dtype != Bi.dtype:
        Bp = Bp.to(Bi.dtype)
    k, n = int(shape[0]), int(shape[1])
    if A.ndim != 2 or A.shape[1] != k:
        raise AssertionError(f'A_data of shape {tuple(A.shape)} does not fit B of shape {tuple(shape)}.')
    if method != 'cutlass':
        raise ValueError
    if tuple(int(b) for b in block_size) != (32, 32):
        raise AssertionError('cutlass based block-sparse mm only support block size (32, 32)')
    if transpose:
        raise NotImplementedError('bcsrmm computes A @ B only (transpose=True is not supported).')
    if A.dtype != Bd.dtype:
        raise AssertionError(f'A_data and B_data must have one dtype, got {A.dtype} and {Bd.dtype}.')
    if A.dtype not in (torch.float32, torch.float64):
        raise TypeError(f'Only support float32 and float64. But we got {A.dtype}.')
    if k % 32 or n % 32:
        raise ValueError(f'shape {tuple(shape)} must be a multiple of the 32 x 32 block size.')
    if Bi.ndim != 2 or Bi.shape[1] != 2:
        raise AssertionError(f'B_indices must have shape (nblk, 2), got {tuple(Bi.shape)}.')
    nblk = int(Bi.shape[0])
    if Bd.numel() != nblk * 32 * 32:
        raise AssertionError(f'B_data has {Bd.numel()} elements, {nblk} blocks of 32 x 32 need {nblk * 1024}.')
    if Bp.ndim != 1 or Bp.shape[0] != n // 32 + 1:
        raise AssertionError(f'B_ptr should have shape ({n // 32 + 1},), but we got {tuple(Bp.shape)}.')
    if needs_ad(A, Bd):
        raise NotImplementedError('bcsrmm has no differentiation rule.')
    A, Bd, Bi, Bp = A.contiguous(), Bd.contiguous(), Bi.contiguous(), Bp.contiguous()
    m = int(A.shape[0])
    out = torch.empty((m, n), dtype=A.dtype, device=A.device)
    rc = _lib.lib().bpx_bcsrmm(_lib.stream_ptr(), _lib.ptr(A), _lib.bpx_dtype(A), _lib.ptr(Bd), _lib.ptr(Bi),
                               _lib.ptr(Bp), _lib.bpx_dtype(Bi), nblk, m, k, n, _lib.ptr(out))
    _lib.check(rc, "bpx_bcsrmm")
    return out
