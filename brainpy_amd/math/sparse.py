"""Sparse CSR matvec with a dense vector (``bm.sparse.csrmv``,
``bm.sparse.csrmv_taichi``) and the CSR/COO/dense converters on the gfx950
library.

References: brainpy/_src/math/sparse/_csr_mv.py:29-123,
sparse/_csr_mv_taichi.py:188-264, sparse/_utils.py:20-120.
"""
from typing import Tuple

import numpy as np
import torch

from .. import _lib
from ._common import acc_dtype, batch_rows, data_tensor, index_tensor, ndim, needs_ad

__all__ = ["csrmv", "csrmv_taichi", "csr_to_dense", "csr_to_coo", "coo_to_csr"]

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
