"""numpy restatement of the COO products (test infrastructure only).

* ``coomv``: ``_coomv_impl`` of brainpy/_src/math/sparse/_coo_mv.py:99-105
  -- rows and columns swapped when transposed, ``data * v[col]`` scattered
  into ``out[row]`` with duplicates summed.
* ``seg_matmul``: sparse/_jax_prim.py:15-52 (left sparse: gather of the
  dense operand's rows by ``cols``, scaled by the values, segment sum by
  ``rows``) and :55-94 (right sparse: gather of the dense operand's columns
  by ``rows``, segment sum by ``cols``), dispatch of :97-166.

Floating point accumulates in float64; int32 operands in int64 and wrap to
int32 at the end, which equals int32 arithmetic modulo 2**32.  Pinned by the
worked examples of the reference's seg_matmul docstring
(sparse/_jax_prim.py:113-140) and by scipy.sparse's COO products.
"""
import numpy as np


def _acc_type(dtype):
    return np.int64 if np.issubdtype(dtype, np.integer) else np.float64


def _finish(out, dtype):
    if np.issubdtype(dtype, np.integer):
        return out.astype(np.int64).astype(dtype)
    return out


def coomv(data, row, col, vector, shape, transpose=False):
    """``A @ v`` (``A.T @ v``) for COO A; scalar data is broadcast."""
    data = np.asarray(data)
    row = np.asarray(row, np.int64)
    col = np.asarray(col, np.int64)
    v = np.asarray(vector)
    acc = _acc_type(v.dtype)
    if transpose:
        row, col = col, row
    out = np.zeros(shape[1] if transpose else shape[0], acc)
    w = np.broadcast_to(data.astype(acc).reshape(-1), row.shape)
    np.add.at(out, row, w * v.astype(acc)[col])
    return _finish(out, v.dtype)


def seg_matmul(A, B):
    """``A @ B`` with one operand a ``{'data', 'index', 'shape'}`` COO dict."""
    if isinstance(A, dict):
        sparse, dense, left = A, np.asarray(B), True
    else:
        sparse, dense, left = B, np.asarray(A), False
    vals = np.asarray(sparse['data'])
    rows, cols = (np.asarray(x, np.int64) for x in sparse['index'])
    m, n = sparse['shape']
    dtype = np.result_type(vals.dtype, dense.dtype)
    acc = _acc_type(dtype)
    w = np.broadcast_to(vals.astype(acc).reshape(-1), rows.shape)
    D = dense.astype(acc)
    if left:
        prod = D[cols] * (w.reshape(-1, 1) if D.ndim == 2 else w)
        out = np.zeros((m,) + D.shape[1:], acc)
        np.add.at(out, rows, prod)
    else:
        if D.ndim == 2:
            prod = (D[:, rows] * w).T
            out = np.zeros((n, D.shape[0]), acc)
            np.add.at(out, cols, prod)
            out = out.T
        else:
            out = np.zeros(n, acc)
            np.add.at(out, cols, D[rows] * w)
    return _finish(out, dtype)
