"""numpy restatement of the block-sparse (BSR) products -- TEST INFRASTRUCTURE.

Only the tests, the smoke check and the bench's CPU baseline use this module
(as the reference the GPU results are compared with).

* ``bsrmv``: cusparse_bcsr_matvec (brainpy/_src/math/sparse/_bsr_mv.py:60-91,
  GPU lowering :118-146 = cuSPARSE bsrmv): for block row i and each block j
  in indptr[i]:indptr[i+1], y[i*bs:(i+1)*bs] += data[j] @ x[col*bs:(col+1)*bs]
  with data[j] the j-th row-major bs x bs block (the layout of the transpose
  rule, :175-199); transposed, y[col*bs:...] += data[j].T @ x[i*bs:...].
  Blocks with a column outside [0, nb) are skipped.
* ``bsrmv_numba_literal``: the reference's Numba CPU loop as written
  (:24-34): it slices block ``indices[j]`` of the data instead of block
  ``j``; it agrees with ``bsrmv`` wherever indices[j] == j.
* ``bsr_value_grad``: the values' cotangent of the transpose rule
  (:185-198).
* ``bcsrmm``: the loop of sparse/_bsr_mm.py:292-307 with the k slice read as
  ki*32 : ki*32 + 32 -- C = A @ B, B[ki*32 + b, ni*32 + a] = B_data[bi, a, b].

Pinned by scipy.sparse.bsr_matrix products (an independent library routine
with the same block layout) and by a dense assembly of B.
"""
import numpy as np


def _blocks(data, bs):
    return np.asarray(data).reshape(-1, bs, bs)


def bsrmv(data, indices, indptr, x, blocksize, shape, transpose=False):
    bs = int(blocksize)
    mb, nb = int(shape[0]), int(shape[1])
    d = _blocks(data, bs).astype(np.float64)
    idx = np.asarray(indices).astype(np.int64)
    ptr = np.asarray(indptr).astype(np.int64)
    x = np.asarray(x, np.float64)
    nnzb = d.shape[0]
    y = np.zeros((nb if transpose else mb) * bs, np.float64)
    ptr_c = np.clip(ptr, 0, nnzb)
    rows = np.repeat(np.arange(mb), np.maximum(np.diff(ptr_c), 0))
    js = np.concatenate([np.arange(ptr_c[i], max(ptr_c[i], ptr_c[i + 1])) for i in range(mb)]) if mb else \
        np.zeros(0, np.int64)
    js = js.astype(np.int64)
    cols = idx[js] if js.size else np.zeros(0, np.int64)
    ok = (cols >= 0) & (cols < nb)
    rows, js, cols = rows[ok], js[ok], cols[ok]
    if transpose:
        xs = x.reshape(mb, bs)[rows]                       # [blocks, p]
        prod = np.einsum('jpq,jp->jq', d[js], xs)
        np.add.at(y.reshape(nb, bs), cols, prod)
    else:
        xs = x.reshape(nb, bs)[cols]                       # [blocks, q]
        prod = np.einsum('jpq,jq->jp', d[js], xs)
        np.add.at(y.reshape(mb, bs), rows, prod)
    return y


def bsrmv_numba_literal(data, indices, indptr, x, blocksize, shape):
    """The reference's Numba CPU loop as written (block ``indices[j]`` of the
    data, sparse/_bsr_mv.py:24-34), for 2-D data of shape (nnzb*bs, bs)."""
    bs = int(blocksize)
    data = np.asarray(data, np.float64).reshape(-1, bs)
    x = np.asarray(x, np.float64)
    out = np.zeros(int(shape[0]) * bs)
    for i in range(int(shape[0])):
        tmp = np.zeros(bs)
        for j in range(int(indptr[i]), int(indptr[i + 1])):
            start = int(indices[j]) * bs
            tmp += data[start:start + bs] @ x[start:start + bs]
        out[i * bs:(i + 1) * bs] = tmp
    return out


def bsr_value_grad(indices, indptr, x, ct, blocksize, shape, transpose=False):
    """d[j][p][q] = ct[i*bs+p] * x[col*bs+q] (transposed: x[i*bs+p] * ct[col*bs+q])."""
    bs = int(blocksize)
    mb, nb = int(shape[0]), int(shape[1])
    idx = np.asarray(indices).astype(np.int64)
    ptr = np.asarray(indptr).astype(np.int64)
    x = np.asarray(x, np.float64)
    ct = np.asarray(ct, np.float64)
    out = np.zeros((idx.shape[0], bs, bs))
    for i in range(mb):
        for j in range(ptr[i], ptr[i + 1]):
            c = idx[j]
            if not 0 <= c < nb:
                continue
            if transpose:
                out[j] = np.outer(x[i * bs:(i + 1) * bs], ct[c * bs:(c + 1) * bs])
            else:
                out[j] = np.outer(ct[i * bs:(i + 1) * bs], x[c * bs:(c + 1) * bs])
    return out


def bcsrmm(A, B_data, B_indices, B_ptr, shape):
    """C = A @ B for B stored by 32-column blocks (sparse/_bsr_mm.py:292-307)."""
    A = np.asarray(A, np.float64)
    k, n = int(shape[0]), int(shape[1])
    Bd = np.asarray(B_data, np.float64).reshape(-1, 32, 32)
    Bi = np.asarray(B_indices).astype(np.int64).reshape(-1, 2)
    Bp = np.asarray(B_ptr).astype(np.int64)
    m = A.shape[0]
    C_T = np.zeros((n, m))
    for ni in range(n // 32):
        for e in range(Bp[ni], Bp[ni + 1]):
            ki, bi = Bi[e]
            if not (0 <= ki < k // 32 and 0 <= bi < Bd.shape[0]):
                continue
            C_T[ni * 32:(ni + 1) * 32] += Bd[bi] @ A[:, ki * 32:(ki + 1) * 32].T
    return C_T.T


def bcsr_dense(B_data, B_indices, B_ptr, shape):
    """Dense B (k x n) from the block layout, element by element."""
    k, n = int(shape[0]), int(shape[1])
    Bd = np.asarray(B_data, np.float64).reshape(-1, 32, 32)
    Bi = np.asarray(B_indices).astype(np.int64).reshape(-1, 2)
    Bp = np.asarray(B_ptr).astype(np.int64)
    B = np.zeros((k, n))
    for ni in range(n // 32):
        for e in range(Bp[ni], Bp[ni + 1]):
            ki, bi = Bi[e]
            for a in range(32):
                for b in range(32):
                    B[ki * 32 + b, ni * 32 + a] += Bd[bi, a, b]
    return B
