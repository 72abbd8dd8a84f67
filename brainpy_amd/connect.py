"""Seeded synthetic connectivity with FixedProb semantics.

``FixedProb(p)`` gives every pre-synaptic row a fixed out-degree
``int(n_post * p)`` of distinct targets (connect/random_conn.py:85-86,
118-128).  The reference draws them with Numba's ``np.random.choice``; the
stream differs here (seeded PCG64), which changes which matrix is built but
not the operator semantics.
"""
import numpy as np


def fixed_prob_csr(m, n, prob, seed, sort=True, dtype=np.int32):
    """(indices, indptr) of an m x n CSR with int(n*prob) distinct sorted
    targets per row."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = int(n * prob)
    if k == 0:
        return np.zeros(0, dtype), np.zeros(m + 1, dtype)
    if k * 4 >= n:
        idx = np.stack([rng.choice(n, k, replace=False) for _ in range(m)])
    else:
        # k distinct values per row: sorted draws from [0, n-k] plus 0..k-1
        idx = np.sort(rng.integers(0, n - k + 1, size=(m, k)), axis=1) + np.arange(k)
        if not sort:
            idx = rng.permuted(idx, axis=1)
    if sort:
        idx = np.sort(idx, axis=1)
    indices = idx.reshape(-1).astype(dtype)
    indptr = (np.arange(m + 1, dtype=np.int64) * k).astype(dtype)
    return indices, indptr


def fixed_prob_csr_device(m, n, prob, seed, device):
    """Same construction drawn directly on the GPU (torch generator), for
    matrices too large to build on the host (1e8 synapses): returns int32
    (indices, indptr) device tensors with int(n*prob) distinct sorted targets
    per row."""
    import torch
    k = int(n * prob)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    indices = torch.empty(m * k, dtype=torch.int32, device=device)
    chunk = max(1, (1 << 26) // max(k, 1))
    ar = torch.arange(k, device=device, dtype=torch.int64)
    for r0 in range(0, m, chunk):
        r1 = min(m, r0 + chunk)
        draw = torch.randint(0, n - k + 1, (r1 - r0, k), generator=g, device=device, dtype=torch.int64)
        draw, _ = torch.sort(draw, dim=1)
        indices[r0 * k:r1 * k] = (draw + ar).to(torch.int32).reshape(-1)
    indptr = torch.arange(m + 1, device=device, dtype=torch.int64).mul_(k).to(torch.int32)
    return indices, indptr


class CSRConn:
    """A two-end connector holding a given CSR (the part of
    connect/base.py's ``TwoEndConnector`` the CSR layers use: ``pre_num``,
    ``post_num`` and ``require('csr')``)."""

    def __init__(self, indices, indptr, pre_num, post_num):
        self.indices, self.indptr = indices, indptr
        self.pre_num, self.post_num = int(pre_num), int(post_num)
        if len(indptr) != self.pre_num + 1:
            raise ValueError(f"indptr must have pre_num + 1 = {self.pre_num + 1} entries, got {len(indptr)}.")

    def require(self, *structures):
        if structures != ("csr",):
            raise ValueError(f"only 'csr' is provided, got {structures}")
        return self.indices, self.indptr


class ConnectorError(Exception):
    """Raised for connector arguments the reference rejects (brainpy.errors.ConnectorError)."""


def fixed_prob_csr_gpu(m, n, prob, seed, include_self=True, allow_multi_conn=False):
    """``FixedProb.build_csr`` on the GPU (``bpx_fixed_prob_csr``): every
    one of the m rows gets ``k = int(n * prob)`` targets drawn uniformly from
    [0, n) -- distinct and ascending, or with replacement and unsorted when
    ``allow_multi_conn`` -- and with ``include_self=False`` (m == n) a row
    drops its own index when drawn (connect/random_conn.py:73-76, 118-128).
    Returns int32 device tensors ``(indices, indptr)``; the seeded stream is
    this library's own (the same seed gives the same matrix)."""
    import torch
    from . import _lib
    m, n = int(m), int(n)
    if not include_self and m != n:
        raise ConnectorError(f'We found pre_num != post_num ({m} != {n}). But `include_self` is set to True.')
    k = int(n * prob)
    dev = _lib.device()
    indices = torch.empty(max(m * k, 1), dtype=torch.int32, device=dev)
    indptr = torch.empty(m + 1, dtype=torch.int32, device=dev)
    ws_bytes = int(_lib.lib().bpx_fixed_prob_workspace(m)) if not include_self else 0
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    rc = _lib.lib().bpx_fixed_prob_csr(_lib.stream_ptr(), m, n, k, int(seed) & 0xFFFFFFFFFFFFFFFF, int(include_self),
                                       int(allow_multi_conn), _lib.ptr(indices), _lib.ptr(indptr), _lib.ptr(ws),
                                       ws_bytes)
    _lib.check(rc, "bpx_fixed_prob_csr")
    total = m * k if (include_self or allow_multi_conn) else int(indptr[-1])
    return indices[:total], indptr


class FixedProb(CSRConn):
    """``FixedProb(prob, include_self=..., allow_multi_conn=..., seed=...)``
    applied to ``(pre_num, post_num)`` (connect/random_conn.py:85-128 with
    ``pre_ratio=1``): every pre-synaptic neuron gets ``int(post_num * prob)``
    targets.  ``backend="gpu"`` builds the CSR on the device with
    :func:`fixed_prob_csr_gpu` (uniform distinct targets, the reference's
    semantics); ``backend="host"`` uses the seeded numpy generator of
    :func:`fixed_prob_csr` (the synthetic inputs of the tests), where
    include_self/allow_multi_conn are not supported."""

    def __init__(self, prob, pre_num, post_num, seed=0, include_self=True, allow_multi_conn=False, backend="host"):
        if not 0.0 <= prob <= 1.0:
            raise ValueError(f"prob must be in [0, 1], got {prob}")
        self.prob, self.seed = float(prob), int(seed)
        self.include_self, self.allow_multi_conn = bool(include_self), bool(allow_multi_conn)
        if backend == "gpu":
            indices, indptr = fixed_prob_csr_gpu(pre_num, post_num, prob, seed, include_self, allow_multi_conn)
        elif backend == "host":
            if not include_self or allow_multi_conn:
                raise ValueError('backend="host" builds distinct targets with self-connections; '
                                 'use backend="gpu" for include_self=False or allow_multi_conn=True')
            indices, indptr = fixed_prob_csr(int(pre_num), int(post_num), prob, seed)
        else:
            raise ValueError(f'backend must be "host" or "gpu", got {backend!r}')
        super().__init__(indices, indptr, pre_num, post_num)

    def build_csr(self):
        """``(indices, indptr)`` (connect/random_conn.py:118-128)."""
        return self.indices, self.indptr
