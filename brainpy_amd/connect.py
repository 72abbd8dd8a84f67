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


class FixedProb(CSRConn):
    """``FixedProb(prob, seed=...)`` applied to ``(pre_num, post_num)``
    (connect/random_conn.py:85-128 with ``pre_ratio=1``, no self-connection
    filter): every pre-synaptic neuron gets ``int(post_num * prob)`` distinct
    targets, drawn here from a seeded PCG64 stream."""

    def __init__(self, prob, pre_num, post_num, seed=0):
        if not 0.0 <= prob <= 1.0:
            raise ValueError(f"prob must be in [0, 1], got {prob}")
        self.prob, self.seed = float(prob), int(seed)
        indices, indptr = fixed_prob_csr(int(pre_num), int(post_num), prob, seed)
        super().__init__(indices, indptr, pre_num, post_num)
