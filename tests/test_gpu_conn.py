"""GPU: FixedProb connectivity built on the device (bpx_fixed_prob_csr;
FixedProb.build_csr, connect/random_conn.py:85-128).

The reference draws rows with Numba's rng.choice (an MT19937 stream), so
the matrix itself is not reproducible: what is pinned is the distribution
-- exactly k = int(n * p) distinct targets per row (k or k - 1 without self
connections), each k-subset equally likely, targets with replacement
uniform -- checked exactly on structure and by chi-square tests on subset
and column frequencies (seeded, so each outcome is fixed)."""
import itertools

import numpy as np
import pytest
import torch

from brainpy_amd.connect import ConnectorError, FixedProb, fixed_prob_csr_gpu

pytestmark = pytest.mark.gpu


def _rows(indices, indptr):
    ind, ptr = indices.cpu().numpy(), indptr.cpu().numpy()
    return [ind[ptr[i]:ptr[i + 1]] for i in range(len(ptr) - 1)]


def _chi2(counts, expected):
    counts = np.asarray(counts, np.float64)
    return float(((counts - expected) ** 2 / expected).sum())


@pytest.mark.parametrize("m,n,p", [(2000, 3000, 0.01), (300, 100, 0.9), (50, 7, 0.5), (10, 1, 1.0)])
def test_rows_distinct_sorted_exact_count(m, n, p):
    k = int(n * p)
    ind, ptr = fixed_prob_csr_gpu(m, n, p, seed=5)
    np.testing.assert_array_equal(ptr.cpu().numpy(), np.arange(m + 1) * k)
    for r in _rows(ind, ptr):
        assert len(r) == k and np.all(np.diff(r) > 0) and r.min(initial=0) >= 0 and r.max(initial=0) < n
    ind2, _ = fixed_prob_csr_gpu(m, n, p, seed=5)
    assert torch.equal(ind, ind2)                                     # same seed, same matrix
    if 0 < k < n:
        ind3, _ = fixed_prob_csr_gpu(m, n, p, seed=6)
        assert not torch.equal(ind, ind3)


@pytest.mark.parametrize("n,k", [(5, 2), (6, 4), (7, 3)])
def test_subsets_uniform(n, k):
    # every k-subset of [0, n) equally likely (k > n/2 takes the complement path)
    m = 20000 * len(list(itertools.combinations(range(n), k))) // 10
    ind, ptr = fixed_prob_csr_gpu(m, n, k / n + 1e-9, seed=11)
    subsets = {c: i for i, c in enumerate(itertools.combinations(range(n), k))}
    counts = np.zeros(len(subsets))
    for r in ind.cpu().numpy().reshape(m, k):
        counts[subsets[tuple(r)]] += 1
    # chi-square, len(subsets) - 1 degrees of freedom, p = 0.001 critical values
    crit = {9: 27.88, 14: 36.12, 34: 65.25}[len(subsets) - 1]
    assert _chi2(counts, m / len(subsets)) < crit


def test_columns_uniform():
    m, n, k = 20000, 200, 20
    ind, _ = fixed_prob_csr_gpu(m, n, k / n, seed=13)
    counts = np.bincount(ind.cpu().numpy(), minlength=n)
    assert _chi2(counts, m * k / n) < 266.39                       # 199 dof, p = 0.001


def test_without_self_connections():
    n, p = 500, 0.2
    k = int(n * p)
    ind, ptr = fixed_prob_csr_gpu(n, n, p, seed=17, include_self=False)
    lens = np.diff(ptr.cpu().numpy())
    assert set(lens.tolist()) <= {k - 1, k}
    for i, r in enumerate(_rows(ind, ptr)):
        assert i not in r and np.all(np.diff(r) > 0)
    # a row drops one target exactly when its own index was drawn: P = k / n
    dropped = int((lens == k - 1).sum())
    assert abs(dropped - n * k / n) < 4 * np.sqrt(n * (k / n) * (1 - k / n))
    with pytest.raises(ConnectorError):
        fixed_prob_csr_gpu(10, 20, 0.5, seed=1, include_self=False)
    c = FixedProb(p, n, n, seed=17, include_self=False, backend="gpu")
    ind2, ptr2 = c.build_csr()
    assert torch.equal(ind2, ind) and torch.equal(ptr2, ptr)


def test_with_replacement():
    m, n, k = 5000, 50, 40
    ind, ptr = fixed_prob_csr_gpu(m, n, k / n, seed=19, allow_multi_conn=True)
    assert int(ptr[-1]) == m * k and ind.numel() == m * k
    v = ind.cpu().numpy()
    assert v.min() >= 0 and v.max() < n
    assert _chi2(np.bincount(v, minlength=n), m * k / n) < 85.35    # 49 dof, p = 0.001
    # repeats within a row occur (they cannot without replacement)
    assert any(len(np.unique(r)) < k for r in v.reshape(m, k)[:100])


def test_headline_size_structure():
    # 100k x 100k, p = 0.01: 1e8 targets, checked on the device
    m = n = 100_000
    ind, ptr = fixed_prob_csr_gpu(m, n, 0.01, seed=2024)
    assert ind.numel() == m * 1000 and int(ptr[-1]) == m * 1000
    rows = ind.view(m, 1000)
    assert bool((rows[:, 1:] > rows[:, :-1]).all()) and int(rows.min()) >= 0 and int(rows.max()) < n
    counts = torch.bincount(ind.long(), minlength=n).double()
    # every column's in-degree ~ Binomial(m, 0.01): mean 1000, sd sqrt(990) = 31.46
    # (the sample sd of 1e5 columns is within 1 % of it)
    assert abs(float(counts.mean()) - 1000.0) < 1e-9
    assert abs(float(counts.std()) / np.sqrt(m * 0.01 * 0.99) - 1.0) < 0.02
