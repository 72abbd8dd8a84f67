"""The oracle's STDP updates (dnn/linear.py:238-275, 650-700) against an
independent vectorised numpy restatement: every synapse is written at most
once per call, so the update of a synapse is a masked elementwise clip."""
import numpy as np
import pytest

from brainpy_amd.connect import fixed_prob_csr
from brainpy_amd.math.event import csr_to_csc
from oracle import orc


def _clip_mm(x, lo, hi):
    return np.minimum(np.maximum(x, np.float32(lo)), np.float32(hi))


@pytest.mark.parametrize("bounds", [(-np.inf, np.inf), (0.0, 1.0), (-0.3, 0.45)])
def test_csr_on_pre_matches_vectorised(bounds):
    m, n = 60, 45
    ind, ptr = fixed_prob_csr(m, n, 0.2, seed=1)
    rng = np.random.default_rng(2)
    w = rng.normal(0.3, 0.3, ind.size).astype(np.float32)
    spike = rng.random(m) < 0.3
    trace = rng.normal(0.0, 0.2, n).astype(np.float32)
    got = orc.csr_on_pre_update(w, ind, ptr, spike, trace, *bounds)
    rows = np.repeat(np.arange(m), np.diff(ptr))
    ref = w.copy()
    sel = spike[rows]
    ref[sel] = _clip_mm(w[sel] + trace[ind[sel]], *bounds)
    np.testing.assert_array_equal(got, ref)
    # no spikes: unchanged; NaN trace propagates like np.maximum/np.minimum
    np.testing.assert_array_equal(orc.csr_on_pre_update(w, ind, ptr, np.zeros(m, bool), trace, 0.0, 1.0), w)
    tr = trace.copy()
    tr[ind[0]] = np.nan
    out = orc.csr_on_pre_update(w, ind, ptr, np.eye(m, dtype=bool)[0], tr, 0.0, 1.0)
    assert np.isnan(out[0])


def test_csc_on_post_matches_vectorised():
    m, n = 50, 70
    ind, ptr = fixed_prob_csr(m, n, 0.15, seed=3)
    rng = np.random.default_rng(4)
    w = rng.normal(0.5, 0.2, ind.size).astype(np.float32)
    spike = rng.random(n) < 0.25
    trace = rng.normal(0.0, 0.3, m).astype(np.float32)
    pre_ids, colptr, w_ids = csr_to_csc(ind, ptr, (m, n), data=np.arange(ind.size))
    got = orc.csc_on_post_update(w, pre_ids, colptr, w_ids, spike, trace, 0.1, 0.9)
    rows = np.repeat(np.arange(m), np.diff(ptr))
    ref = w.copy()
    sel = spike[ind]
    ref[sel] = _clip_mm(w[sel] + trace[rows[sel]], 0.1, 0.9)
    np.testing.assert_array_equal(got, ref)


def test_dense_forms():
    rng = np.random.default_rng(5)
    W = rng.normal(0.0, 1.0, (20, 30)).astype(np.float32)
    sp_pre = rng.random(20) < 0.3
    sp_post = rng.random(30) < 0.3
    tr_post = rng.normal(0.0, 0.5, 30).astype(np.float32)
    tr_pre = rng.normal(0.0, 0.5, 20).astype(np.float32)
    ref = W.copy()
    ref[sp_pre] = np.clip(W[sp_pre] + tr_post, np.float32(-0.5), np.float32(0.8))
    np.testing.assert_array_equal(orc.dense_on_spike(1, W, sp_pre, tr_post, -0.5, 0.8), ref)
    ref = W.copy()
    ref[:, sp_post] = np.clip(W[:, sp_post] + tr_pre[:, None], np.float32(-0.5), np.float32(0.8))
    np.testing.assert_array_equal(orc.dense_on_spike(0, W, sp_post, tr_pre, -0.5, 0.8), ref)


def _dense(w, ind, ptr, shape):
    W = np.full(shape, np.nan, np.float32)
    rows = np.repeat(np.arange(shape[0]), np.diff(ptr))
    W[rows, ind] = w
    return W


@pytest.mark.parametrize("bounds", [(-np.inf, np.inf), (0.0, 1.0), (0.6, 0.2)])
def test_stdp_against_dense_matrix_definition(bounds):
    # The definition of dnn/linear.py:650-700 on the dense weight matrix (no
    # CSR arithmetic): a spiking pre neuron i moves every existing synapse
    # (i, j) by trace[j]; a spiking post neuron j moves every existing
    # synapse (i, j) by trace[i]; both then take
    # np.minimum(np.maximum(x, w_min), w_max) (with w_min > w_max that is
    # w_max).  FixedProb rows hold distinct targets, so (i, j) is one synapse.
    lo, hi = bounds
    m, n = 9, 7
    ind, ptr = fixed_prob_csr(m, n, 0.45, seed=6)
    rng = np.random.default_rng(7)
    w = rng.normal(0.4, 0.3, ind.size).astype(np.float32)
    exists = ~np.isnan(_dense(w, ind, ptr, (m, n)))
    sp_pre = rng.random(m) < 0.5
    tr_post = rng.normal(0.0, 0.4, n).astype(np.float32)
    W = _dense(w, ind, ptr, (m, n))
    upd = exists & sp_pre[:, None]
    W[upd] = np.minimum(np.maximum((W + tr_post[None, :])[upd], np.float32(lo)), np.float32(hi))
    got = _dense(orc.csr_on_pre_update(w, ind, ptr, sp_pre, tr_post, lo, hi), ind, ptr, (m, n))
    np.testing.assert_array_equal(got[exists], W[exists])
    sp_post = rng.random(n) < 0.5
    tr_pre = rng.normal(0.0, 0.4, m).astype(np.float32)
    W = _dense(w, ind, ptr, (m, n))
    upd = exists & sp_post[None, :]
    W[upd] = np.minimum(np.maximum((W + tr_pre[:, None])[upd], np.float32(lo)), np.float32(hi))
    pre_ids, colptr, w_ids = csr_to_csc(ind, ptr, (m, n), data=np.arange(ind.size))
    got = _dense(orc.csc_on_post_update(w, pre_ids, colptr, w_ids, sp_post, tr_pre, lo, hi), ind, ptr, (m, n))
    np.testing.assert_array_equal(got[exists], W[exists])
    if lo > hi:
        assert np.all(got[upd] == np.float32(hi))
