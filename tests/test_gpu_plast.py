"""GPU: STDP weight updates (dnn/linear.py:238-275, 650-700) against the
oracle, bit-exact (each weight is written at most once per call)."""
import numpy as np
import pytest
import torch

from brainpy_amd import dnn
from brainpy_amd.connect import FixedProb, fixed_prob_csr
from brainpy_amd.math.event import csr_to_csc
from oracle import orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("idx_dtype", [np.int32, np.int64])
@pytest.mark.parametrize("spike_kind", ["bool", "float"])
@pytest.mark.parametrize("bounds", [(None, None), (0.0, 1.0)])
def test_csr_on_pre_exact(idx_dtype, spike_kind, bounds):
    m, n = 700, 900
    ind, ptr = fixed_prob_csr(m, n, 0.05, seed=11)
    rng = np.random.default_rng(12)
    w = rng.normal(0.5, 0.3, ind.size).astype(np.float32)
    spike = rng.random(m) < 0.2
    sp = spike if spike_kind == "bool" else spike.astype(np.float32) * 2.0
    trace = rng.normal(0.0, 0.2, n).astype(np.float32)
    got = dnn.csr_on_pre_update(w, ind.astype(idx_dtype), ptr.astype(idx_dtype), sp, trace, *bounds)
    lo = -np.inf if bounds[0] is None else bounds[0]
    hi = np.inf if bounds[1] is None else bounds[1]
    np.testing.assert_array_equal(got.cpu().numpy(), orc.csr_on_pre_update(w, ind, ptr, spike, trace, lo, hi))


def test_csc_on_post_exact_and_in_place():
    m, n = 600, 800
    ind, ptr = fixed_prob_csr(m, n, 0.05, seed=13)
    rng = np.random.default_rng(14)
    w = rng.normal(0.5, 0.3, ind.size).astype(np.float32)
    spike = rng.random(n) < 0.2
    trace = rng.normal(0.0, 0.2, m).astype(np.float32)
    pre_ids, colptr, w_ids = csr_to_csc(ind, ptr, (m, n), data=np.arange(ind.size))
    ref = orc.csc_on_post_update(w, pre_ids, colptr, w_ids, spike, trace, -0.2, 1.1)
    wt = torch.tensor(w, device="cuda")
    out = dnn.csc_on_post_update(wt, pre_ids, colptr, np.asarray(w_ids, np.int32), spike, trace, -0.2, 1.1, out=wt)
    assert out.data_ptr() == wt.data_ptr()
    np.testing.assert_array_equal(wt.cpu().numpy(), ref)


def test_f64_weights():
    m, n = 300, 400
    ind, ptr = fixed_prob_csr(m, n, 0.05, seed=15)
    rng = np.random.default_rng(16)
    w = rng.normal(0.5, 0.3, ind.size)
    spike = rng.random(m) < 0.3
    trace = rng.normal(0.0, 0.2, n)
    got = dnn.csr_on_pre_update(torch.tensor(w, device="cuda"), ind, ptr, spike, torch.tensor(trace), 0.0, 1.0)
    assert got.dtype == torch.float64
    rows = np.repeat(np.arange(m), np.diff(ptr))
    ref = w.copy()
    sel = spike[rows]
    ref[sel] = np.minimum(np.maximum(w[sel] + trace[ind[sel]], 0.0), 1.0)
    np.testing.assert_array_equal(got.cpu().numpy(), ref)


def test_dense_forms_exact():
    rng = np.random.default_rng(17)
    W = rng.normal(0.0, 1.0, (130, 170)).astype(np.float32)
    sp_pre, sp_post = rng.random(130) < 0.3, rng.random(170) < 0.3
    tr_post, tr_pre = rng.normal(0, 0.5, 170).astype(np.float32), rng.normal(0, 0.5, 130).astype(np.float32)
    np.testing.assert_array_equal(dnn.dense_on_pre(W, sp_pre, tr_post, -0.5, 0.8).cpu().numpy(),
                                  orc.dense_on_spike(1, W, sp_pre, tr_post, -0.5, 0.8))
    np.testing.assert_array_equal(dnn.dense_on_post(W, sp_post, tr_pre, -0.5, 0.8).cpu().numpy(),
                                  orc.dense_on_spike(0, W, sp_post, tr_pre, -0.5, 0.8))


def test_layer_stdp_sequence():
    conn = FixedProb(0.05, 400, 300, seed=21)
    rng = np.random.default_rng(22)
    w0 = rng.normal(0.5, 0.1, conn.indices.size).astype(np.float32)
    lay = dnn.EventCSRLinear(conn, w0)
    w_ref = w0.copy()
    pre_ids, colptr, w_ids = csr_to_csc(conn.indices, conn.indptr, (400, 300), data=np.arange(conn.indices.size))
    for step in range(5):
        sp_pre, sp_post = rng.random(400) < 0.2, rng.random(300) < 0.2
        tr_post, tr_pre = rng.normal(0, 0.05, 300).astype(np.float32), rng.normal(0, 0.05, 400).astype(np.float32)
        lay.stdp_update(on_pre={"spike": sp_pre, "trace": tr_post}, on_post={"spike": sp_post, "trace": tr_pre},
                        w_min=0.0, w_max=1.0)
        w_ref = orc.csr_on_pre_update(w_ref, conn.indices, conn.indptr, sp_pre, tr_post, 0.0, 1.0)
        w_ref = orc.csc_on_post_update(w_ref, pre_ids, colptr, w_ids, sp_post, tr_pre, 0.0, 1.0)
    np.testing.assert_array_equal(lay.weight.cpu().numpy(), w_ref)
    with pytest.raises(ValueError, match="scalar"):
        dnn.EventCSRLinear(conn, 0.5).stdp_update(on_pre={"spike": np.ones(400, bool), "trace": np.zeros(300)})
