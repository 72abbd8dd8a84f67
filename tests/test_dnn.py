"""Projection layers (dnn/linear.py:546-647, 805-1280): construction and
argument rules on the CPU; on the GPU, each layer's output equals the
operator it wraps, and batched input equals the rows one by one."""
import numpy as np
import pytest
import torch

from brainpy_amd import dnn
from brainpy_amd.connect import CSRConn, FixedProb


def test_fixed_prob_connector():
    c = FixedProb(0.1, 50, 40, seed=3)
    ind, ptr = c.require("csr")
    assert c.pre_num == 50 and c.post_num == 40
    assert len(ptr) == 51 and ptr[-1] == 50 * 4
    for i in range(50):
        row = ind[ptr[i]:ptr[i + 1]]
        assert len(np.unique(row)) == len(row) == 4
    with pytest.raises(ValueError):
        c.require("coo")
    with pytest.raises(ValueError):
        FixedProb(1.5, 3, 3)


def test_csr_conn_checks_indptr():
    with pytest.raises(ValueError):
        CSRConn(np.zeros(3, np.int32), np.zeros(3, np.int32), 5, 4)


def test_weight_parameter_forms():
    assert dnn._parameter(0.5, 10) == 0.5
    assert dnn._parameter(np.float32(2.0), 10) == 2.0
    t = dnn._parameter(lambda shape: np.ones(shape, np.float32), 7)
    assert tuple(t.shape) == (7,)
    with pytest.raises(ValueError):
        dnn._parameter(np.ones(3, np.float32), 7)


def test_jit_layer_seed_and_shape():
    layer = dnn.JitFPHomoLinear(30, 20, 0.1, 0.5)
    assert 0 <= layer.seed < 100000
    ev = dnn.EventJitFPHomoLinear(30, 20, 0.1, 0.5)
    assert 0 <= ev.seed < 1000000
    assert layer._kw() == {"shape": (20, 30), "transpose": False, "outdim_parallel": True}
    atomic = dnn.JitFPUniformLinear(30, 20, 0.1, 0.0, 1.0, seed=4, atomic=True)
    assert atomic._kw()["outdim_parallel"] is False


@pytest.mark.gpu
def test_csr_layers_vs_oracle():
    # EventCSRLinear / CSRLinear against the oracle's CSR products
    # (dnn/linear.py:630-634 calls event csrmv with shape (pre, post),
    # transpose=True); heterogeneous f32 weights within 1e-5 of each output's
    # sum of |terms|, homogeneous weights on bool spikes exactly
    from oracle import orc
    from tests.helpers import dense_of
    conn = FixedProb(0.05, 300, 200, seed=1)
    w = np.random.default_rng(0).normal(0.5, 0.1, conn.indices.size).astype(np.float32)
    ev = np.random.default_rng(1).random(300) < 0.2
    lay = dnn.EventCSRLinear(conn, w)
    D = dense_of(w, conn.indices, conn.indptr, (300, 200))
    mag = ev.astype(np.float64) @ np.abs(D)
    for _ in range(3):        # the second call builds the plan, the third runs on it
        y = lay(ev).cpu().numpy()
        ref = orc.event_csrmv(w.astype(np.float64), conn.indices, conn.indptr, ev, (300, 200), True)
        assert np.all(np.abs(y - ref) <= 1e-5 * mag + 1e-30)
    homo = dnn.EventCSRLinear(conn, 0.25)
    for _ in range(3):
        np.testing.assert_array_equal(homo(ev).cpu().numpy(),
                                      orc.event_csrmv(np.float32(0.25), conn.indices, conn.indptr, ev, (300, 200),
                                                      True))
    x = np.random.default_rng(2).standard_normal(300).astype(np.float32)
    lin = dnn.CSRLinear(conn, 0.7)
    y = lin(x).cpu().numpy()
    ref = orc.csrmv(np.float64(0.7), conn.indices, conn.indptr, x.astype(np.float64), (300, 200), True)
    magx = np.abs(x.astype(np.float64)) @ np.abs(dense_of(np.float32(0.7), conn.indices, conn.indptr, (300, 200)))
    assert np.all(np.abs(y - ref) <= 1e-5 * magx + 1e-30)
    # batched input: every row against the oracle
    X = np.random.default_rng(3).random((2, 3, 300)) < 0.2
    Y = lay(X).cpu().numpy()
    assert Y.shape == (2, 3, 200)
    for i in range(2):
        for j in range(3):
            ref = orc.event_csrmv(w.astype(np.float64), conn.indices, conn.indptr, X[i, j], (300, 200), True)
            assert np.all(np.abs(Y[i, j] - ref) <= 1e-5 * (X[i, j].astype(np.float64) @ np.abs(D)) + 1e-30)


@pytest.mark.gpu
@pytest.mark.parametrize("atomic", [False, True])
def test_jit_layers_vs_oracle(atomic):
    # JitFP*/EventJitFP* layers (dnn/linear.py:805-1280: shape (num_out,
    # num_in), transpose=False, outdim_parallel = not atomic) against the
    # oracle's jitconn restatement: bool events with homogeneous and uniform
    # weights bit-exact (row-parallel sums in stream order); dense vectors,
    # the atomic scatter and normal weights (device log/sin) within 1e-5 of
    # the term magnitudes
    from oracle import np_ref, orc
    ev = np.random.default_rng(5).random(300) < 0.1
    x = np.random.default_rng(6).standard_normal(300).astype(np.float32)
    clen = np_ref.conn_len(0.05)
    cases = [
        (dnn.EventJitFPHomoLinear(300, 200, 0.05, 0.5, seed=9, atomic=atomic), ev, "homo", 1, 0.5, 0.0),
        (dnn.EventJitFPUniformLinear(300, 200, 0.05, 0.0, 1.0, seed=9, atomic=atomic), ev, "uniform", 1, 0.0, 1.0),
        (dnn.EventJitFPNormalLinear(300, 200, 0.05, 0.5, 0.1, seed=9, atomic=atomic), ev, "normal", 1, 0.5, 0.1),
        (dnn.JitFPHomoLinear(300, 200, 0.05, 0.5, seed=9, atomic=atomic), x, "homo", 0, 0.5, 0.0),
        (dnn.JitFPUniformLinear(300, 200, 0.05, 0.0, 1.0, seed=9, atomic=atomic), x, "uniform", 0, 0.0, 1.0),
        (dnn.JitFPNormalLinear(300, 200, 0.05, 0.5, 0.1, seed=9, atomic=atomic), x, "normal", 0, 0.5, 0.1),
    ]
    for layer, inp, dist, vmode, w0, w1 in cases:
        y = layer(inp).cpu().numpy()
        ref = orc.jitconn_mv(inp, dist, vmode, not atomic, orc.LAYOUT_LANE32, 200, 300, w0, w1, clen, 9)
        exact = (not atomic) and vmode == 1 and dist in ("homo", "uniform")
        if exact:
            np.testing.assert_array_equal(y, ref)
        else:
            mag = orc.jitconn_mv(np.abs(inp).astype(inp.dtype), "homo", vmode, not atomic, orc.LAYOUT_LANE32, 200,
                                 300, 1.0, 0.0, clen, 9)
            wmax = abs(w0) + 6 * abs(w1)
            assert np.all(np.abs(y - ref) <= 1e-5 * wmax * np.abs(mag) + 1e-6), (dist, vmode)
    batch = np.stack([ev, ~ev])
    lay = cases[0][0]
    Y = lay(batch).cpu().numpy()
    for b in range(2):
        ref = orc.jitconn_mv(batch[b], "homo", 1, not atomic, orc.LAYOUT_LANE32, 200, 300, 0.5, 0.0, clen, 9)
        if atomic:
            np.testing.assert_allclose(Y[b], ref, rtol=1e-6, atol=1e-5)
        else:
            np.testing.assert_array_equal(Y[b], ref)


def test_sparse_python_scalar_is_float32():
    # a python float weight is float32 (JAX without x64), like the event op's
    from brainpy_amd.math.sparse import _as_array
    assert _as_array(0.7).dtype == np.float32
    assert _as_array([1.0, 2.0]).dtype == np.float32
    assert _as_array(3).dtype == np.int32
    assert _as_array(np.ones(2)).dtype == np.float64
