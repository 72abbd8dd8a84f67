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
def test_csr_layers_match_operators():
    import brainpy_amd.math as bm
    conn = FixedProb(0.05, 300, 200, seed=1)
    w = np.random.default_rng(0).normal(0.5, 0.1, conn.indices.size).astype(np.float32)
    ev = np.random.default_rng(1).random(300) < 0.2
    lay = dnn.EventCSRLinear(conn, w)
    y = lay(ev)
    ref = bm.event.csrmv(w, conn.indices, conn.indptr, ev, shape=(300, 200), transpose=True)
    # the scatter adds with FP32 atomics: the order (and last bits) vary per call
    assert torch.allclose(y, ref, rtol=1e-6, atol=1e-6)
    x = np.random.default_rng(2).standard_normal(300).astype(np.float32)
    lin = dnn.CSRLinear(conn, 0.7)
    y = lin(x)
    ref = bm.sparse.csrmv(0.7, conn.indices, conn.indptr, x, shape=(300, 200), transpose=True)
    assert torch.allclose(y, ref, rtol=1e-6, atol=1e-5)
    # batched input: rows one by one
    X = np.random.default_rng(3).random((2, 3, 300)) < 0.2
    Y = lay(X)
    assert tuple(Y.shape) == (2, 3, 200)
    for i in range(2):
        for j in range(3):
            assert torch.allclose(Y[i, j], lay(X[i, j]), rtol=1e-6, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("atomic", [False, True])
def test_jit_layers_match_operators(atomic):
    import brainpy_amd.math as bm
    ev = np.random.default_rng(5).random(300) < 0.1
    x = np.random.default_rng(6).standard_normal(300).astype(np.float32)
    kw = dict(shape=(200, 300), transpose=False, outdim_parallel=not atomic)
    cases = [
        (dnn.EventJitFPHomoLinear(300, 200, 0.05, 0.5, seed=9, atomic=atomic), ev,
         bm.jitconn.event_mv_prob_homo(ev, 0.5, 0.05, 9, **kw)),
        (dnn.EventJitFPUniformLinear(300, 200, 0.05, 0.0, 1.0, seed=9, atomic=atomic), ev,
         bm.jitconn.event_mv_prob_uniform(ev, 0.0, 1.0, 0.05, 9, **kw)),
        (dnn.EventJitFPNormalLinear(300, 200, 0.05, 0.5, 0.1, seed=9, atomic=atomic), ev,
         bm.jitconn.event_mv_prob_normal(ev, 0.5, 0.1, 0.05, 9, **kw)),
        (dnn.JitFPHomoLinear(300, 200, 0.05, 0.5, seed=9, atomic=atomic), x,
         bm.jitconn.mv_prob_homo(x, 0.5, 0.05, 9, **kw)),
        (dnn.JitFPUniformLinear(300, 200, 0.05, 0.0, 1.0, seed=9, atomic=atomic), x,
         bm.jitconn.mv_prob_uniform(x, 0.0, 1.0, 0.05, 9, **kw)),
        (dnn.JitFPNormalLinear(300, 200, 0.05, 0.5, 0.1, seed=9, atomic=atomic), x,
         bm.jitconn.mv_prob_normal(x, 0.5, 0.1, 0.05, 9, **kw)),
    ]
    for layer, inp, ref in cases:
        y = layer(inp)
        if atomic:
            assert torch.allclose(y, ref, rtol=1e-5, atol=1e-5)
        else:
            assert torch.equal(y, ref)
    batch = np.stack([ev, ~ev])
    lay = cases[0][0]
    Y = lay(batch)
    assert torch.equal(Y[0], lay(ev)) and torch.equal(Y[1], lay(~ev))


def test_sparse_python_scalar_is_float32():
    # a python float weight is float32 (JAX without x64), like the event op's
    from brainpy_amd.math.sparse import _as_array
    assert _as_array(0.7).dtype == np.float32
    assert _as_array([1.0, 2.0]).dtype == np.float32
    assert _as_array(3).dtype == np.int32
    assert _as_array(np.ones(2)).dtype == np.float64
