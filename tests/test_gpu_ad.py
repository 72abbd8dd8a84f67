"""GPU: differentiation and batching rules of the operators.

Event and sparse csrmv: the rules of event/_csr_matvec.py:522-546 and
sparse/_csr_mv.py:211-244 are the derivatives of linear maps, so they are
checked against torch autograd through the dense FP64 matrix.  jitconn:
with the row layout the reference's rule (matrix reversed, outdim_parallel
flipped) is the exact transpose of the forward matrix, checked against the
dense matrix; with the lane32 layout it is not (the stream span comes from
num_row in both directions), so the gradient is checked against the
reference's rule evaluated by the oracle.  vmap runs the batch row by row,
like the reference's sequential batching rule.
"""
import numpy as np
import pytest
import torch

import brainpy_amd.math as bm
from brainpy_amd.connect import fixed_prob_csr
from oracle import np_ref, orc

pytestmark = pytest.mark.gpu


def _dense(data, indices, indptr, shape):
    A = np.zeros(shape, np.float64)
    for i in range(shape[0]):
        for k in range(indptr[i], indptr[i + 1]):
            A[i, indices[k]] += data[k] if data.size > 1 else data[0]
    return A


@pytest.mark.parametrize("transpose", [True, False])
@pytest.mark.parametrize("homo", [True, False])
def test_event_csrmv_grads(transpose, homo):
    m, n = 300, 200
    indices, indptr = fixed_prob_csr(m, n, 0.05, seed=1)
    rng = np.random.default_rng(2)
    data = np.float32([0.7]) if homo else rng.normal(0.5, 0.1, indices.size).astype(np.float32)
    ev = (rng.random(m if transpose else n) < 0.3).astype(np.float32) * rng.choice([1.0, 2.0], m if transpose else n)
    ev = ev.astype(np.float32)
    ct = rng.standard_normal(n if transpose else m).astype(np.float32)
    d_t = torch.tensor(data, device="cuda", requires_grad=True)
    e_t = torch.tensor(ev, device="cuda", requires_grad=True)
    y = bm.event.csrmv(d_t, indices, indptr, e_t, shape=(m, n), transpose=transpose)
    y.backward(torch.tensor(ct, device="cuda"))
    A = _dense(data.astype(np.float64), indices, indptr, (m, n))
    on = (ev > 0).astype(np.float64)
    np.testing.assert_allclose(y.detach().cpu().numpy(), on @ A if transpose else A @ on, rtol=1e-5, atol=1e-5)
    # events: the plain sparse matvec of data with the cotangent, transpose flipped
    np.testing.assert_allclose(e_t.grad.cpu().numpy(), A @ ct if transpose else A.T @ ct, rtol=1e-5, atol=1e-5)
    rows = np.repeat(np.arange(m), np.diff(indptr))
    if homo:
        P = _dense(np.float64([1.0]), indices, indptr, (m, n))
        ref = np.dot(ct, on @ P if transpose else P @ on)
        np.testing.assert_allclose(d_t.grad.cpu().numpy(), [ref], rtol=1e-5)
    else:
        ref = ev[rows] * ct[indices] if transpose else ev[indices] * ct[rows]
        np.testing.assert_allclose(d_t.grad.cpu().numpy(), ref, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("transpose", [True, False])
def test_sparse_csrmv_grads_match_dense_autograd(transpose):
    m, n = 250, 180
    indices, indptr = fixed_prob_csr(m, n, 0.05, seed=3)
    rng = np.random.default_rng(4)
    data = rng.normal(0.0, 1.0, indices.size).astype(np.float32)
    x = rng.standard_normal(m if transpose else n).astype(np.float32)
    ct = rng.standard_normal(n if transpose else m).astype(np.float32)
    d_t = torch.tensor(data, device="cuda", requires_grad=True)
    x_t = torch.tensor(x, device="cuda", requires_grad=True)
    y = bm.sparse.csrmv(d_t, indices, indptr, x_t, shape=(m, n), transpose=transpose)
    y.backward(torch.tensor(ct, device="cuda"))
    # dense FP64 autograd of the same linear map
    rows = torch.tensor(np.repeat(np.arange(m), np.diff(indptr)), dtype=torch.long)
    cols = torch.tensor(indices, dtype=torch.long)
    d64 = torch.tensor(data, dtype=torch.float64, requires_grad=True)
    x64 = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    A = torch.zeros((m, n), dtype=torch.float64).index_put((rows, cols), d64, accumulate=True)
    y64 = x64 @ A if transpose else A @ x64
    y64.backward(torch.tensor(ct, dtype=torch.float64))
    np.testing.assert_allclose(y.detach().cpu().numpy(), y64.detach().numpy(), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(x_t.grad.cpu().numpy(), x64.grad.numpy(), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(d_t.grad.cpu().numpy(), d64.grad.numpy(), rtol=1e-5, atol=1e-5)
    # scalar weight: <ct, csrmv(1, x)>
    w_t = torch.tensor([0.3], device="cuda", requires_grad=True)
    bm.sparse.csrmv(w_t, indices, indptr, torch.tensor(x, device="cuda"), shape=(m, n),
                    transpose=transpose).backward(torch.tensor(ct, device="cuda"))
    P = torch.zeros((m, n), dtype=torch.float64).index_put((rows, cols), torch.ones(len(cols), dtype=torch.float64),
                                                           accumulate=True)
    ref = float(torch.dot(torch.tensor(ct, dtype=torch.float64),
                          torch.tensor(x, dtype=torch.float64) @ P if transpose else P @ torch.tensor(x, dtype=torch.float64)))
    np.testing.assert_allclose(w_t.grad.cpu().numpy(), [ref], rtol=1e-5)


@pytest.mark.parametrize("od", [True, False])
@pytest.mark.parametrize("dist,w0,w1", [("homo", 0.8, 0.0), ("uniform", -1.0, 1.0)])
def test_jitconn_vector_grad_row_layout_is_transpose(od, dist, w0, w1):
    shape, p, seed = (120, 90), 0.1, 11
    rng = np.random.default_rng(5)
    x = rng.standard_normal(shape[1]).astype(np.float32)
    ct = rng.standard_normal(shape[0]).astype(np.float32)
    x_t = torch.tensor(x, device="cuda", requires_grad=True)
    f = bm.jitconn.mv_prob_homo if dist == "homo" else bm.jitconn.mv_prob_uniform
    args = (w0,) if dist == "homo" else (w0, w1)
    y = f(x_t, *args, p, seed, shape=shape, outdim_parallel=od, layout="row")
    y.backward(torch.tensor(ct, device="cuda"))
    M = np_ref.dense(dist, od, "row", shape[0], shape[1], np.float32(w0), np.float32(w1), np_ref.conn_len(p), seed)
    np.testing.assert_allclose(y.detach().cpu().numpy(), M @ x, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(x_t.grad.cpu().numpy(), M.T @ ct, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("od", [True, False])
def test_jitconn_event_grad_lane32_follows_reference_rule(od):
    # float events: the cotangent is the non-event operator on ct with the
    # matrix reversed and outdim_parallel flipped, evaluated by the oracle
    shape, p, seed = (150, 110), 0.1, 21
    rng = np.random.default_rng(6)
    ev = ((rng.random(shape[1]) < 0.3) * rng.choice([1.0, 0.5], shape[1])).astype(np.float32)
    ct = rng.standard_normal(shape[0]).astype(np.float32)
    e_t = torch.tensor(ev, device="cuda", requires_grad=True)
    y = bm.jitconn.event_mv_prob_uniform(e_t, 0.0, 1.0, p, seed, shape=shape, outdim_parallel=od)
    y.backward(torch.tensor(ct, device="cuda"))
    ref = orc.jitconn_mv(ct, "uniform", 0, not od, orc.LAYOUT_LANE32, shape[1], shape[0], np.float32(0.0),
                         np.float32(1.0), np_ref.conn_len(p), seed)
    np.testing.assert_allclose(e_t.grad.cpu().numpy(), ref, rtol=1e-5, atol=1e-5)


def test_jitconn_homo_weight_grad_rule():
    # the reference's weight rule: sum(op(ct, weight 1, same flags) * vector)
    shape, p, seed = (100, 100), 0.1, 8
    rng = np.random.default_rng(7)
    x = rng.standard_normal(100).astype(np.float32)
    ct = rng.standard_normal(100).astype(np.float32)
    w_t = torch.tensor(0.6, device="cuda", requires_grad=True)
    y = bm.jitconn.mv_prob_homo(torch.tensor(x, device="cuda"), w_t, p, seed, shape=shape)
    y.backward(torch.tensor(ct, device="cuda"))
    row = orc.jitconn_mv(ct, "homo", 0, True, orc.LAYOUT_LANE32, 100, 100, np.float32(1.0), np.float32(0.0),
                         np_ref.conn_len(p), seed)
    np.testing.assert_allclose(w_t.grad.cpu().numpy(), np.sum(row.astype(np.float64) * x), rtol=1e-5)
    with pytest.raises(AssertionError, match="w_low"):
        bm.jitconn.mv_prob_uniform(torch.tensor(x, device="cuda"), torch.tensor(0.0, device="cuda",
                                   requires_grad=True), 1.0, p, seed, shape=shape)


def test_vmap_and_func_grad():
    m, n = 200, 150
    indices, indptr = fixed_prob_csr(m, n, 0.05, seed=9)
    rng = np.random.default_rng(10)
    data = torch.tensor(rng.normal(0.5, 0.1, indices.size).astype(np.float32), device="cuda")
    evs = torch.tensor(rng.random((4, m)) < 0.2, device="cuda")
    Y = torch.func.vmap(lambda e: bm.event.csrmv(data, indices, indptr, e, shape=(m, n), transpose=True))(evs)
    assert tuple(Y.shape) == (4, n)
    for b in range(4):
        ref = bm.event.csrmv(data, indices, indptr, evs[b], shape=(m, n), transpose=True)
        torch.testing.assert_close(Y[b], ref, rtol=1e-6, atol=1e-6)
    xs = torch.tensor(rng.standard_normal((3, n)).astype(np.float32), device="cuda")
    Yj = torch.func.vmap(lambda x: bm.jitconn.mv_prob_homo(x, 0.5, 0.1, 3, shape=(m, n)))(xs)
    for b in range(3):
        assert torch.equal(Yj[b], bm.jitconn.mv_prob_homo(xs[b], 0.5, 0.1, 3, shape=(m, n)))
    x = torch.tensor(rng.standard_normal(n).astype(np.float32), device="cuda")
    g = torch.func.grad(lambda v: bm.sparse.csrmv(data, indices, indptr, v, shape=(m, n)).sum())(x)
    x2 = x.clone().requires_grad_(True)
    bm.sparse.csrmv(data, indices, indptr, x2, shape=(m, n)).sum().backward()
    torch.testing.assert_close(g, x2.grad, rtol=1e-6, atol=1e-6)
