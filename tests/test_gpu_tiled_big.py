"""GPU parity: the tiled plan beyond 64 output tiles (n > 262,144 outputs,
up to 1024 tiles of 4096 = 4,194,304), where the build counts and places
elements through per-(row, tile) cursors in device memory and a call keeps
fewer output copies (8, 4, 2, 1), against the oracle and the FP64 product."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import brainpy_amd.math as bm
from brainpy_amd._lib import BPXError
from brainpy_amd.connect import fixed_prob_csr
from oracle import orc
from tests.helpers import bernoulli, bf16_bits_to_f32, f32_to_bf16_bits

pytestmark = pytest.mark.gpu

RTOL = 1e-5

# (m, n, synapses per row): 72, 152, 248, 368 and 1024 tiles -> 8, 4, 4, 2
# and 1 output copies
BIG = [(300, 262145, 120), (2000, 600000, 200), (500, 1000000, 150), (200, 1500000, 200), (64, 4194304, 600)]


def _csr(m, n, k, seed, sort):
    return fixed_prob_csr(m, n, (k + 0.5) / n, seed, sort=sort)   # exactly k per row


def _mat(w, indices, indptr, shape):
    data = np.full(indices.shape, np.float64(w)) if np.size(w) == 1 else np.array(w, np.float64)
    return sp.csr_matrix((data, indices.astype(np.int64), indptr.astype(np.int64)), shape=shape)


def _check(y, ref, A_abs_colsum):
    err = np.abs(y.astype(np.float64) - ref)
    scale = A_abs_colsum.max() if A_abs_colsum.size else 1.0
    assert (err <= RTOL * np.maximum(np.abs(ref), scale) + 1e-6).all(), float(err.max())


@pytest.mark.parametrize("shape", BIG)
@pytest.mark.parametrize("wkind", ["homo", "f32", "bf16"])
@pytest.mark.parametrize("sort", [True, False])
def test_big_plan_vs_oracle(shape, wkind, sort):
    m, n, k = shape
    indices, indptr = _csr(m, n, k, 17 + m, sort)
    rng = np.random.default_rng(5)
    data = np.float32(1.0) if wkind == "homo" else rng.normal(0.5, 0.1, indices.size).astype(np.float32)
    plan = bm.event.TiledCSR(data, indices, indptr, shape=(m, n),
                             weight_dtype=torch.bfloat16 if wkind == "bf16" else None)
    P, T = plan.plan.tiles, plan.plan.tile_len
    assert P > 64 and P % 8 == 0 and T <= 4096 and P * T >= n and (P - 1) * T < n
    w = bf16_bits_to_f32(f32_to_bf16_bits(data)) if wkind == "bf16" else data
    A = _mat(w, indices, indptr, (m, n))
    colsum = np.asarray(abs(A).sum(axis=0)).ravel()
    for rate in (0.0, 0.05, 0.5, 1.0):
        ev = bernoulli(m, rate, int(rate * 100) + m)
        y = plan.csrmv(ev).cpu().numpy()
        if wkind == "homo":
            # exact integer counts: bit-identical to the oracle
            np.testing.assert_array_equal(y, orc.event_csrmv(data, indices, indptr, ev, (m, n), True))
        else:
            _check(y, A.T @ ev.astype(np.float64), colsum)


@pytest.mark.parametrize("shape", BIG[1:3])
def test_big_plan_dense_matvec_batch_and_gather(shape):
    m, n, k = shape
    indices, indptr = _csr(m, n, k, 41, False)
    rng = np.random.default_rng(6)
    data = rng.normal(0.5, 0.3, indices.size).astype(np.float32)
    plan = bm.event.TiledCSR(data, indices, indptr, shape=(m, n))
    A = _mat(data, indices, indptr, (m, n))
    absA = abs(A)
    # x @ A, a dense vector with zeros (the products rounded to f32 once)
    x = rng.standard_normal(m).astype(np.float32)
    x[rng.random(m) < 0.3] = 0.0
    y = plan.matvec(x).cpu().numpy().astype(np.float64)
    mag = absA.T @ np.abs(x.astype(np.float64))
    assert np.all(np.abs(y - A.T @ x.astype(np.float64)) <= 1e-5 * mag + 1e-30)
    # a batch of event vectors in one launch, each row as the single call
    evs = np.stack([bernoulli(m, r, 90 + i) for i, r in enumerate((0.02, 0.1, 0.6))])
    yb = bm.event.csrmv_batch(data, indices, indptr, evs, shape=(m, n), transpose=True).cpu().numpy()
    colsum = np.asarray(absA.sum(axis=0)).ravel()
    for i in range(evs.shape[0]):
        _check(yb[i], A.T @ evs[i].astype(np.float64), colsum)
    # the row-wise form on the same plan: A @ v
    v = rng.standard_normal(n).astype(np.float32)
    g = plan.gather(v).cpu().numpy().astype(np.float64)
    gmag = absA @ np.abs(v.astype(np.float64))
    assert np.all(np.abs(g - A @ v.astype(np.float64)) <= 1e-5 * gmag + 1e-30)


def test_big_public_path_runs_on_the_plan():
    # bm.event.csrmv(transpose=True) with 1M outputs: the second call builds
    # the plan, later calls hit it; weight 1 gives exact integer counts
    bm.event.clear_plan_cache()
    m, n, k = 500, 1000000, 150
    indices, indptr = _csr(m, n, k, 43, True)
    dev = torch.device("cuda")
    ti, tp = torch.from_numpy(indices).to(dev), torch.from_numpy(indptr).to(dev)
    w = torch.ones(indices.size, dtype=torch.float32, device=dev)
    for s in range(4):
        ev = bernoulli(m, 0.1, 200 + s)
        y = bm.event.csrmv(w, ti, tp, torch.from_numpy(ev).to(dev), shape=(m, n), transpose=True)
        np.testing.assert_array_equal(y.cpu().numpy(), orc.event_csrmv(np.float32(1.0), indices, indptr, ev, (m, n),
                                                                       True))
    info = bm.event.plan_cache_info()
    assert info["plans"] == 1 and info["builds"] == 1 and info["hits"] == 2
    # bm.sparse.csrmv's gather keeps its CSR kernel above 64 tiles
    v = np.random.default_rng(3).standard_normal(n).astype(np.float32)
    for _ in range(2):
        g = bm.sparse.csrmv(w, ti, tp, torch.from_numpy(v).to(dev), shape=(m, n)).cpu().numpy()
    ref = _mat(np.float32(1.0), indices, indptr, (m, n)) @ v.astype(np.float64)
    mag = abs(_mat(np.float32(1.0), indices, indptr, (m, n))) @ np.abs(v.astype(np.float64))
    assert np.all(np.abs(g - ref) <= 1e-5 * mag + 1e-30)
    bm.event.clear_plan_cache()


def test_big_plan_limits():
    # 1024 tiles of 4096 outputs hold n = 4,194,304; one more output is
    # refused, and so are rows averaging fewer than P / 4 synapses
    n = 4194304
    k = 300
    rng = np.random.default_rng(4)
    idx = np.sort(rng.choice(n, k, replace=False)).astype(np.int32)
    idx[0], idx[-1] = 0, n - 1
    ptr = np.array([0, k], np.int32)
    plan = bm.event.TiledCSR(np.float32(1.0), idx, ptr, shape=(1, n))
    assert plan.plan.tiles == 1024 and plan.plan.tile_len == 4096
    y = plan.csrmv(np.ones(1, bool)).cpu().numpy()
    expect = np.zeros(n, np.float32)
    expect[idx] = 1.0
    np.testing.assert_array_equal(y, expect)
    with pytest.raises(BPXError):
        bm.event.TiledCSR(np.float32(1.0), idx, ptr, shape=(1, n + 1))
    # 300,000 outputs take 80 tiles: at least 20 synapses per row on average
    m, n2 = 50, 300000
    i2, p2 = _csr(m, n2, 19, 5, True)
    with pytest.raises(BPXError):
        bm.event.TiledCSR(np.float32(1.0), i2, p2, shape=(m, n2))
    i3, p3 = _csr(m, n2, 20, 5, True)
    assert i3.size == 20 * m
    plan = bm.event.TiledCSR(np.float32(1.0), i3, p3, shape=(m, n2))
    assert plan.plan.tiles == 80
    ev = bernoulli(m, 0.5, 8)
    np.testing.assert_array_equal(plan.csrmv(ev).cpu().numpy(),
                                  orc.event_csrmv(np.float32(1.0), i3, p3, ev, (m, n2), True))


@pytest.mark.parametrize("shape", [(300, 262145, 120), (3000, 1000000, 80)])
def test_listed_equals_scanned(shape):
    # the active-row list (workspace of bpx_tiled_csr_call_workspace bytes)
    # and the scan of every event deal the rows to the output copies
    # differently; each copy's sum is rounded to f32 once before the copies
    # are added, so with integer-valued weights (every partial sum exact)
    # the two are bit-identical; also for a batch, f32 events under both
    # rules, the dense-vector form, and a workspace one byte short (scan)
    import ctypes
    from brainpy_amd import _lib
    m, n, k = shape
    indices, indptr = _csr(m, n, k, 53, False)
    rng = np.random.default_rng(11)
    data = rng.integers(-3, 4, indices.size).astype(np.float32)
    plan = bm.event.TiledCSR(data, indices, indptr, shape=(m, n))
    L = _lib.lib()
    assert plan._list_ws == 256 + 4 * m
    dev = torch.device("cuda")
    B = 3
    evs = torch.from_numpy(np.stack([bernoulli(m, r, 60 + i) for i, r in enumerate((0.01, 0.3, 1.0))])).to(dev)
    for ev, code, rule in ((evs, _lib.BPX_BOOL, _lib.RULE_POSITIVE),
                           (evs.to(torch.float32) * 2 - 1, _lib.BPX_F32, _lib.RULE_POSITIVE),
                           (evs.to(torch.float32) * 2 - 1, _lib.BPX_F32, _lib.RULE_NONZERO)):
        ev = ev.contiguous()
        ys = torch.empty((B, n), dtype=torch.float32, device=dev)
        _lib.check(L.bpx_event_csrmv_tiled_batch(_lib.stream_ptr(), plan._plan_ref, _lib.ptr(ev), code, rule, B, 0,
                                                 _lib.ptr(ys)), "scan")
        wsb = int(L.bpx_tiled_csr_call_workspace(plan._plan_ref, B))
        assert wsb == 256 + 4 * B * m
        for size in (wsb, wsb - 1):
            ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
            yl = torch.full((B, n), 7.0, dtype=torch.float32, device=dev)
            _lib.check(L.bpx_event_csrmv_tiled_ws(_lib.stream_ptr(), plan._plan_ref, _lib.ptr(ev), code, rule, B, 0,
                                                  _lib.ptr(yl), _lib.ptr(ws), size), "listed")
            np.testing.assert_array_equal(yl.cpu().numpy(), ys.cpu().numpy())
        # accumulate onto a non-zero output
        acc = torch.ones((B, n), dtype=torch.float32, device=dev)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        _lib.check(L.bpx_event_csrmv_tiled_ws(_lib.stream_ptr(), plan._plan_ref, _lib.ptr(ev), code, rule, B, 1,
                                              _lib.ptr(acc), _lib.ptr(ws), wsb), "listed accumulate")
        np.testing.assert_array_equal(acc.cpu().numpy(), (ys + 1.0).cpu().numpy())
    # the dense-vector form: 8-byte workspace scans, the larger one lists
    xv = rng.integers(-2, 3, m) * (rng.random(m) < 0.4)
    x = torch.from_numpy(xv.astype(np.float32)).to(dev)
    outs = []
    for wsb in (8, 256 + plan._list_ws):
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        y = torch.empty(n, dtype=torch.float32, device=dev)
        _lib.check(L.bpx_csrmv_tiled(_lib.stream_ptr(), ctypes.byref(plan.plan), _lib.ptr(x), _lib.BPX_F32, 0,
                                     _lib.ptr(y), _lib.ptr(ws), wsb), "dense")
        outs.append(y.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    A = _mat(data, indices, indptr, (m, n))
    np.testing.assert_array_equal(outs[1], A.T @ x.cpu().numpy().astype(np.float64))
    # normal f32 weights: both forms within 1e-5 of the FP64 product
    wn = rng.normal(0.5, 0.2, indices.size).astype(np.float32)
    pn = bm.event.TiledCSR(wn, indices, indptr, shape=(m, n))
    An = _mat(wn, indices, indptr, (m, n))
    colsum = np.asarray(abs(An).sum(axis=0)).ravel()
    ev = evs[1].contiguous()
    ref = An.T @ ev.cpu().numpy().astype(np.float64)
    ys = torch.empty(n, dtype=torch.float32, device=dev)
    _lib.check(L.bpx_event_csrmv_tiled_batch(_lib.stream_ptr(), pn._plan_ref, _lib.ptr(ev), _lib.BPX_BOOL,
                                             _lib.RULE_POSITIVE, 1, 0, _lib.ptr(ys)), "scan")
    _check(ys.cpu().numpy(), ref, colsum)
    _check(pn.csrmv(ev).cpu().numpy(), ref, colsum)
    # plans of at most 64 tiles take no list
    i2, p2 = fixed_prob_csr(100, 100000, 0.01, 3)
    small = bm.event.TiledCSR(np.float32(1.0), i2, p2, shape=(100, 100000))
    assert small._list_ws == 0 and int(L.bpx_tiled_csr_call_workspace(small._plan_ref, 4)) == 0
