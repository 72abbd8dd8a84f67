"""The plan cache behind bm.event.csrmv(..., transpose=True): the second call
on the same (data, indices, indptr) builds a tiled plan and later calls run
on it; in-place writes (torch ops, the STDP kernels) invalidate it; results
stay equal to the oracle's on every call."""
import numpy as np
import pytest
import torch

import brainpy_amd.math as bm
from brainpy_amd import dnn
from oracle import orc
from tests.helpers import bernoulli, dense_of, fixed_prob_csr

pytestmark = pytest.mark.gpu

DEV = "cuda"


@pytest.fixture(autouse=True)
def fresh_cache():
    bm.event.clear_plan_cache()
    bm.event.set_plan_cache(enabled=True, capacity=8)
    yield
    bm.event.clear_plan_cache()
    bm.event.set_plan_cache(enabled=True, capacity=8)


def _inputs(m=2000, n=3000, p=0.02, seed=3):
    idx, ptr = fixed_prob_csr(m, n, p, seed)
    w = np.random.default_rng(seed).normal(0.5, 0.1, idx.size).astype(np.float32)
    return idx, ptr, w


def _check(y, w, idx, ptr, ev, shape):
    D = dense_of(w, idx, ptr, shape)
    on = np.asarray(ev, np.float64)
    ref = on @ D
    assert np.all(np.abs(y.cpu().numpy() - ref) <= 1e-5 * (on @ np.abs(D)) + 1e-30)


def test_second_call_builds_third_hits():
    idx, ptr, w = _inputs()
    t = [torch.from_numpy(a).to(DEV) for a in (w, idx, ptr)]
    ev = torch.from_numpy(bernoulli(2000, 0.1, 5)).to(DEV)
    for k in range(4):
        y = bm.event.csrmv(*t, ev, shape=(2000, 3000), transpose=True)
        _check(y, w, idx, ptr, ev.cpu().numpy(), (2000, 3000))
        info = bm.event.plan_cache_info()
        assert info["builds"] == (0 if k == 0 else 1)
        assert info["hits"] == max(0, k - 1)
    assert bm.event.plan_cache_info()["plans"] == 1


def test_homogeneous_integer_counts_exact_on_plan():
    idx, ptr, _ = _inputs(seed=4)
    ti, tp = torch.from_numpy(idx).to(DEV), torch.from_numpy(ptr).to(DEV)
    ev = bernoulli(2000, 0.3, 6)
    ref = orc.event_csrmv(np.float32(1.0), idx, ptr, ev, (2000, 3000), True)
    for _ in range(3):
        y = bm.event.csrmv(1.0, ti, tp, torch.from_numpy(ev).to(DEV), shape=(2000, 3000), transpose=True)
        np.testing.assert_array_equal(y.cpu().numpy(), ref)
    # a host scalar weight is keyed by value: another value gets its own plan
    for _ in range(3):
        y = bm.event.csrmv(2.5, ti, tp, torch.from_numpy(ev).to(DEV), shape=(2000, 3000), transpose=True)
        np.testing.assert_array_equal(y.cpu().numpy(), ref * np.float32(2.5))
    assert bm.event.plan_cache_info()["plans"] == 2


def test_in_place_writes_invalidate():
    idx, ptr, w = _inputs(seed=5)
    tw, ti, tp = (torch.from_numpy(a).to(DEV) for a in (w, idx, ptr))
    ev = torch.from_numpy(bernoulli(2000, 0.2, 7)).to(DEV)
    for _ in range(3):
        bm.event.csrmv(tw, ti, tp, ev, shape=(2000, 3000), transpose=True)
    tw.mul_(2.0)      # torch in-place op: version bump
    w2 = w * np.float32(2.0)
    for _ in range(3):
        _check(bm.event.csrmv(tw, ti, tp, ev, shape=(2000, 3000), transpose=True), w2, idx, ptr,
               ev.cpu().numpy(), (2000, 3000))
    # the STDP kernel writes the weights in place through its raw pointer
    spike = bernoulli(2000, 0.5, 8)
    trace = np.random.default_rng(9).normal(0.0, 0.05, 3000).astype(np.float32)
    dnn.csr_on_pre_update(tw, ti, tp, spike, trace, 0.0, 2.0, out=tw)
    w3 = orc.csr_on_pre_update(w2, idx, ptr, spike, trace, 0.0, 2.0)
    np.testing.assert_array_equal(tw.cpu().numpy(), w3)
    for _ in range(3):
        _check(bm.event.csrmv(tw, ti, tp, ev, shape=(2000, 3000), transpose=True), w3, idx, ptr,
               ev.cpu().numpy(), (2000, 3000))


def test_layer_steps_run_on_one_plan():
    from brainpy_amd.connect import FixedProb
    conn = FixedProb(0.02, 4000, 5000, seed=2)
    w = np.random.default_rng(1).normal(0.5, 0.1, conn.indices.size).astype(np.float32)
    lay = dnn.EventCSRLinear(conn, w)
    for k in range(6):
        ev = bernoulli(4000, 0.05, 100 + k)
        _check(lay(ev), w, conn.indices, conn.indptr, ev, (4000, 5000))
    info = bm.event.plan_cache_info()
    assert info["builds"] == 1 and info["hits"] == 4


def test_ineligible_and_disabled_run_stateless():
    # rows too short for the plan's 80 tiles (10 synapses per row), f64
    # weights, f64 events and a disabled cache all take the stateless
    # scatter with correct results
    m, n = 50, 300000
    idx, ptr = fixed_prob_csr(m, n, 10.5 / n, 11)
    w = np.random.default_rng(2).normal(0.5, 0.1, idx.size).astype(np.float32)
    tw, ti, tp = (torch.from_numpy(a).to(DEV) for a in (w, idx, ptr))
    ev = torch.from_numpy(bernoulli(m, 0.5, 3)).to(DEV)
    for _ in range(3):
        _check(bm.event.csrmv(tw, ti, tp, ev, shape=(m, n), transpose=True), w, idx, ptr, ev.cpu().numpy(), (m, n))
    info = bm.event.plan_cache_info()
    assert info["plans"] == 0 and info["ineligible"] == 1
    idx, ptr, w = _inputs(seed=6)
    t64 = torch.from_numpy(w.astype(np.float64)).to(DEV)
    ti, tp = torch.from_numpy(idx).to(DEV), torch.from_numpy(ptr).to(DEV)
    ev = torch.from_numpy(bernoulli(2000, 0.2, 4)).to(DEV)
    for _ in range(3):
        y = bm.event.csrmv(t64, ti, tp, ev, shape=(2000, 3000), transpose=True)
        assert y.dtype == torch.float64
    bm.event.set_plan_cache(enabled=False)
    tw = torch.from_numpy(w).to(DEV)
    for _ in range(3):
        _check(bm.event.csrmv(tw, ti, tp, ev, shape=(2000, 3000), transpose=True), w, idx, ptr, ev.cpu().numpy(),
               (2000, 3000))
    assert bm.event.plan_cache_info()["builds"] == 1


def test_capacity_evicts_least_recent():
    bm.event.set_plan_cache(capacity=2)
    mats = [_inputs(m=500, n=700, seed=20 + k) for k in range(3)]
    ts = [[torch.from_numpy(a).to(DEV) for a in (w, idx, ptr)] for idx, ptr, w in mats]
    ev = torch.from_numpy(bernoulli(500, 0.2, 1)).to(DEV)
    for t in ts:
        for _ in range(2):
            bm.event.csrmv(*t, ev, shape=(500, 700), transpose=True)
    assert bm.event.plan_cache_info()["plans"] == 2
    for (idx, ptr, w), t in zip(mats, ts):
        _check(bm.event.csrmv(*t, ev, shape=(500, 700), transpose=True), w, idx, ptr, ev.cpu().numpy(), (500, 700))


def test_sparse_transpose_runs_on_the_plan():
    # bm.sparse.csrmv(..., transpose=True) shares the cache: the second call
    # on the same arrays builds the plan, later calls hit it; every call
    # within 1e-5 of the FP64 product's term magnitudes; f32 and bf16 data
    idx, ptr, w = _inputs(seed=11)
    for wd in (torch.float32, torch.bfloat16):
        bm.event.clear_plan_cache()
        tw = torch.from_numpy(w).to(DEV).to(wd)
        wr = tw.float().cpu().numpy()
        ti, tp = torch.from_numpy(idx).to(DEV), torch.from_numpy(ptr).to(DEV)
        D = dense_of(wr, idx, ptr, (2000, 3000))
        for k in range(4):
            x = np.random.default_rng(k).standard_normal(2000).astype(np.float32)
            x[::3] = 0.0
            y = bm.sparse.csrmv(tw, ti, tp, torch.from_numpy(x).to(DEV), shape=(2000, 3000), transpose=True)
            assert y.dtype == torch.float32
            ref = x.astype(np.float64) @ D
            mag = np.abs(x.astype(np.float64)) @ np.abs(D)
            assert np.all(np.abs(y.cpu().numpy() - ref) <= 1e-5 * mag + 1e-30)
            info = bm.event.plan_cache_info()
            assert info["builds"] == (0 if k == 0 else 1) and info["hits"] == max(0, k - 1)
        # an event call on the same arrays reuses the plan
        ev = torch.from_numpy(bernoulli(2000, 0.1, 8)).to(DEV)
        bm.event.csrmv(tw, ti, tp, ev, shape=(2000, 3000), transpose=True)
        assert bm.event.plan_cache_info()["builds"] == 1


def test_sparse_transpose_nonfinite_on_the_plan_equals_stateless():
    # inf / NaN in the vector: the plan's outputs have the stateless
    # scatter's NaN / +inf / -inf pattern and agree elsewhere
    idx, ptr, w = _inputs(seed=12)
    t = [torch.from_numpy(a).to(DEV) for a in (w, idx, ptr)]
    x = np.random.default_rng(3).standard_normal(2000).astype(np.float32)
    x[5], x[9], x[11] = np.inf, -np.inf, np.nan
    xt = torch.from_numpy(x).to(DEV)
    bm.event.set_plan_cache(enabled=False)
    y0 = bm.sparse.csrmv(*t, xt, shape=(2000, 3000), transpose=True).cpu().numpy()
    bm.event.set_plan_cache(enabled=True)
    for _ in range(3):
        y = bm.sparse.csrmv(*t, xt, shape=(2000, 3000), transpose=True).cpu().numpy()
    assert bm.event.plan_cache_info()["hits"] >= 1
    for f in (np.isnan, np.isposinf, np.isneginf):
        np.testing.assert_array_equal(f(y), f(y0))
    fin = np.isfinite(y0)
    assert (~fin).any() and fin.any()
    np.testing.assert_allclose(y[fin], y0[fin], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("wkind", ["f32", "bf16", "homo"])
@pytest.mark.parametrize("shape", [(2000, 3000), (777, 5000), (3000, 150000)])
def test_plan_gather_dense_and_events(wkind, shape):
    # A @ v on the plan (transpose=False): dense f32 v within 1e-5 of each
    # row's sum of |w v| against FP64; events (bool, f32 by both rules)
    # summing the truthy columns' weights; calls are deterministic
    m, n = shape
    idx, ptr = fixed_prob_csr(m, n, min(0.02, 60.0 / n), 21)
    rng = np.random.default_rng(4)
    w = np.float32(0.7) if wkind == "homo" else rng.normal(0.5, 0.3, idx.size).astype(np.float32)
    plan = bm.event.TiledCSR(w, idx, ptr, shape=shape, weight_dtype=torch.bfloat16 if wkind == "bf16" else None)
    if wkind == "bf16":
        w = torch.from_numpy(w).to(torch.bfloat16).float().numpy()
    D = dense_of(w, idx, ptr, shape).astype(np.float64)
    x = rng.standard_normal(n).astype(np.float32)
    x[rng.random(n) < 0.3] = 0.0
    y = plan.gather(x)
    assert torch.equal(y, plan.gather(x))
    ref, mag = D @ x.astype(np.float64), np.abs(D) @ np.abs(x.astype(np.float64))
    assert np.all(np.abs(y.cpu().numpy() - ref) <= 1e-5 * mag + 1e-30)
    evb = bernoulli(n, 0.1, 5)
    ye = plan.gather(evb, event=True).cpu().numpy()
    on = evb.astype(np.float64)
    assert np.all(np.abs(ye - D @ on) <= 1e-5 * (np.abs(D) @ on) + 1e-30)
    evf = np.where(evb, rng.choice([1.5, -2.0], n), 0.0).astype(np.float32)
    for rule, truthy in ((bm.event._lib.RULE_POSITIVE, evf > 0), (bm.event._lib.RULE_NONZERO, evf != 0)):
        yf = plan.gather(evf, event=True, rule=rule).cpu().numpy()
        on = truthy.astype(np.float64)
        assert np.all(np.abs(yf - D @ on) <= 1e-5 * (np.abs(D) @ on) + 1e-30)


def test_public_gather_paths_run_on_the_plan():
    # bm.sparse.csrmv with transpose=False: the second call on the same
    # arrays builds the plan, later calls hit it; bm.event.csrmv's gather
    # builds the plan of the transposed matrix on its second call (a
    # separate entry), exact integer counts either way
    idx, ptr, w = _inputs(seed=13)
    ti, tp = torch.from_numpy(idx).to(DEV), torch.from_numpy(ptr).to(DEV)
    tw = torch.from_numpy(w).to(DEV)
    D = dense_of(w, idx, ptr, (2000, 3000)).astype(np.float64)
    for k in range(4):
        x = np.random.default_rng(k).standard_normal(3000).astype(np.float32)
        y = bm.sparse.csrmv(tw, ti, tp, torch.from_numpy(x).to(DEV), shape=(2000, 3000))
        assert np.all(np.abs(y.cpu().numpy() - D @ x) <= 1e-5 * (np.abs(D) @ np.abs(x)) + 1e-30)
    info = bm.event.plan_cache_info()
    assert info["builds"] == 1 and info["hits"] == 2
    ev = bernoulli(3000, 0.2, 9)
    ref = orc.event_csrmv(np.float32(1.0), idx, ptr, ev, (2000, 3000), False)
    for _ in range(3):
        y = bm.event.csrmv(1.0, ti, tp, torch.from_numpy(ev).to(DEV), shape=(2000, 3000))
        np.testing.assert_array_equal(y.cpu().numpy(), ref)
    assert bm.event.plan_cache_info()["builds"] == 2
    # the plan's event gather (TiledCSR.gather) gives the same integers
    plan = bm.event.TiledCSR(1.0, ti, tp, shape=(2000, 3000))
    np.testing.assert_array_equal(plan.gather(ev, event=True).cpu().numpy(), ref)


def test_plan_gather_nonfinite_vector_matches_stateless():
    idx, ptr, w = _inputs(seed=14)
    t = [torch.from_numpy(a).to(DEV) for a in (w, idx, ptr)]
    x = np.random.default_rng(2).standard_normal(3000).astype(np.float32)
    x[3], x[40], x[77] = np.inf, -np.inf, np.nan
    xt = torch.from_numpy(x).to(DEV)
    bm.event.set_plan_cache(enabled=False)
    y0 = bm.sparse.csrmv(*t, xt, shape=(2000, 3000)).cpu().numpy()
    bm.event.set_plan_cache(enabled=True)
    for _ in range(3):
        y = bm.sparse.csrmv(*t, xt, shape=(2000, 3000)).cpu().numpy()
    for f in (np.isnan, np.isposinf, np.isneginf):
        np.testing.assert_array_equal(f(y), f(y0))
    fin = np.isfinite(y0)
    assert (~fin).any() and fin.any()
    np.testing.assert_allclose(y[fin], y0[fin], rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("shape", [(2000, 3000), (3000, 700), (1, 5000), (4000, 1)])
@pytest.mark.parametrize("wkind", ["f32", "bf16", "homo"])
def test_event_gather_runs_on_the_transposed_plan(shape, wkind):
    # bm.event.csrmv(..., transpose=False) = A @ events: from the second call
    # on the same arrays it scatters the active columns' entries on the plan
    # of A^T (only their entries are read); against the FP64 dense product
    # of the same (bf16-rounded) weights within 1e-5 of each output's sum of
    # |terms|, for bool events, f32 events under the Numba rule (> 0, csrmv)
    # and the Taichi rule (!= 0, csrmv_taichi), and a vmap over event vectors
    m, n = shape
    idx, ptr = fixed_prob_csr(m, n, 0.05 if n > 10 else 1.0, 21)
    w = np.random.default_rng(22).normal(0.5, 0.3, idx.size).astype(np.float32)
    if wkind == "bf16":
        tw = torch.from_numpy(w).to(DEV).to(torch.bfloat16)
        wref = tw.float().cpu().numpy()
    elif wkind == "homo":
        tw, wref = torch.tensor([0.75], device=DEV), np.full(idx.size, 0.75, np.float32)
    else:
        tw, wref = torch.from_numpy(w).to(DEV), w
    ti, tp = torch.from_numpy(idx).to(DEV), torch.from_numpy(ptr).to(DEV)
    D = dense_of(wref, idx, ptr, shape).astype(np.float64)
    rng = np.random.default_rng(23)
    for k in range(3):
        evb = rng.random(n) < 0.2
        evf = np.where(evb, rng.choice([-1.0, 0.5, 2.0], n), 0.0).astype(np.float32)
        for y, on in ((bm.event.csrmv(tw, ti, tp, torch.from_numpy(evb).to(DEV), shape=shape), evb),
                      (bm.event.csrmv(tw, ti, tp, torch.from_numpy(evf).to(DEV), shape=shape), evf > 0),
                      (bm.event.csrmv_taichi(tw, ti, tp, torch.from_numpy(evf).to(DEV), shape=shape)[0], evf != 0)):
            on = on.astype(np.float64)
            assert y.shape == (m,)
            assert np.all(np.abs(y.cpu().numpy() - D @ on) <= 1e-5 * (np.abs(D) @ on) + 1e-30)
    info = bm.event.plan_cache_info()
    assert info["builds"] == 1 and info["plans"] == 1 and info["hits"] >= 7
    evs = torch.from_numpy(rng.random((5, n)) < 0.3).to(DEV)
    ys = torch.func.vmap(lambda e: bm.event.csrmv(tw, ti, tp, e, shape=shape))(evs)
    ref = evs.cpu().numpy().astype(np.float64) @ D.T
    assert np.all(np.abs(ys.cpu().numpy() - ref) <= 1e-5 * (evs.cpu().numpy() @ np.abs(D).T) + 1e-30)
    assert bm.event.plan_cache_info()["builds"] == 1


def test_event_gather_transposed_plan_integer_exact_and_invalidated():
    # integer-valued weights: every partial sum is exact, so the transposed
    # plan's scatter equals the oracle's row gather bit for bit; an in-place
    # write into the weights builds a new plan from the new values
    idx, ptr = fixed_prob_csr(1500, 2500, 0.03, 31)
    w = np.random.default_rng(32).integers(-4, 5, idx.size).astype(np.float32)
    tw, ti, tp = (torch.from_numpy(a).to(DEV) for a in (w, idx, ptr))
    ev = bernoulli(2500, 0.25, 33)
    ref = orc.event_csrmv(w.astype(np.float64), idx, ptr, ev, (1500, 2500), False)
    for _ in range(3):
        y = bm.event.csrmv(tw, ti, tp, torch.from_numpy(ev).to(DEV), shape=(1500, 2500))
        np.testing.assert_array_equal(y.cpu().numpy(), ref)
    assert bm.event.plan_cache_info()["builds"] == 1
    tw.mul_(2.0)
    for _ in range(3):
        y = bm.event.csrmv(tw, ti, tp, torch.from_numpy(ev).to(DEV), shape=(1500, 2500))
        np.testing.assert_array_equal(y.cpu().numpy(), 2 * ref)
    assert bm.event.plan_cache_info()["builds"] == 2


@pytest.mark.parametrize("bad", [3000, -1])
def test_event_gather_index_outside_the_matrix_stays_stateless(bad):
    # a column index outside [0, n): the stateless gather skips that entry;
    # no transposed plan is built (it could not hold the entry), and every
    # call equals the oracle's gather without it
    idx, ptr = fixed_prob_csr(2000, 3000, 0.02, 35)
    w = np.random.default_rng(36).integers(-4, 5, idx.size).astype(np.float32)
    idx = idx.copy()
    idx[17] = bad
    keep = (idx >= 0) & (idx < 3000)
    rows = np.repeat(np.arange(2000), np.diff(ptr))
    D = np.zeros((2000, 3000))
    np.add.at(D, (rows[keep], idx[keep]), w[keep])
    ev = bernoulli(3000, 0.2, 37)
    ref = D @ ev.astype(np.float64)
    tw, ti, tp = (torch.from_numpy(a).to(DEV) for a in (w, idx, ptr))
    for _ in range(3):
        y = bm.event.csrmv(tw, ti, tp, torch.from_numpy(ev).to(DEV), shape=(2000, 3000))
        np.testing.assert_array_equal(y.cpu().numpy(), ref)
    info = bm.event.plan_cache_info()
    assert info["plans"] == 0 and info["ineligible"] == 1


@pytest.mark.parametrize("shape", [(2000, 3000), (3000, 700)])
@pytest.mark.parametrize("homo", [False, True])
def test_coomv_runs_on_the_plan(shape, homo):
    # bm.sparse.coomv: from the second call on the same (data, row, col) the
    # matrix, ordered by row, runs on a tiled plan (both orientations);
    # unsorted entries with duplicates, against the FP64 dense product
    # within 1e-5 of each output's sum of |terms|
    m, n = shape
    rng = np.random.default_rng(41)
    nse = m * n // 40
    row = rng.integers(0, m, nse).astype(np.int32)
    col = rng.integers(0, n, nse).astype(np.int32)
    row[:50], col[:50] = row[50:100], col[50:100]          # duplicated (row, col) pairs
    data = np.float32([0.6]) if homo else rng.normal(0.2, 1.0, nse).astype(np.float32)
    D = np.zeros(shape, np.float64)
    np.add.at(D, (row, col), np.broadcast_to(data.astype(np.float64), row.shape))
    td, tr, tc = (torch.from_numpy(a).to(DEV) for a in (data, row, col))
    for k in range(3):
        for transpose in (False, True):
            x = rng.standard_normal(m if transpose else n).astype(np.float32)
            y = bm.sparse.coomv(td, tr, tc, torch.from_numpy(x).to(DEV), shape=shape, transpose=transpose)
            A = D.T if transpose else D
            assert np.all(np.abs(y.cpu().numpy() - A @ x) <= 1e-5 * (np.abs(A) @ np.abs(x)) + 1e-30)
    info = bm.event.plan_cache_info()
    assert info["builds"] == 1 and info["plans"] == 1 and info["hits"] == 4


@pytest.mark.parametrize("bad", ["row", "col", "negative col"])
@pytest.mark.parametrize("transpose", [False, True])
def test_coomv_keys_outside_the_matrix_stay_stateless(bad, transpose):
    # a row index outside [0, m) or a column index outside [0, n) leaves the
    # matrix to the stateless kernel (no plan: the plan would drop the entry,
    # where the gather clamps its column as JAX does); the results agree
    # with the first (stateless) call bit for bit
    rng = np.random.default_rng(43)
    row = rng.integers(0, 100, 500).astype(np.int32)
    col = rng.integers(0, 80, 500).astype(np.int32)
    if bad == "row":
        row[7] = 100
    elif bad == "col":
        col[7] = 80
    else:
        col[7] = -1
    data = rng.normal(size=500).astype(np.float32)
    td, tr, tc = (torch.from_numpy(a).to(DEV) for a in (data, row, col))
    x = torch.from_numpy(rng.standard_normal(100 if transpose else 80).astype(np.float32)).to(DEV)
    ys = [bm.sparse.coomv(td, tr, tc, x, shape=(100, 80), transpose=transpose).cpu().numpy() for _ in range(3)]
    info = bm.event.plan_cache_info()
    assert info["plans"] == 0 and info["ineligible"] == 1
    np.testing.assert_array_equal(ys[1], ys[0])
    np.testing.assert_array_equal(ys[2], ys[0])
