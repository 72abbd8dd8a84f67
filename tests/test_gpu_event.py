"""GPU parity: bm.event.csrmv / csrmv_taichi / info and bm.sparse.csrmv
against the oracle and the FP64 closed form, through the C ABI."""
import numpy as np
import pytest
import torch

import brainpy_amd.math as bm
from oracle import orc
from tests.helpers import (bernoulli, bf16_bits_to_f32, dense_of, f32_to_bf16_bits, fixed_prob_csr,
                           random_csr)

pytestmark = pytest.mark.gpu

SHAPES = [(100, 200), (200, 100), (10, 1000), (1000, 10), (1, 1), (257, 129), (3000, 50)]
RTOL = 1e-5   # per-call FP32 current vs FP64, relative (the stated per-step tolerance)


def _tol_check(y, ref, scale):
    err = np.abs(y.astype(np.float64) - ref)
    assert (err <= RTOL * np.maximum(np.abs(ref), scale) + 1e-6).all(), float(err.max())


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("transpose", [True, False])
@pytest.mark.parametrize("homo", [True, False])
@pytest.mark.parametrize("idx_dtype", [np.int32, np.int64])
def test_event_csrmv_vs_fp64(shape, transpose, homo, idx_dtype):
    m, n = shape
    indices, indptr = random_csr(m, n, 0.3, seed=m * 7 + n, dtype=idx_dtype)
    rng = np.random.default_rng(0)
    data = np.float32(1.5) if homo else rng.normal(0.5, 0.1, indices.size).astype(np.float32)
    ev = bernoulli(m if transpose else n, 0.3, 1)
    D = dense_of(data, indices, indptr, shape)
    ref = ev.astype(np.float64) @ D if transpose else D @ ev.astype(np.float64)
    y = bm.event.csrmv(data, indices, indptr, ev, shape=shape, transpose=transpose).cpu().numpy()
    assert y.dtype == np.float32 and y.shape == ref.shape
    scale = np.abs(D).sum(axis=0 if transpose else 1).max() if D.size else 1.0
    _tol_check(y, ref, scale)
    # bool == float 0/1
    y2 = bm.event.csrmv(data, indices, indptr, ev.astype(np.float32), shape=shape, transpose=transpose)
    _tol_check(y2.cpu().numpy(), ref, scale)


@pytest.mark.parametrize("transpose", [True, False])
def test_event_csrmv_homo_integer_exact(transpose):
    # homo weight 1.0 and bool events: all partial sums are exact integers,
    # so the GPU result is bit-identical to the oracle whatever the order
    indices, indptr = fixed_prob_csr(2000, 1500, 0.05, 3)
    ev = bernoulli(2000 if transpose else 1500, 0.2, 4)
    y = bm.event.csrmv(np.float32(1.0), indices, indptr, ev, shape=(2000, 1500), transpose=transpose)
    ref = orc.event_csrmv(np.float32(1.0), indices, indptr, ev, (2000, 1500), transpose)
    np.testing.assert_array_equal(y.cpu().numpy(), ref)


def test_event_rules_and_taichi_twin():
    indices, indptr = fixed_prob_csr(300, 200, 0.1, 5)
    ev = np.zeros(300, np.float32)
    ev[::3] = -1.0
    ev[1::3] = 2.0
    a = bm.event.csrmv(np.float32(1.0), indices, indptr, ev, shape=(300, 200), transpose=True).cpu().numpy()
    b = bm.event.csrmv_taichi(np.float32(1.0), indices, indptr, ev, shape=(300, 200), transpose=True)
    assert isinstance(b, tuple) and len(b) == 1
    np.testing.assert_array_equal(a, orc.event_csrmv(np.float32(1.0), indices, indptr, ev, (300, 200), True,
                                                     rule=orc.RULE_POSITIVE))
    np.testing.assert_array_equal(b[0].cpu().numpy(),
                                  orc.event_csrmv(np.float32(1.0), indices, indptr, ev, (300, 200), True,
                                                  rule=orc.RULE_NONZERO))


@pytest.mark.parametrize("n", [200, 4099, 70000])
@pytest.mark.parametrize("evdt", [np.float32, np.float64])
def test_event_gather_float_events_rules(n, evdt):
    # transpose=False with float events: the gather tests each event with
    # the operator's rule (csrmv: > 0, the Numba kernels; csrmv_taichi: != 0)
    # and never multiplies by it, so signed values, -0.0 and NaN select rows
    # exactly as the oracle does (homo weight 1: exact integer sums); n not a
    # multiple of 32 or 64 leaves partial bitmask words
    m = 300
    indices, indptr = fixed_prob_csr(m, n, min(0.5, 40.0 / n), 6)
    rng = np.random.default_rng(n)
    r = rng.random(n)
    ev = np.zeros(n, evdt)
    ev[r < 0.10] = -1.5
    ev[(r >= 0.10) & (r < 0.25)] = 2.0
    ev[(r >= 0.25) & (r < 0.26)] = -0.0
    ev[r > 0.995] = np.nan
    one = np.float32(1.0)
    a = bm.event.csrmv(one, indices, indptr, ev, shape=(m, n), transpose=False).cpu().numpy()
    np.testing.assert_array_equal(a, orc.event_csrmv(one, indices, indptr, ev, (m, n), False,
                                                     rule=orc.RULE_POSITIVE))
    b = bm.event.csrmv_taichi(one, indices, indptr, ev, shape=(m, n), transpose=False)[0].cpu().numpy()
    np.testing.assert_array_equal(b, orc.event_csrmv(one, indices, indptr, ev, (m, n), False,
                                                     rule=orc.RULE_NONZERO))
    w = rng.normal(0.5, 0.1, indices.size).astype(np.float32)
    y = bm.event.csrmv(w, indices, indptr, ev, shape=(m, n), transpose=False).cpu().numpy()
    D = dense_of(w, indices, indptr, (m, n))
    ref = D @ np.where(ev > 0, 1.0, 0.0)
    _tol_check(y, ref, np.abs(D).sum(axis=1).max())


def test_f64_weights():
    indices, indptr = random_csr(120, 90, 0.2, 6)
    w = np.random.default_rng(1).normal(0.5, 0.1, indices.size)
    ev = bernoulli(120, 0.3, 2)
    y = bm.event.csrmv(w, indices, indptr, ev, shape=(120, 90), transpose=True)
    assert y.dtype == torch.float64
    np.testing.assert_allclose(y.cpu().numpy(), ev @ dense_of(w, indices, indptr, (120, 90)), rtol=1e-12)


@pytest.mark.parametrize("transpose", [True, False])
def test_bf16_weights(transpose):
    m, n = 500, 400
    indices, indptr = fixed_prob_csr(m, n, 0.05, 7)
    w = np.random.default_rng(2).normal(0.5, 0.1, indices.size).astype(np.float32)
    bits = f32_to_bf16_bits(w)
    wb = torch.from_numpy(bits.view(np.int16)).view(torch.bfloat16)
    ev = bernoulli(m if transpose else n, 0.2, 3)
    y = bm.event.csrmv(wb, indices, indptr, ev, shape=(m, n), transpose=transpose).cpu().numpy()
    Db = dense_of(bf16_bits_to_f32(bits), indices, indptr, (m, n))
    ref = ev @ Db if transpose else Db @ ev
    np.testing.assert_allclose(y, ref, rtol=1e-5, atol=1e-5)
    # torch's own bf16 -> f32 conversion agrees with the bit pattern
    np.testing.assert_array_equal(wb.float().numpy(), bf16_bits_to_f32(bits))


def test_edge_cases():
    # empty indices, all-false / all-true events, empty rows, m not a multiple of 64
    shape = (130, 70)
    indices, indptr = random_csr(*shape, 0.05, 8)
    for ev in (np.zeros(130, bool), np.ones(130, bool)):
        y = bm.event.csrmv(np.float32(1.0), indices, indptr, ev, shape=shape, transpose=True).cpu().numpy()
        np.testing.assert_array_equal(y, orc.event_csrmv(np.float32(1.0), indices, indptr, ev, shape, True))
    z = bm.event.csrmv_taichi(np.float32(1.0), np.zeros(0, np.int32), np.zeros(131, np.int32),
                              np.ones(130, bool), shape=shape, transpose=True)
    assert isinstance(z, torch.Tensor) and z.shape == (70,) and float(z.abs().sum()) == 0.0
    y = bm.event.csrmv(np.float32(1.0), np.zeros(0, np.int32), np.zeros(131, np.int32), np.ones(130, bool),
                       shape=shape, transpose=True)
    assert float(y.abs().sum()) == 0.0


def test_validation_errors():
    indices, indptr = fixed_prob_csr(10, 20, 0.2, 1)
    with pytest.raises(ValueError, match="Shape mismatch"):
        bm.event.csrmv(1.0, indices, indptr, np.ones(11, bool), shape=(10, 20), transpose=True)
    with pytest.raises(ValueError, match="size of data"):
        bm.event.csrmv(np.ones(3, np.float32), indices, indptr, np.ones(10, bool), shape=(10, 20), transpose=True)
    with pytest.raises(ValueError, match="int32 or int64"):
        bm.event.csrmv(1.0, indices.astype(np.int16), indptr, np.ones(10, bool), shape=(10, 20), transpose=True)


def test_repeat_calls_reuse_output_semantics():
    # outputs are fully rewritten: two calls give the same values
    indices, indptr = fixed_prob_csr(1000, 800, 0.05, 9)
    ev = bernoulli(1000, 0.1, 5)
    a = bm.event.csrmv(np.float32(1.0), indices, indptr, ev, shape=(1000, 800), transpose=True)
    b = bm.event.csrmv(np.float32(1.0), indices, indptr, ev, shape=(1000, 800), transpose=True)
    assert torch.equal(a, b)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4095, 4096, 4097, 100000, 1000000])
@pytest.mark.parametrize("dtype", [np.bool_, np.float32])
def test_event_info_exact(n, dtype):
    ev = bernoulli(n, 0.02, n + 1).astype(dtype)
    ids, num = bm.event.info(ev)
    rid, rnum = orc.event_info(ev)
    np.testing.assert_array_equal(ids.cpu().numpy(), rid)
    np.testing.assert_array_equal(num.cpu().numpy(), rnum)


def test_event_info_batched_exact():
    ev = np.stack([bernoulli(100000, r, 9) for r in (0.0, 0.001, 0.1, 1.0)])
    ids, num = bm.event.info(ev)
    rid, rnum = orc.event_info(ev)
    np.testing.assert_array_equal(ids.cpu().numpy(), rid)
    np.testing.assert_array_equal(num.cpu().numpy(), rnum)


@pytest.mark.parametrize("n", [16, 4112, 65584, 1000000])
def test_event_info_bool_16_aligned_rows(n):
    # rows of a multiple of 16 bool events take the 16-events-per-load path;
    # a view starting one byte in takes the generic path: both exact
    ev = np.stack([bernoulli(n, r, 31 + k) for k, r in enumerate((0.5, 0.02, 1.0, 0.0))])
    ids, num = bm.event.info(ev)
    rid, rnum = orc.event_info(ev)
    np.testing.assert_array_equal(ids.cpu().numpy(), rid)
    np.testing.assert_array_equal(num.cpu().numpy(), rnum)
    flat = torch.tensor(np.concatenate([[True], ev[1]]), device="cuda")
    ids1, num1 = bm.event.info(flat[1:])
    rid1, rnum1 = orc.event_info(ev[1])
    np.testing.assert_array_equal(ids1.cpu().numpy(), rid1)
    np.testing.assert_array_equal(num1.cpu().numpy(), rnum1)
    # the minimum workspace (tile counts and offsets only) re-reads the events
    # instead of the stored event bits: same result
    import ctypes
    from brainpy_amd import _lib
    L = _lib.lib()
    evt = torch.from_numpy(ev).cuda()
    B = ev.shape[0]
    full = int(L.bpx_event_info_workspace(B, n))
    ntiles = (n + 4095) // 4096
    for wsb in (full, 8 * B * ntiles):
        ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
        ids2 = torch.full((B, n), 5, dtype=torch.int32, device="cuda")
        num2 = torch.full((B,), 5, dtype=torch.int32, device="cuda")
        rc = L.bpx_event_info(_lib.stream_ptr(), _lib.ptr(evt), _lib.BPX_BOOL, B, n, _lib.ptr(ids2), _lib.ptr(num2),
                              _lib.ptr(ws), wsb)
        assert rc == 0
        np.testing.assert_array_equal(ids2.cpu().numpy(), rid)
        np.testing.assert_array_equal(num2.cpu().numpy(), rnum)
    ws = torch.empty(8 * B * ntiles - 1, dtype=torch.uint8, device="cuda")
    assert L.bpx_event_info(_lib.stream_ptr(), _lib.ptr(evt), _lib.BPX_BOOL, B, n, _lib.ptr(ids2), _lib.ptr(num2),
                            _lib.ptr(ws), ws.numel()) == -3          # BPX_ERR_WORKSPACE


def test_fixed_order_bit_exact_vs_serial_scatter():
    m, n = 3000, 2500
    indices, indptr = fixed_prob_csr(m, n, 0.02, 11, sort=False)
    w = np.random.default_rng(4).normal(0.5, 0.1, indices.size).astype(np.float32)
    ev = bernoulli(m, 0.1, 6)
    rows, colptr, cd = bm.event.csr_to_csc(indices, indptr, (m, n), data=w)
    y = bm.event.csrmv_fixed_order(cd, rows, colptr, ev, shape=(m, n)).cpu().numpy()
    np.testing.assert_array_equal(y, orc.event_csrmv(w, indices, indptr, ev, (m, n), True))
    yh = bm.event.csrmv_fixed_order(np.float32(0.6), rows, colptr, ev, shape=(m, n)).cpu().numpy()
    np.testing.assert_array_equal(yh, orc.event_csrmv(np.float32(0.6), indices, indptr, ev, (m, n), True))


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("transpose", [True, False])
@pytest.mark.parametrize("homo", [True, False])
def test_sparse_csrmv_vs_fp64(shape, transpose, homo):
    m, n = shape
    indices, indptr = random_csr(m, n, 0.3, seed=m + 3 * n)
    rng = np.random.default_rng(5)
    data = np.float32(0.7) if homo else rng.standard_normal(indices.size).astype(np.float32)
    x = rng.standard_normal(m if transpose else n).astype(np.float32)
    D = dense_of(data, indices, indptr, shape)
    ref = x.astype(np.float64) @ D if transpose else D @ x.astype(np.float64)
    y = bm.sparse.csrmv(data, indices, indptr, x, shape=shape, transpose=transpose).cpu().numpy()
    scale = (np.abs(D) * np.abs(x.astype(np.float64))[:, None] if transpose else
             np.abs(D) * np.abs(x.astype(np.float64))[None, :]).sum(axis=0 if transpose else 1)
    err = np.abs(y - ref)
    assert (err <= 1e-5 * np.maximum(scale, 1.0)).all(), float(err.max())
    (yt,) = bm.sparse.csrmv_taichi(data, indices, indptr, x, shape=shape, transpose=transpose)
    assert (np.abs(yt.cpu().numpy() - ref) <= 1e-5 * np.maximum(scale, 1.0)).all()


def test_sparse_converters():
    indices, indptr = random_csr(40, 30, 0.2, 12)
    w = np.random.default_rng(6).standard_normal(indices.size).astype(np.float32)
    D = bm.sparse.csr_to_dense(w, indices, indptr, shape=(40, 30)).cpu().numpy()
    np.testing.assert_allclose(D, dense_of(w, indices, indptr, (40, 30)), rtol=1e-7)
    r, c = bm.sparse.csr_to_coo(indices, indptr)
    idx2, ptr2 = bm.sparse.coo_to_csr(r, c, num_row=40)
    np.testing.assert_array_equal(idx2.cpu().numpy(), indices)
    np.testing.assert_array_equal(ptr2.cpu().numpy(), indptr)


@pytest.mark.parametrize("shape", [(100, 200), (1000, 3000), (3000, 50), (257, 129), (5000, 100000),
                                   (2000, 200000), (300, 262144)])
@pytest.mark.parametrize("wkind", ["homo", "f32", "bf16"])
@pytest.mark.parametrize("sort", [True, False])
def test_tiled_plan_vs_oracle(shape, wkind, sort):
    m, n = shape
    p = min(0.5, 50.0 / n)
    indices, indptr = bm_connect_csr(m, n, p, 13, sort)
    rng = np.random.default_rng(7)
    if wkind == "homo":
        data = np.float32(1.0)
    else:
        data = rng.normal(0.5, 0.1, indices.size).astype(np.float32)
    plan = bm.event.TiledCSR(data, indices, indptr, shape=shape,
                             weight_dtype=torch.bfloat16 if wkind == "bf16" else None)
    for rate in (0.0, 0.05, 0.5, 1.0):
        ev = bernoulli(m, rate, int(rate * 100) + m)
        y = plan.csrmv(ev).cpu().numpy()
        if wkind == "homo":
            np.testing.assert_array_equal(y, orc.event_csrmv(data, indices, indptr, ev, shape, True))
        else:
            w = bf16_bits_to_f32(f32_to_bf16_bits(data)) if wkind == "bf16" else data
            D = dense_of(w, indices, indptr, shape)
            ref = ev.astype(np.float64) @ D
            scale = np.abs(D).sum(axis=0).max() if D.size else 1.0
            _tol_check(y, ref, scale)


@pytest.mark.parametrize("shape", [(100, 200), (3000, 50), (5000, 100000)])
@pytest.mark.parametrize("wkind", ["homo", "f32", "bf16"])
@pytest.mark.parametrize("xscale", [1.0, 3e5, 1e-4, 1e-9, 1e-20])
def test_tiled_plan_dense_matvec(shape, wkind, xscale):
    # x @ A for a dense vector with zeros, on the plan, against the FP64
    # dense product within 1e-5 of the sum of term magnitudes (the products
    # w * x are rounded to f32 once, as the plain sparse.csrmv does)
    m, n = shape
    p = min(0.5, 50.0 / n)
    indices, indptr = bm_connect_csr(m, n, p, 23, False)
    rng = np.random.default_rng(8)
    data = np.float32(0.7) if wkind == "homo" else rng.normal(0.5, 0.3, indices.size).astype(np.float32)
    plan = bm.event.TiledCSR(data, indices, indptr, shape=shape,
                             weight_dtype=torch.bfloat16 if wkind == "bf16" else None)
    x = (rng.standard_normal(m) * xscale).astype(np.float32)
    x[rng.random(m) < 0.3] = 0.0
    y = plan.matvec(x).cpu().numpy().astype(np.float64)
    w = bf16_bits_to_f32(f32_to_bf16_bits(data)) if wkind == "bf16" else data
    D = dense_of(w, indices, indptr, shape)
    ref = x.astype(np.float64) @ D
    mag = np.abs(x.astype(np.float64)) @ np.abs(D)
    assert np.all(np.abs(y - ref) <= 1e-5 * mag + 1e-30)
    if wkind == "f32":
        y2 = bm.sparse.csrmv(data, indices, indptr, x, shape=shape, transpose=True).cpu().numpy()
        assert np.all(np.abs(y2 - y) <= 2e-5 * mag + 1e-30)


def test_tiled_plan_output_limit():
    # more than 32 tiles per copy once a tile would not fit the LDS budget:
    # 64 tiles of 4096 outputs hold n = 262,144 whatever the row lengths;
    # one more output takes 72 tiles, which need 18 synapses per row on
    # average, so this two-synapse row is refused
    from brainpy_amd._lib import BPXError
    idx = np.array([0, 262143], np.int32)
    ptr = np.array([0, 2], np.int32)
    plan = bm.event.TiledCSR(np.float32(1.0), idx, ptr, shape=(1, 262144))
    assert plan.plan.tiles == 64 and plan.plan.tile_len == 4096
    y = plan.csrmv(np.ones(1, bool)).cpu().numpy()
    assert y[0] == 1.0 and y[262143] == 1.0 and y.sum() == 2.0
    with pytest.raises(BPXError):
        bm.event.TiledCSR(np.float32(1.0), idx, ptr, shape=(1, 262145))


def test_tiled_plan_dense_matvec_accumulate_and_nonfinite():
    m, n = 2000, 3000
    indices, indptr = bm_connect_csr(m, n, 0.01, 29, True)
    w = np.random.default_rng(9).normal(0.5, 0.1, indices.size).astype(np.float32)
    plan = bm.event.TiledCSR(w, indices, indptr, shape=(m, n))
    x = np.random.default_rng(10).standard_normal(m).astype(np.float32)
    base = torch.full((n,), 2.0, dtype=torch.float32, device="cuda")
    plan.matvec(x, out=base)
    D = dense_of(w, indices, indptr, (m, n))
    np.testing.assert_allclose(base.cpu().numpy(), 2.0 + x.astype(np.float64) @ D, rtol=1e-5, atol=1e-5)
    # non-finite products reach their outputs as IEEE addition does: row 7
    # at +inf gives +inf (or -inf through negative weights) where it has
    # synapses, rows 7 (+inf) and 8 (-inf) sharing a column give NaN, a NaN
    # row gives NaN on its columns; every other output stays finite and
    # equal to the product without those rows
    D64 = D.astype(np.float64)

    def ieee_ref(xv):
        with np.errstate(invalid="ignore", over="ignore"):
            prods = xv.astype(np.float32)[:, None] * D.astype(np.float32)      # the f32 products
            prods[D == 0] = 0.0
            return prods.astype(np.float64).sum(axis=0)

    for special in ([(7, np.inf)], [(7, -np.inf)], [(7, np.nan)], [(7, np.inf), (8, -np.inf)],
                    [(7, np.float32(3e38))]):
        xs = x.copy()
        for r, v in special:
            xs[r] = v
        y = plan.matvec(xs).cpu().numpy().astype(np.float64)
        ref = ieee_ref(xs)
        fin = np.isfinite(ref)
        assert (~fin).any() == any(not np.isfinite(v) for _, v in special)
        np.testing.assert_array_equal(np.isnan(y), np.isnan(ref))
        np.testing.assert_array_equal(np.isposinf(y), np.isposinf(ref))
        np.testing.assert_array_equal(np.isneginf(y), np.isneginf(ref))
        xf = np.where(np.isfinite(xs), xs, 0).astype(np.float64)
        mag = np.abs(xf) @ np.abs(D64)
        assert np.all(np.abs(y[fin] - ref[fin]) <= 1e-5 * mag[fin] + 1e-30)
    assert np.isfinite(plan.matvec(np.zeros(m, np.float32)).cpu().numpy()).all()


def test_tiled_plan_accumulate_and_float_events():
    m, n = 2000, 7000
    indices, indptr = bm_connect_csr(m, n, 0.01, 21, False)
    plan = bm.event.TiledCSR(np.float32(0.5), indices, indptr, shape=(m, n))
    ev = np.zeros(m, np.float32)
    ev[::4] = 1.0
    ev[1::4] = -2.0
    base = torch.full((n,), 3.0, dtype=torch.float32, device="cuda")
    plan.csrmv(ev, out=base)
    ref = np.float32(3.0) + orc.event_csrmv(np.float32(0.5), indices, indptr, ev, (m, n), True)
    np.testing.assert_array_equal(base.cpu().numpy(), ref)


def test_event_csrmv_fixture():
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "event_csrmv_200x100.npz"))
    idx, ptr, w = d["indices"], d["indptr"], d["data"]
    y_t = bm.event.csrmv(w, idx, ptr, d["events_transpose"], shape=(200, 100), transpose=True).cpu().numpy()
    y_g = bm.event.csrmv(w, idx, ptr, d["events_gather"], shape=(200, 100)).cpu().numpy()
    np.testing.assert_allclose(y_t, d["y_transpose"], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(y_g, d["y_gather"], rtol=1e-5, atol=1e-5)
    # the fixed-order mode sums in the serial loop's order: bit-exact
    rows, colptr, w_csc = bm.event.csr_to_csc(idx, ptr, (200, 100), data=w)
    y_fo = bm.event.csrmv_fixed_order(w_csc, rows, colptr, d["events_transpose"], shape=(200, 100)).cpu().numpy()
    np.testing.assert_array_equal(y_fo, d["y_transpose"])


def test_full_size_headline_matrix():
    # the 100k x 100k, p = 0.01 matrix (1e8 synapses, int32) with 10 % bool
    # spikes: with weight 1 every output is an integer count, so the tiled
    # plan, the stateless scatter and gather, and the oracle agree exactly;
    # heterogeneous f32 weights are checked against the oracle's FP64
    # accumulation within 1e-5 of each output's sum of |w|
    from brainpy_amd.connect import fixed_prob_csr_device
    m = n = 100_000
    dev = torch.device("cuda")
    idx, ptr = fixed_prob_csr_device(m, n, 0.01, seed=2024, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(17)
    ev = torch.rand(m, generator=g, device=dev) < 0.1
    idx_h, ptr_h, ev_h = idx.cpu().numpy(), ptr.cpu().numpy(), ev.cpu().numpy()
    one = np.float32(1.0)
    ref = orc.event_csrmv(one, idx_h, ptr_h, ev_h, (m, n), True)
    assert float(ref.sum()) == float(ev_h.sum()) * 1000
    np.testing.assert_array_equal(bm.event.csrmv(one, idx, ptr, ev, shape=(m, n), transpose=True).cpu().numpy(),
                                  ref)
    np.testing.assert_array_equal(bm.event.TiledCSR(one, idx, ptr, shape=(m, n)).csrmv(ev).cpu().numpy(), ref)
    ev_n = torch.rand(n, generator=g, device=dev) < 0.1
    np.testing.assert_array_equal(bm.event.csrmv(one, idx, ptr, ev_n, shape=(m, n)).cpu().numpy(),
                                  orc.event_csrmv(one, idx_h, ptr_h, ev_n.cpu().numpy(), (m, n), False))
    w = torch.empty(idx.numel(), dtype=torch.float32, device=dev).normal_(0.5, 0.1, generator=g)
    w_h = w.cpu().numpy().astype(np.float64)
    ref64 = orc.event_csrmv(w_h, idx_h, ptr_h, ev_h, (m, n), True)
    mag = orc.event_csrmv(np.abs(w_h), idx_h, ptr_h, ev_h, (m, n), True)
    for y in (bm.event.csrmv(w, idx, ptr, ev, shape=(m, n), transpose=True),
              bm.event.TiledCSR(w, idx, ptr, shape=(m, n)).csrmv(ev)):
        err = np.abs(y.cpu().numpy().astype(np.float64) - ref64)
        assert np.all(err <= 1e-5 * mag + 1e-30), float(np.max(err / np.maximum(mag, 1e-30)))


def test_full_size_bf16_weights():
    # the BF16-stored variant of the 100k x 100k, p = 0.01 operator (1e8
    # synapses): normal(0.5, 0.1) weights rounded to BF16, 10 % spikes,
    # accumulated in FP32 by the public transpose path and the tiled plan;
    # within 1e-5 of each output's sum of |w| against the FP64 sum of the
    # same BF16 values
    from brainpy_amd.connect import fixed_prob_csr_device
    m = n = 100_000
    dev = torch.device("cuda")
    idx, ptr = fixed_prob_csr_device(m, n, 0.01, seed=77, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    ev = torch.rand(m, generator=g, device=dev) < 0.1
    wb = torch.empty(idx.numel(), dtype=torch.float32, device=dev).normal_(0.5, 0.1, generator=g).to(torch.bfloat16)
    idx_h, ptr_h, ev_h = idx.cpu().numpy(), ptr.cpu().numpy(), ev.cpu().numpy()
    w_h = wb.float().cpu().numpy().astype(np.float64)
    ref64 = orc.event_csrmv(w_h, idx_h, ptr_h, ev_h, (m, n), True)
    mag = orc.event_csrmv(np.abs(w_h), idx_h, ptr_h, ev_h, (m, n), True)
    for y in (bm.event.csrmv(wb, idx, ptr, ev, shape=(m, n), transpose=True),
              bm.event.TiledCSR(wb, idx, ptr, shape=(m, n)).csrmv(ev)):
        assert y.dtype == torch.float32
        err = np.abs(y.cpu().numpy().astype(np.float64) - ref64)
        assert np.all(err <= 1e-5 * mag + 1e-30), float(np.max(err / np.maximum(mag, 1e-30)))


def bm_connect_csr(m, n, p, seed, sort):
    from brainpy_amd.connect import fixed_prob_csr as fp
    return fp(m, n, p, seed, sort=sort)


def test_tiled_plan_wide_value_ranges_stay_accurate():
    # weights or vector values spanning far more than the fixed-point grid
    # resolves: terms below the grid's floor go to their outputs with FP32
    # atomics, so outputs fed only by tiny terms keep 1e-5 of their own sum
    # of |terms| (on the grid alone they would round to a few units)
    m, n = 2000, 3000
    indices, indptr = bm_connect_csr(m, n, 0.01, 31, True)
    rng = np.random.default_rng(12)
    w = rng.normal(0.5, 0.1, indices.size).astype(np.float32)
    w[indices >= n // 2] *= np.float32(1e-12)          # the right half of the outputs sees only tiny weights
    plan = bm.event.TiledCSR(w, indices, indptr, shape=(m, n))
    D = dense_of(w, indices, indptr, (m, n)).astype(np.float64)
    ev = bernoulli(m, 0.2, 13)
    y = plan.csrmv(ev).cpu().numpy().astype(np.float64)
    ref = ev.astype(np.float64) @ D
    mag = ev.astype(np.float64) @ np.abs(D)
    assert np.all(np.abs(y - ref) <= 1e-5 * mag + 1e-45)
    assert np.abs(y[n // 2:]).max() > 0
    # a dense vector with a few huge entries: outputs away from their rows
    # are sums of O(1) terms only
    w2 = rng.normal(0.5, 0.1, indices.size).astype(np.float32)
    plan2 = bm.event.TiledCSR(w2, indices, indptr, shape=(m, n))
    D2 = dense_of(w2, indices, indptr, (m, n)).astype(np.float64)
    x = rng.standard_normal(m).astype(np.float32)
    x[:10] = np.float32(1e20)
    y2 = plan2.matvec(x).cpu().numpy().astype(np.float64)
    ref2 = x.astype(np.float64) @ D2
    mag2 = np.abs(x.astype(np.float64)) @ np.abs(D2)
    assert np.all(np.abs(y2 - ref2) <= 1e-5 * mag2 + 1e-30)
    small = mag2 < 1e6
    assert small.sum() > n // 2 and np.abs(y2[small]).max() > 0
    # huge and tiny entries only, f32 and homogeneous weights: the outputs
    # that no huge row reaches are sums of products far below the floor of a
    # grid scaled for 1e20, and keep their own 1e-5
    x3 = np.zeros(m, np.float32)
    x3[:10] = np.float32(1e20)
    x3[10:20] = np.float32(-1e-30)
    for plan3, D3 in ((plan2, D2),
                      (bm.event.TiledCSR(np.float32(0.7), indices, indptr, shape=(m, n)),
                       dense_of(np.float32(0.7), indices, indptr, (m, n)).astype(np.float64))):
        y3 = plan3.matvec(x3).cpu().numpy().astype(np.float64)
        ref3 = x3.astype(np.float64) @ D3
        mag3 = np.abs(x3.astype(np.float64)) @ np.abs(D3)
        assert np.all(np.abs(y3 - ref3) <= 1e-5 * mag3 + 1e-45)
        only_tiny = (mag3 > 0) & (mag3 < 1.0)
        assert only_tiny.sum() > 50 and np.all(y3[only_tiny] < 0)
