"""Pins of the event/sparse csrmv and event.info restatements -- CPU only.

Closed forms as in the reference's own tests: event csrmv equals
``events @ dense(A)`` (event/tests/test_event_csrmv.py:66-68), homo equals
heter with ones*w (:59-61), bool equals float 0/1 events (:63-64); sparse
csrmv equals ``dense(A) @ x`` (sparse/tests/test_csrmv.py:49-51);
event.info equals flatnonzero + count (event/tests/test_info.py:34-40).
"""
import numpy as np
import pytest

from oracle import orc
from tests.helpers import bernoulli, dense_of, fixed_prob_csr, random_csr

SHAPES = [(100, 200), (200, 100), (10, 1000), (1000, 10), (1, 1), (257, 129)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("transpose", [True, False])
@pytest.mark.parametrize("homo", [True, False])
@pytest.mark.parametrize("idx_dtype", [np.int32, np.int64])
def test_event_csrmv_closed_form(shape, transpose, homo, idx_dtype):
    m, n = shape
    indices, indptr = random_csr(m, n, 0.3, seed=m * 7 + n, dtype=idx_dtype)
    rng = np.random.default_rng(0)
    data = np.float32(1.5) if homo else rng.standard_normal(indices.size).astype(np.float32)
    ev = bernoulli(m if transpose else n, 0.3, 1)
    D = dense_of(data, indices, indptr, shape)
    ref = ev.astype(np.float64) @ D if transpose else D @ ev.astype(np.float64)
    y = orc.event_csrmv(data, indices, indptr, ev, shape, transpose)
    np.testing.assert_allclose(y, ref, rtol=1e-5, atol=1e-5)
    # bool == float 0/1 events
    y2 = orc.event_csrmv(data, indices, indptr, ev.astype(np.float32), shape, transpose)
    np.testing.assert_array_equal(y, y2)


def test_homo_equals_heter_ones():
    indices, indptr = fixed_prob_csr(300, 200, 0.1, 3)
    ev = bernoulli(300, 0.2, 4)
    a = orc.event_csrmv(np.float32(0.6), indices, indptr, ev, (300, 200), True)
    b = orc.event_csrmv(np.full(indices.size, 0.6, np.float32), indices, indptr, ev, (300, 200), True)
    np.testing.assert_array_equal(a, b)


def test_float_event_rules():
    # Numba: event > 0; Taichi: event != 0
    indices, indptr = fixed_prob_csr(50, 40, 0.2, 5)
    ev = np.zeros(50, np.float32)
    ev[::3] = -1.0
    ev[1::3] = 2.0
    D = dense_of(1.0, indices, indptr, (50, 40))
    pos = orc.event_csrmv(np.float32(1.0), indices, indptr, ev, (50, 40), True, rule=orc.RULE_POSITIVE)
    nz = orc.event_csrmv(np.float32(1.0), indices, indptr, ev, (50, 40), True, rule=orc.RULE_NONZERO)
    np.testing.assert_array_equal(pos, ((ev > 0).astype(np.float64) @ D).astype(np.float32))
    np.testing.assert_array_equal(nz, ((ev != 0).astype(np.float64) @ D).astype(np.float32))


def test_f64_and_bf16():
    indices, indptr = random_csr(80, 90, 0.2, 6)
    w = np.random.default_rng(1).normal(0.5, 0.1, indices.size)
    ev = bernoulli(80, 0.3, 2)
    D = dense_of(w, indices, indptr, (80, 90))
    y = orc.event_csrmv(w, indices, indptr, ev, (80, 90), True)
    assert y.dtype == np.float64
    np.testing.assert_allclose(y, ev @ D, rtol=1e-12)
    from tests.helpers import bf16_bits_to_f32, f32_to_bf16_bits
    bits = f32_to_bf16_bits(w.astype(np.float32))
    yb = orc.event_csrmv(bits, indices, indptr, ev, (80, 90), True, bf16=True)
    Db = dense_of(bf16_bits_to_f32(bits), indices, indptr, (80, 90))
    np.testing.assert_allclose(yb, ev @ Db, rtol=1e-5, atol=1e-5)
    # bf16 storage error against the f32 weights stays within bf16 rounding
    np.testing.assert_allclose(yb, ev @ D, rtol=1e-2, atol=5e-2)


def test_serial_scatter_specialisation_matches_generic():
    indices, indptr = fixed_prob_csr(500, 400, 0.05, 8)
    w = np.random.default_rng(2).normal(0.5, 0.1, indices.size).astype(np.float32)
    ev = bernoulli(500, 0.1, 3)
    a = orc.event_csrmv(w, indices, indptr, ev, (500, 400), True)
    b = orc.event_csrmv_scatter_f32(w, indices, indptr, ev.view(np.uint8), (500, 400))
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("transpose", [True, False])
@pytest.mark.parametrize("homo", [True, False])
def test_sparse_csrmv_closed_form(shape, transpose, homo):
    m, n = shape
    indices, indptr = random_csr(m, n, 0.3, seed=m + 3 * n)
    rng = np.random.default_rng(5)
    data = np.float32(0.7) if homo else rng.standard_normal(indices.size).astype(np.float32)
    x = rng.standard_normal(m if transpose else n).astype(np.float32)
    D = dense_of(data, indices, indptr, shape)
    ref = x.astype(np.float64) @ D if transpose else D @ x.astype(np.float64)
    y = orc.csrmv(data, indices, indptr, x, shape, transpose)
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 4096, 4097, 100000])
@pytest.mark.parametrize("dtype", [np.bool_, np.float32, np.float64])
def test_event_info_closed_form(n, dtype):
    ev = bernoulli(n, 0.1, n + 1).astype(dtype)
    ids, num = orc.event_info(ev)
    nz = np.flatnonzero(ev)
    assert num[0] == nz.size
    np.testing.assert_array_equal(ids[:nz.size], nz)
    assert (ids[nz.size:] == -1).all()


def test_event_info_batched():
    ev = np.stack([bernoulli(1000, r, 9) for r in (0.0, 0.01, 0.5, 1.0)])
    ids, num = orc.event_info(ev)
    for b in range(4):
        nz = np.flatnonzero(ev[b])
        assert num[b] == nz.size
        np.testing.assert_array_equal(ids[b, :nz.size], nz)
