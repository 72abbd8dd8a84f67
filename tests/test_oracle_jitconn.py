"""Pins of the just-in-time connectivity restatement (oracle) -- CPU only.

The connectivity is defined in-tree only by the Taichi kernels
(jitconn/_event_matvec_taichi.py:38-1106, jitconn/_matvec_taichi.py:30-746).
Pinned here by: an independent numpy restatement (brute force on tiny
shapes), the reference's own self-consistency tests
(jitconn/tests/test_event_matvec_taichi.py: same seed -> same result
:53-73, (shape, T) == (shape[::-1], not T) :75-83, linearity in the weight
:152-183), and invariants the loops fix (strictly increasing indices per
stream, mean in-degree, the lane32 skipped index).
"""
import numpy as np
import pytest

from oracle import np_ref, orc

SHAPES = [(100, 200), (2, 1000), (1000, 2), (64, 64), (33, 70)]
PROBS = [0.01, 0.1, 0.5]


def _enum_sorted(dist, od, layout, shape, p, seed, w0=0.3, w1=0.9):
    r, c, w = orc.jitconn_enumerate(dist, od, layout, shape[0], shape[1], w0, w1, np_ref.conn_len(p), seed)
    o = np.lexsort((c, r))
    return r[o], c[o], w[o]


@pytest.mark.parametrize("layout", [orc.LAYOUT_LANE32, orc.LAYOUT_ROW])
@pytest.mark.parametrize("od", [True, False])
@pytest.mark.parametrize("dist", ["homo", "uniform", "normal"])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("p", PROBS)
def test_enumeration_matches_numpy(layout, od, dist, shape, p):
    r1, c1, w1 = _enum_sorted(dist, od, layout, shape, p, 1234)
    f = np_ref.enumerate_lane32 if layout == orc.LAYOUT_LANE32 else np_ref.enumerate_row
    r2, c2, w2 = f(dist, od, shape[0], shape[1], 0.3, 0.9, np_ref.conn_len(p), 1234)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(c1, c2)
    if dist == "normal":
        np.testing.assert_allclose(w1, w2, rtol=1e-5, atol=1e-6)
    else:
        np.testing.assert_array_equal(w1, w2)


def test_conn_len():
    assert np_ref.conn_len(1e-4) == 19999
    assert np_ref.conn_len(0.02) == 99
    assert np_ref.conn_len(0.1) == 19
    assert np_ref.conn_len(0.5) == 3
    assert np_ref.conn_len(1.0) == 1
    # 1/p rounded to float32 before ceil: 1/(1/3) = 3.0000000000000004 -> 3
    assert np_ref.conn_len(1 / 3) == 5


@pytest.mark.parametrize("layout", [orc.LAYOUT_LANE32, orc.LAYOUT_ROW])
@pytest.mark.parametrize("od", [True, False])
@pytest.mark.parametrize("vmode", [0, 1, 2])
@pytest.mark.parametrize("dist", ["homo", "uniform", "normal"])
def test_mv_equals_dense_closed_form(layout, od, vmode, dist):
    shape, p, seed = (97, 150), 0.1, 42
    clen = np_ref.conn_len(p)
    rng = np.random.default_rng(0)
    if vmode == 0:
        vec = rng.standard_normal(shape[1]).astype(np.float32)
    elif vmode == 1:
        vec = rng.random(shape[1]) < 0.3
    else:
        vec = (rng.random(shape[1]) < 0.3).astype(np.float32) * rng.choice([1.0, 2.0, -1.0], shape[1]).astype(np.float32)
    M = np_ref.dense(dist, od, "lane32" if layout == orc.LAYOUT_LANE32 else "row", shape[0], shape[1], 0.3, 0.9,
                     clen, seed)
    y = orc.jitconn_mv(vec, dist, vmode, od, layout, shape[0], shape[1], 0.3, 0.9, clen, seed)
    v64 = vec.astype(np.float64)
    if vmode == 2 and not (layout == orc.LAYOUT_LANE32 and od):
        v64 = (vec != 0).astype(np.float64)          # Taichi `!= 0.` test, value not used (Q4)
    ref = M @ v64
    np.testing.assert_allclose(y, ref, rtol=1e-5, atol=1e-5)


def test_homo_bool_exact_integers():
    # homo weight 1 and bool events: every partial sum is an exact integer
    shape, p = (300, 500), 0.05
    clen = np_ref.conn_len(p)
    ev = np.random.default_rng(1).random(500) < 0.2
    for layout in (orc.LAYOUT_LANE32, orc.LAYOUT_ROW):
        for od in (True, False):
            M = np_ref.dense("homo", od, "lane32" if layout == 0 else "row", shape[0], shape[1], 1.0, 0.0, clen, 7)
            y = orc.jitconn_mv(ev, "homo", 1, od, layout, shape[0], shape[1], 1.0, 0.0, clen, 7)
            np.testing.assert_array_equal(y, (M @ ev).astype(np.float32))


@pytest.mark.parametrize("layout", [orc.LAYOUT_LANE32, orc.LAYOUT_ROW])
def test_determinism_and_seed_sensitivity(layout):
    ev = np.random.default_rng(2).random(400) < 0.3
    a = orc.jitconn_mv(ev, "uniform", 1, True, layout, 300, 400, 0.0, 1.0, 19, 1234)
    b = orc.jitconn_mv(ev, "uniform", 1, True, layout, 300, 400, 0.0, 1.0, 19, 1234)
    c = orc.jitconn_mv(ev, "uniform", 1, True, layout, 300, 400, 0.0, 1.0, 19, 1235)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(a, c)


def test_linearity_in_weight():
    ev = np.random.default_rng(3).random(400) < 0.3
    r1 = orc.jitconn_mv(ev, "homo", 1, True, 0, 300, 400, 1.5, 0.0, 19, 11)
    r2 = orc.jitconn_mv(ev, "homo", 1, True, 0, 300, 400, 3.0, 0.0, 19, 11)
    r3 = orc.jitconn_mv(ev, "homo", 1, True, 0, 300, 400, 4.5, 0.0, 19, 11)
    np.testing.assert_allclose(r1 * 2, r2, rtol=1e-6)
    np.testing.assert_allclose(r1 * 3, r3, rtol=1e-6)


def test_lane32_quirks():
    # Q2: index step*(l+1)-1 is never generated; Q1: with num_col > 32*step-1
    # the tail columns are unreachable (step comes from num_row)
    num_row, num_col, clen = 64, 400, 1
    r, c, _ = orc.jitconn_enumerate(0, True, 0, num_row, num_col, 1.0, 0.0, clen, 3)
    step = max((num_row + 1) >> 5, 1)
    assert step == 2
    reach = set(c.tolist())
    # clen=1 connects every reachable index: lane l covers [2l, 2l] only
    assert reach == set(range(0, 64, 2))
    for l in range(32):
        assert step * (l + 1) - 1 not in reach


def test_row_layout_covers_all_with_clen1():
    r, c, _ = orc.jitconn_enumerate(0, True, 1, 5, 9, 1.0, 0.0, 1, 3)
    assert len(r) == 45     # p=1: every (row, col) exactly once


def test_mean_in_degree():
    # E[#connections per row] ~= num_col * 2/(clen+1) for the row layout
    num_row, num_col, p = 400, 20000, 0.01
    clen = np_ref.conn_len(p)
    r, c, _ = orc.jitconn_enumerate(0, True, 1, num_row, num_col, 1.0, 0.0, clen, 5)
    mean = len(r) / num_row
    assert abs(mean - num_col * 2 / (clen + 1)) < 0.05 * num_col * 2 / (clen + 1)
