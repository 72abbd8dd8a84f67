"""Batched event operators (torch.func.vmap over the events): one launch for
B event vectors (bpx_event_csrmv_batch, bpx_event_csrmv_tiled_batch,
bpx_jitconn_mv_batch), checked per vector against the oracle; row-parallel
jitconn bit-identical to the single-vector call."""
import numpy as np
import pytest
import torch

import brainpy_amd.math as bm
from oracle import np_ref, orc
from tests.helpers import dense_of, fixed_prob_csr

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _events(B, n, rate, seed):
    return np.random.default_rng(seed).random((B, n)) < rate


@pytest.mark.parametrize("B", [1, 5, 32, 37, 64])
@pytest.mark.parametrize("layout", ["lane32", "row"])
@pytest.mark.parametrize("dist", ["homo", "uniform", "normal"])
def test_jitconn_outdim_batch_equals_single_and_oracle(B, layout, dist):
    num_row, num_col, p, seed = 3000, 4000, 0.02, 77
    E = _events(B, num_col, 0.1, B)
    w0, w1 = {"homo": (0.37, 0.0), "uniform": (-0.5, 2.0), "normal": (0.5, 0.1)}[dist]
    fn = {"homo": lambda e: bm.jitconn.event_mv_prob_homo(e, w0, p, seed, shape=(num_row, num_col), layout=layout),
          "uniform": lambda e: bm.jitconn.event_mv_prob_uniform(e, w0, w1, p, seed, shape=(num_row, num_col),
                                                                layout=layout),
          "normal": lambda e: bm.jitconn.event_mv_prob_normal(e, w0, w1, p, seed, shape=(num_row, num_col),
                                                              layout=layout)}[dist]
    Et = torch.from_numpy(E).to(DEV)
    Y = torch.func.vmap(fn)(Et).cpu().numpy()
    assert Y.shape == (B, num_row)
    lay = orc.LAYOUT_LANE32 if layout == "lane32" else orc.LAYOUT_ROW
    for b in range(B):
        single = fn(Et[b]).cpu().numpy()
        np.testing.assert_array_equal(Y[b], single)
        ref = orc.jitconn_mv(E[b], dist, 1, True, lay, num_row, num_col, np.float32(w0), np.float32(w1),
                             np_ref.conn_len(p), seed)
        if dist == "normal":
            np.testing.assert_allclose(Y[b], ref, rtol=1e-5, atol=2e-5)
        else:
            np.testing.assert_array_equal(Y[b], ref)


@pytest.mark.parametrize("B", [3, 40])
@pytest.mark.parametrize("layout", ["lane32", "row"])
def test_jitconn_scatter_batch_vs_oracle(B, layout):
    num_row, num_col, p, seed = 2500, 3500, 0.03, 5
    E = _events(B, num_col, 0.05, 100 + B)
    Et = torch.from_numpy(E).to(DEV)
    Y = torch.func.vmap(lambda e: bm.jitconn.event_mv_prob_homo(e, 1.0, p, seed, shape=(num_row, num_col),
                                                                outdim_parallel=False, layout=layout))(Et)
    Y = Y.cpu().numpy()
    Yu = torch.func.vmap(lambda e: bm.jitconn.event_mv_prob_uniform(e, 0.0, 1.0, p, seed, shape=(num_row, num_col),
                                                                    outdim_parallel=False, layout=layout))(Et)
    Yu = Yu.cpu().numpy()
    lay = orc.LAYOUT_LANE32 if layout == "lane32" else orc.LAYOUT_ROW
    for b in range(B):
        # integer counts: exact in any order
        np.testing.assert_array_equal(Y[b], orc.jitconn_mv(E[b], "homo", 1, False, lay, num_row, num_col, 1.0, 0.0,
                                                           np_ref.conn_len(p), seed))
        np.testing.assert_allclose(Yu[b], orc.jitconn_mv(E[b], "uniform", 1, False, lay, num_row, num_col, 0.0, 1.0,
                                                         np_ref.conn_len(p), seed), rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("wkind", ["homo", "f32", "bf16"])
@pytest.mark.parametrize("B", [2, 33])
def test_event_csrmv_batch_vs_oracle(transpose, wkind, B):
    m, n = 1500, 2500
    idx, ptr = fixed_prob_csr(m, n, 0.02, 9)
    rng = np.random.default_rng(3)
    w = np.float32(1.0) if wkind == "homo" else rng.normal(0.5, 0.1, idx.size).astype(np.float32)
    E = _events(B, m if transpose else n, 0.1, 7 + B)
    wt = torch.tensor([w], device=DEV) if wkind == "homo" else torch.from_numpy(w).to(DEV)
    if wkind == "bf16":
        wt = wt.to(torch.bfloat16)
    ti, tp = torch.from_numpy(idx).to(DEV), torch.from_numpy(ptr).to(DEV)
    Et = torch.from_numpy(E).to(DEV)
    wref = wt.float().cpu().numpy()
    D = dense_of(wref if wkind != "homo" else np.float32(1.0), idx, ptr, (m, n))
    outs = {"stateless": bm.event.csrmv_batch(wt, ti, tp, Et, shape=(m, n), transpose=transpose, stateless=True),
            "vmap": torch.func.vmap(lambda e: bm.event.csrmv(wt, ti, tp, e, shape=(m, n), transpose=transpose))(Et)}
    for name, Y in outs.items():
        Y = Y.cpu().numpy()
        for b in range(B):
            on = E[b].astype(np.float64)
            ref = on @ D if transpose else D @ on
            mag = on @ np.abs(D) if transpose else np.abs(D) @ on
            if wkind == "homo":
                np.testing.assert_array_equal(Y[b], ref.astype(np.float32), err_msg=name)
            else:
                assert np.all(np.abs(Y[b] - ref) <= 1e-5 * mag + 1e-30), name


def test_event_csrmv_vmap_f64_falls_back_to_rows():
    m, n = 300, 400
    idx, ptr = fixed_prob_csr(m, n, 0.05, 2)
    w = np.random.default_rng(1).normal(0.5, 0.1, idx.size)
    E = torch.from_numpy(_events(4, m, 0.2, 3)).to(DEV)
    wt = torch.from_numpy(w).to(DEV)
    Y = torch.func.vmap(lambda e: bm.event.csrmv(wt, idx, ptr, e, shape=(m, n), transpose=True))(E)
    assert Y.dtype == torch.float64
    for b in range(4):
        np.testing.assert_allclose(Y[b].cpu().numpy(), orc.event_csrmv(w, idx, ptr, E[b].cpu().numpy(), (m, n), True),
                                   rtol=1e-12, atol=1e-12)
