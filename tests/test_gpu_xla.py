"""GPU parity of the XLA custom-call adapters (bpx_xla_*), called the way
XLA calls a legacy custom-call target: a stream, a void* array of device
buffers in the reference lowering's operand order followed by the outputs,
and the opaque descriptor.  Results are compared with the oracle."""
import numpy as np
import pytest
import torch

from brainpy_amd import _lib, xla
from oracle import np_ref, orc
from tests.helpers import bernoulli, dense_of, fixed_prob_csr

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _sp():
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("transpose", [True, False])
@pytest.mark.parametrize("homo", [True, False])
@pytest.mark.parametrize("ev_kind", ["bool", "f32"])
@pytest.mark.parametrize("itype", [np.int32, np.int64])
def test_event_csrmv_adapter(transpose, homo, ev_kind, itype):
    m, n = 700, 900
    idx, ptr = fixed_prob_csr(m, n, 0.03, 5)
    idx, ptr = idx.astype(itype), ptr.astype(itype)
    rng = np.random.default_rng(3)
    w = np.array([0.75], np.float32) if homo else rng.normal(0.5, 0.1, idx.size).astype(np.float32)
    k = m if transpose else n
    ev = bernoulli(k, 0.2, 11)
    ev_in = ev if ev_kind == "bool" else np.where(ev, rng.random(k) + 0.1, -rng.random(k)).astype(np.float32)
    t = {name: torch.from_numpy(np.ascontiguousarray(a)).to(DEV) for name, a in
         (("w", w), ("idx", idx), ("ptr", ptr), ("ev", ev_in))}
    out = torch.full((n if transpose else m,), np.nan, dtype=torch.float32, device=DEV)
    op = xla.event_csrmv_descriptor((m, n), idx.size, w.size, _lib.BPX_F32, _lib.bpx_dtype(t["idx"]),
                                    _lib.bpx_dtype(t["ev"]), transpose)
    assert xla.call("bpx_xla_event_csrmv", _sp(), [t["w"], t["idx"], t["ptr"], t["ev"], out], op) == 0
    D = dense_of(w, idx, ptr, (m, n))
    on = (ev_in > 0).astype(np.float64)
    ref = on @ D if transpose else D @ on
    mag = (on @ np.abs(D)) if transpose else (np.abs(D) @ on)
    assert np.all(np.abs(out.cpu().numpy() - ref) <= 1e-5 * mag + 1e-30)
    np.testing.assert_allclose(out.cpu().numpy(), orc.event_csrmv(w, idx, ptr, ev_in, (m, n), transpose),
                               rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("transpose", [True, False])
def test_csrmv_adapter(transpose):
    m, n = 500, 800
    idx, ptr = fixed_prob_csr(m, n, 0.05, 9)
    rng = np.random.default_rng(4)
    w = rng.normal(0.0, 1.0, idx.size).astype(np.float32)
    x = rng.normal(0.0, 1.0, m if transpose else n).astype(np.float32)
    tw, ti, tp, tx = (torch.from_numpy(a).to(DEV) for a in (w, idx, ptr, x))
    out = torch.empty(n if transpose else m, dtype=torch.float32, device=DEV)
    op = xla.csrmv_descriptor((m, n), idx.size, w.size, _lib.BPX_F32, _lib.BPX_I32, _lib.BPX_F32, transpose)
    assert xla.call("bpx_xla_csrmv", _sp(), [tw, ti, tp, tx, out], op) == 0
    ref = orc.csrmv(w, idx, ptr, x, (m, n), transpose)
    D = dense_of(w, idx, ptr, (m, n))
    mag = np.abs(x) @ np.abs(D) if transpose else np.abs(D) @ np.abs(x)
    assert np.all(np.abs(out.cpu().numpy() - ref) <= 1e-5 * mag + 1e-30)


@pytest.mark.parametrize("batch,n", [(1, 100000), (3, 5000), (2, 17)])
def test_event_info_adapter(batch, n):
    ev = np.stack([bernoulli(n, 0.1, 40 + b) for b in range(batch)])
    te = torch.from_numpy(ev).to(DEV)
    ids = torch.empty((batch, n), dtype=torch.int32, device=DEV)
    num = torch.empty(batch, dtype=torch.int32, device=DEV)
    op, wsb = xla.event_info_descriptor(n, batch, _lib.BPX_BOOL)
    ws = torch.empty(max(wsb, 4), dtype=torch.uint8, device=DEV)
    assert xla.call("bpx_xla_event_info", _sp(), [te, ids, num, ws], op) == 0
    for b in range(batch):
        rid, rnum = orc.event_info(ev[b])
        np.testing.assert_array_equal(ids[b].cpu().numpy(), rid)
        assert int(num[b]) == int(rnum[0])


@pytest.mark.parametrize("dist", ["homo", "uniform", "normal"])
@pytest.mark.parametrize("od", [True, False])
@pytest.mark.parametrize("event", [True, False])
@pytest.mark.parametrize("workspace", [True, False])
def test_jitconn_adapter(dist, od, event, workspace):
    num_row, num_col, p, seed = 3000, 2000, 0.05, 1234
    code = {"homo": _lib.DIST_HOMO, "uniform": _lib.DIST_UNIFORM, "normal": _lib.DIST_NORMAL}[dist]
    w0, w1 = {"homo": (1.0, 0.0), "uniform": (-0.5, 1.5), "normal": (0.5, 0.1)}[dist]
    if event:
        v = bernoulli(num_col, 0.1, 8)
    else:
        v = np.random.default_rng(2).normal(0.0, 1.0, num_col).astype(np.float32)
    tv = torch.from_numpy(v).to(DEV)
    clen = np_ref.conn_len(p)
    wt = [torch.tensor([w0], dtype=torch.float32, device=DEV)]
    if dist != "homo":
        wt.append(torch.tensor([w1], dtype=torch.float32, device=DEV))
    tc = torch.tensor([clen], dtype=torch.int32, device=DEV)
    ts = torch.tensor([seed], dtype=torch.int64, device=DEV).to(torch.int32)
    out = torch.full((num_row,), np.nan, dtype=torch.float32, device=DEV)
    op, wsb = xla.jitconn_descriptor(num_row, num_col, code, event, _lib.bpx_dtype(tv), od, workspace=workspace)
    bufs = [tv] + wt + [tc, ts, out]
    if wsb:
        bufs.append(torch.empty(wsb, dtype=torch.uint8, device=DEV))
    assert xla.call("bpx_xla_jitconn_mv", _sp(), bufs, op) == 0
    ref = orc.jitconn_mv(v, dist, 1 if event else 0, od, orc.LAYOUT_LANE32, num_row, num_col, w0, w1, clen, seed)
    y = out.cpu().numpy()
    if dist == "homo" and event:
        np.testing.assert_array_equal(y, ref)
    else:
        mag = np.abs(orc.jitconn_mv(np.abs(v).astype(v.dtype), dist if dist == "homo" else "homo", 1 if event else 0,
                                    od, orc.LAYOUT_LANE32, num_row, num_col, 1.0, 0.0, clen, seed))
        bound = (abs(w0) + 6 * abs(w1) + abs(w1)) * mag
        assert np.all(np.abs(y - ref) <= 1e-5 * bound + 1e-6)
