"""GPU parity: the fused LIF/COBA step and the COBA network.

The neuron update uses only correctly rounded operations (and the
library's deterministic exp), so it is bit-identical to the oracle; the
fixed-order network reproduces the oracle's spike raster exactly.
"""
import ctypes

import numpy as np
import pytest
import torch

from brainpy_amd import _lib
from brainpy_amd.dyn import EICOBANet, expon_coef
from oracle import orc

pytestmark = pytest.mark.gpu

PARAMS = (-60.0, -50.0, -60.0, 20.0, 5.0, 1.0)


def _gpu_step(V, tl, sp, gs, tau_g, E_g, c_g, I, t, dt, ltc):
    L = _lib.lib()
    ng = len(gs)
    gp = (ctypes.c_void_p * max(ng, 1))(*[g.data_ptr() for g in gs])
    arr = lambda xs: (ctypes.c_float * max(len(xs), 1))(*xs)
    rc = L.bpx_lif_coba_step(_lib.stream_ptr(), V.numel(), int(ltc), ng, _lib.ptr(V), _lib.ptr(tl), _lib.ptr(sp),
                             ctypes.cast(gp, ctypes.POINTER(ctypes.c_void_p)), ctypes.cast(arr(tau_g), ctypes.c_void_p),
                             ctypes.cast(arr(E_g), ctypes.c_void_p), ctypes.cast(arr(c_g), ctypes.c_void_p), I, None,
                             *PARAMS, t, dt)
    _lib.check(rc, "bpx_lif_coba_step")


@pytest.mark.parametrize("ltc", [True, False])
def test_lif_step_bit_exact(ltc):
    rng = np.random.default_rng(0)
    n = 5000
    V = rng.uniform(-60, -49, n).astype(np.float32)
    tl = np.where(rng.random(n) < 0.2, np.float32(0.5), np.float32(-1e7)).astype(np.float32)
    sp = np.zeros(n, np.uint8)
    ge = rng.uniform(0, 2, n).astype(np.float32)
    gi = rng.uniform(0, 5, n).astype(np.float32)
    c = [expon_coef(5.0, 0.1), expon_coef(10.0, 0.1)]
    assert c[0] == orc.expon_coef(5.0, 0.1) and c[1] == orc.expon_coef(10.0, 0.1)
    d = torch.device("cuda")
    tV, ttl, tsp, tge, tgi = (torch.from_numpy(x.copy()).to(d) for x in (V, tl, sp, ge, gi))
    for k in range(5):
        t = np.float32(k + 10) * np.float32(0.1)
        orc.lif_coba_step(V, tl, sp, [ge, gi], [5.0, 10.0], [0.0, -80.0], c, 20.0, PARAMS, t, 0.1, ltc=ltc)
        _gpu_step(tV, ttl, tsp, [tge, tgi], [5.0, 10.0], [0.0, -80.0], c, 20.0, t, np.float32(0.1), ltc)
    for a, b in ((V, tV), (tl, ttl), (sp, tsp), (ge, tge), (gi, tgi)):
        np.testing.assert_array_equal(a, b.cpu().numpy())


def _oracle_net(net, steps, snaps=None):
    """Serial CPU run of the same network: csrmv scatter in ascending row
    order into a zeroed output, g += out, then the neuron step.  `snaps`
    (a dict keyed by step) receives, for those steps, the state before the
    projections and the conductances after them."""
    V = {"E": net.V0[:net.n_exc].copy(), "I": net.V0[net.n_exc:].copy()}
    tl = {p: np.full(n, -1e7, np.float32) for p, n in (("E", net.n_exc), ("I", net.n_inh))}
    sp = {p: np.zeros(n, np.uint8) for p, n in (("E", net.n_exc), ("I", net.n_inh))}
    g = {p: [np.zeros(n, np.float32), np.zeros(n, np.float32)] for p, n in (("E", net.n_exc), ("I", net.n_inh))}
    raster = np.zeros((steps, net.n_exc + net.n_inh), np.uint8)
    for i in range(steps):
        t = np.float32(i) * np.float32(net.dt)
        if snaps is not None and i in snaps:
            snaps[i] = {"sp": {p: sp[p].copy() for p in sp}, "g": {p: [x.copy() for x in g[p]] for p in g}}
        for name, pre, post, _, k in net.proj_spec:
            w, idx, ptr, shape = net.csr[name]
            out = orc.event_csrmv(np.float32(w), idx, ptr, sp[pre].view(bool), shape, True)
            g[post][k] += out
        if snaps is not None and i in snaps:
            snaps[i]["g_after"] = {p: [x.copy() for x in g[p]] for p in g}
        for p in ("E", "I"):
            orc.lif_coba_step(V[p], tl[p], sp[p], g[p], net.tau_g, net.E_g, net.c_g, net.inp, net.params, t,
                              net.dt, ltc=net.ltc)
        raster[i, :net.n_exc] = sp["E"]
        raster[i, net.n_exc:] = sp["I"]
    return raster, V


@pytest.mark.parametrize("ltc", [True, False])
def test_coba_raster_fixed_order_100_steps(ltc):
    net = EICOBANet(seed=0, ltc=ltc, mode="fixed_order")
    raster = net.run(100).cpu().numpy()
    ref, Vref = _oracle_net(net, 100)
    assert ref.sum() > 0
    np.testing.assert_array_equal(raster, ref)
    np.testing.assert_array_equal(net.V["E"].cpu().numpy(), Vref["E"])


def test_coba_atomic_mode_close():
    # the atomic scatter sums in scheduling order: conductances agree to
    # float32 rounding, the raster stays close over a short window
    net = EICOBANet(seed=1, mode="atomic")
    raster = net.run(50).cpu().numpy()
    ref, _ = _oracle_net(net, 50)
    assert abs(int(raster.sum()) - int(ref.sum())) <= max(5, 0.05 * ref.sum())


def test_coba_atomic_mode_per_step_conductance():
    # the atomic mode's projections (g += scatter of the spikes, FP32
    # atomics straight into g) from the oracle's own state at several steps:
    # each conductance within 1e-6 relative of the oracle's serial
    # g += (scatter into zeros), i.e. float32 rounding of the order only
    net = EICOBANet(seed=1, mode="atomic")
    snaps = {5: None, 20: None, 35: None, 49: None}
    _oracle_net(net, 50, snaps)
    dev = net.dev
    delivered = 0
    for i, st in snaps.items():
        for p in ("E", "I"):
            net.spike[p] = torch.from_numpy(st["sp"][p]).to(dev)
            for k in range(2):
                net.g[p][k] = torch.from_numpy(st["g"][p][k]).to(dev)
        net._project()
        for p in ("E", "I"):
            for k in range(2):
                got = net.g[p][k].cpu().numpy()
                want = st["g_after"][p][k]
                delivered += int(np.count_nonzero(want != st["g"][p][k]))
                np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6 * float(np.abs(want).max() + 1e-30))
    assert delivered > 0
