"""Pins of the LIF/Expon/COBA exp_euler restatement -- CPU only.

Closed forms (dyn/neurons/lif.py:461-500, dyn/synapses/abstract_models.py:
121-134, integrators/ode/exponential.py:357-362):
* exp_euler is exact on the linear Expon ODE: g(t+dt) = g e^{-dt/tau};
* a LIF neuron at V_rest with constant input I reaches V_th at
  t* = tau ln(RI / (RI - (V_th - V_rest))) = 20 ln 2 ms for I = 20;
* after a spike at t_s the neuron is held while (t - t_s) <= tau_ref in
  float32 time t = f32(i) * f32(dt) (49-51 steps for tau_ref = 5 ms);
* the deterministic expf is within 2 ulp of exp.
"""
import math

import numpy as np
import pytest

from oracle import orc

PARAMS = (-60.0, -50.0, -60.0, 20.0, 5.0, 1.0)   # V_rest, V_th, V_reset, tau, tau_ref, R


def _ulp_err(a, b):
    a = np.float32(a)
    return abs(float(a) - b) / float(np.spacing(np.float32(abs(b))))


def test_expf_accuracy():
    xs = np.concatenate([np.linspace(-87, 88, 20001), np.linspace(-0.2, 0.0, 5001)]).astype(np.float32)
    worst = max(_ulp_err(orc.expf(x), math.exp(float(x))) for x in xs[::7])
    assert worst <= 2.0


def test_expon_decay_closed_form():
    dt = np.float32(0.1)
    for tau in (5.0, 10.0):
        c = orc.expon_coef(tau, dt)
        # g + c*(-g/tau) == g*exp(-dt/tau) up to float32 rounding
        factor = 1.0 - float(c) / tau
        assert abs(factor - math.exp(-0.1 / tau)) < 2e-6


def _run_single(I, steps, ltc=True):
    V = np.full(1, -60.0, np.float32)
    t_last = np.full(1, -1e7, np.float32)
    sp = np.zeros(1, np.uint8)
    raster = []
    for i in range(steps):
        t = np.float32(i) * np.float32(0.1)
        orc.lif_coba_step(V, t_last, sp, [], [], [], [], I, PARAMS, t, 0.1, ltc=ltc)
        raster.append(bool(sp[0]))
    return np.array(raster)


@pytest.mark.parametrize("ltc", [True, False])
def test_threshold_time_closed_form(ltc):
    raster = _run_single(20.0, 400, ltc)
    first = int(np.flatnonzero(raster)[0])
    t_star = 20.0 * math.log(2.0)        # 13.86 ms
    # V reaches threshold inside step `first` (t = first*dt is the step's time)
    assert abs(first * 0.1 - t_star) <= 0.1 + 1e-6


def test_refractory_length():
    raster = _run_single(40.0, 1000)
    spikes = np.flatnonzero(raster)
    isi = np.diff(spikes)
    # held for every step with (t - t_s) <= 5 ms in f32 time; rounding of
    # t = f32(i)*f32(0.1) makes that 49, 50 or 51 steps, then integration
    # restarts from V_reset
    ts = [np.float32(i) * np.float32(0.1) for i in range(1000)]
    for s0, s1 in zip(spikes[:-1], spikes[1:]):
        held = [i for i in range(s0 + 1, s1) if (ts[i] - ts[s0]) <= np.float32(5.0)]
        assert len(held) in (49, 50, 51)
        assert s1 > max(held)
    # with I=40 the free charge time is tau*ln(40/30) = 5.75 ms
    assert all(abs(d * 0.1 - (5.0 + 20 * math.log(40 / 30))) < 0.25 for d in isi)


def test_f32_vs_f64_twin():
    rng = np.random.default_rng(0)
    n = 4000
    V = rng.uniform(-60, -50, n).astype(np.float32)
    tl = np.full(n, -1e7, np.float32)
    sp = np.zeros(n, np.uint8)
    ge = rng.uniform(0, 2, n).astype(np.float32)
    gi = rng.uniform(0, 5, n).astype(np.float32)
    V64, tl64, ge64, gi64 = V.astype(np.float64), tl.astype(np.float64), ge.astype(np.float64), gi.astype(np.float64)
    sp64 = np.zeros(n, np.uint8)
    c = [orc.expon_coef(5.0, 0.1), orc.expon_coef(10.0, 0.1)]
    orc.lif_coba_step(V, tl, sp, [ge, gi], [5.0, 10.0], [0.0, -80.0], c, 20.0, PARAMS, 1.0, 0.1)
    orc.lif_coba_step_f64(V64, tl64, sp64, [ge64, gi64], [5.0, 10.0], [0.0, -80.0], 20.0, PARAMS, 1.0, 0.1)
    np.testing.assert_allclose(ge, ge64, rtol=1e-6)
    np.testing.assert_allclose(gi, gi64, rtol=1e-6)
    same = sp == sp64
    np.testing.assert_allclose(V[same], V64[same], rtol=0, atol=2e-4)
