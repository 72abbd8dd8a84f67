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


# ---------------------------------------------------------------------------
# Conductance-coupled update, pinned by the closed form of a linear ODE.
# With the conductances held constant (c_g = 0 switches the Expon decay
# off), the membrane equation of LifLTC.derivative (lif.py:254-256) with the
# COBA currents I = I_ext + sum_k g_k (E_k - V) (outputs.py:53-54) is the
# linear ODE  dV/dt = a V + b,  a = -(1 + R sum g)/tau,
# b = (V_rest + R (I_ext + sum g_k E_k))/tau, and exponential Euler
# (exponential.py:357-362) is exact on it:
#     V(t+dt) = V_inf + (V - V_inf) e^{a dt},  V_inf = -b/a.
# LifRef (lif.py:550-556) sums the currents once with the old V and
# integrates dV/dt = (-V + V_rest + R I_tot)/tau:
#     V(t+dt) = V_rest + R I_tot + (V - V_rest - R I_tot) e^{-dt/tau}.
# A wrong sign of g (E - V), a wrong reversal potential or the wrong linear
# term moves the result far outside the f32 tolerance (checked below).
# ---------------------------------------------------------------------------
E_COBA = (0.0, -80.0)      # E_e, E_i of examples/dynamics_simulation/COBA.py


def _closed_form(V, ge, gi, I, R, tau, dt, ltc, E=E_COBA, V_rest=-60.0):
    V, ge, gi = (np.asarray(x, np.float64) for x in (V, ge, gi))
    if ltc:
        gs = ge + gi
        a = -(1.0 + R * gs) / tau
        v_inf = (V_rest + R * (I + ge * E[0] + gi * E[1])) / (1.0 + R * gs)
        return v_inf + (V - v_inf) * np.exp(a * dt)
    i_tot = I + ge * (E[0] - V) + gi * (E[1] - V)
    v_inf = V_rest + R * i_tot
    return v_inf + (V - v_inf) * np.exp(-dt / tau)


def _one_step(V, ge, gi, I, R, dt, ltc, c=(0.0, 0.0)):
    V1 = V.copy()
    tl = np.full(V.size, -1e7, np.float32)
    sp = np.zeros(V.size, np.uint8)
    g = [ge.copy(), gi.copy()]
    params = (-60.0, -50.0, -60.0, 20.0, 5.0, R)
    orc.lif_coba_step(V1, tl, sp, g, [5.0, 10.0], list(E_COBA), list(c), I, params, 1.0, dt, ltc=ltc)
    return V1, sp, g


@pytest.mark.parametrize("ltc", [True, False])
@pytest.mark.parametrize("R", [1.0, 2.0])
@pytest.mark.parametrize("dt", [0.1, 1.0])
def test_conductance_step_closed_form(ltc, R, dt):
    rng = np.random.default_rng(11)
    n = 3000
    V = rng.uniform(-75.0, -52.0, n).astype(np.float32)
    ge = rng.uniform(0.0, 3.0, n).astype(np.float32)
    gi = rng.uniform(0.0, 2.0, n).astype(np.float32)
    I = 20.0
    V1, sp, g = _one_step(V, ge, gi, I, R, dt, ltc)
    np.testing.assert_array_equal(g[0], ge)        # c_g = 0: conductances unchanged
    ref = _closed_form(V, ge, gi, I, R, 20.0, dt, ltc)
    quiet = (sp == 0) & (ref < -50.01)
    assert quiet.sum() > n // 2
    err = np.abs(V1[quiet] - ref[quiet])
    tol = 6e-5 if dt == 0.1 else 2e-4          # a few f32 ulp of |V| <= 80 over the step's operations
    assert err.max() <= tol
    # plausible mistakes are far outside the tolerance
    wrong = [
        _closed_form(V, ge, gi, I, R, 20.0, dt, ltc, E=(0.0, -70.0)),       # wrong E_i
        _closed_form(V, gi, ge, I, R, 20.0, dt, ltc),                        # E/I conductances swapped
        _closed_form(V, -ge, -gi, I, R, 20.0, dt, ltc),                      # sign of g (E - V)
        _closed_form(V, ge, gi, I, R, 20.0, dt, not ltc),                    # LTC vs old-V currents
    ]
    for k, w in enumerate(wrong):
        if k == 3 and dt == 0.1:
            # the two forms differ at second order in dt: resolvable, but
            # only by ~10x the tolerance at dt = 0.1
            assert np.abs(V1[quiet] - w[quiet]).max() > 5 * tol
        else:
            assert np.abs(V1[quiet] - w[quiet]).max() > 50 * tol, k


def test_conductance_decays_before_the_update():
    # with the decay on, the update uses g e^{-dt/tau_g} (the Expon update
    # runs first, abstract_models.py:127-131, then the neuron, lif.py:461)
    rng = np.random.default_rng(12)
    n = 2000
    V = rng.uniform(-70.0, -55.0, n).astype(np.float32)
    ge = rng.uniform(0.0, 3.0, n).astype(np.float32)
    gi = rng.uniform(0.0, 2.0, n).astype(np.float32)
    c = (orc.expon_coef(5.0, 0.1), orc.expon_coef(10.0, 0.1))
    V1, sp, g = _one_step(V, ge, gi, 20.0, 1.0, 0.1, True, c)
    de, di = math.exp(-0.1 / 5.0), math.exp(-0.1 / 10.0)
    np.testing.assert_allclose(g[0], ge * de, rtol=2e-6)
    np.testing.assert_allclose(g[1], gi * di, rtol=2e-6)
    quiet = sp == 0
    ref = _closed_form(V, ge * de, gi * di, 20.0, 1.0, 20.0, 0.1, True)
    assert np.abs(V1[quiet] - ref[quiet]).max() <= 6e-5
    stale = _closed_form(V, ge, gi, 20.0, 1.0, 20.0, 0.1, True)          # g before its decay
    assert np.abs(V1[quiet] - stale[quiet]).max() > 20 * 6e-5


def test_network_step_order_closed_form():
    # The network oracle's composition (examples/dynamics_simulation/COBA.py:
    # 58-109, align_post.py:271-275): a spike of step i reaches the targets'
    # conductance at step i + 1, excitatory pre neurons feed g_exc (E = 0),
    # inhibitory ones g_inh (E = -80), each g decays before the neuron update.
    # Three neurons: 0 (E) and 2 (I) spike at step 0 (V0 = -40 > V_th) and
    # project onto neuron 1 (E, at rest); neuron 1 is followed in closed form.
    indptr = np.array([0, 1, 1, 2], np.int32)
    indices = np.array([1, 1], np.int32)
    V0 = np.array([-40.0, -60.0, -40.0], np.float32)
    params = (-60.0, -50.0, -60.0, 20.0, 5.0, 1.0)
    we, wi, I, dt = 0.6, 6.7, 20.0, 0.1
    de = 1.0 - float(orc.expon_coef(5.0, dt)) / 5.0
    di = 1.0 - float(orc.expon_coef(10.0, dt)) / 10.0
    v, ge, gi = -60.0, 0.0, 0.0
    for steps in range(1, 8):
        tot, raster, Vf = orc.coba_net_run(2, 1, indptr, indices, we, wi, I, V0, steps, params, return_v=True)
        if steps == 1:
            assert raster[0].tolist() == [1, 0, 1]
        else:
            assert raster[steps - 1].tolist() == [0, 0, 0]
            if steps == 2:
                ge, gi = we, wi          # delivered one step after the spikes
        ge, gi = ge * de, gi * di
        v = float(_closed_form([v], [ge], [gi], I, 1.0, 20.0, dt, True)[0])
        assert abs(Vf[1] - v) <= 1e-4, (steps, Vf[1], v)
    # the same with the populations' conductances swapped is far off
    v_bad = -60.0
    ge, gi = 0.0, 0.0
    for steps in range(1, 8):
        if steps == 2:
            ge, gi = wi, we
        ge, gi = ge * de, gi * di
        v_bad = float(_closed_form([v_bad], [ge], [gi], I, 1.0, 20.0, dt, True)[0])
    assert abs(Vf[1] - v_bad) > 0.05
