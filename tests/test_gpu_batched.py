"""GPU parity of the batched network step (bpx_net_step).

* just-in-time networks: the spike raster equals the oracle's serial run
  (lane32 event-driven jitconn, outdim_parallel=False, then the neuron step)
  bit for bit;
* CSR networks: equals the serial CSR oracle of the EICOBA model;
* a network gives the same raster alone and inside a batch of other
  networks, with and without HIP-graph capture.
"""
import numpy as np
import pytest
import torch

from brainpy_amd.dyn import EICOBANet
from brainpy_amd.dyn.batched import BatchedCOBA
from brainpy_amd.math.jitconn import conn_len
from oracle import orc

pytestmark = pytest.mark.gpu

PARAMS = (-60.0, -50.0, -60.0, 20.0, 5.0, 1.0)


def _jit_specs(sizes, seed0=100):
    rng = np.random.default_rng(seed0)
    specs = []
    for k, N in enumerate(sizes):
        ne = int(N * 0.8)
        specs.append(dict(n_exc=ne, n_inh=N - ne, seed=seed0 + k, w_exc=float(rng.uniform(0.3, 0.9)),
                          w_inh=float(rng.uniform(4, 9)), I_ext=float(rng.uniform(15, 25)), in_degree=80))
    return specs


def _oracle_jit(batch, k, steps, delays=(1, 1, 1, 1)):
    sp = batch.specs[k]
    d = batch.descs[k]
    ne, ni = sp["n_exc"], sp["n_inh"]
    V = batch.V0[batch.net_slice(k)].copy()
    VE, VI = V[:ne].copy(), V[ne:].copy()
    tl = {"E": np.full(ne, -1e7, np.float32), "I": np.full(ni, -1e7, np.float32)}
    s = {"E": np.zeros(ne, np.uint8), "I": np.zeros(ni, np.uint8)}
    g = {"E": [np.zeros(ne, np.float32), np.zeros(ne, np.float32)],
         "I": [np.zeros(ni, np.float32), np.zeros(ni, np.float32)]}
    Vs = {"E": VE, "I": VI}
    size = {"E": ne, "I": ni}
    prm = batch.params
    c_g = [np.float32(prm.c_e), np.float32(prm.c_i)]
    raster = np.zeros((steps, ne + ni), np.uint8)
    # (projection index, pre, post, conductance group, weight)
    projs = [(0, "E", "E", 0, d.w_exc), (1, "E", "I", 0, d.w_exc), (2, "I", "E", 1, d.w_inh),
             (3, "I", "I", 1, d.w_inh)]
    hist = {"E": [], "I": []}     # spikes of steps 0, 1, ... per population
    for i in range(steps):
        t = np.float32(i) * np.float32(prm.dt)
        for q, pre, post, kind, w in projs:
            # the projection with delay D reads the spikes of step i - D
            src = i - max(int(delays[q]), 1)
            if src < 0:
                continue
            out = orc.jitconn_mv(hist[pre][src].view(bool), "homo", 1, False, orc.LAYOUT_LANE32, size[post],
                                 size[pre], np.float32(w), 0.0, d.clen[q], d.seed[q])
            g[post][kind] += out
        for p in ("E", "I"):
            orc.lif_coba_step(Vs[p], tl[p], s[p], g[p], [prm.tau_e, prm.tau_i], [prm.E_e, prm.E_i], c_g, d.I_ext,
                              PARAMS, t, prm.dt, ltc=bool(prm.ltc))
        raster[i, :ne] = s["E"]
        raster[i, ne:] = s["I"]
        hist["E"].append(s["E"].copy())
        hist["I"].append(s["I"].copy())
    return raster


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("ltc", [True, False])
def test_jit_batch_vs_oracle(ltc, fused):
    batch = BatchedCOBA(_jit_specs([1000, 1500, 4000]), ltc=ltc)
    assert batch.fused
    raster = batch.run(100, raster=True, fused=fused).cpu().numpy()
    for k in range(3):
        ref = _oracle_jit(batch, k, 100)
        assert ref.sum() > 0
        np.testing.assert_array_equal(raster[:, batch.net_slice(k)], ref)


@pytest.mark.parametrize("fused", [False, True])
def test_alone_equals_in_mix_and_graph(fused):
    specs = _jit_specs([1200, 3000, 800, 5000, 2000], seed0=7)
    mix = BatchedCOBA(specs)
    r_mix = mix.run(60, raster=True, fused=fused).cpu().numpy()
    for k in (0, 3):
        alone = BatchedCOBA([specs[k]], V_init=mix.V0[mix.net_slice(k)])
        r_alone = alone.run(60, raster=True, fused=fused).cpu().numpy()
        np.testing.assert_array_equal(r_mix[:, mix.net_slice(k)], r_alone)
    # graph capture: 3 replays of 20 steps == 60 eager steps
    g = BatchedCOBA(specs)
    g.capture(20, fused=fused)
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
    np.testing.assert_array_equal(g.V.cpu().numpy(), mix.V.cpu().numpy())
    np.testing.assert_array_equal(g.t_last.cpu().numpy(), mix.t_last.cpu().numpy())
    assert int(g.events.sum()) == int(mix.events.sum()) > 0


@pytest.mark.parametrize("kind", ["jit", "csr"])
def test_tile_sizes_same_raster_and_autotune_restores(kind):
    # every tile size (neurons per block) gives the same raster: counts are
    # integers, so the block decomposition cannot change a result
    if kind == "jit":
        specs = _jit_specs([700, 2600, 1500], seed0=31)
    else:
        net = EICOBANet(seed=5, mode="fixed_order")
        csrs = {name: (net.csr[name][1], net.csr[name][2]) for name in ("E2E", "E2I", "I2E", "I2I")}
        specs = [dict(n_exc=net.n_exc, n_inh=net.n_inh, seed=5, w_exc=0.6, w_inh=6.7, I_ext=20.0, csr=csrs)]
    ref = BatchedCOBA(specs, tile=256)
    r_ref = ref.run(50, raster=True, fused=False).cpu().numpy()
    assert r_ref.sum() > 0
    for tile in (64, 128, 512, 1024):
        b = BatchedCOBA(specs, tile=tile, V_init=ref.V0)
        np.testing.assert_array_equal(b.run(50, raster=True, fused=False).cpu().numpy(), r_ref)
    # autotune runs candidates from a snapshot and restores it: 20 steps,
    # tune, 30 more steps == 50 steps
    b = BatchedCOBA(specs, V_init=ref.V0)
    r1 = b.run(20, raster=True, fused=False).cpu().numpy()
    times = b.autotune(steps=5, repeats=1)
    tiles = {k: v for k, v in times.items() if k != "fused"}
    assert set(tiles) == {64, 128, 256, 512, 1024} and b.tile == min(tiles, key=tiles.get)
    assert ("fused" in times) == b.fused
    assert b.use_fused == (b.fused and times["fused"] < tiles[b.tile])
    r2 = b.run(30, raster=True, fused=False).cpu().numpy()
    np.testing.assert_array_equal(np.concatenate([r1, r2]), r_ref)
    np.testing.assert_array_equal(b.V.cpu().numpy(), ref.V.cpu().numpy())
    assert int(b.events.sum()) == int(ref.events.sum())
    # a new batch of the same kind, count and size takes the tuned tile
    assert BatchedCOBA(specs, V_init=ref.V0).tile == b.tile


def test_bf16_weights_rounded():
    specs = _jit_specs([2000], seed0=3)
    b = BatchedCOBA(specs, bf16=True)
    w = np.array([b.descs[0].w_exc, b.descs[0].w_inh], np.float32)
    assert (w.view(np.uint32) & 0xFFFF == 0).all()
    b.run(20)


@pytest.mark.parametrize("fused", [False, True])
def test_csr_network_vs_eicoba_oracle(fused):
    net = EICOBANet(seed=5, mode="fixed_order")
    csrs = {name: (net.csr[name][1], net.csr[name][2]) for name in ("E2E", "E2I", "I2E", "I2I")}
    spec = dict(n_exc=net.n_exc, n_inh=net.n_inh, seed=5, w_exc=0.6, w_inh=6.7, I_ext=20.0, csr=csrs)
    b = BatchedCOBA([spec], V_init=net.V0)
    assert b.fused
    raster = b.run(100, raster=True, fused=fused).cpu().numpy()
    ref = net.run(100).cpu().numpy()
    assert ref.sum() > 0
    np.testing.assert_array_equal(raster, ref)


def test_coba4000_fixture_first_100_steps():
    # the committed 100-step raster of the 4000-neuron network: the op-by-op
    # fixed-order path and the batched step reproduce it bit for bit
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "coba4000_raster100.npz"))
    ref = np.unpackbits(d["raster_bits"], axis=1)[:, :4000]
    net = EICOBANet(seed=0, mode="fixed_order")
    np.testing.assert_array_equal(net.V0, d["V0"])
    np.testing.assert_array_equal(net.run(100).cpu().numpy(), ref)
    csrs = {name: (net.csr[name][1], net.csr[name][2]) for name in ("E2E", "E2I", "I2E", "I2I")}
    spec = dict(n_exc=3200, n_inh=800, seed=0, w_exc=0.6, w_inh=6.7, I_ext=20.0, csr=csrs)
    for fused in (False, True):
        b = BatchedCOBA([spec], V_init=net.V0)
        np.testing.assert_array_equal(b.run(100, raster=True, fused=fused).cpu().numpy(), ref)


def test_plain_path_equals_batched():
    from brainpy_amd.dyn.batched import PlainCOBA
    specs = _jit_specs([1500, 3000], seed0=11)
    b = BatchedCOBA(specs, bf16=True)
    rb = b.run(80, raster=True).cpu().numpy()
    for k in range(2):
        p = PlainCOBA(specs[k], V_init=b.V0[b.net_slice(k)], bf16=True)
        rp = p.run(80, raster=True).cpu().numpy()
        np.testing.assert_array_equal(rb[:, b.net_slice(k)], rp)


def test_plain_path_graph_equals_eager():
    # the plain op-by-op path captured in HIP graphs (device step counter,
    # bpx_lif_coba_step_dev) continues in time on every replay: after 3 x 40
    # replayed steps the state equals 120 eager steps bit for bit
    from brainpy_amd.dyn.batched import PlainCOBA
    for sp in _jit_specs([1500, 4000], seed0=21):
        a = PlainCOBA(sp)
        a.run(120)
        b = PlainCOBA(sp)
        b.capture(40)
        for _ in range(3):
            b.replay()
        torch.cuda.synchronize()
        assert b.i == a.i == 120
        for p in ("E", "I"):
            assert torch.equal(a.V[p], b.V[p]) and torch.equal(a.t_last[p], b.t_last[p])
            assert torch.equal(a.g[p][0], b.g[p][0]) and torch.equal(a.g[p][1], b.g[p][1])
        assert int(a.t_last["E"].max()) > 0          # spikes happened after the first window


def _oracle_served(req, srv, steps):
    """The oracle's serial run of one served request (its own seeds, BF16
    weights, V0 from PCG64(seed + 7)): spike count, events delivered and
    final potentials."""
    from brainpy_amd.dyn.batched import bf16_round
    prm = srv.params
    ne, ni = req.n_exc, req.n_inh
    rng = np.random.Generator(np.random.PCG64(req.seed + 7))
    V0 = rng.uniform(prm.V_rest, prm.V_th, req.n).astype(np.float32)
    Vs = {"E": V0[:ne].copy(), "I": V0[ne:].copy()}
    size = {"E": ne, "I": ni}
    tl = {p: np.full(size[p], -1e7, np.float32) for p in "EI"}
    s = {p: np.zeros(size[p], np.uint8) for p in "EI"}
    g = {p: [np.zeros(size[p], np.float32), np.zeros(size[p], np.float32)] for p in "EI"}
    we, wi = np.float32(bf16_round(req.w_exc)), np.float32(bf16_round(req.w_inh))
    clen = conn_len(min(80.0 / req.n, 1.0))
    projs = [(0, "E", "E", 0, we), (1, "E", "I", 0, we), (2, "I", "E", 1, wi), (3, "I", "I", 1, wi)]
    c_g = [np.float32(prm.c_e), np.float32(prm.c_i)]
    pending = None        # conductance delivered by the previous step's spikes
    spikes = events = 0
    for i in range(steps):
        t = np.float32(i) * np.float32(prm.dt)
        if pending is not None:
            for post, kind, out in pending:
                g[post][kind] += out
        for p in "EI":
            orc.lif_coba_step(Vs[p], tl[p], s[p], g[p], [prm.tau_e, prm.tau_i], [prm.E_e, prm.E_i], c_g, req.I_ext,
                              PARAMS, t, prm.dt, ltc=bool(prm.ltc))
        spikes += int(s["E"].sum()) + int(s["I"].sum())
        # this step's spikes are delivered (and counted as events) now and
        # reach the conductances at the next step
        pending = []
        for q, pre, post, kind, w in projs:
            seed = (req.seed * 4 + q) & 0xFFFFFFFF
            ev = s[pre].view(bool)
            pending.append((post, kind, orc.jitconn_mv(ev, "homo", 1, False, orc.LAYOUT_LANE32, size[post],
                                                       size[pre], w, 0.0, clen, seed)))
            events += int(round(float(orc.jitconn_mv(ev, "homo", 1, False, orc.LAYOUT_LANE32, size[post], size[pre],
                                                     1.0, 0.0, clen, seed).sum())))
    return spikes, events, np.concatenate([Vs["E"], Vs["I"]])


def _served_mix(seed=5, count=8, lo=2000, hi=9000, smin=40, smax=160):
    from brainpy_amd.dyn.serve import Request
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        reqs.append(Request(k, n, seed=1000 + k, w_exc=rng.uniform(0.3, 0.9), w_inh=rng.uniform(4, 9),
                            I_ext=rng.uniform(15, 25), steps=int(rng.integers(smin, smax)), arrival=0.001 * k,
                            keep_state=True))
    return reqs


def test_served_mix_vs_oracle():
    # every request of a served mix (pools smaller than the sum, so some
    # queue and join later with their own time origin) equals the oracle's
    # serial run of that request alone: spike count, synaptic events and
    # final potentials bit for bit
    from brainpy_amd.dyn.serve import DONE, SimServer
    reqs = _served_mix()
    srv = SimServer(capacity=1 << 14, initial_capacity=1 << 12)   # grows 4k -> 16k, some requests wait
    done = {r.rid: r for r in srv.serve(reqs)}
    assert len(done) == len(reqs) and all(r.status == DONE for r in done.values())
    assert max(r.admitted for r in done.values()) > min(r.completed for r in done.values())   # queueing happened
    for r in reqs:
        spikes, events, V = _oracle_served(r, srv, r.steps)
        assert (r.spikes, r.events) == (spikes, events), r.rid
        np.testing.assert_array_equal(r.V_final, V)


def test_served_failed_slot_is_isolated():
    # a NaN written into one running network's potential: the health check
    # stops that network (status failed) and every other request's results
    # are bit-identical to a run without the fault
    from brainpy_amd.dyn.serve import DONE, FAILED, SimServer
    clean = {r.rid: r for r in SimServer(capacity=1 << 16, health_every=8).serve(_served_mix(seed=9))}
    srv = SimServer(capacity=1 << 16, health_every=8)
    reqs = _served_mix(seed=9)
    victim = reqs[2]

    def poison(s):
        # ten steps after the victim joined (requests arrive over a few ms)
        if victim.status == "active" and s.step_idx - victim.t0 == 10:
            s.ex.inject_fault(victim, where=victim.n // 2)
    done = {r.rid: r for r in srv.serve(reqs, on_step=poison)}
    assert done[2].status == FAILED and 10 < done[2].steps_done <= 18
    for k, r in done.items():
        if k != 2:
            assert r.status == DONE
            assert (r.spikes, r.events) == (clean[k].spikes, clean[k].events)
            np.testing.assert_array_equal(r.V_final, clean[k].V_final)


def test_served_cancel_running_keeps_others():
    from brainpy_amd.dyn.serve import CANCELLED, DONE, SimServer
    clean = {r.rid: r for r in SimServer(capacity=1 << 16).serve(_served_mix(seed=11))}
    srv = SimServer(capacity=1 << 16)
    reqs = _served_mix(seed=11)

    def cancel(s):
        # fifteen steps after request 1 joined (requests arrive over a few ms)
        if reqs[1].status == "active" and s.step_idx - reqs[1].t0 == 15:
            assert s.cancel(1)
    done = {r.rid: r for r in srv.serve(reqs, on_step=cancel)}
    assert done[1].status == CANCELLED and done[1].steps_done == 15
    for k, r in done.items():
        if k != 1:
            assert r.status == DONE and r.spikes == clean[k].spikes
            np.testing.assert_array_equal(r.V_final, clean[k].V_final)


def _twin_error(specs, bf16, steps, stored_weights=False):
    """FP64 twin fed with the run's spike counts:
    g64 = (g64 + count * w) * exp(-dt / tau), w the unrounded weights (or the
    weights the run stores, with `stored_weights`).  Returns the max over
    networks and steps of max|I32 - I64| / max|I64| (synaptic current) and
    the per-network max relative conductance error, E and I."""
    import math
    from brainpy_amd.dyn.batched import bf16_round
    b = BatchedCOBA(specs, bf16=bf16)
    d, T = b.dev, b.total

    def wcol(key):
        vals = [bf16_round(s[key]) if (bf16 and stored_weights) else float(np.float32(s[key])) for s in specs]
        return torch.cat([torch.full((n,), v, dtype=torch.float64) for v, n in zip(vals, b.sizes)]).to(d)

    we, wi = wcol("w_exc"), wcol("w_inh")
    de, di = math.exp(-0.1 / 5.0), math.exp(-0.1 / 10.0)
    ge = torch.zeros(T, dtype=torch.float64, device=d)
    gi = torch.zeros(T, dtype=torch.float64, device=d)
    worst = 0.0
    g_err = np.zeros((len(specs), 2))
    for _ in range(steps):
        base = (b.i % b.n_buf) * 2 * T
        ce = b.counts[base:base + T].double()
        ci = b.counts[base + T:base + 2 * T].double()
        V = b.V.double()
        b.step()
        ge = (ge + ce * we) * de
        gi = (gi + ci * wi) * di
        I64 = ge * (0.0 - V) + gi * (-80.0 - V)
        I32 = b.g_exc.double() * (0.0 - V) + b.g_inh.double() * (-80.0 - V)
        for k in range(len(specs)):
            sl = b.net_slice(k)
            den = float(I64[sl].abs().max())
            if den > 0:
                worst = max(worst, float((I32[sl] - I64[sl]).abs().max()) / den)
            for j, (g32, g64) in enumerate(((b.g_exc, ge), (b.g_inh, gi))):
                m = g64[sl] > 0
                if bool(m.any()):
                    r = float(((g32[sl].double() - g64[sl]).abs()[m] / g64[sl][m]).max())
                    g_err[k, j] = max(g_err[k, j], r)
    return worst, g_err, b


@pytest.mark.gpu
def test_current_error_vs_f64_twin():
    """FP32 arithmetic: per-step synaptic currents within 1e-5 relative of the
    FP64 twin (FP32 weights, and BF16-stored weights against a twin using the
    same stored weights).  Against the unrounded weights the BF16 run's
    conductances are off by exactly each network's weight rounding (to 1e-5)."""
    from brainpy_amd.dyn.batched import bf16_round
    specs = _jit_specs((1500, 4000, 9000, 2500))
    e32, _, b = _twin_error(specs, False, 200)
    assert int(b.spike_count.sum()) > 0
    assert e32 <= 1e-5, e32
    e16s, _, _ = _twin_error(specs, True, 200, stored_weights=True)
    assert e16s <= 1e-5, e16s
    _, g_err, _ = _twin_error(specs, True, 200)
    for k, s in enumerate(specs):
        for j, key in enumerate(("w_exc", "w_inh")):
            w = float(np.float32(s[key]))
            delta = abs(bf16_round(s[key]) - w) / w
            assert abs(g_err[k, j] - delta) <= 1e-5, (k, key, g_err[k, j], delta)


def test_jit_delays_vs_oracle_and_mix():
    # per-projection delays (E2E, E2I, I2E, I2I) in steps, from ms as
    # VarDelay computes them; networks with different delays in one batch
    specs = _jit_specs([1200, 2500, 1800], seed0=31)
    specs[0]["delay"] = [0.3, 0.7, 1.0, 0.0]       # 3, 7, 10, 1 steps
    specs[1]["delay"] = 2.5                        # 25 steps everywhere
    batch = BatchedCOBA(specs)
    assert batch.n_buf == 26
    raster = batch.run(90, raster=True).cpu().numpy()
    for k, d in ((0, (3, 7, 10, 1)), (1, (25, 25, 25, 25)), (2, (1, 1, 1, 1))):
        ref = _oracle_jit(batch, k, 90, delays=d)
        assert ref.sum() > 0
        np.testing.assert_array_equal(raster[:, batch.net_slice(k)], ref)
    alone = BatchedCOBA([specs[0]], V_init=batch.V0[batch.net_slice(0)])
    np.testing.assert_array_equal(alone.run(90, raster=True).cpu().numpy(), raster[:, batch.net_slice(0)])


def test_csr_delays_vs_oracle_and_graph():
    net = EICOBANet(seed=9, mode="fixed_order")
    csrs = {name: (net.csr[name][1], net.csr[name][2]) for name in ("E2E", "E2I", "I2E", "I2I")}
    spec = dict(n_exc=net.n_exc, n_inh=net.n_inh, seed=9, w_exc=0.6, w_inh=6.7, I_ext=20.0, csr=csrs,
                delay=[0.5, 0.2, 1.2, 0.1])
    b = BatchedCOBA([spec], V_init=net.V0)
    raster = b.run(120, raster=True).cpu().numpy()
    from brainpy_amd.dyn.batched import combined_csr
    idx, ptr = combined_csr(csrs, net.n_exc, net.n_inh)
    _, ref = orc.coba_net_run(net.n_exc, net.n_inh, ptr, idx, 0.6, 6.7, 20.0, net.V0, 120, PARAMS,
                              delays=[5, 2, 12, 1])
    assert ref.sum() > 0
    np.testing.assert_array_equal(raster, ref)
    g = BatchedCOBA([spec], V_init=net.V0)
    g.capture(40)
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(g.V.cpu().numpy(), b.V.cpu().numpy())


def _state(b):
    return [t.cpu().numpy().copy() for t in (b.V, b.t_last, b.g_exc, b.g_inh, b.counts, b.events, b.spike_count,
                                             b.spikes)]


@pytest.mark.parametrize("kind", ["jit", "csr"])
def test_fused_run_equals_step_launches(kind):
    # the multi-step launch (one block per network, counts in LDS) leaves the
    # same raster, state, count ring, event and spike totals as one launch per
    # step, with delays, interleaved with step launches, and under capture
    if kind == "jit":
        specs = _jit_specs([600, 2700, 1300], seed0=41)
        specs[0]["delay"] = [0.3, 0.1, 0.2, 0.0]      # 3, 1, 2, 1 steps
    else:
        net = EICOBANet(seed=3, mode="fixed_order")
        csrs = {name: (net.csr[name][1], net.csr[name][2]) for name in ("E2E", "E2I", "I2E", "I2I")}
        specs = [dict(n_exc=net.n_exc, n_inh=net.n_inh, seed=3, w_exc=0.6, w_inh=6.7, I_ext=20.0, csr=csrs,
                      delay=[0.2, 0.1, 0.0, 0.2])]     # 2, 1, 1, 2 steps
    ref = BatchedCOBA(specs, record=True)
    assert ref.n_buf == (4 if kind == "jit" else 3) and ref.fused
    r_ref = ref.run(90, raster=True, fused=False).cpu().numpy()
    assert r_ref.sum() > 0
    b = BatchedCOBA(specs, record=True, V_init=ref.V0)
    parts = [b.run(30, raster=True, fused=True), b.run(20, raster=True, fused=False),
             b.run(40, raster=True, fused=True)]
    np.testing.assert_array_equal(torch.cat(parts).cpu().numpy(), r_ref)
    for x, y in zip(_state(b), _state(ref)):
        np.testing.assert_array_equal(x, y)
    assert b.i == ref.i == 90
    # without a raster a recording batch keeps the last step's flags
    ref.run(7, fused=False)
    b.run(7, fused=True)
    for x, y in zip(_state(b), _state(ref)):
        np.testing.assert_array_equal(x, y)
    # captured multi-step launch: 3 replays of 30 steps
    g = BatchedCOBA(specs, record=True, V_init=ref.V0)
    g.capture(30, fused=True)
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
    g.run(7, fused=False)
    for x, y in zip(_state(g), _state(ref)):
        np.testing.assert_array_equal(x, y)


def test_fused_run_at_the_lds_limit_and_fallback():
    # a network of exactly bpx_net_run_max_neurons(2) neurons runs fused; a
    # batch with a larger network keeps per-step launches
    from brainpy_amd import _lib
    cap = int(_lib.lib().bpx_net_run_max_neurons(2))
    assert cap == 5456
    specs = _jit_specs([cap, 900], seed0=77)
    b = BatchedCOBA(specs)
    assert b.fused and b.n_buf == 2
    ref = BatchedCOBA(specs, V_init=b.V0)
    r = b.run(40, raster=True, fused=True).cpu().numpy()
    np.testing.assert_array_equal(r, ref.run(40, raster=True, fused=False).cpu().numpy())
    assert r.sum() > 0
    big = BatchedCOBA(_jit_specs([cap + 1, 900], seed0=77))
    assert not big.fused
    with pytest.raises(ValueError):
        big.run_fused(5)
    big.run(5)      # per-step launches


def test_fused_run_reports_a_network_above_max_neurons():
    # bpx_net_run called directly with max_neurons below a network's size:
    # that network is left unchanged and BPX_ERR_SHAPE is written to the
    # status word; the others advance as usual
    import ctypes
    from brainpy_amd import _lib
    specs = _jit_specs([3000, 1200], seed0=31)
    b = BatchedCOBA(specs)
    ref = BatchedCOBA(specs, V_init=b.V0)
    V_big = b.V[b.net_slice(0)].clone()
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    L = _lib.lib()
    rc = L.bpx_net_run(_lib.stream_ptr(), _lib.ptr(b.d_descs), 2, 2000, b.max_in_degree, b.kind, b.total,
                       _lib.ptr(b.V), _lib.ptr(b.t_last), _lib.ptr(b.g_exc), _lib.ptr(b.g_inh), _lib.ptr(b.counts),
                       b.n_buf, None, None, _lib.ptr(b.events), _lib.ptr(b.spike_count), ctypes.byref(b.params), 0,
                       None, 30, _lib.ptr(status))
    assert rc == 0
    torch.cuda.synchronize()
    assert int(status[0]) == -2
    assert torch.equal(b.V[b.net_slice(0)], V_big)
    ref.run(30, fused=False)
    assert torch.equal(b.V[b.net_slice(1)], ref.V[ref.net_slice(1)])


def test_coba4000_5000_steps_every_path_vs_oracle():
    # long horizon: the 4000-neuron network for 5000 steps (0.5 s of model
    # time), per-step launches, the multi-step launch and a captured graph of
    # multi-step launches against the serial CPU restatement, spike for spike
    from brainpy_amd.dyn.batched import combined_csr
    net = EICOBANet(seed=0, mode="atomic")
    csrs = {name: (net.csr[name][1], net.csr[name][2]) for name in ("E2E", "E2I", "I2E", "I2I")}
    spec = dict(n_exc=3200, n_inh=800, seed=0, w_exc=0.6, w_inh=6.7, I_ext=20.0, csr=csrs)
    idx, ptr = combined_csr(csrs, 3200, 800)
    steps = 5000
    total, ref = orc.coba_net_run(3200, 800, ptr, idx, 0.6, 6.7, 20.0, net.V0, steps, PARAMS)
    assert total > 20000
    for fused in (False, True):
        b = BatchedCOBA([spec], V_init=net.V0)
        np.testing.assert_array_equal(b.run(steps, raster=True, fused=fused).cpu().numpy(), ref)
        assert int(b.spike_count.sum()) == total
    g = BatchedCOBA([spec], V_init=net.V0, record=True)
    g.capture(500, fused=True)
    for _ in range(steps // 500):
        g.replay()
    torch.cuda.synchronize()
    assert int(g.spike_count.sum()) == total
    np.testing.assert_array_equal(g.spikes.cpu().numpy(), ref[-1])


def test_served_full_size_range_vs_oracle():
    # the served configuration's size range, 4k to 1M neurons (a 1M-neuron
    # request forced in), with in-degree 80: every request equals the
    # oracle's serial run alone, bit for bit
    from brainpy_amd.dyn.serve import DONE, Request, SimServer
    reqs = _served_mix(seed=13, count=5, lo=4000, hi=1_000_000, smin=30, smax=60)
    reqs.append(Request(len(reqs), 1_000_000, seed=1999, w_exc=0.6, w_inh=6.7, I_ext=20.0, steps=50,
                        arrival=0.001 * len(reqs), keep_state=True))
    srv = SimServer(capacity=1 << 22)
    done = {r.rid: r for r in srv.serve(reqs)}
    assert all(r.status == DONE for r in done.values())
    for r in reqs:
        spikes, events, V = _oracle_served(r, srv, r.steps)
        assert spikes > 0
        assert (r.spikes, r.events) == (spikes, events), r.rid
        np.testing.assert_array_equal(r.V_final, V)


@pytest.mark.parametrize("graph", [False, True])
def test_c4_scale_batch_vs_oracle(graph):
    # the batched configuration's shape: 256 networks, sizes log-uniform in
    # 1k..64k, 4:1 E:I, jitconn in-degree 80 (above the one-launch run's
    # size limit, so one launch per step, eager or replayed from a captured
    # graph); a sample of networks (the smallest, the largest and four
    # others) equals the oracle's serial run
    rng = np.random.default_rng(256)
    sizes = [int(np.exp(v)) for v in rng.uniform(np.log(1000), np.log(64000), 256)]
    sizes[0], sizes[1] = 1000, 64000
    batch = BatchedCOBA(_jit_specs(sizes, seed0=500))
    assert not batch.fused
    steps = 40
    if graph:
        batch.capture(10, fused=False)
        for _ in range(steps // 10):
            batch.replay()
        torch.cuda.synchronize()
        ref_batch = BatchedCOBA(_jit_specs(sizes, seed0=500))
        raster = ref_batch.run(steps, raster=True, fused=False).cpu().numpy()
        np.testing.assert_array_equal(batch.V.cpu().numpy(), ref_batch.V.cpu().numpy())
    else:
        raster = batch.run(steps, raster=True, fused=False).cpu().numpy()
    pick = [0, 1] + [int(k) for k in rng.choice(np.arange(2, 256), 4, replace=False)]
    for k in pick:
        ref = _oracle_jit(batch, k, steps)
        assert ref.sum() > 0
        np.testing.assert_array_equal(raster[:, batch.net_slice(k)], ref)
