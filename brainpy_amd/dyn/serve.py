"""Served simulation: continuous batching of COBA network requests on one GPU.

Requests (network size, seed, weights, input, number of steps) arrive over
time.  The server keeps every admitted network's state resident in device
pools and advances all of them together, one ``bpx_net_step`` launch per
step.  A request is admitted as soon as its state fits the HBM budget and a
free range of the pools; it joins the batch at the next step with its own
time origin (``t0_step``), so its trajectory is the same as when run alone;
when its last step is done its results are read back and its range is freed
for queued requests.

Admission control: a request whose state can never fit (more neurons than
the pools can ever hold, or more bytes than the budget) is rejected at
submission with a reason; queued requests are admitted in arrival order.
The pools grow with the admitted state, up to the budget, so the budget is
what binds.  A request can be cancelled while queued or running.  Every
``health_every`` steps the running networks' state is checked for non-finite
values; a network that went bad is stopped and reported as failed, and the
others continue unchanged (networks never read each other's state).

The scheduling logic is independent of the device: ``SimServer`` drives an
executor (``DeviceExecutor`` here, a host stand-in in the tests).

Per-request results: status, spike count, synaptic events delivered, final
membrane potentials (optional), arrival / admission / completion times.
"""
import ctypes
import time
from collections import deque

import numpy as np
import torch

from .. import _lib
from ..math.jitconn import conn_len
from .batched import TILE, NetDesc, NetParams, _register, bf16_round
from .coba import expon_coef

# device bytes per neuron: V, t_last, g_exc, g_inh (f32), 4 count words, spike flag
BYTES_PER_NEURON = 4 * 4 + 4 * 4 + 1

QUEUED, ACTIVE, DONE, REJECTED, CANCELLED, FAILED = "queued", "active", "done", "rejected", "cancelled", "failed"


class Request:
    __slots__ = ("rid", "n_exc", "n_inh", "seed", "w_exc", "w_inh", "I_ext", "in_degree", "steps", "arrival",
                 "admitted", "completed", "off", "t0", "spikes", "events", "V_final", "keep_state", "index",
                 "status", "reason", "steps_done")

    def __init__(self, rid, n, seed, w_exc, w_inh, I_ext, steps, in_degree=80, arrival=0.0, keep_state=False):
        self.rid = rid
        self.n_exc = int(round(n * 0.8))
        self.n_inh = int(n) - self.n_exc
        self.seed, self.w_exc, self.w_inh, self.I_ext = int(seed), float(w_exc), float(w_inh), float(I_ext)
        self.in_degree = in_degree
        self.steps = int(steps)
        self.arrival = arrival
        self.admitted = self.completed = None
        self.off = self.t0 = None
        self.spikes = self.events = None
        self.V_final = None
        self.keep_state = keep_state
        self.index = None
        self.status = QUEUED
        self.reason = None
        self.steps_done = 0

    @property
    def n(self):
        return self.n_exc + self.n_inh


class _Pool:
    """First-fit allocator of contiguous neuron ranges."""

    def __init__(self, capacity):
        self.capacity = capacity
        self.free = [(0, capacity)] if capacity > 0 else []

    def alloc(self, n):
        for k, (s, L) in enumerate(self.free):
            if L >= n:
                if L == n:
                    self.free.pop(k)
                else:
                    self.free[k] = (s + n, L - n)
                return s
        return None

    def release(self, s, n):
        self.free.append((s, n))
        self.free.sort()
        merged = []
        for a, L in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + L)
            else:
                merged.append((a, L))
        self.free = merged

    def grow(self, capacity):
        """Extend the pool to `capacity` neurons (new range at the end)."""
        if capacity > self.capacity:
            self.release(self.capacity, capacity - self.capacity)
            self.capacity = capacity


class DeviceExecutor:
    """Resident state of the admitted networks in device pools and the
    batched step launch (bpx_net_step, jitconn projections).  The pools
    start at `initial` neurons and grow (doubling, state copied) up to
    `max_capacity`."""

    def __init__(self, initial, max_capacity, params, bf16, V_rest, V_th):
        self.dev = _lib.device()
        self.L = _register()
        self.params = params
        self.bf16 = bf16
        self.V_rest, self.V_th = V_rest, V_th
        self.max_capacity = int(max_capacity)
        self.capacity = 0
        self._alloc(min(int(initial), self.max_capacity))
        self.max_slots = 4096
        d = self.dev
        self.events = torch.zeros(self.max_slots, dtype=torch.int64, device=d)
        self.spike_count = torch.zeros(self.max_slots, dtype=torch.int32, device=d)
        self.slot_of = None
        self._dirty = True

    def _alloc(self, C):
        d, old = self.dev, self.capacity
        V = torch.zeros(C, dtype=torch.float32, device=d)
        t_last = torch.full((C,), -1e7, dtype=torch.float32, device=d)
        g_exc = torch.zeros(C, dtype=torch.float32, device=d)
        g_inh = torch.zeros(C, dtype=torch.float32, device=d)
        counts = torch.zeros(4 * C, dtype=torch.int32, device=d)
        if old:
            V[:old] = self.V
            t_last[:old] = self.t_last
            g_exc[:old] = self.g_exc
            g_inh[:old] = self.g_inh
            counts.view(4, C)[:, :old] = self.counts.view(4, old)
        self.V, self.t_last, self.g_exc, self.g_inh, self.counts = V, t_last, g_exc, g_inh, counts
        self.capacity = C
        self._dirty = True

    def ensure(self, need):
        """Grow the pools to hold at least `need` neurons (False if that is
        beyond max_capacity)."""
        if need <= self.capacity:
            return True
        if need > self.max_capacity:
            return False
        C = max(self.capacity, 1)
        while C < need:
            C *= 2
        self._alloc(min(C, self.max_capacity))
        return True

    def init(self, req):
        sl = slice(req.off, req.off + req.n)
        rng = np.random.Generator(np.random.PCG64(req.seed + 7))
        self.V[sl] = torch.from_numpy(rng.uniform(self.V_rest, self.V_th, req.n).astype(np.float32)).to(self.dev)
        self.t_last[sl] = -1e7
        self.g_exc[sl] = 0.0
        self.g_inh[sl] = 0.0
        C = self.capacity
        for b in range(4):
            self.counts[b * C + req.off:b * C + req.off + req.n] = 0
        self.events[req.index] = 0
        self.spike_count[req.index] = 0
        self._dirty = True

    def changed(self):
        self._dirty = True

    def _rebuild(self, active):
        n = len(active)
        descs = (NetDesc * max(n, 1))()
        tiles = []
        for k, req in enumerate(active):
            d = descs[k]
            d.off, d.n_exc, d.n_inh = req.off, req.n_exc, req.n_inh
            w_e, w_i = (bf16_round(req.w_exc), bf16_round(req.w_inh)) if self.bf16 else (req.w_exc, req.w_inh)
            d.w_exc, d.w_inh, d.I_ext = w_e, w_i, req.I_ext
            c = conn_len(min(float(req.in_degree) / req.n, 1.0))
            for q in range(4):
                d.seed[q] = (req.seed * 4 + q) & 0xFFFFFFFF
                d.clen[q] = c
            d.t0_step = req.t0
            self.L.bpx_net_prepare(ctypes.byref(d))
            for s in range(0, req.n, TILE):
                tiles += [k, s]
        self.d_descs = torch.frombuffer(bytearray(bytes(descs)), dtype=torch.uint8).to(self.dev)
        self.d_tiles = torch.tensor(tiles if tiles else [0, 0], dtype=torch.int32, device=self.dev)
        self.n_tiles = len(tiles) // 2
        # per-network counters live at the request's slot index
        self.slot_of = torch.tensor([r.index for r in active] or [0], dtype=torch.int64, device=self.dev)
        self.ev_scratch = torch.zeros(max(n, 1), dtype=torch.int64, device=self.dev)
        self.sc_scratch = torch.zeros(max(n, 1), dtype=torch.int32, device=self.dev)
        self._dirty = False

    def _fold(self):
        if self.slot_of is not None:
            self.events.index_add_(0, self.slot_of, self.ev_scratch)
            self.spike_count.index_add_(0, self.slot_of, self.sc_scratch)
            self.ev_scratch.zero_()
            self.sc_scratch.zero_()

    def step(self, active, step_idx):
        if self._dirty:
            self._fold()
            self._rebuild(active)
        if self.n_tiles:
            rc = self.L.bpx_net_step(_lib.stream_ptr(), _lib.ptr(self.d_descs), _lib.ptr(self.d_tiles),
                                     self.n_tiles, TILE, 0, self.capacity, _lib.ptr(self.V), _lib.ptr(self.t_last),
                                     _lib.ptr(self.g_exc), _lib.ptr(self.g_inh), _lib.ptr(self.counts), 2, None,
                                     _lib.ptr(self.ev_scratch), _lib.ptr(self.sc_scratch), ctypes.byref(self.params),
                                     step_idx, None)
            _lib.check(rc, "bpx_net_step")

    def read(self, reqs):
        """(events, spikes, V or None) of each request (synchronises)."""
        self._fold()
        torch.cuda.synchronize()
        ev = self.events.cpu().numpy()
        sc = self.spike_count.cpu().numpy()
        out = []
        for r in reqs:
            V = self.V[r.off:r.off + r.n].cpu().numpy() if r.keep_state else None
            out.append((int(ev[r.index]), int(sc[r.index]), V))
        return out

    def unhealthy(self, active):
        """Requests whose potential or conductances hold a non-finite value:
        one pass over the pools and two gathers at the networks' ranges."""
        if not active:
            return []
        C = self.capacity
        bad = ~(torch.isfinite(self.V) & torch.isfinite(self.g_exc) & torch.isfinite(self.g_inh))
        csum = torch.zeros(C + 1, dtype=torch.int64, device=self.dev)
        torch.cumsum(bad.to(torch.int64), 0, out=csum[1:])
        lo = torch.tensor([r.off for r in active], dtype=torch.int64, device=self.dev)
        hi = torch.tensor([r.off + r.n for r in active], dtype=torch.int64, device=self.dev)
        nbad = (csum[hi] - csum[lo]).cpu().numpy()
        return [r for r, b in zip(active, nbad) if b > 0]

    def sync(self):
        torch.cuda.synchronize()

    def inject_fault(self, req, where=0):
        """Test hook: write a NaN into the potential of neuron `where` of a
        running request (a stand-in for a corrupted slot)."""
        self.V[req.off + where] = float("nan")


class SimServer:
    """Continuous-batching server of just-in-time-connectivity COBA networks.

    hbm_budget: bytes the resident state of admitted requests may use; the
    device pools grow up to it (capacity, if given, caps them lower).
    bf16: round the homogeneous weights to bfloat16.  max_active=1 makes a
    plain FIFO queue (one network at a time).  health_every: steps between
    checks of the running networks' state.
    """

    def __init__(self, capacity=None, hbm_budget=200 * (1 << 30), bf16=True, dt=0.1, ltc=True,
                 V_rest=-60.0, V_th=-50.0, V_reset=-60.0, tau=20.0, tau_ref=5.0, R=1.0, tau_e=5.0, tau_i=10.0,
                 E_e=0.0, E_i=-80.0, max_ahead=16, max_active=None, health_every=64, initial_capacity=1 << 20,
                 executor=None):
        self.budget = int(hbm_budget)
        cap = self.budget // BYTES_PER_NEURON
        if capacity is not None:
            cap = min(cap, int(capacity))
        self.max_capacity = cap
        self.params = NetParams(V_rest=V_rest, V_th=V_th, V_reset=V_reset, tau=tau, tau_ref=tau_ref, R=R,
                                tau_e=tau_e, tau_i=tau_i, E_e=E_e, E_i=E_i, c_e=float(expon_coef(tau_e, dt)),
                                c_i=float(expon_coef(tau_i, dt)), dt=dt, ltc=int(ltc))
        if executor is None:
            executor = DeviceExecutor(min(initial_capacity, cap), cap, self.params, bf16, V_rest, V_th)
        self.ex = executor
        self.pool = _Pool(self.ex.capacity)
        self.used_bytes = 0
        self.queue = deque()
        self.active = []
        self.done = []
        self.step_idx = 0
        self.max_ahead = max_ahead
        self.max_active = max_active
        self.health_every = health_every
        self.free_slots = list(range(self.ex.max_slots - 1, -1, -1))
        self._by_rid = {}

    @property
    def capacity(self):
        return self.ex.capacity

    # ------------------------------------------------------------------
    def _finish(self, req, status, reason=None, now=None):
        req.status = status
        req.reason = reason
        req.completed = time.perf_counter() if now is None else now
        self.done.append(req)

    def submit(self, req):
        """Queue a request, or reject it at once (status 'rejected' with a
        reason) when it can never be served."""
        need = req.n * BYTES_PER_NEURON
        if req.n_exc < 1 or req.n_inh < 1 or req.steps < 1:
            self._finish(req, REJECTED, "needs at least one excitatory and one inhibitory neuron and one step")
        elif need > self.budget:
            self._finish(req, REJECTED, f"state of {need} bytes exceeds the HBM budget of {self.budget} bytes")
        elif req.n > self.max_capacity:
            self._finish(req, REJECTED, f"{req.n} neurons exceed the pool limit of {self.max_capacity}")
        else:
            req.status = QUEUED
            self.queue.append(req)
            self._by_rid[req.rid] = req
            return True
        return False

    def cancel(self, rid):
        """Cancel a queued or running request; its partial results (steps
        done, events, spikes) are kept.  Returns False if it is not known or
        already finished."""
        req = self._by_rid.get(rid)
        if req is None or req.status not in (QUEUED, ACTIVE):
            return False
        if req.status == QUEUED:
            self.queue.remove(req)
            self._finish(req, CANCELLED, "cancelled while queued")
        else:
            self._stop([req], CANCELLED, "cancelled while running")
        return True

    def _place(self, req):
        """First-fit range for `req`, growing the pools (up to the budget's
        capacity) when no free range is long enough."""
        off = self.pool.alloc(req.n)
        while off is None and self.ex.capacity < self.max_capacity:
            if not self.ex.ensure(min(self.max_capacity, max(2 * self.ex.capacity, self.ex.capacity + req.n))):
                break
            self.pool.grow(self.ex.capacity)
            off = self.pool.alloc(req.n)
        return off

    def _admit(self, now):
        while self.queue and self.free_slots:
            if self.max_active is not None and len(self.active) >= self.max_active:
                break
            req = self.queue[0]
            need = req.n * BYTES_PER_NEURON
            if self.used_bytes + need > self.budget:
                break
            off = self._place(req)
            if off is None:
                break
            self.queue.popleft()
            req.off, req.t0, req.admitted = off, self.step_idx, now
            req.index = self.free_slots.pop()
            req.status = ACTIVE
            self.used_bytes += need
            self.ex.init(req)
            self.active.append(req)

    def _stop(self, reqs, status, reason=None):
        results = self.ex.read(reqs)
        now = time.perf_counter()
        for r, (ev, sc, V) in zip(reqs, results):
            r.events, r.spikes = ev, sc
            r.steps_done = self.step_idx - r.t0
            if r.keep_state:
                r.V_final = V
            self.pool.release(r.off, r.n)
            self.used_bytes -= r.n * BYTES_PER_NEURON
            self.free_slots.append(r.index)
            self.active.remove(r)
            self._finish(r, status, reason, now)
        self.ex.changed()

    def _retire(self):
        finished = [r for r in self.active if self.step_idx - r.t0 >= r.steps]
        if finished:
            self._stop(finished, DONE)

    def _check_health(self):
        bad = self.ex.unhealthy(self.active)
        if bad:
            self._stop(bad, FAILED, f"non-finite state by step {self.step_idx}")

    def step(self):
        self.ex.step(self.active, self.step_idx)
        self.step_idx += 1
        if self.health_every and self.step_idx % self.health_every == 0:
            self._check_health()

    def serve(self, requests, time_limit=None, cancel_at=None, on_step=None):
        """Run the arrival process of `requests` (arrival in seconds from
        start) until every request is finished (done, rejected, cancelled or
        failed) or `time_limit` seconds pass.  `cancel_at`: {rid: step}
        cancels a request once the server reaches that step; `on_step(self)`
        is called after every step (monitoring).  Returns the finished
        requests."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        cancel_at = dict(cancel_at or {})
        t_start = time.perf_counter()
        since_sync = 0
        while pending or self.queue or self.active:
            if time_limit is not None and time.perf_counter() - t_start > time_limit:
                break
            now = time.perf_counter()
            while pending and pending[0].arrival <= now - t_start:
                r = pending.popleft()
                r.arrival = t_start + r.arrival
                self.submit(r)
            self._admit(now)
            for rid, at in list(cancel_at.items()):
                if self.step_idx >= at and self.cancel(rid):
                    del cancel_at[rid]
            if not self.active:
                if self.queue:
                    # nothing runs and the queue head still does not fit:
                    # it never will (the pools are empty)
                    head = self.queue.popleft()
                    self._finish(head, REJECTED, "does not fit the empty pools")
                    continue
                if pending:
                    time.sleep(max(0.0, min(0.001, pending[0].arrival - (time.perf_counter() - t_start))))
                continue
            self.step()
            if on_step is not None:
                on_step(self)
            since_sync += 1
            self._retire()
            if since_sync >= self.max_ahead:
                self.ex.sync()
                since_sync = 0
        self.ex.sync()
        self.wall = time.perf_counter() - t_start
        return self.done
