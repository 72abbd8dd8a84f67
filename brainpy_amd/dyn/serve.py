"""Served simulation: continuous batching of COBA network requests on one GPU.

Requests (network size, seed, weights, input, number of steps) arrive over
time.  The server keeps every admitted network's state resident in device
pools and advances all of them together, one ``bpx_net_step`` launch per
step.  A request is admitted as soon as its state fits both the HBM budget
and a free range of the pools; it joins the batch at the next step with its
own time origin (``t0_step``), so its trajectory is the same as when run
alone; when its last step is done its results are read back and its slot is
freed for queued requests.

Per-request results: spike count, synaptic events delivered, final membrane
potentials (optional), arrival / admission / completion times.
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


class Request:
    __slots__ = ("rid", "n_exc", "n_inh", "seed", "w_exc", "w_inh", "I_ext", "in_degree", "steps", "arrival",
                 "admitted", "completed", "off", "t0", "spikes", "events", "V_final", "keep_state", "index")

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

    @property
    def n(self):
        return self.n_exc + self.n_inh


class _Pool:
    """First-fit allocator of contiguous neuron ranges."""

    def __init__(self, capacity):
        self.free = [(0, capacity)]

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


class SimServer:
    """Continuous-batching server of just-in-time-connectivity COBA networks.

    capacity: neurons the device pools hold; hbm_budget: bytes the resident
    state of admitted requests may use (admission control); bf16: round the
    homogeneous weights to bfloat16.
    """

    def __init__(self, capacity=1 << 24, hbm_budget=200 * (1 << 30), bf16=True, dt=0.1, ltc=True,
                 V_rest=-60.0, V_th=-50.0, V_reset=-60.0, tau=20.0, tau_ref=5.0, R=1.0, tau_e=5.0, tau_i=10.0,
                 E_e=0.0, E_i=-80.0, max_ahead=16, max_active=None):
        self.dev = _lib.device()
        self.L = _register()
        self.capacity = int(capacity)
        self.budget = int(hbm_budget)
        self.bf16 = bf16
        self.V_rest, self.V_th = V_rest, V_th
        self.params = NetParams(V_rest=V_rest, V_th=V_th, V_reset=V_reset, tau=tau, tau_ref=tau_ref, R=R,
                                tau_e=tau_e, tau_i=tau_i, E_e=E_e, E_i=E_i, c_e=float(expon_coef(tau_e, dt)),
                                c_i=float(expon_coef(tau_i, dt)), dt=dt, ltc=int(ltc))
        d, C = self.dev, self.capacity
        self.V = torch.zeros(C, dtype=torch.float32, device=d)
        self.t_last = torch.full((C,), -1e7, dtype=torch.float32, device=d)
        self.g_exc = torch.zeros(C, dtype=torch.float32, device=d)
        self.g_inh = torch.zeros(C, dtype=torch.float32, device=d)
        self.counts = torch.zeros(4 * C, dtype=torch.int32, device=d)
        self.pool = _Pool(C)
        self.used_bytes = 0
        self.queue = deque()
        self.active = []
        self.done = []
        self.step_idx = 0
        self.max_ahead = max_ahead
        self.max_active = max_active          # 1 = a plain FIFO queue (no batching)
        self._dirty = True
        self.max_slots = 4096
        self.events = torch.zeros(self.max_slots, dtype=torch.int64, device=d)
        self.spike_count = torch.zeros(self.max_slots, dtype=torch.int32, device=d)
        self.free_slots = list(range(self.max_slots - 1, -1, -1))
        self.slot_of = None           # device slot index of each network of the current batch

    # ------------------------------------------------------------------
    def submit(self, req):
        self.queue.append(req)

    def _admit(self, now):
        admitted = False
        while self.queue and self.free_slots:
            if self.max_active is not None and len(self.active) >= self.max_active:
                break
            req = self.queue[0]
            need = req.n * BYTES_PER_NEURON
            if self.used_bytes + need > self.budget:
                break
            off = self.pool.alloc(req.n)
            if off is None:
                break
            self.queue.popleft()
            req.off, req.t0, req.admitted = off, self.step_idx, now
            req.index = self.free_slots.pop()
            self.used_bytes += need
            sl = slice(off, off + req.n)
            rng = np.random.Generator(np.random.PCG64(req.seed + 7))
            self.V[sl] = torch.from_numpy(rng.uniform(self.V_rest, self.V_th, req.n).astype(np.float32)).to(self.dev)
            self.t_last[sl] = -1e7
            self.g_exc[sl] = 0.0
            self.g_inh[sl] = 0.0
            C = self.capacity
            for b in range(4):
                self.counts[b * C + off:b * C + off + req.n] = 0
            self.events[req.index] = 0
            self.spike_count[req.index] = 0
            self.active.append(req)
            admitted = True
        if admitted:
            self._dirty = True

    def _rebuild(self):
        n = len(self.active)
        descs = (NetDesc * max(n, 1))()
        tiles = []
        for k, req in enumerate(self.active):
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
        self.slot_of = torch.tensor([r.index for r in self.active] or [0], dtype=torch.int64, device=self.dev)
        self.ev_scratch = torch.zeros(max(n, 1), dtype=torch.int64, device=self.dev)
        self.sc_scratch = torch.zeros(max(n, 1), dtype=torch.int32, device=self.dev)
        self._dirty = False

    def _fold_counters(self):
        if self.slot_of is not None:
            self.events.index_add_(0, self.slot_of, self.ev_scratch)
            self.spike_count.index_add_(0, self.slot_of, self.sc_scratch)
            self.ev_scratch.zero_()
            self.sc_scratch.zero_()

    def _retire(self, now):
        finished = [r for r in self.active if self.step_idx - r.t0 >= r.steps]
        if not finished:
            return
        self._fold_counters()
        torch.cuda.synchronize()
        now = time.perf_counter()
        ev = self.events.cpu().numpy()
        sc = self.spike_count.cpu().numpy()
        for r in finished:
            r.completed = now
            r.events = int(ev[r.index])
            r.spikes = int(sc[r.index])
            if r.keep_state:
                r.V_final = self.V[r.off:r.off + r.n].cpu().numpy()
            self.pool.release(r.off, r.n)
            self.used_bytes -= r.n * BYTES_PER_NEURON
            self.free_slots.append(r.index)
            self.active.remove(r)
            self.done.append(r)
        self._dirty = True

    def step(self):
        if self._dirty:
            self._fold_counters()
            self._rebuild()
        if self.n_tiles:
            rc = self.L.bpx_net_step(_lib.stream_ptr(), _lib.ptr(self.d_descs), _lib.ptr(self.d_tiles), self.n_tiles,
                                     TILE, 0, self.capacity, _lib.ptr(self.V), _lib.ptr(self.t_last), _lib.ptr(self.g_exc),
                                     _lib.ptr(self.g_inh), _lib.ptr(self.counts), 2, None, _lib.ptr(self.ev_scratch),
                                     _lib.ptr(self.sc_scratch), ctypes.byref(self.params), self.step_idx, None)
            _lib.check(rc, "bpx_net_step")
        self.step_idx += 1

    def serve(self, requests, time_limit=None):
        """Run the arrival process of `requests` (sorted by `arrival`, in
        seconds from start) to completion.  Returns the completed requests."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t_start = time.perf_counter()
        since_sync = 0
        while pending or self.queue or self.active:
            now = time.perf_counter()
            while pending and pending[0].arrival <= now - t_start:
                r = pending.popleft()
                r.arrival = t_start + r.arrival
                self.submit(r)
            self._admit(now)
            if not self.active:
                if pending:
                    time.sleep(max(0.0, min(0.001, pending[0].arrival - (time.perf_counter() - t_start))))
                continue
            self.step()
            since_sync += 1
            self._retire(now)
            if since_sync >= self.max_ahead:
                torch.cuda.synchronize()
                since_sync = 0
            if time_limit is not None and time.perf_counter() - t_start > time_limit:
                break
        torch.cuda.synchronize()
        self.wall = time.perf_counter() - t_start
        return self.done
