"""Batched E/I COBA networks: many independent networks advanced by one
launch per step (``bpx_net_step``), or, for networks of a few thousand
neurons, by one launch for many steps (``bpx_net_run``).

Each network is the ``EICOBA_PostAlign`` model of
examples/dynamics_simulation/COBA.py:58-109 (LifRefLTC or LifRef neurons,
Expon + COBA synapses, homogeneous weights, optional projection delays)
with either just-in-time connectivity (``bm.jitconn.event_mv_prob_homo`` with
``outdim_parallel=False``: every spiking neuron regenerates its targets from
LFSR88 streams, nothing stored) or a stored CSR.  Per step the kernel
updates all neurons (adding the conductance delivered by the spikes of the
step one delay earlier) and then delivers this step's spikes as integer
counts into the step they reach, which makes
every network's trajectory exact and independent of what else runs in the
batch.
"""
import ctypes
import math

import numpy as np
import torch

from .. import _lib
from ..math.jitconn import conn_len
from .coba import expon_coef

NET_JIT, NET_CSR = 0, 1
TILE = 256                                  # default neurons per block of the step launch
TILE_CHOICES = (64, 128, 256, 512, 1024)
_TUNED = {}                                 # (kind, networks, size class) -> tile chosen by autotune()
_TUNED_FUSED = {}                           # same key -> autotune() found the multi-step launch faster


class NetDesc(ctypes.Structure):
    """Mirror of bpx_net_desc (include/bpx.h)."""
    _fields_ = [("off", ctypes.c_int64), ("n_exc", ctypes.c_int32), ("n_inh", ctypes.c_int32),
                ("seed", ctypes.c_uint32 * 4), ("clen", ctypes.c_int32 * 4), ("div", (ctypes.c_uint32 * 4) * 4),
                ("w_exc", ctypes.c_float), ("w_inh", ctypes.c_float), ("I_ext", ctypes.c_float),
                ("t0_step", ctypes.c_int32), ("delay", ctypes.c_int32 * 4), ("indptr", ctypes.c_void_p),
                ("indices", ctypes.c_void_p)]


class NetParams(ctypes.Structure):
    """Mirror of bpx_net_params (include/bpx.h)."""
    _fields_ = [(k, ctypes.c_float) for k in ("V_rest", "V_th", "V_reset", "tau", "tau_ref", "R", "tau_e", "tau_i",
                                              "E_e", "E_i", "c_e", "c_i", "dt")] + [("ltc", ctypes.c_int32)]


def _register():
    L = _lib.lib()
    if not hasattr(L, "_bpx_net_registered"):
        L.bpx_net_prepare.argtypes = [ctypes.POINTER(NetDesc)]
        L.bpx_net_prepare.restype = None
        L.bpx_net_step.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.POINTER(NetParams), ctypes.c_int64, ctypes.c_void_p]
        L.bpx_net_step.restype = ctypes.c_int
        L.bpx_net_advance.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.bpx_net_advance.restype = ctypes.c_int
        L.bpx_net_run_max_neurons.argtypes = [ctypes.c_int]
        L.bpx_net_run_max_neurons.restype = ctypes.c_int64
        L.bpx_net_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(NetParams), ctypes.c_int64,
                                  ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.bpx_net_run.restype = ctypes.c_int
        L._bpx_net_registered = True
    return L


def bf16_round(x):
    """Round a float32 to the nearest bfloat16 value (ties to even)."""
    u = np.array([x], np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return float(r.astype(np.uint32).view(np.float32)[0])


def delay_steps(delay, dt):
    """Delay in steps of the four projections (E2E, E2I, I2E, I2I) from a
    delay in ms (one value or four; None for no delay): math.ceil(delay / dt)
    as VarDelay.register_entry computes it (delay.py:263-266), at least 1 --
    a projection without delay, or with delay 0, reads the pre's spikes of
    the previous step, as delay 1 does."""
    if delay is None:
        return [1, 1, 1, 1]
    vals = list(delay) if isinstance(delay, (list, tuple, np.ndarray)) else [delay] * 4
    if len(vals) != 4:
        raise ValueError("delay must be one value or four (E2E, E2I, I2E, I2I)")
    out = []
    for v in vals:
        if v is None:
            out.append(1)
            continue
        if v < 0:
            raise ValueError(f"delay must be non-negative, got {v}")
        out.append(max(int(math.ceil(float(v) / float(dt))), 1))
    return out


def combined_csr(csrs, n_exc, n_inh):
    """One row per neuron over network-local post ids from the four
    projection CSRs {name: (indices, indptr)} (E2E, E2I, I2E, I2I)."""
    rows = []
    for p in range(n_exc + n_inh):
        parts = []
        if p < n_exc:
            for name, off in (("E2E", 0), ("E2I", n_exc)):
                idx, ptr = csrs[name]
                parts.append(idx[ptr[p]:ptr[p + 1]].astype(np.int64) + off)
        else:
            q = p - n_exc
            for name, off in (("I2E", 0), ("I2I", n_exc)):
                idx, ptr = csrs[name]
                parts.append(idx[ptr[q]:ptr[q + 1]].astype(np.int64) + off)
        rows.append(np.concatenate(parts))
    indptr = np.zeros(n_exc + n_inh + 1, np.int64)
    np.cumsum([len(r) for r in rows], out=indptr[1:])
    return np.concatenate(rows).astype(np.int32), indptr.astype(np.int32)


def max_in_degree(indices, indptr, n_exc, n_inh):
    """Largest number of synapses one neuron receives from one population
    (E rows, then I rows) of a combined network CSR, duplicates counted and
    out-of-range targets (which delivery skips) ignored: the largest count a
    neuron can take in one step."""
    N = n_exc + n_inh
    best = 0
    for rows in (indices[indptr[0]:indptr[n_exc]], indices[indptr[n_exc]:indptr[N]]):
        rows = np.asarray(rows, np.int64)
        rows = rows[(rows >= 0) & (rows < N)]
        if rows.size:
            best = max(best, int(np.bincount(rows, minlength=N).max()))
    return best


class BatchedCOBA:
    """A batch of independent COBA networks sharing neuron parameters.

    ``specs``: list of dicts with keys ``n_exc``, ``n_inh``, ``seed``,
    ``w_exc``, ``w_inh``, ``I_ext`` and either ``in_degree`` (just-in-time
    connectivity with p = in_degree / N for all four projections) or ``csr``
    = {E2E, E2I, I2E, I2I: (indices, indptr)}, and optionally ``delay`` (ms;
    one value or four, see :func:`delay_steps`).  All specs must use the same
    kind.  ``bf16=True`` rounds the homogeneous weights to bfloat16.
    ``tile``: neurons per block of the step launch (64 ... 1024); by default
    the value :meth:`autotune` found for a batch of the same kind, count and
    size, else 256.
    """

    def __init__(self, specs, V_rest=-60.0, V_th=-50.0, V_reset=-60.0, tau=20.0, tau_ref=5.0, R=1.0, tau_e=5.0,
                 tau_i=10.0, E_e=0.0, E_i=-80.0, dt=0.1, ltc=True, bf16=False, record=False, V_init=None,
                 tile=None):
        self.dev = _lib.device()
        L = _register()
        self.kind = NET_CSR if "csr" in specs[0] else NET_JIT
        self.specs = specs
        n = len(specs)
        self.descs = (NetDesc * n)()
        self.sizes = []
        self._keep = []
        off = 0
        max_delay = 1
        in_deg = 0
        for k, sp in enumerate(specs):
            d = self.descs[k]
            ne, ni = int(sp["n_exc"]), int(sp["n_inh"])
            N = ne + ni
            d.off, d.n_exc, d.n_inh = off, ne, ni
            w_e, w_i = float(sp["w_exc"]), float(sp["w_inh"])
            if bf16:
                w_e, w_i = bf16_round(w_e), bf16_round(w_i)
            d.w_exc, d.w_inh, d.I_ext = w_e, w_i, float(sp["I_ext"])
            if self.kind == NET_JIT:
                p = float(sp["in_degree"]) / N
                c = conn_len(min(p, 1.0))
                base = int(sp["seed"])
                for q in range(4):
                    d.seed[q] = (base * 4 + q) & 0xFFFFFFFF
                    d.clen[q] = c
                in_deg = max(in_deg, ne, ni)    # each projection draws a pair at most once
            else:
                idx, ptr = combined_csr(sp["csr"], ne, ni)
                in_deg = max(in_deg, max_in_degree(idx, ptr, ne, ni))
                ti = torch.from_numpy(idx).to(self.dev)
                tp = torch.from_numpy(ptr).to(self.dev)
                self._keep += [ti, tp]
                d.indptr, d.indices = tp.data_ptr(), ti.data_ptr()
                for q in range(4):
                    d.clen[q] = 1
            for q, steps in enumerate(delay_steps(sp.get("delay"), dt)):
                d.delay[q] = steps
            max_delay = max(max_delay, *d.delay)
            L.bpx_net_prepare(ctypes.byref(d))
            self.sizes.append(N)
            off += N
        self.total = off
        self.n_buf = max_delay + 1             # count buffers: one per step of the longest delay, plus one
        # the multi-step launch (one block per network, counts in LDS) serves
        # batches whose networks and in-degrees fit it
        self.max_in_degree = in_deg
        self.fused = (max(self.sizes) <= L.bpx_net_run_max_neurons(self.n_buf) and in_deg <= 65535)
        self.d_descs = torch.frombuffer(bytearray(bytes(self.descs)), dtype=torch.uint8).to(self.dev)
        self.set_tile(tile if tile is not None else _TUNED.get(self._tune_key(), TILE))
        self.use_fused = self.fused and _TUNED_FUSED.get(self._tune_key(), False)
        self.params = NetParams(V_rest=V_rest, V_th=V_th, V_reset=V_reset, tau=tau, tau_ref=tau_ref, R=R,
                                tau_e=tau_e, tau_i=tau_i, E_e=E_e, E_i=E_i, c_e=float(expon_coef(tau_e, dt)),
                                c_i=float(expon_coef(tau_i, dt)), dt=dt, ltc=int(ltc))
        self.dt = np.float32(dt)
        if V_init is None:
            parts = []
            for sp, N in zip(specs, self.sizes):
                rng = np.random.Generator(np.random.PCG64(int(sp["seed"]) + 7))
                parts.append(rng.uniform(V_rest, V_th, N).astype(np.float32))
            V_init = np.concatenate(parts)
        self.V0 = np.asarray(V_init, np.float32)
        self.record = record
        self.reset()

    def _tune_key(self):
        # batches of one connectivity kind, network count and total size
        # (to a factor of two) share a tuned tile size
        return (self.kind, len(self.specs), int(self.total).bit_length())

    def set_tile(self, tile):
        """Neurons per block (= threads per block and LDS spike-list length)
        of the step launch.  Invalidates a captured graph."""
        if tile not in TILE_CHOICES:
            raise ValueError(f"tile must be one of {TILE_CHOICES}, got {tile}")
        tiles = []
        for k, N in enumerate(self.sizes):
            for s in range(0, N, tile):
                tiles += [k, s]
        self.tile = tile
        self.tiles = tiles
        self.d_tiles = torch.tensor(tiles, dtype=torch.int32, device=self.dev)
        self.n_tiles = len(tiles) // 2
        self.graph = None

    def autotune(self, candidates=TILE_CHOICES, steps=200, repeats=2):
        """Pick the tile size (neurons per block) with the shortest step
        kernel on this batch: each candidate runs `steps` steps `repeats`
        times from a snapshot of the state that is restored afterwards (the
        batch continues exactly where it was), and the step kernel's device
        time is read from the library's kernel timer (HIP events around each
        launch on its stream; host launch overhead is not counted, as under
        graph replay).  The window should span the run's regimes: from a
        fresh state the first steps carry the synchronised initial spikes
        (delivery-heavy, favouring larger tiles) and later steps the steady
        state.  The choice is kept for later batches of the same kind, count
        and size.  A batch that fits the multi-step launch also times it
        over the same window (key "fused"); when it is faster, ``capture``
        uses it by default.  Returns {tile or "fused": kernel microseconds
        per step}."""
        state = [self.V, self.t_last, self.g_exc, self.g_inh, self.counts, self.spikes, self.events,
                 self.spike_count, self.d_step]
        saved = [t.clone() for t in state]
        i0 = self.i
        times = {}
        for tile in candidates:
            self.set_tile(tile)
            best = float("inf")
            for _ in range(repeats):
                for t, s in zip(state, saved):
                    t.copy_(s)
                self.i = i0
                torch.cuda.synchronize()
                _lib.timer_reset()
                _lib.timer_enable(True)
                try:
                    for _ in range(steps):
                        self.step()
                    torch.cuda.synchronize()
                finally:
                    _lib.timer_enable(False)
                ms, launches = _lib.timer_read(_lib.KID["net"])
                best = min(best, ms * 1e3 / max(launches, 1))
            times[tile] = best
        if self.fused:
            # the multi-step launch over the same window, per step
            best = float("inf")
            for _ in range(repeats):
                for t, s in zip(state, saved):
                    t.copy_(s)
                self.i = i0
                torch.cuda.synchronize()
                _lib.timer_reset()
                _lib.timer_enable(True)
                try:
                    self.run_fused(steps)
                    torch.cuda.synchronize()
                finally:
                    _lib.timer_enable(False)
                ms, _ = _lib.timer_read(_lib.KID["net_run"])
                best = min(best, ms * 1e3 / steps)
            times["fused"] = best
        for t, s in zip(state, saved):
            t.copy_(s)
        self.i = i0
        choice = min((k for k in times if k != "fused"), key=times.get)
        self.set_tile(choice)
        self.use_fused = self.fused and times["fused"] < times[choice]
        _TUNED[self._tune_key()] = choice
        _TUNED_FUSED[self._tune_key()] = self.use_fused
        self.tune_times = times
        return times

    def reset(self):
        """Initial state (in place, so a captured graph stays valid)."""
        d = self.dev
        T = self.total
        if getattr(self, "V", None) is None:
            self.V = torch.empty(T, dtype=torch.float32, device=d)
            self.t_last = torch.empty(T, dtype=torch.float32, device=d)
            self.g_exc = torch.empty(T, dtype=torch.float32, device=d)
            self.g_inh = torch.empty(T, dtype=torch.float32, device=d)
            self.counts = torch.empty(self.n_buf * 2 * T, dtype=torch.int32, device=d)
            self.spikes = torch.empty(T, dtype=torch.uint8, device=d)
            self.events = torch.empty(len(self.specs), dtype=torch.int64, device=d)
            self.spike_count = torch.empty(len(self.specs), dtype=torch.int32, device=d)
            self.d_step = torch.empty(1, dtype=torch.int64, device=d)      # base step of graph replays
            self.graph = None
        self.V.copy_(torch.from_numpy(self.V0))
        self.t_last.fill_(-1e7)
        self.g_exc.zero_()
        self.g_inh.zero_()
        self.counts.zero_()
        self.spikes.zero_()
        self.events.zero_()
        self.spike_count.zero_()
        self.d_step.zero_()
        self.i = 0

    def step(self, spikes_out=None, count=True, device_offset=None):
        """One step of every network.  `spikes_out` (uint8 [total]) receives
        this step's spike flags.  With `device_offset` k the step index is the
        device base step + k, as graph replays need."""
        L = _lib.lib()
        sp = spikes_out if spikes_out is not None else (self.spikes if self.record else None)
        s = _lib.stream_ptr()
        dev = device_offset is not None
        rc = L.bpx_net_step(s, _lib.ptr(self.d_descs), _lib.ptr(self.d_tiles), self.n_tiles, self.tile, self.kind,
                            self.total, _lib.ptr(self.V), _lib.ptr(self.t_last), _lib.ptr(self.g_exc),
                            _lib.ptr(self.g_inh), _lib.ptr(self.counts), self.n_buf, _lib.ptr(sp),
                            _lib.ptr(self.events) if count else None, _lib.ptr(self.spike_count) if count else None,
                            ctypes.byref(self.params), device_offset if dev else self.i,
                            _lib.ptr(self.d_step) if dev else None)
        _lib.check(rc, "bpx_net_step")
        self.i += 1

    def run_fused(self, steps, raster=None, device_offset=None):
        """Advance `steps` steps in one launch (``bpx_net_run``: one block per
        network); same results as `steps` calls of :meth:`step`.  `raster`
        (uint8 [steps, total]) receives every step's spike flags; without it
        a recording batch keeps the last step's flags in ``spikes``, as
        :meth:`step` does."""
        if not self.fused:
            raise ValueError("the multi-step launch needs networks of at most "
                             f"{_lib.lib().bpx_net_run_max_neurons(self.n_buf)} neurons and in-degrees below 65536")
        if raster is not None and (raster.dtype != torch.uint8 or not raster.is_contiguous()
                                   or tuple(raster.shape) != (steps, self.total)):
            raise ValueError(f"raster must be a contiguous uint8 tensor of shape ({steps}, {self.total})")
        L = _lib.lib()
        dev = device_offset is not None
        rc = L.bpx_net_run(_lib.stream_ptr(), _lib.ptr(self.d_descs), len(self.specs), max(self.sizes),
                           self.max_in_degree, self.kind, self.total, _lib.ptr(self.V), _lib.ptr(self.t_last),
                           _lib.ptr(self.g_exc), _lib.ptr(self.g_inh), _lib.ptr(self.counts), self.n_buf,
                           _lib.ptr(raster), _lib.ptr(self.spikes) if (self.record and raster is None) else None, _lib.ptr(self.events),
                           _lib.ptr(self.spike_count), ctypes.byref(self.params),
                           device_offset if dev else self.i, _lib.ptr(self.d_step) if dev else None, steps, None)
        _lib.check(rc, "bpx_net_run")
        self.i += steps

    def run(self, steps, raster=False, fused=None):
        """Advance `steps` steps; with `raster` returns uint8 [steps, total].
        `fused` (default: when the batch fits) runs them as one launch."""
        out = torch.empty((steps, self.total), dtype=torch.uint8, device=self.dev) if raster else None
        if (self.fused if fused is None else fused) and steps > 0:
            self.run_fused(steps, out)
            return out
        for k in range(steps):
            self.step(out[k] if raster else None)
        return out

    def capture(self, steps, fused=None):
        """Capture `steps` steps in one HIP graph (one batched launch per step
        at offsets 0..steps-1 from the device base step, or with `fused` one
        multi-step launch, then one launch that advances the base by
        `steps`); replay() advances the batch by `steps` steps per call,
        continuing in time.  `fused` defaults to the autotuner's finding
        (per-step launches until it has run): replayed, the per-step
        launches of one network are as fast as the multi-step launch, whose
        single block per network is bound by the neuron update's arithmetic
        on one compute unit, while a batch of hundreds of small networks
        runs faster as multi-step launches."""
        fused = self.use_fused if fused is None else fused
        self.d_step.fill_(self.i)
        torch.cuda.synchronize()
        i0 = self.i
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                if fused:
                    self.run_fused(steps, device_offset=0)
                else:
                    for k in range(steps):
                        self.step(device_offset=k)
                _lib.check(_lib.lib().bpx_net_advance(_lib.stream_ptr(), _lib.ptr(self.d_step), steps),
                           "bpx_net_advance")
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        # capture records the launches without running them
        self.d_step.fill_(i0)
        self.i = i0
        self.graph = g
        self._graph_steps = steps
        return g

    def replay(self):
        self.graph.replay()
        self.i += self._graph_steps

    def net_slice(self, k):
        off = int(self.descs[k].off)
        return slice(off, off + self.sizes[k])


class PlainCOBA:
    """One jitconn COBA network stepped operator by operator, the way the
    reference's step executes: four ``bm.jitconn.event_mv_prob_homo``
    projections (outdim_parallel=False) into fresh outputs, ``g += out``,
    then the E and I population updates -- ~14 launches per step.  FP32
    weights unless ``bf16``.  Used as the per-network baseline of the
    batched path; its raster equals the batched one."""

    def __init__(self, spec, V_init=None, bf16=False, dt=0.1, ltc=True, V_rest=-60.0, V_th=-50.0, V_reset=-60.0,
                 tau=20.0, tau_ref=5.0, R=1.0, tau_e=5.0, tau_i=10.0, E_e=0.0, E_i=-80.0):
        self.dev = _lib.device()
        ne, ni = int(spec["n_exc"]), int(spec["n_inh"])
        self.ne, self.ni = ne, ni
        N = ne + ni
        w_e, w_i = float(spec["w_exc"]), float(spec["w_inh"])
        if bf16:
            w_e, w_i = bf16_round(w_e), bf16_round(w_i)
        self.clen = conn_len(min(float(spec["in_degree"]) / N, 1.0))
        base = int(spec["seed"])
        self.projs = [(0, "E", "E", 0, w_e), (1, "E", "I", 0, w_e), (2, "I", "E", 1, w_i), (3, "I", "I", 1, w_i)]
        self.seeds = [(base * 4 + q) & 0xFFFFFFFF for q in range(4)]
        self.I_ext = float(spec["I_ext"])
        if V_init is None:
            rng = np.random.Generator(np.random.PCG64(base + 7))
            V_init = rng.uniform(V_rest, V_th, N).astype(np.float32)
        V_init = np.asarray(V_init, np.float32)
        d = self.dev
        self.size = {"E": ne, "I": ni}
        self.V = {"E": torch.from_numpy(V_init[:ne].copy()).to(d), "I": torch.from_numpy(V_init[ne:].copy()).to(d)}
        self.t_last = {p: torch.full((n,), -1e7, dtype=torch.float32, device=d) for p, n in self.size.items()}
        self.spike = {p: torch.zeros(n, dtype=torch.uint8, device=d) for p, n in self.size.items()}
        self.g = {p: [torch.zeros(n, dtype=torch.float32, device=d) for _ in range(2)] for p, n in self.size.items()}
        self.out = {p: torch.empty(n, dtype=torch.float32, device=d) for p, n in self.size.items()}
        self.tau_g = (ctypes.c_float * 2)(tau_e, tau_i)
        self.E_g = (ctypes.c_float * 2)(E_e, E_i)
        self.c_g = (ctypes.c_float * 2)(float(expon_coef(tau_e, dt)), float(expon_coef(tau_i, dt)))
        self.params = (V_rest, V_th, V_reset, tau, tau_ref, R)
        self.dt = np.float32(dt)
        self.ltc = int(ltc)
        self.i = 0

    def _launch_step(self, local=None):
        """The launches of one step: `local` None takes the time from the
        host (t = i dt), otherwise from the device step counter plus
        `local` (for captured graphs)."""
        L = _lib.lib()
        s = _lib.stream_ptr()
        for q, pre, post, kind, w in self.projs:
            out = self.out[post]
            rc = L.bpx_jitconn_mv(s, _lib.ptr(self.spike[pre]), _lib.BPX_BOOL, 1, _lib.DIST_HOMO, w, 0.0, self.clen,
                                  self.seeds[q], self.size[post], self.size[pre], 0, _lib.LAYOUT_LANE32,
                                  _lib.ptr(out), None, 0)
            _lib.check(rc, "bpx_jitconn_mv")
            self.g[post][kind].add_(out)
        for p in ("E", "I"):
            gp = (ctypes.c_void_p * 2)(self.g[p][0].data_ptr(), self.g[p][1].data_ptr())
            args = (s, self.size[p], self.ltc, 2, _lib.ptr(self.V[p]), _lib.ptr(self.t_last[p]),
                    _lib.ptr(self.spike[p]), ctypes.cast(gp, ctypes.POINTER(ctypes.c_void_p)),
                    ctypes.cast(self.tau_g, ctypes.c_void_p), ctypes.cast(self.E_g, ctypes.c_void_p),
                    ctypes.cast(self.c_g, ctypes.c_void_p), self.I_ext, None, *self.params)
            if local is None:
                rc = L.bpx_lif_coba_step(*args, np.float32(self.i) * self.dt, self.dt)
            else:
                rc = L.bpx_lif_coba_step_dev(*args, _lib.ptr(self.d_step), local, self.dt)
            _lib.check(rc, "bpx_lif_coba_step")

    def step(self):
        self._launch_step()
        self.i += 1

    def capture(self, steps):
        """Record `steps` steps (about 14 launches each) in one HIP graph --
        the counterpart of the reference's compiled scan over steps, with
        the step index on the device; replay() advances `steps` steps."""
        _register()
        self.d_step = torch.full((1,), self.i, dtype=torch.int64, device=self.dev)
        g = torch.cuda.CUDAGraph()
        cs = torch.cuda.Stream()
        cs.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(cs):
            with torch.cuda.graph(g, stream=cs):
                for k in range(steps):
                    self._launch_step(local=k)
                _lib.check(_lib.lib().bpx_net_advance(_lib.stream_ptr(), _lib.ptr(self.d_step), steps),
                           "bpx_net_advance")
        torch.cuda.current_stream().wait_stream(cs)
        self.graph, self.graph_steps = g, steps

    def replay(self):
        self.graph.replay()
        self.i += self.graph_steps

    def run(self, steps, raster=False):
        out = torch.empty((steps, self.ne + self.ni), dtype=torch.uint8, device=self.dev) if raster else None
        for k in range(steps):
            self.step()
            if raster:
                out[k, :self.ne] = self.spike["E"]
                out[k, self.ne:] = self.spike["I"]
        return out
