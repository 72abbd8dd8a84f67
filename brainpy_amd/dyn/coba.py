"""E/I COBA network of LIF neurons on the gfx950 library.

Mirrors ``EICOBA_PostAlign`` of examples/dynamics_simulation/COBA.py:58-109
(LifRefLTC or LifRef neurons, FullProjAlignPostMg projections with
EventCSRLinear + Expon + COBA, delay None).  Step i, in the reference order
(dyn/projections/align_post.py:271-275, dynsys.py:370-383):

1. t = f32(i) * f32(dt)                                 (runners.py:624)
2. projections E2I, I2I, I2E, E2E:
       g_post += event_csrmv(w, spikes_pre(step i-1), transpose=True)
3. population E then I: Expon decay of each g (registration order:
   excitatory then inhibitory), then the LIF exp_euler update
   (one fused kernel, bpx_lif_coba_step).

``mode='fixed_order'`` accumulates every conductance in ascending
pre-synaptic order (CSC gather), bit-identical to the serial CPU scatter;
``mode='atomic'`` uses the event-driven scatter with FP32 atomics.
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from ..connect import fixed_prob_csr
from ..math import event as ev_ops


def expon_coef(tau, dt):
    """c = dt * exprel(dt * (-(1/tau))) in float32: the Expon exp_euler factor
    of `g + c*(-g/tau)` (integrators/ode/exponential.py:357-362), computed by
    the library's host code with its deterministic exp."""
    return np.float32(_lib.lib().bpx_expon_coef(np.float32(tau), np.float32(dt)))


class EICOBANet:
    """E/I COBA LIF network (default: the 4000-neuron benchmark network,
    3200 E + 800 I, FixedProb p=0.02, we=0.6, wi=6.7)."""

    def __init__(self, n_exc=3200, n_inh=800, prob=0.02, we=0.6, wi=6.7, tau_e=5.0, tau_i=10.0, E_e=0.0,
                 E_i=-80.0, V_rest=-60.0, V_th=-50.0, V_reset=-60.0, tau=20.0, tau_ref=5.0, R=1.0, inp=20.0,
                 dt=0.1, seed=0, ltc=True, mode="fixed_order", V_init=None):
        self.dev = _lib.device()
        self.n_exc, self.n_inh = n_exc, n_inh
        self.dt = np.float32(dt)
        self.ltc = bool(ltc)
        self.mode = mode
        self.inp = float(inp)
        self.params = tuple(np.float32(x) for x in (V_rest, V_th, V_reset, tau, tau_ref, R))
        # projections in call order: (name, pre, post, weight, synapse group in post)
        self.proj_spec = [("E2I", "E", "I", we, 0), ("I2I", "I", "I", wi, 1),
                          ("I2E", "I", "E", wi, 1), ("E2E", "E", "E", we, 0)]
        size = {"E": n_exc, "I": n_inh}
        self.csr = {}
        for k, (name, pre, post, w, _) in enumerate(self.proj_spec):
            idx, ptr = fixed_prob_csr(size[pre], size[post], prob, seed * 7919 + k + 1)
            self.csr[name] = (np.float32(w), idx, ptr, (size[pre], size[post]))
        self.tau_g = [np.float32(tau_e), np.float32(tau_i)]
        self.E_g = [np.float32(E_e), np.float32(E_i)]
        self.c_g = [expon_coef(tau_e, dt), expon_coef(tau_i, dt)]
        rng = np.random.Generator(np.random.PCG64(seed))
        if V_init is None:
            V_init = rng.uniform(V_rest, V_th, n_exc + n_inh).astype(np.float32)
        self.V0 = np.asarray(V_init, np.float32)
        self._build_device()
        self.reset()

    def _build_device(self):
        self.dcsr = {}
        for name, (w, idx, ptr, shape) in self.csr.items():
            wt = torch.tensor([w], dtype=torch.float32, device=self.dev)
            if self.mode == "fixed_order":
                rows, colptr, _ = ev_ops.csr_to_csc(idx, ptr, shape)
                self.dcsr[name] = (wt, torch.from_numpy(rows).to(self.dev), torch.from_numpy(colptr).to(self.dev),
                                   shape)
            else:
                self.dcsr[name] = (wt, torch.from_numpy(idx).to(self.dev), torch.from_numpy(ptr).to(self.dev), shape)

    def reset(self):
        d = self.dev
        self.V = {"E": torch.from_numpy(self.V0[:self.n_exc].copy()).to(d),
                  "I": torch.from_numpy(self.V0[self.n_exc:].copy()).to(d)}
        self.t_last = {p: torch.full((n,), -1e7, dtype=torch.float32, device=d)
                       for p, n in (("E", self.n_exc), ("I", self.n_inh))}
        self.spike = {p: torch.zeros(n, dtype=torch.uint8, device=d) for p, n in (("E", self.n_exc), ("I", self.n_inh))}
        self.g = {p: [torch.zeros(n, dtype=torch.float32, device=d) for _ in range(2)]
                  for p, n in (("E", self.n_exc), ("I", self.n_inh))}
        self.i = 0

    def _project(self):
        for name, pre, post, _, k in self.proj_spec:
            w, a, b, shape = self.dcsr[name]
            spikes = self.spike[pre].view(torch.bool)
            if self.mode == "fixed_order":
                ev_ops.csrmv_fixed_order(w, a, b, spikes, shape=shape, out=self.g[post][k])
            else:
                ev_ops.csrmv_add(self.g[post][k], w, a, b, spikes, shape=shape)

    def _neurons(self, t):
        L = _lib.lib()
        s = _lib.stream_ptr()
        tau_g = (ctypes.c_float * 2)(*self.tau_g)
        E_g = (ctypes.c_float * 2)(*self.E_g)
        c_g = (ctypes.c_float * 2)(*self.c_g)
        for p in ("E", "I"):
            gp = (ctypes.c_void_p * 2)(self.g[p][0].data_ptr(), self.g[p][1].data_ptr())
            rc = L.bpx_lif_coba_step(s, self.V[p].numel(), int(self.ltc), 2, _lib.ptr(self.V[p]),
                                     _lib.ptr(self.t_last[p]), _lib.ptr(self.spike[p]),
                                     ctypes.cast(gp, ctypes.POINTER(ctypes.c_void_p)),
                                     ctypes.cast(tau_g, ctypes.c_void_p), ctypes.cast(E_g, ctypes.c_void_p),
                                     ctypes.cast(c_g, ctypes.c_void_p), self.inp, None, *self.params, t, self.dt)
            _lib.check(rc, "bpx_lif_coba_step")

    def step(self):
        t = np.float32(self.i) * self.dt
        self._project()
        self._neurons(t)
        self.i += 1

    def run(self, steps, record=True):
        """Advance `steps` steps; returns the spike raster [steps, N] (uint8
        torch tensor on the GPU) when `record`."""
        raster = torch.empty((steps, self.n_exc + self.n_inh), dtype=torch.uint8, device=self.dev) if record else None
        for k in range(steps):
            self.step()
            if record:
                raster[k, :self.n_exc] = self.spike["E"]
                raster[k, self.n_exc:] = self.spike["I"]
        return raster

    def synaptic_events(self, raster):
        """Number of (active pre spike, target) pairs delivered by a raster
        (spikes of step i are delivered at step i+1)."""
        r = raster.to(torch.int64)
        ke = int(self.csr["E2E"][1].size // self.n_exc) + int(self.csr["E2I"][1].size // self.n_exc)
        ki = int(self.csr["I2E"][1].size // self.n_inh) + int(self.csr["I2I"][1].size // self.n_inh)
        return int(r[:-1, :self.n_exc].sum()) * ke + int(r[:-1, self.n_exc:].sum()) * ki
