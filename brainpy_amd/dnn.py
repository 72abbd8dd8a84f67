"""Synaptic projection layers over the event-driven operators.

The communication layers of dnn/linear.py that the reference's models use
to turn pre-synaptic activity into post-synaptic input: ``CSRLinear`` /
``EventCSRLinear`` (:546-647) over ``bm.sparse.csrmv`` / ``bm.event.csrmv``
with ``shape=(pre_num, post_num)``, and the just-in-time connectivity layers
``JitFP{Homo,Uniform,Normal}Linear`` (:805-1041) and
``EventJitFP{Homo,Uniform,Normal}Linear`` (:1044-1280) over
``bm.jitconn.(event_)mv_prob_*`` with ``shape=(num_out, num_in)`` and
``outdim_parallel = not atomic``.  Inputs with leading batch dimensions are
flattened to rows and each row goes through the operator (the reference maps
the operator over rows with ``jax.vmap``, which its primitives run as a
sequential scan).
"""
import numpy as np
import torch

from . import _lib
from .math import event as _event
from .math import jitconn as _jit
from .math import sparse as _sparse


def _parameter(weight, size):
    """init.parameter(weight, (size,)) for the forms the layers accept: a
    scalar stays a scalar, an array must have `size` entries, a callable is
    called with the shape (initialize/generic.py)."""
    if callable(weight):
        weight = weight((size,))
    if isinstance(weight, (int, float, np.floating, np.integer)):
        return float(weight)
    t = torch.as_tensor(weight)
    if t.ndim == 0:
        return float(t)
    if t.numel() != size:
        raise ValueError(f"The shape of weight should be ({size},), got {tuple(t.shape)}.")
    return t


def _rows(x, fn):
    """Apply `fn` to a 1-D input, or to each row of the input flattened to
    (-1, last) and reshape back."""
    if x.ndim == 1:
        return fn(x)
    if x.ndim == 0:
        raise ValueError("input must have at least one dimension")
    lead = tuple(x.shape[:-1])
    flat = x.reshape(-1, x.shape[-1])
    ys = [fn(flat[i]) for i in range(flat.shape[0])]
    y = torch.stack(ys) if ys else torch.empty((0,), device=_lib.device())
    return y.reshape(lead + (y.shape[-1],))


class _CSRLayer:
    def __init__(self, conn, weight, sharding=None, mode=None, name=None, transpose=True):
        if sharding is not None:
            raise AssertionError("Currently this model does not support sharding.")
        if not hasattr(conn, "require"):
            raise AssertionError("conn must be a two-end connector")
        self.conn = conn
        self.name = name
        self.mode = mode
        self.transpose = transpose
        indices, indptr = conn.require("csr")
        dev = _lib.device()
        self.indices = torch.as_tensor(indices).to(dev)
        self.indptr = torch.as_tensor(indptr).to(dev)
        w = _parameter(weight, int(self.indices.numel()))
        self.weight = w.to(dev) if isinstance(w, torch.Tensor) else w

    @property
    def shape(self):
        return (self.conn.pre_num, self.conn.post_num)

    def __call__(self, x):
        return self.update(x)


class CSRLinear(_CSRLayer):
    """y = x @ M with M a CSR matrix (dnn/linear.py:546-597)."""

    def __init__(self, conn, weight, sharding=None, mode=None, name=None, method="cusparse", transpose=True):
        super().__init__(conn, weight, sharding=sharding, mode=mode, name=name, transpose=transpose)
        self.method = method

    def update(self, x):
        x = torch.as_tensor(x, device=_lib.device())
        return _rows(x, lambda v: _sparse.csrmv(self.weight, self.indices, self.indptr, v, shape=self.shape,
                                                transpose=self.transpose, method=self.method))


class EventCSRLinear(_CSRLayer):
    """y = spikes @ M with M a CSR matrix (dnn/linear.py:600-647)."""

    def update(self, x):
        x = torch.as_tensor(x, device=_lib.device())
        return _rows(x, lambda v: _event.csrmv(self.weight, self.indices, self.indptr, v, shape=self.shape,
                                               transpose=self.transpose))


class _JitLayer:
    _seed_high = 100000

    def __init__(self, num_in, num_out, prob, seed=None, sharding=None, mode=None, name=None, transpose=False,
                 atomic=False):
        self.prob = prob
        self.sharding = sharding
        self.transpose = transpose
        self.seed = np.random.randint(0, self._seed_high) if seed is None else seed
        self.atomic = atomic
        self.num_in = num_in
        self.num_out = num_out
        self.mode = mode
        self.name = name

    def _kw(self):
        return dict(shape=(self.num_out, self.num_in), transpose=self.transpose, outdim_parallel=not self.atomic)

    def update(self, x):
        x = torch.as_tensor(x, device=_lib.device())
        return _rows(x, self._mv)

    def __call__(self, x):
        return self.update(x)


class JitFPHomoLinear(_JitLayer):
    """y = x @ M, M with just-in-time connectivity of probability `prob` and
    one weight (dnn/linear.py:805-881)."""

    def __init__(self, num_in, num_out, prob, weight, seed=None, sharding=None, mode=None, name=None,
                 transpose=False, atomic=False):
        super().__init__(num_in, num_out, prob, seed, sharding, mode, name, transpose, atomic)
        self.weight = weight

    def _mv(self, v):
        return _jit.mv_prob_homo(v, self.weight, self.prob, self.seed, **self._kw())


class JitFPUniformLinear(_JitLayer):
    """Uniform weights in [w_low, w_high) (dnn/linear.py:884-961)."""

    def __init__(self, num_in, num_out, prob, w_low, w_high, seed=None, sharding=None, mode=None, name=None,
                 transpose=False, atomic=False):
        super().__init__(num_in, num_out, prob, seed, sharding, mode, name, transpose, atomic)
        self.w_low, self.w_high = w_low, w_high

    def _mv(self, v):
        return _jit.mv_prob_uniform(v, self.w_low, self.w_high, self.prob, self.seed, **self._kw())


class JitFPNormalLinear(_JitLayer):
    """Normal weights N(w_mu, w_sigma) (dnn/linear.py:964-1041)."""

    def __init__(self, num_in, num_out, prob, w_mu, w_sigma, seed=None, sharding=None, mode=None, name=None,
                 transpose=False, atomic=False):
        super().__init__(num_in, num_out, prob, seed, sharding, mode, name, transpose, atomic)
        self.w_mu, self.w_sigma = w_mu, w_sigma

    def _mv(self, v):
        return _jit.mv_prob_normal(v, self.w_mu, self.w_sigma, self.prob, self.seed, **self._kw())


class EventJitFPHomoLinear(JitFPHomoLinear):
    """Spike input, one weight (dnn/linear.py:1044-1120)."""
    _seed_high = 1000000

    def _mv(self, v):
        return _jit.event_mv_prob_homo(v, self.weight, self.prob, self.seed, **self._kw())


class EventJitFPUniformLinear(JitFPUniformLinear):
    """Spike input, uniform weights (dnn/linear.py:1123-1200)."""
    _seed_high = 100000

    def _mv(self, v):
        return _jit.event_mv_prob_uniform(v, self.w_low, self.w_high, self.prob, self.seed, **self._kw())


class EventJitFPNormalLinear(JitFPNormalLinear):
    """Spike input, normal weights (dnn/linear.py:1203-1280)."""
    _seed_high = 100000

    def _mv(self, v):
        return _jit.event_mv_prob_normal(v, self.w_mu, self.w_sigma, self.prob, self.seed, **self._kw())
