"""Synaptic projection layers over the event-driven operators.

The communication layers of dnn/linear.py that the reference's models use
to turn pre-synaptic activity into post-synaptic input: ``CSRLinear`` /
``EventCSRLinear`` (:546-647) over ``bm.sparse.csrmv`` / ``bm.event.csrmv``
with ``shape=(pre_num, post_num)``, and the just-in-time connectivity layers
``JitFP{Homo,Uniform,Normal}Linear`` (:805-1041) and
``EventJitFP{Homo,Uniform,Normal}Linear`` (:1044-1280) over
``bm.jitconn.(event_)mv_prob_*`` with ``shape=(num_out, num_in)`` and
``outdim_parallel = not atomic``; the STDP weight updates of the same file
(``csr_on_pre_update``, ``csc_on_post_update``, ``dense_on_pre/post``,
``stdp_update``).  Inputs with leading batch dimensions are
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


def _bounds(w_min, w_max):
    return (-np.inf if w_min is None else float(w_min)), (np.inf if w_max is None else float(w_max))


def _plast_args(w, spike, trace, out):
    w = _lib.to_device(w)
    if w.dtype not in (torch.float32, torch.float64):
        w = w.to(torch.float32)
    sp = _lib.to_device(spike)
    if sp.dtype not in (torch.bool, torch.float32, torch.float64):
        sp = sp.to(torch.float32)
    tr = _lib.to_device(trace).to(w.dtype).contiguous()
    if out is None:
        out = torch.empty_like(w)
    elif out.dtype != w.dtype or out.shape != w.shape or not out.is_contiguous():
        raise ValueError("out must be a contiguous tensor of the weights' shape and dtype")
    return w, sp, tr, out


def csr_on_pre_update(w, indices, indptr, spike, trace, w_min=None, w_max=None, *, out=None):
    """STDP on pre-synaptic spikes over a CSR: for every spiking pre neuron i
    and synapse k of row i, w[k] = min(max(w[k] + trace[indices[k]], w_min),
    w_max) (dnn/linear.py:650-676).  Returns the new weights (`out`, which
    may be `w` itself for an in-place update)."""
    w, sp, tr, out = _plast_args(w, spike, trace, out)
    idx = _lib.to_device(indices)
    ptr = _lib.to_device(indptr).to(idx.dtype)
    if idx.dtype not in (torch.int32, torch.int64):
        idx, ptr = idx.to(torch.int32), ptr.to(torch.int32)
    if w.numel() != idx.numel():
        raise ValueError(f'The shape of weight should be the same as the shape of sparse weight {tuple(idx.shape)}.')
    if ptr.numel() != sp.numel() + 1:
        raise ValueError(f'indptr must have {sp.numel() + 1} entries (one per pre neuron + 1).')
    lo, hi = _bounds(w_min, w_max)
    rc = _lib.lib().bpx_csr_on_pre_update(_lib.stream_ptr(), _lib.ptr(w), _lib.bpx_dtype(w), _lib.ptr(idx),
                                          _lib.ptr(ptr), _lib.bpx_dtype(idx), idx.numel(), _lib.ptr(sp),
                                          _lib.bpx_dtype(sp), sp.numel(), _lib.ptr(tr), tr.numel(), lo, hi,
                                          _lib.ptr(out))
    _lib.check(rc, "bpx_csr_on_pre_update")
    _lib.written(out)
    return out


def csc_on_post_update(w, post_ids, indptr, w_ids, spike, trace, w_min=None, w_max=None, *, out=None):
    """STDP on post-synaptic spikes over the CSC view of a CSR (from
    ``brainpy_amd.math.event.csr_to_csc`` with data = arange(nnz)): for every
    spiking post neuron j and entry k of column j, w[w_ids[k]] is moved by
    trace[post_ids[k]] (the pre id) and clipped (dnn/linear.py:679-700)."""
    w, sp, tr, out = _plast_args(w, spike, trace, out)
    pre = _lib.to_device(post_ids)
    if pre.dtype not in (torch.int32, torch.int64):
        pre = pre.to(torch.int32)
    cp = _lib.to_device(indptr).to(pre.dtype)
    wid = _lib.to_device(w_ids).to(pre.dtype)
    if w.numel() != pre.numel() or wid.numel() != pre.numel():
        raise ValueError('w, post_ids and w_ids must have one entry per synapse.')
    if cp.numel() != sp.numel() + 1:
        raise ValueError(f'indptr must have {sp.numel() + 1} entries (one per post neuron + 1).')
    lo, hi = _bounds(w_min, w_max)
    rc = _lib.lib().bpx_csc_on_post_update(_lib.stream_ptr(), _lib.ptr(w), _lib.bpx_dtype(w), _lib.ptr(pre),
                                           _lib.ptr(cp), _lib.ptr(wid), _lib.bpx_dtype(pre), pre.numel(),
                                           _lib.ptr(sp), _lib.bpx_dtype(sp), sp.numel(), _lib.ptr(tr), tr.numel(),
                                           lo, hi, _lib.ptr(out))
    _lib.check(rc, "bpx_csc_on_post_update")
    _lib.written(out)
    return out


def _dense_on_spike(fn, weight, spike, trace, w_min, w_max, out, rows):
    w, sp, tr, out = _plast_args(weight, spike, trace, out)
    if w.ndim != 2:
        raise ValueError('weight must be a 2D [num_pre, num_post] matrix.')
    npre, npost = w.shape
    if sp.numel() != (npre if rows else npost) or tr.numel() != (npost if rows else npre):
        raise ValueError(f'spike / trace sizes do not match the weight shape {tuple(w.shape)}.')
    lo, hi = _bounds(w_min, w_max)
    rc = getattr(_lib.lib(), fn)(_lib.stream_ptr(), _lib.ptr(w), _lib.bpx_dtype(w), _lib.ptr(sp), _lib.bpx_dtype(sp),
                                 npre, npost, _lib.ptr(tr), lo, hi, _lib.ptr(out))
    _lib.check(rc, fn)
    return out


def dense_on_pre(weight, spike, trace, w_min, w_max, *, out=None):
    """Rows of spiking pre neurons += trace (over post), clipped
    (dnn/linear.py:238-254)."""
    return _dense_on_spike("bpx_dense_on_pre", weight, spike, trace, w_min, w_max, out, True)


def dense_on_post(weight, spike, trace, w_min, w_max, *, out=None):
    """Columns of spiking post neurons += trace (over pre), clipped
    (dnn/linear.py:257-275)."""
    return _dense_on_spike("bpx_dense_on_post", weight, spike, trace, w_min, w_max, out, False)


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

    def stdp_update(self, on_pre=None, on_post=None, w_min=None, w_max=None):
        """Event-driven STDP on the layer's weights (dnn/linear.py:517-543):
        `on_pre` / `on_post` = {'spike': ..., 'trace': ...}; the CSC view for
        post-synaptic updates is built once."""
        if not isinstance(self.weight, torch.Tensor):
            raise ValueError('When using STDP to update synaptic weights, the weight cannot be a scalar.')
        if tuple(self.weight.shape) != tuple(self.indices.shape):
            raise ValueError(f'The shape of weight should be the same as the shape of sparse weight '
                             f'{tuple(self.weight.shape)}.')
        if on_pre is not None:
            self.weight = csr_on_pre_update(self.weight, self.indices, self.indptr, on_pre['spike'],
                                            on_pre['trace'], w_min, w_max, out=self.weight)
        if on_post is not None:
            if not hasattr(self, '_pre_ids'):
                rows, colptr, order = _event.csr_to_csc(self.indices, self.indptr, self.shape,
                                                        data=np.arange(self.indices.numel()))
                dev = _lib.device()
                self._pre_ids = torch.from_numpy(rows).to(dev)
                self._post_indptr = torch.from_numpy(colptr).to(dev)
                self.w_indices = torch.from_numpy(np.asarray(order, np.int32)).to(dev)
            self.weight = csc_on_post_update(self.weight, self._pre_ids, self._post_indptr, self.w_indices,
                                             on_post['spike'], on_post['trace'], w_min, w_max, out=self.weight)

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
