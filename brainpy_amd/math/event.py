"""Event-driven operators: ``bm.event.csrmv``, ``bm.event.csrmv_taichi``,
``bm.event.info`` on the gfx950 library.

Signatures, validation and error messages follow
``brainpy/_src/math/event/_csr_matvec.py:38-110``,
``event/_csr_matvec_taichi.py:344-452`` and
``event/_info_collection.py:24-43``.  Inputs may be numpy arrays, python
scalars or torch tensors; results are torch tensors on the GPU.
"""
import collections
import ctypes
import threading
from typing import Tuple

import numpy as np
import torch

from .. import _lib
from ._common import acc_dtype, batch_rows, data_tensor, index_tensor, ndim, needs_ad

__all__ = ["csrmv", "csrmv_taichi", "info", "csrmv_fixed_order", "csr_to_csc", "csrmv_add", "TiledCSR",
           "plan_cache_info", "clear_plan_cache", "set_plan_cache", "csrmv_batch"]


def _check(data, indices, indptr, events, shape, transpose, allowed_index, index_msg):
    if ndim(data) == 1:
        if data.shape[0] not in [1, indices.shape[0]]:
            raise ValueError('The size of data should be 1 or be consistent with indices.'
                             f'But we got {tuple(data.shape)} != {tuple(indices.shape)}, {tuple(data.shape)} != 1.')
    else:
        raise ValueError('data should be a scalar or 1D vector. '
                         f'But we got {ndim(data)}-D array.')
    if ndim(indices) != 1:
        raise ValueError('indices should be a 1D vector with integer type.')
    if ndim(indptr) != 1:
        raise ValueError('indptr should be a 1D vector with integer type.')
    if indices.dtype not in allowed_index:
        raise ValueError(index_msg.format('indices'))
    if indptr.dtype not in allowed_index:
        raise ValueError(index_msg.format('indptr'))
    if ndim(events) != 1:
        raise ValueError('events should be a 1D vector.')
    if len(shape) != 2:
        raise ValueError('shape should be a length-2 tuple.')
    if transpose:
        if events.shape[0] != shape[0]:
            raise ValueError(f'Shape mismatch, vec ({events.shape[0]},) @ mat {shape}.')
    else:
        if events.shape[0] != shape[1]:
            raise ValueError(f'Shape mismatch, mat {shape} @ vec ({events.shape[0]},).')


_PUBLIC_INDEX = {torch.int32, torch.int64, np.dtype(np.int32), np.dtype(np.int64), np.dtype(np.uint32),
                 np.dtype(np.uint64)}
_TAICHI_INDEX = _PUBLIC_INDEX | {torch.int8, torch.int16, torch.uint8, np.dtype(np.int8), np.dtype(np.int16),
                                 np.dtype(np.uint8), np.dtype(np.uint16)}


def _as_array(x):
    """Arrays pass through; python scalars and sequences take JAX's default
    dtypes without x64 (float -> float32, int -> int32)."""
    if isinstance(x, (torch.Tensor, np.ndarray)):
        return x
    a = np.asarray(x)
    if a.dtype == np.float64:
        a = a.astype(np.float32)
    elif a.dtype == np.int64:
        a = a.astype(np.int32)
    return a


def _events_tensor(events):
    ev = _lib.to_device(events)
    if ev.dtype == torch.bool or ev.dtype in (torch.float32, torch.float64):
        return ev
    return ev.to(torch.float32)


def _launch(d, idx, ptr, ev, shape, transpose, rule, host_w=None, use_plan=True):
    # raw pointers go to the kernel: views are made dense first
    d, idx, ptr, ev = d.contiguous(), idx.contiguous(), ptr.contiguous(), ev.contiguous()
    m, n = int(shape[0]), int(shape[1])
    if ptr.shape[0] != m + 1:
        raise ValueError(f'indptr should have shape ({m + 1},), but we got {tuple(ptr.shape)}.')
    if use_plan:
        # transpose: the plan of A scatters the active rows; the gather form
        # runs on the plan of A^T, whose active rows (the events' columns)
        # it scatters, reading only their entries instead of every row
        plan = _PLANS.lookup(d, idx, ptr, m, n, ev, host_w, transposed=not transpose)
        if plan is not None:
            return plan._exec(ev, rule, None)
    out = torch.empty(n if transpose else m, dtype=acc_dtype(d.dtype), device=ev.device)
    rc = _lib.lib().bpx_event_csrmv(_lib.stream_ptr(), _lib.ptr(d), _lib.bpx_dtype(d), d.shape[0], _lib.ptr(idx),
                                    _lib.ptr(ptr), _lib.bpx_dtype(idx), idx.shape[0], _lib.ptr(ev),
                                    _lib.bpx_dtype(ev), rule, m, n, int(bool(transpose)), _lib.ptr(out))
    _lib.check(rc, "bpx_event_csrmv")
    return out


class _PlanCache:
    """Tiled plans (:class:`TiledCSR`) of the matrices that ``csrmv`` (event
    or sparse, either orientation) is called with repeatedly; the event
    gather (transpose=False) uses a plan of the transposed matrix instead,
    a separate entry.

    The reference's layers call the operator statelessly every step with
    the same arrays (dnn/linear.py:630-634); the plan replaces one global
    atomic per synaptic event by LDS accumulation, so the second call on the
    same (data, indices, indptr) builds a plan and later calls run on it.
    A key is the storage address, dtype, length and version counter of each
    array (a homogeneous weight given on the host is keyed by its value):
    writing into an array in place bumps its version, so the next call
    builds a new plan from the new contents.  An entry holds references to
    its arrays, so their addresses cannot be reused by other tensors while
    it lives.  Least recently used entries are dropped beyond ``capacity``
    plans or ``max_bytes`` of plan memory.  Eligible: f32 or bf16 weights,
    bool or f32 events (or an f32 vector), shapes the plan layout accepts;
    everything else runs the stateless kernels.  Results equal the stateless
    kernels' within f32 summation order (the scatter sums each output copy
    exactly in fixed point; the gather adds per-tile row partials in tile
    order)."""

    def __init__(self, capacity=8, max_bytes=16 << 30):
        self._lock = threading.Lock()
        self._plans = collections.OrderedDict()   # key -> (TiledCSR or None, refs)
        self._seen = collections.OrderedDict()    # keys seen once
        self.capacity = capacity
        self.max_bytes = max_bytes
        self.enabled = True
        self.builds = 0
        self.hits = 0

    @staticmethod
    def _key(d, idx, ptr, m, n, host_w, transposed):
        dk = (("value", float(host_w)) if host_w is not None
              else (d.data_ptr(), d._version))
        return (dk, d.dtype, d.shape[0], idx.data_ptr(), idx._version, idx.dtype, idx.shape[0],
                ptr.data_ptr(), ptr._version, m, n, d.device.index, transposed)

    def lookup(self, d, idx, ptr, m, n, ev, host_w, transposed=False, coo=False):
        """The plan of A (m x n), or with ``transposed`` the plan of A^T
        (n x m), whose scatter of active rows is the event gather A @ ev.
        With ``coo``, ``ptr`` and ``idx`` are the row and column arrays of a
        COO matrix, ordered by row into CSR for the plan."""
        if not self.enabled or d.dtype not in (torch.float32, torch.bfloat16) or \
                ev.dtype not in (torch.bool, torch.float32) or idx.shape[0] == 0:
            return None
        key = self._key(d, idx, ptr, m, n, host_w, ("coo" if coo else "csr", transposed))
        with self._lock:
            ent = self._plans.get(key)
            if ent is not None:
                self._plans.move_to_end(key)
                self.hits += 1
                return ent[0]
            if key not in self._seen:
                self._seen[key] = True
                while len(self._seen) > 4 * self.capacity:
                    self._seen.popitem(last=False)
                return None
            del self._seen[key]
        try:
            if coo:
                csr = _coo_to_csr(d, ptr, idx, m, n)
                plan = None if csr is None else TiledCSR(
                    *csr, shape=(m, n), weight_dtype=d.dtype if d.shape[0] > 1 else None)
            elif transposed:
                tr = _transpose_csr(d, idx, ptr, m, n)
                plan = None if tr is None else TiledCSR(
                    *tr, shape=(n, m), weight_dtype=d.dtype if d.shape[0] > 1 else None)
            else:
                plan = TiledCSR(d, idx, ptr, shape=(m, n), weight_dtype=d.dtype if d.shape[0] > 1 else None)
        except _lib.BPXError as e:
            if "inconsistent shapes" not in str(e):
                raise
            plan = None     # outside the plan's limits: stateless from now on
        with self._lock:
            self.builds += 1
            self._plans[key] = (plan, (d, idx, ptr))
            self._evict()
        return plan

    def _evict(self):
        def total():
            return sum(p.mem.numel() for p, _ in self._plans.values() if p is not None)
        while len(self._plans) > self.capacity or (len(self._plans) > 1 and total() > self.max_bytes):
            self._plans.popitem(last=False)

    def clear(self):
        with self._lock:
            self._plans.clear()
            self._seen.clear()
            self.builds = 0
            self.hits = 0

    def info(self):
        with self._lock:
            return {"plans": sum(p is not None for p, _ in self._plans.values()),
                    "ineligible": sum(p is None for p, _ in self._plans.values()),
                    "bytes": sum(p.mem.numel() for p, _ in self._plans.values() if p is not None),
                    "builds": self.builds, "hits": self.hits, "capacity": self.capacity,
                    "max_bytes": self.max_bytes, "enabled": self.enabled}


_PLANS = _PlanCache()


def _coo_to_csr(d, row, col, m, n):
    """CSR arrays (data, indices, indptr) of an m x n COO matrix, entries
    ordered by row (a stable sort, so each row keeps its entries' order);
    None when a row index lies outside [0, m) or a column index outside
    [0, n) (the stateless COO kernel keeps JAX's semantics for those: output
    keys out of shape dropped, gather indices clamped)."""
    if int(row.min()) < 0 or int(row.max()) >= m or int(col.min()) < 0 or int(col.max()) >= n:
        return None
    rows, order = torch.sort(row, stable=True)
    return (d[order] if d.shape[0] > 1 else d).contiguous(), col[order].contiguous(), _sorted_offsets(rows, m,
                                                                                                   col.dtype)


def _sorted_offsets(keys, k, dtype):
    """Offsets (k + 1) of the runs of the sorted keys 0..k-1: entry r is the
    number of keys below r (one binary search per offset)."""
    bounds = torch.arange(k + 1, device=keys.device, dtype=keys.dtype)
    return torch.searchsorted(keys, bounds, out_int32=dtype == torch.int32).to(dtype)


def _transpose_csr(d, idx, ptr, m, n):
    """The CSR arrays of A^T (n x m) from those of A (m x n), on the device:
    entries ordered by column (a stable sort, so each column's rows stay
    ascending), A^T's indices the rows, its weights permuted alike; None
    when an index lies outside [0, n) (the stateless kernels skip such
    entries)."""
    nnz = idx.shape[0]
    if int(idx.min()) < 0 or int(idx.max()) >= n:
        return None
    counts = (ptr[1:] - ptr[:-1]).long()
    rows = torch.repeat_interleave(torch.arange(m, device=idx.device, dtype=idx.dtype), counts,
                                   output_size=nnz)
    cols, order = torch.sort(idx, stable=True)
    t_idx = rows[order]
    t_d = d[order] if d.shape[0] > 1 else d
    return t_d.contiguous(), t_idx.contiguous(), _sorted_offsets(cols, n, idx.dtype)


def plan_cache_info():
    """Counters of the plan cache behind ``csrmv(..., transpose=True)``."""
    return _PLANS.info()


def clear_plan_cache():
    """Drop every cached plan (and the references it holds) and reset the
    counters."""
    _PLANS.clear()


def set_plan_cache(enabled=None, capacity=None, max_bytes=None):
    """Switch the plan cache on or off, or change its limits."""
    with _PLANS._lock:
        if enabled is not None:
            _PLANS.enabled = bool(enabled)
        if capacity is not None:
            _PLANS.capacity = int(capacity)
        if max_bytes is not None:
            _PLANS.max_bytes = int(max_bytes)
        _PLANS._evict()


class _EventCSRMV(torch.autograd.Function):
    """Differentiation rules of event/_csr_matvec.py:522-546: the tangent in
    `data` is the same operator on the tangent; the cotangent of a scalar
    weight is <ct, csrmv(1)>, of heterogeneous weights events[row] * ct[col]
    (transpose) or events[col] * ct[row]; float events take the plain sparse
    matvec of `data` with the cotangent and `transpose` flipped."""

    @staticmethod
    def forward(d, ev, idx, ptr, shape, transpose, rule):
        return _launch(d, idx, ptr, ev, shape, transpose, rule)

    @staticmethod
    def setup_context(ctx, inputs, output):
        d, ev, idx, ptr, shape, transpose, rule = inputs
        ctx.save_for_backward(d, ev, idx, ptr)
        ctx.shape, ctx.transpose, ctx.rule = shape, transpose, rule

    @staticmethod
    def backward(ctx, ct):
        from .sparse import _CSRMV, csr_to_coo
        d, ev, idx, ptr = ctx.saved_tensors
        gd = gev = None
        if ctx.needs_input_grad[0]:
            if d.shape[0] == 1:
                ones = torch.ones(1, dtype=d.dtype, device=d.device)
                base = _EventCSRMV.apply(ones, ev, idx, ptr, ctx.shape, ctx.transpose, ctx.rule)
                gd = torch.dot(ct.to(base.dtype), base).reshape(1)
            else:
                row, col = csr_to_coo(idx, ptr)
                e = ev.to(ct.dtype)
                gd = e[row.long()] * ct[col.long()] if ctx.transpose else e[col.long()] * ct[row.long()]
            gd = gd.to(d.dtype)
        if ctx.needs_input_grad[1]:
            gev = _CSRMV.apply(d, ct.to(acc_dtype(d.dtype)), idx, ptr, ctx.shape, not ctx.transpose).to(ev.dtype)
        return gd, gev, None, None, None, None, None

    @staticmethod
    def vmap(info, in_dims, d, ev, idx, ptr, shape, transpose, rule):
        if in_dims[0] is None and in_dims[2] is None and in_dims[3] is None and in_dims[1] is not None:
            out = _launch_batch(d, ev.movedim(in_dims[1], 0), idx, ptr, shape, transpose, rule)
            if out is not None:
                return out, 0
        return batch_rows(info, in_dims, (d, ev, idx, ptr, shape, transpose, rule), _EventCSRMV.apply)


def _launch_batch(d, ev, idx, ptr, shape, transpose, rule):
    """csrmv over a batch of event vectors ev [B, len] in one launch: on the
    cached tiled plan for transpose=True when the matrix has one (a batch is
    a repeated use), else bpx_event_csrmv_batch (each row read once for all
    vectors).  None when the dtypes are outside the batched kernels (f64)."""
    if d.dtype not in (torch.float32, torch.bfloat16) or ev.dtype not in (torch.bool, torch.float32):
        return None
    d, idx, ptr, ev = d.contiguous(), idx.contiguous(), ptr.contiguous(), ev.contiguous()
    m, n = int(shape[0]), int(shape[1])
    B = ev.shape[0]
    if ptr.shape[0] != m + 1:
        raise ValueError(f'indptr should have shape ({m + 1},), but we got {tuple(ptr.shape)}.')
    if B >= 1:
        plan = _PLANS.lookup(d, idx, ptr, m, n, ev, None, transposed=not transpose)
        if plan is None:
            # a batch counts as repeated use
            plan = _PLANS.lookup(d, idx, ptr, m, n, ev, None, transposed=not transpose)
        if plan is not None:
            out = torch.empty((B, n if transpose else m), dtype=torch.float32, device=ev.device)
            return plan._exec_batch(ev, rule, B, False, out)
    L = _lib.lib()
    ev_len = m if transpose else n
    out = torch.empty((B, n if transpose else m), dtype=torch.float32, device=ev.device)
    wsb = int(L.bpx_batch_workspace(ev_len))
    ws = torch.empty(max(wsb, 4), dtype=torch.uint8, device=ev.device)
    rc = L.bpx_event_csrmv_batch(_lib.stream_ptr(), _lib.ptr(d), _lib.bpx_dtype(d), d.shape[0], _lib.ptr(idx),
                                 _lib.ptr(ptr), _lib.bpx_dtype(idx), idx.shape[0], _lib.ptr(ev), _lib.bpx_dtype(ev),
                                 rule, B, m, n, int(bool(transpose)), _lib.ptr(out), _lib.ptr(ws), ws.numel())
    _lib.check(rc, "bpx_event_csrmv_batch")
    return out


def _run(data, indices, indptr, events, shape, transpose, rule):
    host_w = float(data[0]) if isinstance(data, np.ndarray) and data.shape == (1,) else None
    d = data_tensor(data)
    idx = index_tensor(indices)
    ptr = index_tensor(indptr)
    if ptr.dtype != idx.dtype:
        ptr = ptr.to(idx.dtype)
    ev = _events_tensor(events)
    if d.dtype == torch.float16:
        d = d.to(torch.float32)
    if needs_ad(d, ev):
        return _EventCSRMV.apply(d, ev, idx, ptr, (int(shape[0]), int(shape[1])), bool(transpose), rule)
    return _launch(d, idx, ptr, ev, shape, transpose, rule, host_w)


_EV_CODE = {torch.bool: _lib.BPX_BOOL, torch.float32: _lib.BPX_F32}
_FAST_W = (torch.float32, torch.bfloat16)
_FAST_I = (torch.int32, torch.int64)
_wrapped = torch._C._functorch.is_functorch_wrapped_tensor


def _fast_transpose(data, indices, indptr, events, shape):
    """``csrmv(..., transpose=True)`` for device tensors that need neither
    conversion nor differentiation -- the per-step call of a projection
    layer.  Returns None whenever anything is unusual (other types, dtypes,
    devices, views, gradients, inconsistent sizes); the general path then
    checks the arguments and raises the reference's errors."""
    T = torch.Tensor
    if type(data) is not T or type(indices) is not T or type(indptr) is not T or type(events) is not T:
        return None
    if not (data.is_cuda and events.is_cuda) or len(shape) != 2:
        return None
    if data.dtype not in _FAST_W or events.dtype not in _EV_CODE or indices.dtype not in _FAST_I or \
            indptr.dtype is not indices.dtype:
        return None
    if data.dim() != 1 or indices.dim() != 1 or indptr.dim() != 1 or events.dim() != 1:
        return None
    m, n = int(shape[0]), int(shape[1])
    nd = data.shape[0]
    if (nd != 1 and nd != indices.shape[0]) or events.shape[0] != m or indptr.shape[0] != m + 1:
        return None
    if not (data.is_contiguous() and indices.is_contiguous() and indptr.is_contiguous() and events.is_contiguous()):
        return None
    if torch.is_grad_enabled() and (data.requires_grad or events.requires_grad):
        return None
    if _wrapped(data) or _wrapped(events):
        return None
    dev = events.get_device()
    if data.get_device() != dev or indices.get_device() != dev or indptr.get_device() != dev or \
            dev != torch.cuda.current_device():
        return None
    plan = _PLANS.lookup(data, indices, indptr, m, n, events, None)
    if plan is None:
        return _launch(data, indices, indptr, events, shape, True, _lib.RULE_POSITIVE, use_plan=False)
    return plan._exec(events, _lib.RULE_POSITIVE, None)


def csrmv(data, indices, indptr, events, *, shape: Tuple[int, int], transpose: bool = False) -> torch.Tensor:
    """Product of a sparse CSR matrix and a dense event vector.

    ``transpose=True`` computes ``events @ A`` event-driven: only rows with a
    truthy event are read.  Bool events are truthy when set; float events
    when ``> 0`` (the Numba rule of event/_csr_matvec.py:349).  Weights may
    be f32, f64 or bf16 (bf16 is stored at half the bytes and accumulated in
    f32).  Returns ``(shape[1] if transpose else shape[0],)``.
    Reference: brainpy/_src/math/event/_csr_matvec.py:38-110.
    """
    if transpose is True:
        out = _fast_transpose(data, indices, indptr, events, shape)
        if out is not None:
            return out
    data, indices, indptr, events = (_as_array(x) for x in (data, indices, indptr, events))
    if ndim(data) == 0:
        data = data.reshape(1)
    _check(data, indices, indptr, events, shape, transpose, _PUBLIC_INDEX,
           '{} should be a 1D vector with int32 or int64 type.')
    return _run(data, indices, indptr, events, shape, transpose, _lib.RULE_POSITIVE)


def csrmv_batch(data, indices, indptr, events, *, shape: Tuple[int, int], transpose: bool = False,
                stateless: bool = False) -> torch.Tensor:
    """``csrmv`` of every row of ``events`` ([B, len]) in one launch -- what
    ``torch.func.vmap(csrmv)`` over the events runs.  ``stateless`` skips the
    plan cache (transpose=True then runs bpx_event_csrmv_batch)."""
    d = data_tensor(_as_array(data))
    if d.dtype == torch.float16:
        d = d.to(torch.float32)
    idx = index_tensor(_as_array(indices))
    ptr = index_tensor(_as_array(indptr))
    if ptr.dtype != idx.dtype:
        ptr = ptr.to(idx.dtype)
    ev = _events_tensor(_as_array(events))
    if ev.ndim != 2:
        raise ValueError('events should be a 2D [batch, length] array.')
    _check(d, idx, ptr, ev[0], shape, transpose, _PUBLIC_INDEX, '{} should be a 1D vector with int32 or int64 type.')
    if stateless:
        with_cache = _PLANS.enabled
        _PLANS.enabled = False
    try:
        out = _launch_batch(d, ev, idx, ptr, shape, transpose, _lib.RULE_POSITIVE)
    finally:
        if stateless:
            _PLANS.enabled = with_cache
    if out is None:
        out = torch.stack([_launch(d, idx, ptr, ev[b], shape, transpose, _lib.RULE_POSITIVE)
                           for b in range(ev.shape[0])])
    return out


def csrmv_taichi(data, indices, indptr, events, *, shape: Tuple[int, int], transpose: bool = False):
    """Taichi twin of :func:`csrmv` (float events truthy when ``!= 0``);
    returns a 1-tuple like the reference, or a bare zero vector when
    ``indices`` is empty.  Reference: event/_csr_matvec_taichi.py:344-452."""
    data, indices, indptr, events = (_as_array(x) for x in (data, indices, indptr, events))
    if ndim(data) == 0:
        data = data.reshape(1)
    _check(data, indices, indptr, events, shape, transpose, _TAICHI_INDEX,
           '{} should be a 1D vector with int8, int16, int32, int64, uint8, uint16, uint32 or uint64 type.')
    if indices.shape[0] == 0:
        d = data_tensor(data)
        return torch.zeros(shape[1] if transpose else shape[0], dtype=acc_dtype(d.dtype), device=_lib.device())
    return (_run(data, indices, indptr, events, shape, transpose, _lib.RULE_NONZERO),)


def csr_to_csc(indices, indptr, shape, data=None):
    """CSR -> CSC with ascending row indices inside every column (stable
    counting sort on the host; a one-time set-up step).  Returns numpy
    ``(csc_rows int32, csc_colptr int32, csc_data or None)``."""
    indices = np.asarray(indices.cpu() if isinstance(indices, torch.Tensor) else indices).astype(np.int64)
    indptr = np.asarray(indptr.cpu() if isinstance(indptr, torch.Tensor) else indptr).astype(np.int64)
    m, n = shape
    rows = np.repeat(np.arange(m, dtype=np.int64), np.diff(indptr))
    order = np.argsort(indices, kind="stable")
    colptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(indices, minlength=n), out=colptr[1:])
    csc_data = None
    if data is not None:
        d = np.asarray(data.cpu() if isinstance(data, torch.Tensor) else data)
        csc_data = d if d.size == 1 else d[order]
    return rows[order].astype(np.int32), colptr.astype(np.int32), csc_data


def csrmv_fixed_order(csc_data, csc_rows, csc_colptr, events, *, shape: Tuple[int, int],
                      rule: int = _lib.RULE_POSITIVE, out: torch.Tensor = None) -> torch.Tensor:
    """Deterministic ``events @ A`` (transpose event csrmv) from the CSC form
    of :func:`csr_to_csc`: every output is summed in ascending pre-synaptic
    order, bit-identical to the serial CPU scatter of
    event/_csr_matvec.py:321-340 regardless of scheduling.  With ``out``
    given, the sums are added into it (``out += events @ A``)."""
    d = data_tensor(csc_data).to(torch.float32)
    rows = _lib.to_device(csc_rows, torch.int32)
    colptr = _lib.to_device(csc_colptr, torch.int32)
    ev = _events_tensor(events)
    if ev.dtype == torch.float64:
        ev = ev.to(torch.float32)
    m, n = int(shape[0]), int(shape[1])
    accumulate = out is not None
    if out is None:
        out = torch.empty(n, dtype=torch.float32, device=ev.device)
    elif out.dtype != torch.float32 or out.shape != (n,) or not out.is_contiguous():
        raise ValueError(f'out must be a contiguous float32 vector of shape ({n},).')
    rc = _lib.lib().bpx_event_csrmv_fixed_order(_lib.stream_ptr(), _lib.ptr(d), _lib.BPX_F32, d.shape[0],
                                                _lib.ptr(rows), _lib.ptr(colptr), rows.shape[0], _lib.ptr(ev),
                                                _lib.bpx_dtype(ev), rule, m, n, int(accumulate), _lib.ptr(out))
    _lib.check(rc, "bpx_event_csrmv_fixed_order")
    return out


def csrmv_add(out, data, indices, indptr, events, *, shape: Tuple[int, int], rule: int = _lib.RULE_POSITIVE):
    """``out += events @ A`` (transpose event csrmv accumulated in place,
    FP32 atomics); the conductance update ``g += csrmv(...)`` of
    dyn/synapses/abstract_models.py:133-134 without a temporary."""
    d = data_tensor(data)
    idx = index_tensor(indices)
    ptr = index_tensor(indptr)
    if ptr.dtype != idx.dtype:
        ptr = ptr.to(idx.dtype)
    ev = _events_tensor(events)
    if ev.dtype == torch.float64:
        ev = ev.to(torch.float32)
    m, n = int(shape[0]), int(shape[1])
    if out.dtype != torch.float32 or out.shape != (n,) or not out.is_contiguous():
        raise ValueError(f'out must be a contiguous float32 vector of shape ({n},).')
    rc = _lib.lib().bpx_event_csrmv_add(_lib.stream_ptr(), _lib.ptr(d), _lib.bpx_dtype(d), d.shape[0], _lib.ptr(idx),
                                        _lib.ptr(ptr), _lib.bpx_dtype(idx), idx.shape[0], _lib.ptr(ev),
                                        _lib.bpx_dtype(ev), rule, m, n, _lib.ptr(out))
    _lib.check(rc, "bpx_event_csrmv_add")
    return out


def info(events):
    """Collect event information: ``(ids, num)`` with ``ids`` the ascending
    indices of truthy events padded with -1 and ``num`` their count (shape
    ``(1,)``; batched ``[B, N]`` input gives ``[B, N]`` and ``[B]``).
    Reference: brainpy/_src/math/event/_info_collection.py:24-43, 54-66."""
    ev = _lib.to_device(events)
    if ev.ndim not in (1, 2):
        raise TypeError('Only support 1D boolean vector.')
    if ev.dtype not in (torch.bool, torch.float32, torch.float64, torch.int32):
        ev = ev.to(torch.float32)
    batched = ev.ndim == 2
    b, n = (ev.shape[0], ev.shape[1]) if batched else (1, ev.shape[0])
    ids = torch.empty((b, n), dtype=torch.int32, device=ev.device)
    num = torch.empty(b, dtype=torch.int32, device=ev.device)
    L = _lib.lib()
    wsb = L.bpx_event_info_workspace(b, n)
    ws = torch.empty(max(int(wsb), 4), dtype=torch.uint8, device=ev.device)
    rc = L.bpx_event_info(_lib.stream_ptr(), _lib.ptr(ev), _lib.bpx_dtype(ev), b, n, _lib.ptr(ids), _lib.ptr(num),
                          _lib.ptr(ws), ws.numel())
    _lib.check(rc, "bpx_event_info")
    if batched:
        return ids, num
    return ids.reshape(n), num.reshape(1)


class TiledCSR:
    """A CSR matrix re-laid out once for the event-driven scatter
    (bpx_tiled_csr_build): output tiles in the LDS of every CU, one copy per
    XCD, 16-bit tile-local column indices.  ``plan.csrmv(events)`` equals
    ``csrmv(data, indices, indptr, events, shape=shape, transpose=True)``
    but replaces one global atomic per synaptic event by LDS accumulation and
    one coalesced flush per tile.  ``weight_dtype`` may be ``torch.float32``
    or ``torch.bfloat16`` (heterogeneous weights stored at half the bytes and
    accumulated in f32); homogeneous ``data`` (size 1) stores no weights."""

    def __init__(self, data, indices, indptr, *, shape: Tuple[int, int], weight_dtype=None):
        d = data_tensor(_as_array(data))
        idx = index_tensor(_as_array(indices))
        ptr = index_tensor(_as_array(indptr))
        if ptr.dtype != idx.dtype:
            ptr = ptr.to(idx.dtype)
        m, n = int(shape[0]), int(shape[1])
        if ptr.shape[0] != m + 1:
            raise ValueError(f'indptr should have shape ({m + 1},), but we got {tuple(ptr.shape)}.')
        if d.shape[0] not in (1, idx.shape[0]):
            raise ValueError('The size of data should be 1 or be consistent with indices.')
        if d.dtype == torch.float16:
            d = d.to(torch.float32)
        homo = d.shape[0] == 1
        if homo:
            wcode = _lib.TILED_HOMO
        elif weight_dtype in (None, torch.float32):
            wcode = _lib.BPX_F32
        elif weight_dtype == torch.bfloat16:
            wcode = _lib.BPX_BF16
        else:
            raise TypeError(f'weight_dtype must be float32 or bfloat16, got {weight_dtype}')
        L = _lib.lib()
        self.plan = _lib.TiledCSRPlan()
        nbytes = ctypes.c_size_t()
        _lib.check(L.bpx_tiled_csr_layout(m, n, idx.shape[0], wcode, ctypes.byref(self.plan), ctypes.byref(nbytes)),
                   "bpx_tiled_csr_layout")
        self.mem = torch.empty(int(nbytes.value), dtype=torch.uint8, device=d.device)
        _lib.check(L.bpx_tiled_csr_build(_lib.stream_ptr(), _lib.ptr(d), _lib.bpx_dtype(d), d.shape[0], _lib.ptr(idx),
                                         _lib.ptr(ptr), _lib.bpx_dtype(idx), ctypes.byref(self.plan),
                                         _lib.ptr(self.mem), self.mem.numel()), "bpx_tiled_csr_build")
        self.shape = (m, n)
        self.weight_dtype = wcode
        self._plan_ref = ctypes.byref(self.plan)
        self._fn = L.bpx_event_csrmv_tiled
        # plans of more than 64 tiles list a call's active rows once, in a
        # workspace of this many bytes per event vector (0: the call scans)
        self._list_ws = int(L.bpx_tiled_csr_call_workspace(self._plan_ref, 1))

    def csrmv(self, events, *, out: torch.Tensor = None, rule: int = _lib.RULE_POSITIVE) -> torch.Tensor:
        """``events @ A`` (a new vector), or ``out += events @ A`` when
        ``out`` is given."""
        ev = _events_tensor(events)
        if ev.dtype == torch.float64:
            ev = ev.to(torch.float32)
        m, n = self.shape
        if ev.ndim != 1 or ev.shape[0] != m:
            raise ValueError(f'Shape mismatch, vec ({ev.shape[0]},) @ mat {self.shape}.')
        if out is not None and (out.dtype != torch.float32 or out.shape != (n,) or not out.is_contiguous()):
            raise ValueError(f'out must be a contiguous float32 vector of shape ({n},).')
        return self._exec(ev.contiguous(), rule, out)

    def _exec(self, ev, rule, out):
        """Launch on a contiguous bool/f32 event vector of length m, on the
        current stream of its device."""
        accumulate = out is not None
        if out is None:
            out = torch.empty(self.shape[1], dtype=torch.float32, device=ev.device)
        if self._list_ws:
            return self._exec_batch(ev, rule, 1, accumulate, out)
        rc = self._fn(_lib.raw_stream(ev.get_device()), self._plan_ref, ev.data_ptr(),
                      _EV_CODE[ev.dtype], rule, accumulate, out.data_ptr())
        if rc != 0:
            _lib.check(rc, "bpx_event_csrmv_tiled")
        return out

    def _exec_batch(self, ev, rule, batch, accumulate, out):
        """`batch` contiguous event vectors [batch][m] into out [batch][n],
        with the active-row list workspace when the plan takes one."""
        L = _lib.lib()
        wsb = int(L.bpx_tiled_csr_call_workspace(self._plan_ref, batch)) if self._list_ws else 0
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=ev.device)
        rc = L.bpx_event_csrmv_tiled_ws(_lib.raw_stream(ev.get_device()), self._plan_ref, ev.data_ptr(),
                                        _EV_CODE[ev.dtype], rule, batch, int(bool(accumulate)), out.data_ptr(),
                                        ws.data_ptr(), wsb)
        if rc != 0:
            _lib.check(rc, "bpx_event_csrmv_tiled_ws")
        return out

    def _gather(self, v, event, rule):
        """A @ v (transpose=False) on the plan: bool/f32 events (event) or a
        dense f32 vector, contiguous, of length n."""
        m = self.shape[0]
        out = torch.empty(m, dtype=torch.float32, device=v.device)
        L = _lib.lib()
        wsb = int(L.bpx_tiled_csr_gather_workspace(self._plan_ref))
        ws = torch.empty(max(wsb, 4), dtype=torch.uint8, device=v.device)
        code = _EV_CODE[v.dtype] if event else _lib.BPX_F32
        _lib.check(L.bpx_tiled_csr_gather(_lib.stream_ptr(), self._plan_ref, _lib.ptr(v), code, int(bool(event)), rule,
                                          _lib.ptr(out), _lib.ptr(ws), ws.numel()), "bpx_tiled_csr_gather")
        return out

    def gather(self, vec, *, event: bool = False, rule: int = _lib.RULE_POSITIVE) -> torch.Tensor:
        """``A @ vec``: the transpose=False form on the plan (each column tile
        of vec staged in LDS, per-tile row partials added in tile order).
        ``event=True`` treats vec as events (bool, or f32 tested with
        ``rule``) and sums the weights of truthy columns, as
        ``bm.event.csrmv``; otherwise vec is a dense f32 vector, as
        ``bm.sparse.csrmv``."""
        v = _events_tensor(vec) if event else _lib.to_device(_as_array(vec))
        if event and v.dtype == torch.float64:
            v = v.to(torch.float32)
        if not event and v.dtype != torch.float32:
            v = v.to(torch.float32)
        v = v.contiguous()
        n = self.shape[1]
        if v.ndim != 1 or v.shape[0] != n:
            raise ValueError(f'Shape mismatch, mat {self.shape} @ vec ({v.shape[0]},).')
        return self._gather(v, event, rule)

    def matvec(self, x, *, out: torch.Tensor = None) -> torch.Tensor:
        """``x @ A`` for a dense float32 vector x -- ``bm.sparse.csrmv(data,
        indices, indptr, x, shape=shape, transpose=True)`` on the plan: the
        rows with x != 0 stream their segments and every product ``w *
        x[row]`` (rounded to f32) is summed exactly in fixed point, with no
        global atomic per synapse.  ``out += x @ A`` when ``out`` is given.
        Products that are inf or NaN (non-finite x, f32 overflow) reach
        their outputs as IEEE addition would: an infinity, or NaN for a NaN
        term or infinities of both signs."""
        xt = _lib.to_device(_as_array(x))
        if xt.dtype != torch.float32:
            xt = xt.to(torch.float32)
        xt = xt.contiguous()
        m, n = self.shape
        if xt.ndim != 1 or xt.shape[0] != m:
            raise ValueError(f'Shape mismatch, vec ({xt.shape[0]},) @ mat {self.shape}.')
        accumulate = out is not None
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=xt.device)
        elif out.dtype != torch.float32 or out.shape != (n,) or not out.is_contiguous():
            raise ValueError(f'out must be a contiguous float32 vector of shape ({n},).')
        # max and min |x| of this call (8 bytes), and the list of the rows
        # with x != 0 for plans of more than 64 tiles
        wsb = 256 + self._list_ws if self._list_ws else 8
        ws = torch.empty(wsb, dtype=torch.uint8, device=xt.device)
        rc = _lib.lib().bpx_csrmv_tiled(_lib.stream_ptr(), ctypes.byref(self.plan), _lib.ptr(xt),
                                        _lib.BPX_F32, int(accumulate), _lib.ptr(out), _lib.ptr(ws), wsb)
        _lib.check(rc, "bpx_csrmv_tiled")
        return out
