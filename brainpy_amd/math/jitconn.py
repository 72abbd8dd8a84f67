"""Just-in-time connectivity operators (``bm.jitconn``) on the gfx950
library: ``event_mv_prob_{homo,uniform,normal}``,
``mv_prob_{homo,uniform,normal}`` and their ``*_taichi`` twins.

The matrix M (``shape``, ``conn_prob``, ``seed``) is regenerated inside the
kernel from LFSR88 streams and never stored.  The connectivity is the one
the reference defines in-tree -- the Taichi kernels of
``jitconn/_event_matvec_taichi.py`` and ``jitconn/_matvec_taichi.py`` -- and
is reproduced bit-exactly:

* ``layout='lane32'`` (default): the Taichi GPU kernels, 32 streams per row
  (``outdim_parallel=True``) or per column;
* ``layout='row'``: the Taichi CPU kernels, one stream per row/column.

The public brainpylib names (``event_mv_prob_homo`` ...) draw their matrix
from brainpylib's own RNG, which is not in the reference tree; here they
share the Taichi-defined connectivity (their signature, dtype rules and
``clen``/seed handling are kept: ``jitconn/_event_matvec.py:32-126``).
"""
from typing import Optional, Tuple

import numpy as np
import torch

from .. import _lib
from ._common import batch_rows, needs_ad

__all__ = [
    "event_mv_prob_homo", "event_mv_prob_uniform", "event_mv_prob_normal",
    "mv_prob_homo", "mv_prob_uniform", "mv_prob_normal",
    "event_mv_prob_homo_taichi", "event_mv_prob_uniform_taichi", "event_mv_prob_normal_taichi",
    "mv_prob_homo_taichi", "mv_prob_uniform_taichi", "mv_prob_normal_taichi",
    "conn_len", "get_conn_matrix",
]

_LAYOUTS = {"lane32": _lib.LAYOUT_LANE32, "row": _lib.LAYOUT_ROW}


def conn_len(conn_prob) -> int:
    """``ceil(1/p)*2 - 1``: 1/p from a python float, rounded to float32 by
    ``jnp.ceil`` (x64 disabled), cast to int32
    (jitconn/_event_matvec_taichi.py:401-402; _event_matvec.py:44-45)."""
    p = float(np.asarray(conn_prob.cpu() if isinstance(conn_prob, torch.Tensor) else conn_prob).reshape(-1)[0])
    if not (0.0 < p <= 1.0):
        raise ValueError(f'conn_prob must be in (0, 1], but we got {p}.')
    return int(np.ceil(np.float32(1.0 / p)) * np.float32(2) - np.float32(1))


def _seed(seed) -> int:
    if seed is None:
        seed = int(np.random.randint(0, int(1e8)))
    s = np.asarray(seed.cpu() if isinstance(seed, torch.Tensor) else seed).reshape(-1)
    return int(s[0]) & 0xFFFFFFFF


def _scalar(x) -> float:
    a = np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x).reshape(-1)
    if a.size != 1:
        raise ValueError('weight must be a 1D scalar.')
    return float(a[0])


def _general_checking(vector, shape, outdim_parallel, transpose):
    """jitconn/_matvec_taichi.py:174-207; returns (mat_shape, out_len)."""
    if vector.ndim != 1:
        raise ValueError('vector should be a 1D vector.')
    if len(shape) != 2:
        raise ValueError('shape should be a length-2 tuple.')
    if not isinstance(outdim_parallel, bool):
        raise ValueError('outdim_parallel must be boolean value.')
    if not isinstance(transpose, bool):
        raise ValueError('transpose must be boolean value.')
    if transpose:
        if vector.shape[0] != shape[0]:
            raise ValueError(f'Shape mismatch, vec {tuple(vector.shape)} @ mat {shape}.')
        shape = shape[::-1]
    else:
        if vector.shape[0] != shape[1]:
            raise ValueError(f'Shape mismatch, mat {shape} @ vec ({vector.shape[0]},).')
    return (int(shape[0]), int(shape[1]))


def _launch(v, event, dist, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout):
    v = v.contiguous()         # raw pointer to the kernel: views are made dense first
    out = torch.empty(num_row, dtype=torch.float32, device=v.device)
    L = _lib.lib()
    # bool events take the workspace in both forms: the event bitmask of the
    # row-parallel kernel; for the scatter form (outdim_parallel=False) the
    # per-tile connection lists of its two passes
    if not event:
        wsb = 0
    elif outdim_parallel:
        wsb = int(L.bpx_jitconn_workspace(num_row, num_col))
    else:
        wsb = int(L.bpx_jitconn_workspace_ex(num_row, num_col, clen, dist))
    ws = torch.empty(max(wsb, 4), dtype=torch.uint8, device=v.device) if wsb else None
    rc = L.bpx_jitconn_mv(_lib.stream_ptr(), _lib.ptr(v), _lib.bpx_dtype(v), int(event), dist,
                          np.float32(w0), np.float32(w1), clen, seed, num_row, num_col,
                          int(outdim_parallel), _LAYOUTS[layout], _lib.ptr(out), _lib.ptr(ws), wsb)
    _lib.check(rc, "bpx_jitconn_mv")
    return out


class _JitMV(torch.autograd.Function):
    """Differentiation rules of the Taichi primitives
    (jitconn/_matvec_taichi.py:138-172, 480-496; _event_matvec_taichi.py:
    284-300): the cotangent of the vector (or of float events) is the
    non-event operator on the cotangent with the matrix shape reversed and
    `outdim_parallel` flipped; the cotangent of a homogeneous weight is
    sum(op(ct, 1) * vector) with the same flags (the reference's rule, which
    reads the cotangent as an input of the same operator, so it needs a
    square matrix).  Uniform and normal weight parameters are not
    differentiable, as in the reference.  `w` is the homogeneous weight as a
    tensor when it takes part in differentiation, else None."""

    @staticmethod
    def forward(v, w, event, dist, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout):
        if w is not None:
            w0 = float(w.detach().reshape(-1)[0])
        return _launch(v, event, dist, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout)

    @staticmethod
    def setup_context(ctx, inputs, output):
        v, w, event, dist, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout = inputs
        ctx.save_for_backward(v, w)
        ctx.args = (dist, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout)

    @staticmethod
    def backward(ctx, ct):
        v, w = ctx.saved_tensors
        dist, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout = ctx.args
        gv = gw = None
        ct = ct.to(torch.float32)
        if ctx.needs_input_grad[0]:
            gv = _JitMV.apply(ct, w, False, dist, w0, w1, clen, seed, num_col, num_row, not outdim_parallel,
                              layout).to(v.dtype)
        if ctx.needs_input_grad[1]:
            if num_row != num_col:
                raise ValueError(f'Shape mismatch, mat {(num_row, num_col)} @ vec ({num_row},).')
            row = _launch(ct, False, dist, 1.0, w1, clen, seed, num_row, num_col, outdim_parallel, layout)
            gw = torch.sum(row * v.to(torch.float32)).reshape(w.shape).to(w.dtype)
        return gv, gw, None, None, None, None, None, None, None, None, None, None

    @staticmethod
    def vmap(info, in_dims, *args):
        v, w = args[0], args[1]
        if in_dims[0] is not None and all(d is None for d in in_dims[1:]) and args[2] and v.dtype == torch.bool:
            dist, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout = args[3:]
            if w is not None:
                w0 = float(w.detach().reshape(-1)[0])
            return _launch_batch(v.movedim(in_dims[0], 0), dist, w0, w1, clen, seed, num_row, num_col,
                                 outdim_parallel, layout), 0
        return batch_rows(info, in_dims, args, _JitMV.apply)


def _launch_batch(v, dist, w0, w1, clen, seed, num_row, num_col, outdim_parallel, layout):
    """Event operator over bool events v [B, num_col] in one launch
    (bpx_jitconn_mv_batch): the connections are generated once for the
    whole batch."""
    v = v.contiguous()
    B = v.shape[0]
    out = torch.empty((B, num_row), dtype=torch.float32, device=v.device)
    L = _lib.lib()
    wsb = int(L.bpx_batch_workspace(num_col))
    ws = torch.empty(max(wsb, 4), dtype=torch.uint8, device=v.device)
    rc = L.bpx_jitconn_mv_batch(_lib.stream_ptr(), _lib.ptr(v), _lib.BPX_BOOL, dist, np.float32(w0), np.float32(w1),
                                clen, seed, B, num_row, num_col, int(outdim_parallel), _LAYOUTS[layout],
                                _lib.ptr(out), _lib.ptr(ws), ws.numel())
    _lib.check(rc, "bpx_jitconn_mv_batch")
    return out


def _weight_arg(x, name, differentiable):
    """A weight parameter as (float value, tensor or None): the tensor is
    kept only when it takes part in differentiation."""
    if isinstance(x, torch.Tensor) and needs_ad(x):
        if not differentiable:
            raise AssertionError(f'Cannot differentiate through {name}.')
        t = x.reshape(-1)
        if t.numel() != 1:
            raise ValueError('weight must be a 1D scalar.')
        return 0.0, t
    return _scalar(x), None


def _run(vector, event, dist, w0, w1, conn_prob, seed, shape, transpose, outdim_parallel, layout, w_t=None):
    v = _lib.to_device(vector)
    if event:
        if v.dtype not in (torch.bool, torch.float16, torch.float32, torch.float64):
            raise AssertionError('events must be bool or float valued.')
        if v.dtype != torch.bool and v.dtype != torch.float32:
            v = v.to(torch.float32)
    else:
        if v.dtype not in (torch.float16, torch.float32, torch.float64):
            raise AssertionError('vector must be float valued.')
        if v.dtype != torch.float32:
            v = v.to(torch.float32)
    num_row, num_col = _general_checking(v, tuple(shape), outdim_parallel, transpose)
    if layout not in _LAYOUTS:
        raise ValueError(f'layout must be one of {tuple(_LAYOUTS)}, but we got {layout}.')
    clen = conn_len(conn_prob)
    sd = _seed(seed)
    if w_t is not None or needs_ad(v):
        if w_t is not None:
            w_t = w_t.to(device=v.device, dtype=torch.float32)
        return _JitMV.apply(v, w_t, bool(event), dist, float(w0), float(w1), clen, sd, num_row, num_col,
                            bool(outdim_parallel), layout)
    return _launch(v, event, dist, w0, w1, clen, sd, num_row, num_col, outdim_parallel, layout)


def event_mv_prob_homo(events, weight, conn_prob, seed: Optional[int] = None, *, shape: Tuple[int, int],
                       transpose: bool = False, outdim_parallel: bool = True, layout: str = "lane32"):
    """y = M @ events with a homogeneous weight at every generated connection.
    Reference: jitconn/_event_matvec.py:32-59, _event_matvec_taichi.py:340-408."""
    w, wt = _weight_arg(weight, 'weight', True)
    return _run(events, True, _lib.DIST_HOMO, w, 0.0, conn_prob, seed, shape, transpose, outdim_parallel, layout,
                wt)


def event_mv_prob_uniform(events, w_low, w_high, conn_prob, seed: Optional[int] = None, *, shape: Tuple[int, int],
                          transpose: bool = False, outdim_parallel: bool = True, layout: str = "lane32"):
    """y = M @ events with U[w_low, w_high) weights.
    Reference: jitconn/_event_matvec.py:64-92, _event_matvec_taichi.py:751-824."""
    return _run(events, True, _lib.DIST_UNIFORM, _weight_arg(w_low, 'w_low', False)[0],
                _weight_arg(w_high, 'w_high', False)[0], conn_prob, seed, shape,
                transpose, outdim_parallel, layout)


def event_mv_prob_normal(events, w_mu, w_sigma, conn_prob, seed: Optional[int] = None, *, shape: Tuple[int, int],
                         transpose: bool = False, outdim_parallel: bool = True, layout: str = "lane32"):
    """y = M @ events with N(w_mu, w_sigma) weights.
    Reference: jitconn/_event_matvec.py:97-126, _event_matvec_taichi.py:1168-1241."""
    return _run(events, True, _lib.DIST_NORMAL, _weight_arg(w_mu, 'w_mu', False)[0],
                _weight_arg(w_sigma, 'w_sigma', False)[0], conn_prob, seed, shape,
                transpose, outdim_parallel, layout)


def mv_prob_homo(vector, weight, conn_prob, seed: Optional[int] = None, *, shape: Tuple[int, int],
                 transpose: bool = False, outdim_parallel: bool = True, layout: str = "lane32"):
    """y = M @ vector (dense float vector), homogeneous weight.
    Reference: jitconn/_matvec.py:27-100, _matvec_taichi.py:242-318."""
    w, wt = _weight_arg(weight, 'weight', True)
    return _run(vector, False, _lib.DIST_HOMO, w, 0.0, conn_prob, seed, shape, transpose, outdim_parallel, layout,
                wt)


def mv_prob_uniform(vector, w_low, w_high, conn_prob, seed: Optional[int] = None, *, shape: Tuple[int, int],
                    transpose: bool = False, outdim_parallel: bool = True, layout: str = "lane32"):
    """Reference: jitconn/_matvec.py:103-181, _matvec_taichi.py:527-600."""
    return _run(vector, False, _lib.DIST_UNIFORM, _weight_arg(w_low, 'w_low', False)[0],
                _weight_arg(w_high, 'w_high', False)[0], conn_prob, seed, shape,
                transpose, outdim_parallel, layout)


def mv_prob_normal(vector, w_mu, w_sigma, conn_prob, seed: Optional[int] = None, *, shape: Tuple[int, int],
                   transpose: bool = False, outdim_parallel: bool = True, layout: str = "lane32"):
    """Reference: jitconn/_matvec.py:184-260, _matvec_taichi.py:814-887."""
    return _run(vector, False, _lib.DIST_NORMAL, _weight_arg(w_mu, 'w_mu', False)[0],
                _weight_arg(w_sigma, 'w_sigma', False)[0], conn_prob, seed, shape,
                transpose, outdim_parallel, layout)


event_mv_prob_homo_taichi = event_mv_prob_homo
event_mv_prob_uniform_taichi = event_mv_prob_uniform
event_mv_prob_normal_taichi = event_mv_prob_normal
mv_prob_homo_taichi = mv_prob_homo
mv_prob_uniform_taichi = mv_prob_uniform
mv_prob_normal_taichi = mv_prob_normal


def get_conn_matrix(conn_prob, seed, *, shape, transpose=False, outdim_parallel=True, layout="lane32"):
    """Dense 0/1 float32 matrix (of the op's own orientation) of the generated
    connectivity, recovered on the GPU with one-hot probes of
    :func:`mv_prob_homo` (weight 1): column c of M is M @ e_c."""
    num_in = shape[0] if transpose else shape[1]
    num_out = shape[1] if transpose else shape[0]
    dev = _lib.device()
    cols = []
    for c in range(num_in):
        e = torch.zeros(num_in, dtype=torch.float32, device=dev)
        e[c] = 1.0
        cols.append(mv_prob_homo(e, 1.0, conn_prob, seed, shape=shape, transpose=transpose,
                                 outdim_parallel=outdim_parallel, layout=layout))
    return torch.stack(cols, dim=1).reshape(num_out, num_in)
