"""Argument normalisation shared by the operator mirrors."""
import numpy as np
import torch

from .. import _lib

_INT_TORCH = (torch.int8, torch.int16, torch.int32, torch.int64, torch.uint8)
_FLOAT_TORCH = (torch.float16, torch.bfloat16, torch.float32, torch.float64)


def np_or_torch_dtype(x):
    if isinstance(x, torch.Tensor):
        return x.dtype
    if isinstance(x, np.ndarray):
        return x.dtype
    return None


def ndim(x):
    if isinstance(x, (torch.Tensor, np.ndarray)):
        return x.ndim
    return np.ndim(x)


def is_int_dtype(dt):
    if isinstance(dt, torch.dtype):
        return dt in _INT_TORCH
    return dt is not None and np.issubdtype(dt, np.integer)


def index_tensor(x):
    """indices / indptr -> int32 or int64 device tensor (uint32 and wider
    unsigned become int64; narrower ints become int32)."""
    t = _lib.to_device(x)
    if t.dtype in (torch.int32, torch.int64):
        return t
    if t.dtype in (torch.int8, torch.int16, torch.uint8):
        return t.to(torch.int32)
    return t.to(torch.int64)


def data_tensor(data):
    """Synaptic weights -> 1-D float device tensor (python floats are f32,
    the JAX default without x64)."""
    t = _lib.to_device(data)
    if t.ndim == 0:
        t = t.reshape(1)
    if not t.dtype.is_floating_point:
        t = t.to(torch.float32)
    return t


def acc_dtype(data_dtype):
    return torch.float64 if data_dtype == torch.float64 else torch.float32


def needs_ad(*xs):
    """True when an input takes part in differentiation or a torch.func
    transform (it requires grad, or it is a functorch-wrapped tensor): the
    operator then runs through its autograd Function; otherwise it launches
    directly."""
    wrapped = torch._C._functorch.is_functorch_wrapped_tensor
    for x in xs:
        if isinstance(x, torch.Tensor) and ((x.requires_grad and torch.is_grad_enabled()) or wrapped(x)):
            return True
    return False


def batch_rows(info, in_dims, args, fn):
    """vmap rule shared by the operators: run the batch one row at a time
    (the reference's batching rule for these primitives is a sequential scan
    over the batch, op_register/utils.py:16-37) and stack along dim 0."""
    outs = []
    for b in range(info.batch_size):
        row = [a.select(d, b) if (d is not None and isinstance(a, torch.Tensor)) else a for a, d in zip(args, in_dims)]
        outs.append(fn(*row))
    if isinstance(outs[0], tuple):
        return tuple(torch.stack(o) for o in zip(*outs)), tuple(0 for _ in outs[0])
    return torch.stack(outs), 0
