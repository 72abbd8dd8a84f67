"""Loader of ``libbpx.so`` (the gfx950 C-ABI library) and tensor plumbing.

There is no CPU fallback: every operator of this package runs on the HIP
library, and loading fails loudly when the library or the GPU is missing.
PyTorch is used only to hold device memory and to supply the current stream.
"""
import ctypes
import os

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbpx.so")

# dtype codes of include/bpx.h
BPX_BOOL, BPX_F32, BPX_F64, BPX_I32, BPX_I64, BPX_BF16, BPX_F16 = 0, 1, 2, 3, 4, 5, 6
RULE_POSITIVE, RULE_NONZERO = 0, 1
DIST_HOMO, DIST_UNIFORM, DIST_NORMAL = 0, 1, 2
LAYOUT_LANE32, LAYOUT_ROW = 0, 1

_TORCH2BPX = {
    torch.bool: BPX_BOOL, torch.float32: BPX_F32, torch.float64: BPX_F64, torch.int32: BPX_I32,
    torch.int64: BPX_I64, torch.bfloat16: BPX_BF16, torch.float16: BPX_F16,
}

_lib = None


class TiledCSRPlan(ctypes.Structure):
    """Mirror of bpx_tiled_csr (include/bpx.h)."""
    _fields_ = [("m", ctypes.c_int64), ("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("tiles", ctypes.c_int32),
                ("tile_len", ctypes.c_int32), ("weight_dtype", ctypes.c_int32), ("fx_shift", ctypes.c_int32),
                ("coarse", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("capacity", ctypes.c_int64), ("rowtile_offset", ctypes.c_int64), ("aux_offset", ctypes.c_int64),
                ("lidx_offset", ctypes.c_int64),
                ("weights_offset", ctypes.c_int64), ("homo_offset", ctypes.c_int64), ("mem", ctypes.c_void_p)]


TILED_HOMO = -1

# function name -> (argtypes, restype)
_vp, _i64, _i32, _u32, _f32, _sz = (ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32,
                                     ctypes.c_float, ctypes.c_size_t)
SIGNATURES = {
    "bpx_status_string": ([_i32], ctypes.c_char_p),
    "bpx_version": ([], ctypes.c_char_p),
    "bpx_event_info_workspace": ([_i64, _i64], _sz),
    "bpx_event_info": ([_vp, _vp, _i32, _i64, _i64, _vp, _vp, _vp, _sz], _i32),
    "bpx_event_csrmv": ([_vp, _vp, _i32, _i64, _vp, _vp, _i32, _i64, _vp, _i32, _i32, _i64, _i64, _i32, _vp], _i32),
    "bpx_csrmv": ([_vp, _vp, _i32, _i64, _vp, _vp, _i32, _i64, _vp, _i32, _i64, _i64, _i32, _vp], _i32),
    "bpx_event_csrmv_fixed_order": ([_vp, _vp, _i32, _i64, _vp, _vp, _i64, _vp, _i32, _i32, _i64, _i64, _i32, _vp],
                                    _i32),
    "bpx_event_csrmv_add": ([_vp, _vp, _i32, _i64, _vp, _vp, _i32, _i64, _vp, _i32, _i32, _i64, _i64, _vp], _i32),
    "bpx_jitconn_mv": ([_vp, _vp, _i32, _i32, _i32, _f32, _f32, ctypes.c_int32, _u32, _i64, _i64, _i32, _i32, _vp,
                        _vp, _sz], _i32),
    "bpx_jitconn_workspace": ([_i64, _i64], _sz),
    "bpx_jitconn_workspace_ex": ([_i64, _i64, ctypes.c_int32, ctypes.c_int], _sz),
    "bpx_lif_coba_step": ([_vp, _i64, _i32, _i32, _vp, _vp, _vp, ctypes.POINTER(_vp), _vp, _vp, _vp, _f32, _vp,
                           _f32, _f32, _f32, _f32, _f32, _f32, _f32, _f32], _i32),
    "bpx_csr_on_pre_update": ([_vp, _vp, _i32, _vp, _vp, _i32, _i64, _vp, _i32, _i64, _vp, _i64,
                               ctypes.c_double, ctypes.c_double, _vp], _i32),
    "bpx_csc_on_post_update": ([_vp, _vp, _i32, _vp, _vp, _vp, _i32, _i64, _vp, _i32, _i64, _vp, _i64,
                                ctypes.c_double, ctypes.c_double, _vp], _i32),
    "bpx_dense_on_pre": ([_vp, _vp, _i32, _vp, _i32, _i64, _i64, _vp, ctypes.c_double, ctypes.c_double, _vp], _i32),
    "bpx_dense_on_post": ([_vp, _vp, _i32, _vp, _i32, _i64, _i64, _vp, ctypes.c_double, ctypes.c_double, _vp], _i32),
    "bpx_coomv": ([_vp, _vp, _i32, _i64, _vp, _vp, _i32, _i64, _vp, _i32, _i64, _i64, _i32, _vp], _i32),
    "bpx_coo_matmat": ([_vp, _vp, _i32, _i64, _vp, _vp, _i32, _i64, _vp, _i32, _i64, _i64, _i64, _i32, _vp], _i32),
    "bpx_coo_sddmm": ([_vp, _vp, _vp, _i32, _i64, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64, _i64, _i32, _vp],
                      _i32),
    "bpx_bsrmv": ([_vp, _vp, _i32, _vp, _vp, _i32, _i64, _i64, _vp, _i64, _i64, _i32, _vp], _i32),
    "bpx_bsr_sddmm": ([_vp, _vp, _vp, _i32, _i64, _i64, _vp, _vp, _i32, _i64, _i64, _vp], _i32),
    "bpx_bcsrmm": ([_vp, _vp, _i32, _vp, _vp, _vp, _i32, _i64, _i64, _i64, _i64, _vp], _i32),
    "bpx_fixed_prob_workspace": ([_i64], _sz),
    "bpx_fixed_prob_csr": ([_vp, _i64, _i64, _i64, ctypes.c_uint64, _i32, _i32, _vp, _vp, _vp, _sz], _i32),
    "bpx_tiled_csr_layout": ([_i64, _i64, _i64, _i32, ctypes.POINTER(TiledCSRPlan), ctypes.POINTER(_sz)], _i32),
    "bpx_tiled_csr_build": ([_vp, _vp, _i32, _i64, _vp, _vp, _i32, ctypes.POINTER(TiledCSRPlan), _vp, _sz], _i32),
    "bpx_event_csrmv_tiled": ([_vp, ctypes.POINTER(TiledCSRPlan), _vp, _i32, _i32, _i32, _vp], _i32),
    "bpx_csrmv_tiled": ([_vp, ctypes.POINTER(TiledCSRPlan), _vp, _i32, _i32, _vp, _vp, _sz], _i32),
    "bpx_tiled_csr_gather_workspace": ([ctypes.POINTER(TiledCSRPlan)], _sz),
    "bpx_tiled_csr_gather": ([_vp, ctypes.POINTER(TiledCSRPlan), _vp, _i32, _i32, _i32, _vp, _vp, _sz], _i32),
    "bpx_event_csrmv_tiled_batch": ([_vp, ctypes.POINTER(TiledCSRPlan), _vp, _i32, _i32, _i64, _i32, _vp], _i32),
    "bpx_tiled_csr_call_workspace": ([ctypes.POINTER(TiledCSRPlan), _i64], _sz),
    "bpx_event_csrmv_tiled_ws": ([_vp, ctypes.POINTER(TiledCSRPlan), _vp, _i32, _i32, _i64, _i32, _vp, _vp, _sz],
                                 _i32),
    "bpx_batch_workspace": ([_i64], _sz),
    "bpx_event_csrmv_batch": ([_vp, _vp, _i32, _i64, _vp, _vp, _i32, _i64, _vp, _i32, _i32, _i64, _i64, _i64, _i32,
                               _vp, _vp, _sz], _i32),
    "bpx_jitconn_mv_batch": ([_vp, _vp, _i32, _i32, _f32, _f32, ctypes.c_int32, _u32, _i64, _i64, _i64, _i32, _i32,
                              _vp, _vp, _sz], _i32),
    "bpx_expon_coef": ([_f32, _f32], _f32),
    "bpx_lif_coba_step_dev": ([_vp, _i64, _i32, _i32, _vp, _vp, _vp, ctypes.POINTER(_vp), _vp, _vp, _vp, _f32, _vp,
                               _f32, _f32, _f32, _f32, _f32, _f32, _vp, _i64, _f32], _i32),
    "bpx_timer_enable": ([_i32], _i32),
    "bpx_timer_read": ([_i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)], _i32),
    "bpx_timer_reset": ([], _i32),
}


class BPXError(RuntimeError):
    pass


def lib():
    """The loaded libbpx.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BPXError(f"{LIB_PATH} is missing: build it with `python -m brainpy_amd.csrc.build` "
                           "(brainpy_amd has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (argtypes, restype) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _lib = L
    return _lib


def check(rc, op):
    if rc != 0:
        raise BPXError(f"{op} failed: {lib().bpx_status_string(rc).decode()} ({rc})")


_gpu_ok = False


def device():
    global _gpu_ok
    if not _gpu_ok:
        if not torch.cuda.is_available():
            raise BPXError("brainpy_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback")
        _gpu_ok = True
    return torch.device("cuda", torch.cuda.current_device())


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_ptr(dev=None):
    """The current HIP stream of `dev` (default: the current device) as a
    pointer, the stream every launch of this package goes to."""
    if _raw_stream is not None:
        idx = torch.cuda.current_device() if dev is None else torch.device(dev).index
        if idx is None:
            idx = torch.cuda.current_device()
        return ctypes.c_void_p(_raw_stream(idx))
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def raw_stream(index):
    """The current stream of device `index` as an int (the fast form of
    stream_ptr for launches on a known device)."""
    if _raw_stream is not None:
        return _raw_stream(index)
    return torch.cuda.current_stream(index).cuda_stream


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def bpx_dtype(t):
    try:
        return _TORCH2BPX[t.dtype]
    except KeyError:
        raise TypeError(f"unsupported dtype {t.dtype}")


def to_device(x, dtype=None):
    """Array-like -> contiguous torch tensor on the current GPU (no copy when
    it already is one)."""
    dev = device()
    if isinstance(x, torch.Tensor):
        t = x
    elif isinstance(x, np.ndarray):
        if x.dtype == np.uint32:
            x = x.astype(np.int64)
        elif x.dtype == np.uint64:
            x = x.astype(np.int64)
        t = torch.from_numpy(np.ascontiguousarray(x))
    elif isinstance(x, (bool, np.bool_)):
        t = torch.tensor(x, dtype=torch.bool)
    elif isinstance(x, (int, np.integer)):
        t = torch.tensor(x, dtype=torch.int32)
    elif isinstance(x, (float, np.floating)):
        t = torch.tensor(x, dtype=torch.float32)
    else:
        t = torch.as_tensor(np.asarray(x))
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if t.device != dev:
        t = t.to(dev, non_blocking=False)
    return t.contiguous()


def written(t):
    """Record that a kernel wrote into `t` through its raw pointer: bumps
    the tensor's version counter as an in-place torch op would, so caches
    keyed on it (the event csrmv plan cache) see the new contents."""
    torch.autograd.graph.increment_version(t)


def timer_enable(on=True):
    check(lib().bpx_timer_enable(int(on)), "bpx_timer_enable")


def timer_reset():
    check(lib().bpx_timer_reset(), "bpx_timer_reset")


def timer_read(kernel_id):
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    check(lib().bpx_timer_read(kernel_id, ctypes.byref(ms), ctypes.byref(n)), "bpx_timer_read")
    return ms.value, n.value


KID = dict(event_scatter=0, event_gather=1, csr_scatter=2, csr_gather=3, jit_outdim=4, jit_scatter=5, info=6,
           lif=7, fixed_order=8, tiled_list=9, net=10, tiled=11, coo=12, bsr=13, net_run=14, tiled_gather=15, jit_emit=16)
