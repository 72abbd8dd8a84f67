"""Host side of the XLA custom-call adapters (``bpx_xla_*`` in libbpx.so).

The reference registers its GPU kernels as legacy XLA custom calls
(``xla_client.register_custom_call_target(name, capsule, platform)``,
brainpy/_src/dependency_check.py:75-79) and passes shapes in an opaque
string built by brainpylib (``build_double_size_descriptor(m, n)``,
event/_csr_matvec.py:485; ``build_nonzero_descriptor(n, batch)``,
event/_info_collection.py:100).  With this library in brainpylib's place a
lowering registers :func:`capsule` of the adapter under the reference's
target names and builds the opaque with the ``*_descriptor`` functions here
(they also carry the dtypes and lengths the kernels need).

jax is not required by this module: it only packs descriptors and makes the
PyCapsules; :func:`call` invokes an adapter directly with a ``void*`` buffer
array (the way XLA does), which the tests use.
"""
import ctypes

from . import _lib

__all__ = ["XlaDesc", "event_csrmv_descriptor", "csrmv_descriptor", "event_info_descriptor", "jitconn_descriptor",
           "capsule", "call", "last_status", "TARGETS"]

XLA_MAGIC = 0x31585042

# adapter -> reference custom-call targets it stands in for
TARGETS = {
    "bpx_xla_event_csrmv": "event_csrmv_{homo,heter}_{float,double}_{int,long}{_bool,}{_transpose,}",
    "bpx_xla_csrmv": "csrmv_{homo,heter}_vector_{float,double}_{int,long}",
    "bpx_xla_event_info": "nonzero_{bool,int,long,float,double}",
    "bpx_xla_jitconn_mv": "gpu_jit_{event_,}csrmv_{atomic_,}prob_{homo,uniform,normal}_v2_*",
}


class XlaDesc(ctypes.Structure):
    """Mirror of bpx_xla_desc (include/bpx.h)."""
    _fields_ = [("magic", ctypes.c_uint32), ("data_dtype", ctypes.c_int32), ("index_dtype", ctypes.c_int32),
                ("vec_dtype", ctypes.c_int32), ("rule", ctypes.c_int32), ("transpose", ctypes.c_int32),
                ("dist", ctypes.c_int32), ("outdim_parallel", ctypes.c_int32), ("layout", ctypes.c_int32),
                ("event", ctypes.c_int32), ("m", ctypes.c_int64), ("n", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("data_len", ctypes.c_int64), ("batch", ctypes.c_int64), ("ws_bytes", ctypes.c_int64)]


def _pack(**kw):
    d = XlaDesc(magic=XLA_MAGIC)
    for k, v in kw.items():
        setattr(d, k, int(v))
    return ctypes.string_at(ctypes.addressof(d), ctypes.sizeof(d))


def event_csrmv_descriptor(shape, nnz, data_len, data_dtype, index_dtype, event_dtype, transpose,
                           rule=_lib.RULE_POSITIVE):
    """Opaque of bpx_xla_event_csrmv: the matrix shape (as brainpylib's
    double-size descriptor), len(indices), len(data) (1 = homogeneous) and
    the bpx dtype codes of data, indices/indptr and events."""
    return _pack(m=shape[0], n=shape[1], nnz=nnz, data_len=data_len, data_dtype=data_dtype,
                 index_dtype=index_dtype, vec_dtype=event_dtype, transpose=bool(transpose), rule=rule)


def csrmv_descriptor(shape, nnz, data_len, data_dtype, index_dtype, vector_dtype, transpose):
    """Opaque of bpx_xla_csrmv (sparse csrmv with a dense vector)."""
    return _pack(m=shape[0], n=shape[1], nnz=nnz, data_len=data_len, data_dtype=data_dtype,
                 index_dtype=index_dtype, vec_dtype=vector_dtype, transpose=bool(transpose))


def event_info_descriptor(n, batch, event_dtype):
    """Opaque of bpx_xla_event_info (brainpylib's nonzero descriptor plus the
    event dtype) and the byte size of its workspace output."""
    ws = int(_lib.lib().bpx_event_info_workspace(batch, n))
    return _pack(n=n, batch=batch, vec_dtype=event_dtype, ws_bytes=ws), ws


def jitconn_descriptor(num_row, num_col, dist, event, vec_dtype, outdim_parallel, layout=_lib.LAYOUT_LANE32,
                       weight_dtype=_lib.BPX_F32, workspace=True):
    """Opaque of bpx_xla_jitconn_mv for the matrix (num_row, num_col) after
    the transpose swap (the reference's (n_row, n_col) descriptor), and the
    byte size of its optional bitmask workspace output (0 when not used)."""
    ws = int(_lib.lib().bpx_jitconn_workspace(num_row, num_col)) if workspace else 0
    return _pack(m=num_row, n=num_col, dist=dist, event=bool(event), vec_dtype=vec_dtype,
                 outdim_parallel=bool(outdim_parallel), layout=layout, data_dtype=weight_dtype, ws_bytes=ws), ws


def _fn(name):
    if name not in TARGETS:
        raise ValueError(f"unknown adapter {name}")
    return getattr(_lib.lib(), name)


def capsule(name):
    """PyCapsule of the adapter `name` as XLA's register_custom_call_target
    expects it (capsule name "xla._CUSTOM_CALL_TARGET")."""
    fn = _fn(name)
    addr = ctypes.cast(fn, ctypes.c_void_p).value
    new = ctypes.pythonapi.PyCapsule_New
    new.restype = ctypes.py_object
    new.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]
    return new(addr, b"xla._CUSTOM_CALL_TARGET", None)


def call(name, stream, buffers, opaque):
    """Invoke adapter `name` the way XLA does: a stream, an array of device
    pointers (operands then outputs) and the opaque bytes.  Returns the
    adapter's status (bpx_xla_last_status)."""
    fn = _fn(name)
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    fn.restype = None
    arr = (ctypes.c_void_p * len(buffers))(*[b if isinstance(b, int) or b is None else b.data_ptr()
                                              for b in buffers])
    fn(stream, ctypes.cast(arr, ctypes.c_void_p), opaque, len(opaque))
    return last_status()


def last_status():
    L = _lib.lib()
    L.bpx_xla_last_status.restype = ctypes.c_int
    return int(L.bpx_xla_last_status())
