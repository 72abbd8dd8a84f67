"""The XLA custom-call adapters' descriptor plumbing, without a GPU: the
host-side descriptor must agree with the C struct (a wrong field order or
size would misread shapes), and invalid descriptors are rejected with a
status before anything is launched."""
import ctypes

import pytest

from brainpy_amd import _lib, xla

SHAPE, DTYPE, ARG = -2, -1, -5


def test_descriptor_size_and_magic():
    assert ctypes.sizeof(xla.XlaDesc) == 88
    op = xla.event_csrmv_descriptor((3, 4), 5, 5, _lib.BPX_F32, _lib.BPX_I32, _lib.BPX_BOOL, True)
    assert len(op) == 88 and op[:4] == b"BPX1"


def _csr(**over):
    kw = dict(shape=(4, 5), nnz=10, data_len=10, data_dtype=_lib.BPX_F32, index_dtype=_lib.BPX_I32,
              event_dtype=_lib.BPX_BOOL, transpose=True)
    kw.update(over)
    return xla.event_csrmv_descriptor(**kw)


@pytest.mark.parametrize("opaque,status", [
    (b"\0" * 88, ARG),                     # no magic
    (b"BPX1" + b"\0" * 80, ARG),           # short
    (_csr(shape=(-1, 5)), SHAPE),          # m < 0
    (_csr(shape=(4, -5)), SHAPE),          # n < 0
    (_csr(data_len=3), SHAPE),             # data_len neither 1 nor nnz
    (_csr(rule=7), ARG),                   # rule
    (_csr(), ARG),                         # null buffers with work to do
])
def test_event_csrmv_adapter_rejects(opaque, status):
    assert xla.call("bpx_xla_event_csrmv", None, [None] * 5, opaque) == status


def test_csrmv_and_info_adapters_reject():
    op = xla.csrmv_descriptor((4, 5), 10, 10, _lib.BPX_F32, _lib.BPX_I32, _lib.BPX_F64, False)
    assert xla.call("bpx_xla_csrmv", None, [None] * 5, op) == DTYPE      # vector dtype != f32
    op = xla.csrmv_descriptor((4, 5), 10, 2, _lib.BPX_F32, _lib.BPX_I32, _lib.BPX_F32, False)
    assert xla.call("bpx_xla_csrmv", None, [None] * 5, op) == SHAPE
    op, ws = xla.event_info_descriptor(100, 70000, _lib.BPX_BOOL)
    assert xla.call("bpx_xla_event_info", None, [None] * 4, op) == SHAPE  # batch > 65535
    op, ws = xla.event_info_descriptor(100, 2, _lib.BPX_BOOL)
    assert ws == int(_lib.lib().bpx_event_info_workspace(2, 100)) > 0
    assert xla.call("bpx_xla_event_info", None, [None] * 4, op) == ARG    # null buffers


def test_capsule():
    cap = xla.capsule("bpx_xla_event_csrmv")
    get = ctypes.pythonapi.PyCapsule_GetPointer
    get.restype = ctypes.c_void_p
    get.argtypes = [ctypes.py_object, ctypes.c_char_p]
    addr = get(cap, b"xla._CUSTOM_CALL_TARGET")
    assert addr == ctypes.cast(_lib.lib().bpx_xla_event_csrmv, ctypes.c_void_p).value
    with pytest.raises(ValueError):
        xla.capsule("bpx_event_csrmv")
