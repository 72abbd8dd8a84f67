"""bpx_event_info_noalloc (the entry point that allocates nothing: its
workspace is a static device array) against bpx_event_info with a caller
workspace, through the C ABI: two calls in a row, the second smaller, so a
stale scratch from the first call would show; and the size limit."""
import ctypes
import os

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "brainpy_amd", "libbpx.so")
WS_ERR = -3


def _lib():
    L = ctypes.CDLL(LIB)
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    L.bpx_event_info_noalloc.argtypes = [vp, vp, ctypes.c_int, i64, i64, vp, vp]
    L.bpx_event_info.argtypes = [vp, vp, ctypes.c_int, i64, i64, vp, vp, vp, ctypes.c_size_t]
    L.bpx_event_info_workspace.argtypes = [i64, i64]
    L.bpx_event_info_workspace.restype = ctypes.c_size_t
    return L


def test_noalloc_rejects_beyond_scratch_without_launch():
    L = _lib()
    # 65535 rows of 100,000 events need 65535 * 25 * 8 bytes > 4 MiB
    assert L.bpx_event_info_workspace(65535, 100000) > (4 << 20)
    assert L.bpx_event_info_noalloc(None, 1, 0, 65535, 100000, 1, 1) == WS_ERR


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.bool, torch.float32])
def test_noalloc_equals_workspace_entry_point(dtype):
    L = _lib()
    dev = torch.device("cuda")
    sp = torch.cuda.current_stream().cuda_stream
    code = {torch.bool: 0, torch.float32: 1}[dtype]
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    # a large call, then a smaller one: the second must not see the first's scratch
    # (the 5e7-event row's fast-path bits exceed the scratch: the no-allocation
    # call takes the minimum-workspace path, the workspace call the fast path)
    for batch, n, rate in ((3, 2_000_000, 0.1), (1, 50_000_000, 0.01), (2, 50_000, 0.3), (1, 4_096 * 7 + 5, 0.02)):
        ev = (torch.rand((batch, n), generator=g, device=dev) < rate).to(dtype)
        ids_a = torch.full((batch, n), 7, dtype=torch.int32, device=dev)
        num_a = torch.full((batch,), 7, dtype=torch.int32, device=dev)
        ids_b = torch.full((batch, n), 9, dtype=torch.int32, device=dev)
        num_b = torch.full((batch,), 9, dtype=torch.int32, device=dev)
        wsb = int(L.bpx_event_info_workspace(batch, n))
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        assert L.bpx_event_info_noalloc(sp, ev.data_ptr(), code, batch, n, ids_a.data_ptr(), num_a.data_ptr()) == 0
        assert L.bpx_event_info(sp, ev.data_ptr(), code, batch, n, ids_b.data_ptr(), num_b.data_ptr(),
                                ws.data_ptr(), wsb) == 0
        torch.cuda.synchronize()
        assert torch.equal(ids_a, ids_b) and torch.equal(num_a, num_b)
        # and both equal the serial definition (ascending ids, then -1)
        evh = ev.cpu().numpy() != 0
        for b in range(batch):
            nz = np.flatnonzero(evh[b])
            assert int(num_a[b]) == nz.size
            row = ids_a[b].cpu().numpy()
            np.testing.assert_array_equal(row[:nz.size], nz)
            assert np.all(row[nz.size:] == -1)
