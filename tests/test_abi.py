"""The C-ABI library loads and exports every entry point include/bpx.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bpx.h")
LIB = os.path.join(ROOT, "brainpy_amd", "libbpx.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bpx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("bpx_event_csrmv", "bpx_csrmv", "bpx_event_info", "bpx_jitconn_mv", "bpx_lif_coba_step",
                 "bpx_event_csrmv_fixed_order", "bpx_status_string"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbpx.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbpx.so not built")
def test_measurement_probes_stay_out_of_the_product_library():
    lib = ctypes.CDLL(LIB)
    for name in ("bpx_coo_probe", "bpx_copy", "bpx_probe_coo", "bpx_probe_copy"):
        assert not hasattr(lib, name), name
    probe = os.path.join(ROOT, "brainpy_amd", "libbpx_probe.so")
    if os.path.exists(probe):
        P = ctypes.CDLL(probe)
        f = P.bpx_probe_coo
        f.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_int64]
        assert f(None, None, None, None, None, 1, 10, 1, None, 256) == -2       # nse % 4
        assert f(None, None, None, None, None, 1, 8, 1, None, 100) == -2        # out_len % 256
        assert f(None, None, None, None, None, 1, 8, 3, None, 256) == -2        # mode
        assert f(None, None, None, None, None, 1, 8, 1, None, 256) == -5        # null arrays
        g = P.bpx_probe_copy
        g.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        assert g(None, None, None, -1) == -5 and g(None, None, None, 0) == 0


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbpx.so not built")
def test_host_only_entry_points():
    lib = ctypes.CDLL(LIB)
    lib.bpx_status_string.restype = ctypes.c_char_p
    assert lib.bpx_status_string(0) == b"ok"
    assert b"dtype" in lib.bpx_status_string(-1)
    lib.bpx_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.bpx_version()
    lib.bpx_event_info_workspace.restype = ctypes.c_size_t
    lib.bpx_event_info_workspace.argtypes = [ctypes.c_int64, ctypes.c_int64]
    # tile counts and offsets (8 B per 4096-event tile and row), then the
    # bool fast path's event bits (512 B per tile and row)
    assert lib.bpx_event_info_workspace(1, 4096) == 8 + 512
    assert lib.bpx_event_info_workspace(2, 4097) == 32 + 2048
    # jitconn: bitmask or active-column list; the two-pass scatter's lists
    # (1M x 1M, clen 19999: 8 row tiles of 3907 rows per stream range, 2048
    # lists of 8064 connections: 2 B each, +4 B with weights)
    lib.bpx_jitconn_workspace.restype = ctypes.c_size_t
    lib.bpx_jitconn_workspace.argtypes = [ctypes.c_int64, ctypes.c_int64]
    lib.bpx_jitconn_workspace_ex.restype = ctypes.c_size_t
    lib.bpx_jitconn_workspace_ex.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int]
    n = 1_000_000
    assert lib.bpx_jitconn_workspace(n, n) == 256 + 4 * n
    cap = int(131072 * 3907 * 2.0 / 20000 * 0.125 * 1.25 + 64.0)
    rows_off = (256 + 2048 * 4 + 255) // 256 * 256
    w_off = (rows_off + 2048 * cap * 2 + 255) // 256 * 256
    assert lib.bpx_jitconn_workspace_ex(n, n, 19999, 0) == w_off
    assert lib.bpx_jitconn_workspace_ex(n, n, 19999, 1) == w_off + 2048 * cap * 4
    # no two-pass form: an unknown distribution, or lists beyond 1 GiB
    assert lib.bpx_jitconn_workspace_ex(n, n, 19999, 7) == 256 + 4 * n
    assert lib.bpx_jitconn_workspace_ex(n, n, 1, 0) == 256 + 4 * n


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbpx.so not built")
def test_argument_validation_without_launch():
    # invalid arguments are rejected before anything touches the device
    lib = ctypes.CDLL(LIB)
    f = lib.bpx_jitconn_mv
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                  ctypes.c_float, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                  ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    assert f(None, None, 1, 1, 0, 1.0, 0.0, 0, 1, 10, 10, 1, 0, None, None, 0) == -5       # clen < 1
    assert f(None, None, 1, 1, 7, 1.0, 0.0, 3, 1, 10, 10, 1, 0, None, None, 0) == -5       # bad dist
    assert f(None, None, 2, 1, 0, 1.0, 0.0, 3, 1, 10, 10, 1, 0, None, None, 0) == -1       # f64 events
    g = lib.bpx_event_csrmv
    g.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                  ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
    assert g(None, None, 1, 5, None, None, 3, 7, None, 0, 0, 3, 3, 1, None) == -2  # data_len mismatch
    assert g(None, None, 1, 1, None, None, 3, 7, None, 0, 9, 3, 3, 1, None) == -5  # bad rule


def test_fastmod_exact():
    # the u32 division by an invariant divisor used for `bits % clen`
    # (brainpy_amd/csrc/bpx_rng.h) is exact: compile its host side and compare
    import subprocess
    import tempfile
    src = r'''
#include <stdio.h>
#include <stdint.h>
#include "bpx_rng.h"
int main(){
  uint32_t ds[] = {1u,2u,3u,7u,19u,99u,199u,19999u,65535u,65536u,1000003u,2147483647u,2147483648u,4294967295u};
  uint64_t x = 88172645463325252ull; long bad = 0;
  for (unsigned di = 0; di < sizeof(ds)/4; ++di) {
    bpx::DivU32 dv = bpx::DivU32::make(ds[di]);
    for (int i = 0; i < 200000; ++i) {
      x ^= x << 13; x ^= x >> 7; x ^= x << 17;
      uint32_t n = (uint32_t)x;
      if (i < 64) n = i; if (i >= 64 && i < 128) n = 0xFFFFFFFFu - (i - 64);
      if (dv.mod(n) != n % ds[di] || dv.div(n) != n / ds[di]) ++bad;
    }
  }
  printf("%ld\n", bad); return bad != 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.cpp")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "brainpy_amd", "csrc"), c, "-o", exe])
        out = subprocess.run([exe], capture_output=True, text=True)
        assert out.returncode == 0 and out.stdout.strip() == "0", out.stdout


def test_short_divisor_exact():
    # the short-form division q = mulhi(n, m) >> s (bpx_rng.h DivShort) is
    # exact for every divisor it accepts: random numerators, both ends of the
    # u32 range and the neighbours of multiples of d, for every d in 2..3000
    # and a spread of larger ones (19999 = conn_len(1e-4) among them)
    import subprocess
    import tempfile
    src = r'''
#include <stdio.h>
#include <stdint.h>
#include "bpx_rng.h"
int main(){
  uint64_t x = 88172645463325252ull; long bad = 0, accepted = 0, tested = 0;
  for (uint32_t d = 2; d < 3000000u; d = d < 3000u ? d + 1 : d * 3 / 2 + 7) {
    bpx::DivShort ds;
    uint32_t list[] = {d, 19999u, 99999u};
    for (uint32_t dd : list) {
      if (!bpx::DivShort::make(dd, &ds)) continue;
      ++accepted;
      for (int i = 0; i < 4000; ++i) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        uint32_t n = (uint32_t)x;
        if (i < 32) n = i;
        else if (i < 64) n = 0xFFFFFFFFu - (i - 32);
        else if (i < 128) { uint64_t k = (x >> 32) % (0xFFFFFFFFull / dd); n = (uint32_t)(k * dd + (i & 1 ? 0 : dd - 1)); }
        ++tested;
        uint32_t q = ds.div(n);
        uint32_t r = n + (uint32_t)((int64_t)q * -(int64_t)dd);   // the device's 24-bit multiply-add
        if (q != n / dd) ++bad;
        if (bpx::DivShort::mad24_ok(dd) && r != n % dd) ++bad;
        if (ds.mod24(n) != n % dd) ++bad;
      }
    }
  }
  printf("%ld %ld %ld\n", bad, accepted, tested); return bad != 0 || accepted < 1000;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.cpp")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "brainpy_amd", "csrc"), c, "-o", exe])
        out = subprocess.run([exe], capture_output=True, text=True)
        assert out.returncode == 0 and out.stdout.split()[0] == "0", out.stdout


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbpx.so not built")
def test_argument_validation_of_coo_conn_and_plan_entry_points():
    # shape, dtype and argument errors come back as status codes before
    # anything is launched (no GPU needed)
    from brainpy_amd import _lib
    L = _lib.lib()
    SHAPE, DTYPE, ARG, WS = -2, -1, -5, -3
    F32, I32, F64, BF16 = _lib.BPX_F32, _lib.BPX_I32, _lib.BPX_F64, _lib.BPX_BF16
    assert L.bpx_coomv(None, None, F32, 1, None, None, I32, 10, None, F32, -1, 5, 0, None) == SHAPE
    assert L.bpx_coomv(None, None, F32, 3, None, None, I32, 10, None, F32, 4, 5, 0, None) == SHAPE   # data_len
    assert L.bpx_coomv(None, None, F32, 1, None, None, I32, 10, None, F64, 4, 5, 0, None) == DTYPE   # vector dtype
    assert L.bpx_coomv(None, None, BF16, 1, None, None, I32, 10, None, BF16, 4, 5, 0, None) == DTYPE
    assert L.bpx_coo_matmat(None, None, F32, 1, None, None, I32, 10, None, F32, 4, 5, -1, 1, None) == SHAPE
    assert L.bpx_coo_matmat(None, None, F32, 1, None, None, I32, 10, None, I32, 4, 5, 2, 1, None) == DTYPE
    assert L.bpx_coo_sddmm(None, None, None, I32, -1, None, 1, 1, 0, None, 1, 1, 0, 1, F32, None) == SHAPE
    assert L.bpx_fixed_prob_csr(None, 10, 5, 6, 1, 1, 0, None, None, None, 0) == SHAPE         # k > n
    assert L.bpx_fixed_prob_csr(None, 10, 20, 5, 1, 0, 0, None, None, None, 0) == SHAPE        # no self, m != n
    assert L.bpx_fixed_prob_csr(None, 10, 2_000_000, 5, 1, 1, 0, None, None, None, 0) == ARG   # null outputs
    assert L.bpx_fixed_prob_workspace(7) == 28
    assert L.bpx_csrmv_tiled(None, None, None, F32, 0, None, None, 0) == ARG
    plan = _lib.TiledCSRPlan()
    plan.mem = 1
    assert L.bpx_csrmv_tiled(None, ctypes.byref(plan), None, F64, 0, None, None, 0) == DTYPE
    assert L.bpx_csrmv_tiled(None, ctypes.byref(plan), None, F32, 0, None, None, 0) == WS
    # the workspace holds max and min |x|: 8 bytes (checked before any launch)
    assert L.bpx_csrmv_tiled(None, ctypes.byref(plan), None, F32, 0, None, ctypes.c_void_p(256), 4) == WS
    # the active-row list workspace: none up to 64 tiles; beyond, the
    # counters (a 256-byte multiple) then [batch][m] rows
    plan.m, plan.tiles = 10, 64
    assert L.bpx_tiled_csr_call_workspace(ctypes.byref(plan), 1) == 0
    plan.tiles = 72
    assert L.bpx_tiled_csr_call_workspace(ctypes.byref(plan), 1) == 256 + 40
    assert L.bpx_tiled_csr_call_workspace(ctypes.byref(plan), 3) == 256 + 120
    assert L.bpx_tiled_csr_call_workspace(ctypes.byref(plan), 0) == 0
    assert L.bpx_tiled_csr_call_workspace(None, 1) == 0
    BOOL = _lib.BPX_BOOL
    ws = ctypes.c_void_p(256)
    assert L.bpx_event_csrmv_tiled_ws(None, None, None, BOOL, 0, 1, 0, None, ws, 296) == ARG
    assert L.bpx_event_csrmv_tiled_ws(None, ctypes.byref(plan), None, BOOL, 7, 1, 0, None, ws, 296) == ARG  # rule
    assert L.bpx_event_csrmv_tiled_ws(None, ctypes.byref(plan), None, F64, 0, 1, 0, None, ws, 296) == DTYPE
    assert L.bpx_event_csrmv_tiled_ws(None, ctypes.byref(plan), None, BOOL, 0, -1, 0, None, ws, 296) == SHAPE
    assert L.bpx_event_csrmv_tiled_ws(None, ctypes.byref(plan), None, BOOL, 0, 1, 0, None, ws, 296) == ARG  # no out


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbpx.so not built")
def test_argument_validation_of_bsr_entry_points():
    from brainpy_amd import _lib
    L = _lib.lib()
    SHAPE, DTYPE, ARG = -2, -1, -5
    F32, I32, BF16 = _lib.BPX_F32, _lib.BPX_I32, _lib.BPX_BF16
    assert L.bpx_bsrmv(None, None, F32, None, None, I32, 4, 0, None, 2, 2, 0, None) == SHAPE    # blocksize 0
    assert L.bpx_bsrmv(None, None, F32, None, None, I32, 4, 4, None, -1, 2, 0, None) == SHAPE   # mb < 0
    assert L.bpx_bsrmv(None, None, F32, None, None, I32, 4, 4, None, 2, 2, 0, None) == ARG      # null arrays
    dummy = ctypes.c_void_p(16)   # never dereferenced: the null vector is rejected first
    assert L.bpx_bsrmv(None, dummy, F32, dummy, dummy, I32, 4, 4, None, 2, 2, 0, dummy) == ARG  # null vector
    assert L.bpx_bsr_sddmm(None, None, None, I32, 4, 0, None, None, F32, 2, 2, None) == SHAPE
    assert L.bpx_bsr_sddmm(None, None, None, I32, 4, 4, None, None, F32, 2, 2, None) == ARG
    assert L.bpx_bcsrmm(None, None, F32, None, None, None, I32, 1, 8, 48, 64, None) == SHAPE    # k % 32
    assert L.bpx_bcsrmm(None, None, F32, None, None, None, I32, 1, 8, 64, 40, None) == SHAPE    # n % 32
    assert L.bpx_bcsrmm(None, None, F32, None, None, None, I32, 1, 8, 64, 64, None) == ARG      # null arrays
    assert L.bpx_bcsrmm(None, None, BF16, None, None, None, I32, 0, 0, 64, 64, None) == 0       # empty: no work


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbpx.so not built")
def test_null_arrays_are_rejected_before_launch():
    # an entry point with work to do and a null array returns BPX_ERR_ARG
    # instead of launching a kernel that would read through it
    from brainpy_amd import _lib
    L = _lib.lib()
    ARG = -5
    F32, I32, BOOL = _lib.BPX_F32, _lib.BPX_I32, _lib.BPX_BOOL
    d = ctypes.c_void_p(16)   # stand-in for a device pointer; never dereferenced
    assert L.bpx_event_csrmv(None, d, F32, 1, d, d, I32, 10, None, BOOL, 1, 4, 5, 1, d) == ARG    # events
    assert L.bpx_event_csrmv(None, d, F32, 1, d, d, I32, 10, d, BOOL, 1, 4, 5, 1, None) == ARG    # out
    assert L.bpx_csrmv(None, d, F32, 1, None, d, I32, 10, d, F32, 4, 5, 0, d) == ARG              # indices
    assert L.bpx_coomv(None, d, F32, 1, d, d, I32, 10, None, F32, 4, 5, 0, d) == ARG              # vector
    assert L.bpx_coo_matmat(None, d, F32, 1, d, d, I32, 10, d, F32, 4, 5, 3, 1, None) == ARG      # out
    assert L.bpx_event_info(None, None, BOOL, 1, 100, d, d, d, 1 << 20) == ARG                    # events
    assert L.bpx_jitconn_mv(None, None, BOOL, 1, 0, 1.0, 0.0, 3, 1, 10, 10, 1, 0, d, None, 0) == ARG


@pytest.mark.skipif(not os.path.exists(LIB), reason="libbpx.so not built")
def test_argument_validation_of_net_run():
    # the multi-step network launch: LDS limit per neuron and per in-degree,
    # rejected before launch; no steps or no networks is no work
    from brainpy_amd.dyn.batched import NetParams, _register
    L = _register()
    SHAPE, ARG = -2, -5
    assert L.bpx_net_run_max_neurons(2) == 5456       # (65536 - 64) / (4 * (n_buf + 1))
    assert L.bpx_net_run_max_neurons(26) == 606
    assert L.bpx_net_run_max_neurons(1) == 0
    d = ctypes.c_void_p(16)   # never dereferenced
    prm = NetParams()

    def run(nets=d, n_nets=1, max_n=4000, in_deg=80, n_buf=2, steps=10, p=ctypes.byref(prm), V=d):
        return L.bpx_net_run(None, nets, n_nets, max_n, in_deg, 1, 4000, V, d, d, d, d, n_buf, None, None, None,
                             None, p, 0, None, steps, None)

    assert run(nets=None) == ARG
    assert run(p=None) == ARG
    assert run(n_buf=1) == ARG
    assert run(steps=-1) == ARG
    assert run(max_n=5457) == SHAPE
    assert run(max_n=0) == SHAPE
    assert run(in_deg=65536) == SHAPE
    assert run(V=None) == ARG
    assert run(steps=0) == 0
    assert run(n_nets=0) == 0
    tiles = ctypes.c_void_p(32)
    assert L.bpx_net_step(None, d, tiles, 16, 256, 1, 4000, None, d, d, d, d, 2, None, None, None,
                          ctypes.byref(prm), 0, None) == ARG                       # null state
