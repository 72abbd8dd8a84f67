"""Benchmark of the event-driven CSR scatter (bm.event.csrmv, transpose=True).

Workload (one "step" = one operator call over one spike vector):
  100,000 x 100,000 CSR, FixedProb p = 0.01 (1000 targets per row,
  nnz = 1e8, int32 indices), Bernoulli spike vectors at 0.1 / 1 / 5 / 10 %,
  weights homogeneous 1.0, heterogeneous normal(0.5, 0.1) f32, and the same
  heterogeneous weights stored in bf16 (accumulated in f32).
  Headline: 10 % spikes, heterogeneous f32 weights.  All inputs are
  synthetic, seeded and resident in HBM before the timed region.

Metric: synaptic events/s = (sum over active rows of their row length) per
second of whole-operator wall time (zero-fill + scatter), plus the scatter
kernel's algorithmic HBM bytes / its average device duration (HIP events on
the launch stream) against the 8 TB/s MI355X peak.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-sweep] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_MFMA_TFLOPS = 157.3   # dense f32 matrix-core peak
M = N = 100_000
PROB = 0.01
K_ROW = int(N * PROB)
RATES = (0.001, 0.01, 0.05, 0.10)
HEADLINE = ("heter_f32", 0.10)
NV = 16      # spike vectors per rate, taken in turn by the timed calls
TILES = 32   # output tiles per copy on MI355X (CUs per XCD); set from the plan at run time


_PROBE = None


def probe_lib():
    """libbpx_probe.so: measurement probes (a plain copy for counter
    calibration, the COO/gather access-pattern ceilings), built beside the
    product library but not part of it."""
    global _PROBE
    if _PROBE is None:
        import ctypes
        P = ctypes.CDLL(os.path.join(ROOT, "brainpy_amd", "libbpx_probe.so"))
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        P.bpx_probe_copy.argtypes = [vp, vp, vp, i64]
        P.bpx_probe_coo.argtypes = [vp, vp, vp, vp, vp, i64, i64, ctypes.c_int, vp, i64]
        _PROBE = P
    return _PROBE


def algorithmic_bytes(active_rows, kind):
    """Algorithmic bytes of one event csrmv transpose call, as the operator
    receives its inputs (whatever the kernel's own layout): events (1 B
    each), one 8-byte offset pair per active row, the 4-byte column indices
    and the weights of active rows, one 4-byte write of each output."""
    w_bytes = {"homo": 0, "heter_f32": 4, "heter_bf16": 2}[kind]
    return M * 1 + active_rows * 8 + active_rows * K_ROW * (4 + w_bytes) + N * 4


def layout_bytes(active_rows, kind, path="csr"):
    """Bytes the kernel itself must read and write once in its own layout:
    the plain CSR path as above; the tiled plan reads 2-byte tile-local
    indices and one offset pair per active row and tile."""
    if path == "csr":
        return algorithmic_bytes(active_rows, kind)
    w_bytes = {"homo": 0, "heter_f32": 4, "heter_bf16": 2}[kind]
    return M * 1 + active_rows * 8 * TILES + active_rows * K_ROW * (2 + w_bytes) + N * 4


INT_PEAK_TOPS = 39.3     # 256 CUs x 64 lanes x 2.4 GHz, 32-bit integer VALU
OPS_PER_DRAW = 24        # LFSR88 next (18) + 2 xor + modulo + 3 add/compare


def bench_jitconn(L, _lib, torch, steps, warmup, cpu=True):
    """1M x 1M just-in-time connectivity, p = 1e-4 (clen 19999), seed 42,
    2 % Bernoulli bool spikes; homo / uniform [0,1] / normal (0.5, 0.1);
    outdim_parallel True (row-parallel, the default) and False (event-driven
    scatter).  Work = LFSR88 draws, counted exactly from the generated
    connections: per stream 1 initial draw + 1 per connection (+1 uniform /
    +2 normal weight draws per connection).  With `cpu`, each configuration
    is also run once by the oracle's serial C restatement on the host
    (1 thread) on the same events, and the outputs are compared."""
    import ctypes
    from brainpy_amd.math.jitconn import conn_len
    n = 1_000_000
    dev = _lib.device()
    g = torch.Generator(device=dev)
    g.manual_seed(42)
    ev = torch.rand(n, generator=g, device=dev) < 0.02
    ones = torch.ones(n, dtype=torch.float32, device=dev)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    clen = conn_len(1e-4)
    sp = _lib.stream_ptr()
    wsb = int(L.bpx_jitconn_workspace(n, n))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    # the scatter form's two passes take per-tile connection lists (the
    # public call's workspace); with the smaller workspace it runs one pass
    wsb2 = max(int(L.bpx_jitconn_workspace_ex(n, n, clen, d)) for d in (0, 1, 2))
    ws2 = torch.empty(wsb2, dtype=torch.uint8, device=dev)
    res = []
    for od in (1, 0):
        # exact connection counts of this layout: M @ ones (all streams) and
        # M @ events (streams of active columns, for the scatter form)
        L.bpx_jitconn_mv(sp, _lib.ptr(ones), _lib.BPX_F32, 0, 0, 1.0, 0.0, clen, 42, n, n, od, 0, _lib.ptr(out),
                         None, 0)
        conns_all = float(out.double().sum())
        L.bpx_jitconn_mv(sp, _lib.ptr(ev), _lib.BPX_BOOL, 1, 0, 1.0, 0.0, clen, 42, n, n, od, 0, _lib.ptr(out),
                         _lib.ptr(ws), wsb)
        events = float(out.double().sum())       # synaptic events of this call
        n_active = int(ev.sum())
        streams = n * 32 if od else n_active * 32
        conns = conns_all if od else events
        for dist, (w0, w1), extra in ((0, (1.0, 0.0), 0), (1, (0.0, 1.0), 1), (2, (0.5, 0.1), 2)):
            def call(two_pass=True):
                wp, wn = (ws2, wsb2) if (two_pass and not od) else (ws, wsb)
                rc = L.bpx_jitconn_mv(sp, _lib.ptr(ev), _lib.BPX_BOOL, 1, dist, w0, w1, clen, 42, n, n, od, 0,
                                      _lib.ptr(out), _lib.ptr(wp), wn)
                if rc != 0:
                    raise RuntimeError(f"bpx_jitconn_mv failed ({rc})")

            def device_time(two_pass=True):
                # device seconds per call: the kernel (row-parallel), or the
                # scatter form's list pass plus its sum pass
                _lib.timer_reset()
                _lib.timer_enable(True)
                for _ in range(steps):
                    call(two_pass)
                torch.cuda.synchronize()
                _lib.timer_enable(False)
                ms, cnt = _lib.timer_read(_lib.KID["jit_outdim"] if od else _lib.KID["jit_scatter"])
                emit = _lib.timer_read(_lib.KID["jit_emit"])[0] if not od else 0.0
                split[two_pass] = (emit / 1e3 / max(cnt, 1), ms / 1e3 / max(cnt, 1))
                return (ms + emit) / 1e3 / max(cnt, 1)
            split = {}
            for _ in range(warmup):
                call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                call()
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t0) / steps
            kern = device_time()
            one_pass = None if od else device_time(False)
            draws = streams + conns * (1 + extra)
            ops = draws * OPS_PER_DRAW
            r = {"dist": ["homo", "uniform", "normal"][dist], "outdim_parallel": bool(od), "us_per_call": wall * 1e6,
                 "kernel_us": kern * 1e6, "events_per_call": events, "events_per_s": events / wall,
                 "draws": draws, "int_tops": ops / kern / 1e12, "int_frac": ops / kern / 1e12 / INT_PEAK_TOPS}
            if one_pass is not None:
                r["one_pass_kernel_us"] = one_pass * 1e6
                r["list_pass_us"], r["sum_pass_us"] = (x * 1e6 for x in split[True])
            if cpu:
                # the oracle's serial restatement of the same lane32 kernel
                # (the reference's Taichi loop order), one call, 1 thread
                from oracle import orc
                ev_h = ev.cpu().numpy()
                t0 = time.perf_counter()
                y_cpu = orc.jitconn_mv(ev_h, dist, 1, od, orc.LAYOUT_LANE32, n, n, np.float32(w0), np.float32(w1),
                                       clen, 42)
                t_cpu = time.perf_counter() - t0
                call()                           # the two-pass form's output for the scatter form
                y_gpu = out.cpu().numpy()
                r["cpu_port_s"] = t_cpu
                r["cpu_port_events_per_s"] = events / t_cpu
                r["gpu_equals_cpu"] = bool(np.array_equal(y_gpu, y_cpu))
                r["gpu_vs_cpu_max_abs_diff"] = float(np.max(np.abs(y_gpu.astype(np.float64) - y_cpu)))
            res.append(r)
            print(f"# jitconn od={od} {r['dist']:7s} {r['us_per_call']:9.1f} us/call kernel {r['kernel_us']:9.1f} us "
                  f"{r['int_tops']:6.2f} Tops ({100*r['int_frac']:.1f}%) {r['events_per_s']:.3e} ev/s"
                  + (f" (list {r['list_pass_us']:.1f} + sum {r['sum_pass_us']:.1f} us; one pass "
                     f"{r['one_pass_kernel_us']:.1f} us)" if one_pass is not None else "")
                  + (f" | cpu port {r['cpu_port_s']:.3f} s/call (1 thread), equal={r['gpu_equals_cpu']} "
                     f"max|diff|={r['gpu_vs_cpu_max_abs_diff']:.2e}" if cpu else ""),
                  file=sys.stderr, flush=True)
    return {"workload": "jitconn_event_mv_prob_1Mx1M_p1e-4_seed42_rate0.02", "results": res}


def _time_calls(_lib, torch, call, kid, steps, warmup):
    """(wall seconds per call, device seconds per launch of kernel `kid`)."""
    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        call()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    _lib.timer_reset()
    _lib.timer_enable(True)
    for _ in range(steps):
        call()
    torch.cuda.synchronize()
    _lib.timer_enable(False)
    ms, cnt = _lib.timer_read(kid)
    return wall, ms / 1e3 / max(cnt, 1)


def bench_coo(_lib, torch, steps=50, warmup=5):
    """COO products on FixedProb matrices (row-sorted, as a CSR expands):
    coomv on 100,000 x 100,000, p = 0.01 (nse 1e8), f32 values, int32
    indices, both orientations; seg_matmul of 100,000 x 100,000, p = 0.001
    (nse 1e7) with a dense (100,000, 64) f32 operand.  Algorithmic HBM bytes:
    12 per entry (row, col, value) plus one read of the dense operand and one
    write (and zero-fill) of the output."""
    from brainpy_amd.connect import fixed_prob_csr_device
    import ctypes  # noqa: F401
    dev = _lib.device()
    L = _lib.lib()
    sp = _lib.stream_ptr()
    kid = _lib.KID["coo"]
    res = []
    n = 100_000
    idx, ptr = fixed_prob_csr_device(n, n, 0.01, seed=2024, device=dev)
    nse = idx.numel()
    row = torch.repeat_interleave(torch.arange(n, device=dev, dtype=torch.int32), ptr[1:] - ptr[:-1])
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    w = torch.empty(nse, dtype=torch.float32, device=dev).normal_(0.5, 0.1, generator=g)
    v = torch.rand(n, generator=g, device=dev)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    for transpose in (0, 1):
        def call():
            rc = L.bpx_coomv(sp, _lib.ptr(w), _lib.BPX_F32, nse, _lib.ptr(row), _lib.ptr(idx), _lib.BPX_I32, nse,
                             _lib.ptr(v), _lib.BPX_F32, n, n, transpose, _lib.ptr(out))
            if rc != 0:
                raise RuntimeError(f"bpx_coomv failed ({rc})")
        wall, kern = _time_calls(_lib, torch, call, kid, steps, warmup)
        alg = nse * 12 + n * 4 + n * 4
        r = {"op": "coomv", "transpose": bool(transpose), "nse": nse, "us_per_call": wall * 1e6,
             "kernel_us": kern * 1e6, "alg_bytes": alg, "gbs": alg / kern / 1e9,
             "frac": alg / kern / 1e9 / HBM_PEAK_GBS, "entries_per_s": nse / wall}
        res.append(r)
        print(f"# coomv transpose={transpose} nse={nse:.3g} {r['us_per_call']:8.1f} us/call kernel "
              f"{r['kernel_us']:8.1f} us {r['gbs']:7.0f} GB/s ({100 * r['frac']:.1f}%)", file=sys.stderr, flush=True)
    # ceilings of the access pattern on the same arrays (no scan, no atomics)
    probe_out = torch.empty(32768 * 256, dtype=torch.float32, device=dev)
    ceil = {}
    for mode, name in ((0, "read_row_col_value"), (1, "read_col_gather"), (2, "read_all_gather")):
        def pcall():
            rc = probe_lib().bpx_probe_coo(sp, _lib.ptr(row), _lib.ptr(idx), _lib.ptr(w), _lib.ptr(v), n, nse, mode,
                                 _lib.ptr(probe_out), probe_out.numel())
            if rc != 0:
                raise RuntimeError(f"bpx_probe_coo failed ({rc})")
        for _ in range(3):
            pcall()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
        for _ in range(10):
            pcall()
        ev1.record()
        ev1.synchronize()
        ceil[name] = ev0.elapsed_time(ev1) * 1e3 / 10
    res[0]["gather_ceiling_us"] = ceil
    res[0]["frac_of_read_gather_ceiling"] = ceil["read_all_gather"] / res[0]["kernel_us"]
    print("# coomv ceilings (probe): " + ", ".join(f"{k} {u:.1f} us" for k, u in ceil.items())
          + f"; coomv at {100 * res[0]['frac_of_read_gather_ceiling']:.0f}% of read+gather",
          file=sys.stderr, flush=True)
    # the public bm.sparse.coomv on the same arrays: from its second call on
    # them the matrix runs on a tiled plan (ordered by row); bytes as passed
    import brainpy_amd.math as bm
    for transpose in (0, 1):
        # first call stateless; the second builds the plan (transpose=0) and
        # runs on it, timed as the plan build (sort by row, offsets, tiling)
        bm.sparse.coomv(w, row, idx, v, shape=(n, n), transpose=bool(transpose))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bm.sparse.coomv(w, row, idx, v, shape=(n, n), transpose=bool(transpose))
        torch.cuda.synchronize()
        second_ms = (time.perf_counter() - t0) * 1e3
        hits0 = bm.event.plan_cache_info()["hits"]

        def pubcall():
            bm.sparse.coomv(w, row, idx, v, shape=(n, n), transpose=bool(transpose))
        wall, kern = _time_calls(_lib, torch, pubcall, _lib.KID["tiled"] if transpose else _lib.KID["tiled_gather"],
                                 steps, warmup)
        if bm.event.plan_cache_info()["hits"] - hits0 < steps:
            raise RuntimeError("bm.sparse.coomv did not run on its cached plan")
        alg = nse * 12 + n * 4 + n * 4
        r = {"op": "bm.sparse.coomv (public path, cached plan)", "transpose": bool(transpose), "nse": nse,
             "us_per_call": wall * 1e6, "kernel_us": kern * 1e6, "alg_bytes": alg, "gbs": alg / kern / 1e9,
             "frac": alg / kern / 1e9 / HBM_PEAK_GBS, "entries_per_s": nse / wall}
        if transpose == 0:
            r["plan_build_ms"] = second_ms
        res.append(r)
        print(f"# bm.sparse.coomv transpose={transpose} (public path, cached plan) {r['us_per_call']:8.1f} us/call "
              f"kernel {r['kernel_us']:8.1f} us {r['gbs']:7.0f} GB/s ({100 * r['frac']:.1f}%)"
              + (f"; plan built in the second call, {second_ms:.1f} ms" if transpose == 0 else ""), file=sys.stderr,
              flush=True)
    bm.event.clear_plan_cache()
    del row, w, idx, ptr, probe_out
    idx, ptr = fixed_prob_csr_device(n, n, 0.001, seed=2025, device=dev)
    nse = idx.numel()
    row = torch.repeat_interleave(torch.arange(n, device=dev, dtype=torch.int32), ptr[1:] - ptr[:-1])
    w = torch.empty(nse, dtype=torch.float32, device=dev).normal_(0.0, 1.0, generator=g)
    k = 64
    D = torch.rand((n, k), generator=g, device=dev)
    out2 = torch.empty((n, k), dtype=torch.float32, device=dev)

    def call2():
        rc = L.bpx_coo_matmat(sp, _lib.ptr(w), _lib.BPX_F32, nse, _lib.ptr(row), _lib.ptr(idx), _lib.BPX_I32, nse,
                              _lib.ptr(D), _lib.BPX_F32, n, n, k, 1, _lib.ptr(out2))
        if rc != 0:
            raise RuntimeError(f"bpx_coo_matmat failed ({rc})")
    wall, kern = _time_calls(_lib, torch, call2, kid, steps, warmup)
    alg = nse * 12 + n * k * 4 * 3
    gathered = nse * k * 4
    r = {"op": "seg_matmul", "k": k, "nse": nse, "us_per_call": wall * 1e6, "kernel_us": kern * 1e6,
         "alg_bytes": alg, "gbs": alg / kern / 1e9, "frac": alg / kern / 1e9 / HBM_PEAK_GBS,
         "gathered_bytes": gathered, "gather_gbs": gathered / kern / 1e9,
         "gflops": 2 * nse * k / kern / 1e9}
    res.append(r)
    print(f"# seg_matmul k={k} nse={nse:.3g} {r['us_per_call']:8.1f} us/call kernel {r['kernel_us']:8.1f} us "
          f"{r['gbs']:7.0f} GB/s alg ({100 * r['frac']:.1f}%), gathers {r['gather_gbs']:7.0f} GB/s, "
          f"{r['gflops']:.0f} GFLOP/s", file=sys.stderr, flush=True)
    return {"workload": "coomv_100kx100k_p0.01_f32 / seg_matmul_100kx100k_p0.001_k64_f32", "results": res}


def bench_forms(_lib, torch, steps=30, warmup=3):
    """The other forms of the operators on the headline matrix (100,000 x
    100,000 CSR, p = 0.01, nnz 1e8, int32, heterogeneous f32 weights):
    bm.sparse.csrmv with a dense f32 vector (gather, and transpose = scatter
    with atomics) and bm.event.csrmv without transpose (gather of 10 %
    Bernoulli bool events).  The gathers read a random 4-byte (or 1-byte)
    element per synapse, so the probe's read-col-and-gather time on the
    same indices is reported as their ceiling."""
    from brainpy_amd.connect import fixed_prob_csr_device
    dev = _lib.device()
    L = _lib.lib()
    sp = _lib.stream_ptr()
    n = 100_000
    idx, ptr = fixed_prob_csr_device(n, n, 0.01, seed=2024, device=dev)
    nnz = idx.numel()
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    w = torch.empty(nnz, dtype=torch.float32, device=dev).normal_(0.5, 0.1, generator=g)
    x = torch.rand(n, generator=g, device=dev)
    ev = torch.rand(n, generator=g, device=dev) < 0.1
    ev_f = ev.to(torch.float32)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    res = []
    legs = (("sparse.csrmv", 0, x, _lib.BPX_F32, _lib.KID["csr_gather"], 8),
            ("sparse.csrmv", 1, x, _lib.BPX_F32, _lib.KID["csr_scatter"], 8),
            ("event.csrmv", 0, ev, _lib.BPX_BOOL, _lib.KID["event_gather"], 8),
            ("event.csrmv", 0, ev_f, _lib.BPX_F32, _lib.KID["event_gather"], 8))
    for name, transpose, vec, vdt, kid, per in legs:
        def call():
            if name == "sparse.csrmv":
                rc = L.bpx_csrmv(sp, _lib.ptr(w), _lib.BPX_F32, nnz, _lib.ptr(idx), _lib.ptr(ptr), _lib.BPX_I32, nnz,
                                 _lib.ptr(vec), vdt, n, n, transpose, _lib.ptr(out))
            else:
                rc = L.bpx_event_csrmv(sp, _lib.ptr(w), _lib.BPX_F32, nnz, _lib.ptr(idx), _lib.ptr(ptr), _lib.BPX_I32,
                                       nnz, _lib.ptr(vec), vdt, _lib.RULE_POSITIVE, n, n, transpose, _lib.ptr(out))
            if rc != 0:
                raise RuntimeError(f"{name} failed ({rc})")
        wall, kern = _time_calls(_lib, torch, call, kid, steps, warmup)
        alg = nnz * per + (n + 1) * 4 + n * 4 + n * (1 if vdt == _lib.BPX_BOOL else 4)
        r = {"op": name, "transpose": bool(transpose), "events": "bool" if vdt == _lib.BPX_BOOL else "f32",
             "nnz": nnz, "us_per_call": wall * 1e6,
             "kernel_us": kern * 1e6, "alg_bytes": alg, "gbs": alg / kern / 1e9,
             "frac": alg / kern / 1e9 / HBM_PEAK_GBS, "synapses_per_s": nnz / wall}
        res.append(r)
        tag = f"{name} ({'bool' if vdt == _lib.BPX_BOOL else 'f32'} vector)" + (" stateless" if transpose else "")
        print(f"# {tag:26s} transpose={transpose} {r['us_per_call']:8.1f} us/call kernel {r['kernel_us']:8.1f} us "
              f"{r['gbs']:7.0f} GB/s ({100 * r['frac']:.1f}%) {r['synapses_per_s']:.3e} synapses/s",
              file=sys.stderr, flush=True)
    # the public bm.sparse.csrmv(..., transpose=True) on the same arrays: its
    # plan cache builds the tiled plan on the second call, then every call
    # streams the 1e8 synapses once at 2 + 4 bytes with no global atomic per
    # synapse (algorithmic bytes: the operator's inputs as passed)
    import brainpy_amd.math as bm
    xs = torch.rand(n, generator=g, device=dev) + 0.5
    for _ in range(2):
        bm.sparse.csrmv(w, idx, ptr, xs, shape=(n, n), transpose=True)
    hits0 = bm.event.plan_cache_info()["hits"]

    def pcall():
        bm.sparse.csrmv(w, idx, ptr, xs, shape=(n, n), transpose=True)
    wall, kern = _time_calls(_lib, torch, pcall, _lib.KID["tiled"], steps, warmup)
    if bm.event.plan_cache_info()["hits"] - hits0 < steps:
        raise RuntimeError("bm.sparse.csrmv(transpose=True) did not run on its cached plan")
    alg = nnz * 8 + (n + 1) * 4 + n * 4 + n * 4
    r = {"op": "bm.sparse.csrmv transpose (cached plan)", "transpose": True, "nnz": nnz, "us_per_call": wall * 1e6,
         "kernel_us": kern * 1e6, "alg_bytes": alg, "gbs": alg / kern / 1e9, "frac": alg / kern / 1e9 / HBM_PEAK_GBS,
         "synapses_per_s": nnz / wall}
    res.append(r)
    print(f"# bm.sparse.csrmv transpose=1 (public path, cached plan) {r['us_per_call']:8.1f} us/call kernel "
          f"{r['kernel_us']:8.1f} us {r['gbs']:7.0f} GB/s ({100 * r['frac']:.1f}%) {r['synapses_per_s']:.3e} "
          f"synapses/s", file=sys.stderr, flush=True)
    # the public sparse gather (transpose=False) on the same cached plan:
    # each column tile of the vector staged in LDS
    for name, vec, kw in (("bm.sparse.csrmv transpose=0 (public path, cached plan)", xs, {}),):
        fn = bm.sparse.csrmv
        for _ in range(2):
            fn(w, idx, ptr, vec, shape=(n, n), **kw)

        def gcall():
            fn(w, idx, ptr, vec, shape=(n, n), **kw)
        wall, kern = _time_calls(_lib, torch, gcall, _lib.KID["tiled_gather"], steps, warmup)
        alg = nnz * 8 + (n + 1) * 4 + n * 4 + n * (1 if vec.dtype == torch.bool else 4)
        r = {"op": name, "transpose": False, "nnz": nnz, "us_per_call": wall * 1e6, "kernel_us": kern * 1e6,
             "alg_bytes": alg, "gbs": alg / kern / 1e9, "frac": alg / kern / 1e9 / HBM_PEAK_GBS,
             "synapses_per_s": nnz / wall}
        res.append(r)
        print(f"# {name} {r['us_per_call']:8.1f} us/call kernel {r['kernel_us']:8.1f} us {r['gbs']:7.0f} GB/s "
              f"({100 * r['frac']:.1f}%) {r['synapses_per_s']:.3e} synapses/s", file=sys.stderr, flush=True)
    # the public event gather, bm.event.csrmv(..., transpose=False), on the
    # cached plan of the transposed matrix: each call scatters the entries of
    # its active columns only (16 Bernoulli(0.1) vectors in turn, so the
    # entries come from HBM); bytes as an event-driven operator reads them:
    # 1 per event, an 8-byte offset pair per active column, 8 per synaptic
    # event and 4 per output
    evs = [torch.rand(n, generator=g, device=dev) < 0.1 for _ in range(16)]
    bm.event.csrmv(w, idx, ptr, evs[0], shape=(n, n))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bm.event.csrmv(w, idx, ptr, evs[1], shape=(n, n))      # builds the plan of A^T
    torch.cuda.synchronize()
    build_t_ms = (time.perf_counter() - t0) * 1e3
    hits0 = bm.event.plan_cache_info()["hits"]
    cnt = [0]

    def ecall():
        bm.event.csrmv(w, idx, ptr, evs[cnt[0] % 16], shape=(n, n))
        cnt[0] += 1
    wall, kern = _time_calls(_lib, torch, ecall, _lib.KID["tiled"], 64, 16)
    if bm.event.plan_cache_info()["hits"] - hits0 < 64:
        raise RuntimeError("bm.event.csrmv(transpose=False) did not run on its cached transposed plan")
    cols = sum(int(e.sum()) for e in evs) / 16
    events = cols * (nnz / n)
    alg = n + cols * 8 + events * 8 + n * 4
    r = {"op": "bm.event.csrmv transpose=0 (public path, cached plan of A^T)", "transpose": False,
         "events": "bool", "nnz": nnz, "active_columns": cols, "events_per_call": events,
         "us_per_call": wall * 1e6, "kernel_us": kern * 1e6, "alg_bytes_event_driven": alg,
         "gbs": alg / kern / 1e9, "frac": alg / kern / 1e9 / HBM_PEAK_GBS, "events_per_s": events / wall,
         "plan_build_ms": build_t_ms}
    res.append(r)
    print(f"# bm.event.csrmv transpose=0 (public path, cached plan of A^T, 16 event vectors) {r['us_per_call']:8.1f} "
          f"us/call kernel {r['kernel_us']:8.1f} us {r['gbs']:7.0f} GB/s event-driven ({100 * r['frac']:.1f}%) "
          f"{r['events_per_s']:.3e} events/s; plan of A^T built in the second call, {build_t_ms:.1f} ms",
          file=sys.stderr, flush=True)
    bm.event.clear_plan_cache()
    del xs, evs
    probe_out = torch.empty(32768 * 256, dtype=torch.float32, device=dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for k in range(13):
        if k == 3:
            ev0.record()
        rc = probe_lib().bpx_probe_coo(sp, None, _lib.ptr(idx), None, _lib.ptr(x), n, nnz, 1, _lib.ptr(probe_out),
                                       probe_out.numel())
        if rc != 0:
            raise RuntimeError(f"bpx_probe_coo failed ({rc})")
    ev1.record()
    ev1.synchronize()
    ceiling = ev0.elapsed_time(ev1) * 1e3 / 10
    for r in res:
        if not r["transpose"]:
            r["read_col_gather_ceiling_us"] = ceiling
    print(f"# gather ceiling (read col + gather, probe mode 1): {ceiling:.1f} us", file=sys.stderr,
          flush=True)
    del idx, ptr, w
    # bm.event.info: count, scan and write kernels; bytes = one read of the
    # bool events + the int32 ids row (active ids, then -1) + the counts
    info = []
    for batch, m in ((1, 100_000), (256, 100_000), (1, 100_000_000)):
        evb = torch.rand((batch, m), generator=g, device=dev) < 0.1
        ids = torch.empty((batch, m), dtype=torch.int32, device=dev)
        num = torch.empty(batch, dtype=torch.int32, device=dev)
        wsb = int(L.bpx_event_info_workspace(batch, m))
        ws = torch.empty(max(wsb, 1), dtype=torch.uint8, device=dev)

        def icall():
            rc = L.bpx_event_info(sp, _lib.ptr(evb), _lib.BPX_BOOL, batch, m, _lib.ptr(ids), _lib.ptr(num),
                                  _lib.ptr(ws), wsb)
            if rc != 0:
                raise RuntimeError(f"bpx_event_info failed ({rc})")
        wall, kern = _time_calls(_lib, torch, icall, _lib.KID["info"], steps, warmup)
        alg = batch * m * 5 + batch * 4
        r = {"op": "event.info", "batch": batch, "n": m, "us_per_call": wall * 1e6, "write_kernel_us": kern * 1e6,
             "alg_bytes": alg, "gbs": alg / wall / 1e9, "frac": alg / wall / 1e9 / HBM_PEAK_GBS,
             "events_per_s": batch * m / wall}
        info.append(r)
        print(f"# event.info batch={batch} n={m:.0e} {r['us_per_call']:8.1f} us/call (write kernel "
              f"{r['write_kernel_us']:.1f} us) {r['gbs']:7.0f} GB/s ({100 * r['frac']:.1f}%)", file=sys.stderr,
              flush=True)
        del evb, ids, ws
    return {"workload": "csrmv_forms_100kx100k_p0.01_heter_f32 / event_info", "results": res, "info": info}


def bench_bign(_lib, torch, steps=30, warmup=3):
    """The transpose event csrmv beyond 262,144 outputs, where the plan has
    more than 64 output tiles (fewer output copies per call): the public
    bm.event.csrmv(transpose=True) on its cached plan against the stateless
    scatter on the same arrays, heterogeneous f32 weights, bool events.
    Workloads: 100,000 x 1,000,000 with 1000 synapses per row (nnz 1e8) at
    10 % events, and 1,000,000 x 1,000,000 with 100 per row (nnz 1e8) at
    2 %.  Algorithmic bytes as the headline's: m + 8 per active row + 8 per
    synaptic event + 4n."""
    from brainpy_amd.connect import fixed_prob_csr_device
    import brainpy_amd.math as bm
    dev = _lib.device()
    L = _lib.lib()
    sp = _lib.stream_ptr()
    res = []
    for m, n, k, rate in ((100_000, 1_000_000, 1000, 0.1), (1_000_000, 1_000_000, 100, 0.02)):
        idx, ptr = fixed_prob_csr_device(m, n, (k + 0.5) / n, seed=77, device=dev)
        nnz = idx.numel()
        g = torch.Generator(device=dev)
        g.manual_seed(13)
        w = torch.empty(nnz, dtype=torch.float32, device=dev).normal_(0.5, 0.1, generator=g)
        ev = torch.rand(m, generator=g, device=dev) < rate
        active = int(ev.sum())
        out = torch.empty(n, dtype=torch.float32, device=dev)
        alg = m + 8 * active + 8 * active * k + 4 * n
        bm.event.clear_plan_cache()
        bm.event.csrmv(w, idx, ptr, ev, shape=(m, n), transpose=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bm.event.csrmv(w, idx, ptr, ev, shape=(m, n), transpose=True)      # builds the plan
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - t0) * 1e3
        info = bm.event.plan_cache_info()
        if info["plans"] != 1:
            raise RuntimeError(f"no plan for {m} x {n}: {info}")
        hits0 = info["hits"]

        def pcall():
            bm.event.csrmv(w, idx, ptr, ev, shape=(m, n), transpose=True)
        wall, kern = _time_calls(_lib, torch, pcall, _lib.KID["tiled"], steps, warmup)
        if bm.event.plan_cache_info()["hits"] - hits0 < steps:
            raise RuntimeError("bm.event.csrmv(transpose=True) did not run on its cached plan")
        y_plan = bm.event.csrmv(w, idx, ptr, ev, shape=(m, n), transpose=True).clone()

        def scall():
            rc = L.bpx_event_csrmv(sp, _lib.ptr(w), _lib.BPX_F32, nnz, _lib.ptr(idx), _lib.ptr(ptr), _lib.BPX_I32,
                                   nnz, _lib.ptr(ev), _lib.BPX_BOOL, _lib.RULE_POSITIVE, m, n, 1, _lib.ptr(out))
            if rc != 0:
                raise RuntimeError(f"bpx_event_csrmv failed ({rc})")
        swall, skern = _time_calls(_lib, torch, scall, _lib.KID["event_scatter"], steps, warmup)
        diff = float((y_plan - out).abs().max())
        plan = next(iter(bm.event._PLANS._plans.values()))[0]
        r = {"m": m, "n": n, "nnz": nnz, "rate": rate, "events": active * k, "tiles": plan.plan.tiles,
             "tile_len": plan.plan.tile_len, "plan_build_ms": build_ms, "plan_bytes": plan.mem.numel(),
             "us_per_call": wall * 1e6, "kernel_us": kern * 1e6, "alg_bytes": alg, "gbs": alg / kern / 1e9,
             "frac": alg / kern / 1e9 / HBM_PEAK_GBS, "stateless_us_per_call": swall * 1e6,
             "stateless_kernel_us": skern * 1e6, "speedup_kernel": skern / kern, "max_abs_diff_vs_stateless": diff}
        res.append(r)
        print(f"# event csrmv transpose {m}x{n} k={k} rate={rate}: plan ({r['tiles']} tiles of {r['tile_len']}, "
              f"built in {build_ms:.1f} ms, {r['plan_bytes'] / 2**20:.0f} MiB) {r['us_per_call']:.1f} us/call, kernel "
              f"{r['kernel_us']:.1f} us = {r['gbs']:.0f} GB/s ({100 * r['frac']:.1f}%); stateless kernel "
              f"{r['stateless_kernel_us']:.1f} us ({r['speedup_kernel']:.1f}x); max|diff| {diff:.2e}",
              file=sys.stderr, flush=True)
        bm.event.clear_plan_cache()
        del idx, ptr, w, ev, out, y_plan, plan
        torch.cuda.empty_cache()
    return {"workload": "event_csrmv_transpose_large_n", "results": res}


def bench_conn(_lib, torch, reps=5):
    """FixedProb.build_csr of the headline matrix (100,000 x 100,000,
    p = 0.01, 1e8 targets): the device builder (bpx_fixed_prob_csr: uniform
    distinct sorted targets), the torch-based input generator the other
    benchmarks use, and the host numpy builder on a 2,000-row sample scaled
    to all rows (1 thread; the reference draws rows with Numba's rng.choice,
    connect/random_conn.py:92-105)."""
    from brainpy_amd.connect import fixed_prob_csr, fixed_prob_csr_device, fixed_prob_csr_gpu
    dev = _lib.device()
    n = 100_000
    res = {}
    for name, fn in (("device_builder", lambda: fixed_prob_csr_gpu(n, n, 0.01, seed=2024)),
                     ("torch_generator", lambda: fixed_prob_csr_device(n, n, 0.01, seed=2024, device=dev))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            out = fn()
        torch.cuda.synchronize()
        res[name + "_s"] = (time.perf_counter() - t0) / reps
        del out
    sample = 2000
    t0 = time.perf_counter()
    fixed_prob_csr(sample, n, 0.01, seed=2024)
    res["host_numpy_s_scaled"] = (time.perf_counter() - t0) * n / sample
    res["synapses"] = n * 1000
    res["device_synapses_per_s"] = n * 1000 / res["device_builder_s"]
    print(f"# conn FixedProb 100k x 100k p=0.01: device builder {res['device_builder_s'] * 1e3:.1f} ms "
          f"({res['device_synapses_per_s']:.3e} synapses/s), torch generator {res['torch_generator_s'] * 1e3:.1f} ms, "
          f"host numpy (scaled from {sample} rows) {res['host_numpy_s_scaled']:.1f} s", file=sys.stderr, flush=True)
    return {"workload": "fixed_prob_build_csr_100kx100k_p0.01", **res}


def bench_bsr(_lib, torch, steps=30, warmup=3):
    """Block-sparse products.  cusparse_bcsr_matvec: 100,000 x 100,000 as
    6250 x 6250 blocks of 16 x 16, 62 blocks per block row (FixedProb over
    blocks, p = 0.01), f32, int32 indices, both orientations; algorithmic
    bytes = the block values (1 KiB per block), block indices and pointers,
    one read of x and one write of y.  bcsrmm: A (4096 x 4096) @ B (4096 x
    4096) with 12 of the 128 32 x 32 blocks of each column block present,
    f32 on MFMA; work = 2 * m * 1024 flops per block."""
    from brainpy_amd.connect import fixed_prob_csr_device
    dev = _lib.device()
    L = _lib.lib()
    sp = _lib.stream_ptr()
    kid = _lib.KID["bsr"]
    res = []
    bs, mb = 16, 6250
    idx, ptr = fixed_prob_csr_device(mb, mb, 0.01, seed=7, device=dev)
    nnzb = idx.numel()
    g = torch.Generator(device=dev)
    g.manual_seed(9)
    data = torch.empty(nnzb * bs * bs, dtype=torch.float32, device=dev).normal_(0.0, 1.0, generator=g)
    x = torch.rand(mb * bs, generator=g, device=dev)
    y = torch.empty(mb * bs, dtype=torch.float32, device=dev)
    for transpose in (0, 1):
        def call():
            rc = L.bpx_bsrmv(sp, _lib.ptr(data), _lib.BPX_F32, _lib.ptr(idx), _lib.ptr(ptr), _lib.BPX_I32, nnzb, bs,
                             _lib.ptr(x), mb, mb, transpose, _lib.ptr(y))
            if rc != 0:
                raise RuntimeError(f"bpx_bsrmv failed ({rc})")
        wall, kern = _time_calls(_lib, torch, call, kid, steps, warmup)
        alg = nnzb * (bs * bs * 4 + 4) + (mb + 1) * 4 + 2 * mb * bs * 4
        r = {"op": "bsrmv", "transpose": bool(transpose), "blocksize": bs, "nnzb": nnzb, "us_per_call": wall * 1e6,
             "kernel_us": kern * 1e6, "alg_bytes": alg, "gbs": alg / kern / 1e9,
             "frac": alg / kern / 1e9 / HBM_PEAK_GBS}
        res.append(r)
        print(f"# bsrmv transpose={transpose} bs={bs} nnzb={nnzb} {r['us_per_call']:8.1f} us/call kernel "
              f"{r['kernel_us']:8.1f} us {r['gbs']:7.0f} GB/s ({100 * r['frac']:.1f}%)", file=sys.stderr, flush=True)
    del data, x, y, idx, ptr
    m = k = n = 4096
    cidx, cptr = fixed_prob_csr_device(n // 32, k // 32, 0.1, seed=11, device=dev)
    nblk = cidx.numel()
    bidx = torch.stack([cidx, torch.arange(nblk, device=dev, dtype=torch.int32)], 1).contiguous()
    bdata = torch.empty(nblk * 1024, dtype=torch.float32, device=dev).normal_(0.0, 1.0, generator=g)
    A = torch.empty(m * k, dtype=torch.float32, device=dev).normal_(0.0, 1.0, generator=g)
    C = torch.empty(m * n, dtype=torch.float32, device=dev)

    def mm():
        rc = L.bpx_bcsrmm(sp, _lib.ptr(A), _lib.BPX_F32, _lib.ptr(bdata), _lib.ptr(bidx), _lib.ptr(cptr),
                          _lib.BPX_I32, nblk, m, k, n, _lib.ptr(C))
        if rc != 0:
            raise RuntimeError(f"bpx_bcsrmm failed ({rc})")
    wall, kern = _time_calls(_lib, torch, mm, kid, steps, warmup)
    flops = 2.0 * m * nblk * 1024
    r = {"op": "bcsrmm", "m": m, "k": k, "n": n, "blocks": nblk, "us_per_call": wall * 1e6, "kernel_us": kern * 1e6,
         "flops": flops, "tflops": flops / kern / 1e12, "frac": flops / kern / 1e12 / FP32_MFMA_TFLOPS}
    res.append(r)
    print(f"# bcsrmm {m}x{k} @ {k}x{n} ({nblk} blocks) {r['us_per_call']:8.1f} us/call kernel {r['kernel_us']:8.1f} us "
          f"{r['tflops']:6.1f} TFLOP/s ({100 * r['frac']:.1f}% of f32 MFMA)", file=sys.stderr, flush=True)
    return {"workload": "bsrmv_100kx100k_bs16_p0.01_f32 + bcsrmm_4096_f32_10pct_blocks", "results": res}


def bench_stdp(_lib, torch, steps=30, warmup=3):
    """STDP weight updates on the headline matrix (100,000 x 100,000 CSR,
    p = 0.01, f32 weights, int32 indices), 10 % Bernoulli spikes, in place:
    csr_on_pre_update (rows of spiking pre neurons; per synapse 4 B index,
    4 B weight read and write, a trace gather) and csc_on_post_update
    (columns of spiking post neurons through the CSC view: 4 B pre id and
    4 B weight id per synapse, the weight read and written at scattered
    positions)."""
    from brainpy_amd.connect import fixed_prob_csr_device
    from brainpy_amd import dnn
    dev = _lib.device()
    n = 100_000
    idx, ptr = fixed_prob_csr_device(n, n, 0.01, seed=2024, device=dev)
    nnz = idx.numel()
    g = torch.Generator(device=dev)
    g.manual_seed(21)
    w = torch.empty(nnz, dtype=torch.float32, device=dev).uniform_(0.2, 0.8, generator=g)
    trace = torch.empty(n, dtype=torch.float32, device=dev).normal_(0.0, 1e-3, generator=g)
    spk = torch.rand(n, generator=g, device=dev) < 0.1
    rows = torch.repeat_interleave(torch.arange(n, device=dev, dtype=torch.int32), ptr[1:] - ptr[:-1])
    order = torch.argsort(idx, stable=True)
    pre_ids = rows[order].contiguous()
    w_ids = order.to(torch.int32)
    colptr = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    colptr[1:] = torch.cumsum(torch.bincount(idx.long(), minlength=n), 0).to(torch.int32)
    del rows, order
    act = int(spk.sum())
    syn_rows = int((ptr[1:] - ptr[:-1])[spk].sum())
    syn_cols = int((colptr[1:] - colptr[:-1])[spk].sum())
    res = []
    for name, call, per, syn in (
            ("csr_on_pre_update", lambda: dnn.csr_on_pre_update(w, idx, ptr, spk, trace, 0.0, 1.0, out=w), 12,
             syn_rows),
            ("csc_on_post_update", lambda: dnn.csc_on_post_update(w, pre_ids, colptr, w_ids, spk, trace, 0.0, 1.0,
                                                                  out=w), 16, syn_cols)):
        for _ in range(warmup):
            call()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            call()
        e1.record()
        e1.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / steps
        alg = syn * per + n + (n + 1) * 4
        r = {"op": name, "active": act, "synapses": syn, "us_per_call": us, "alg_bytes": alg,
             "gbs": alg / us / 1e3, "frac": alg / us / 1e3 / HBM_PEAK_GBS, "synapses_per_s": syn / us * 1e6}
        res.append(r)
        print(f"# stdp {name:20s} {us:8.1f} us/call {r['gbs']:7.0f} GB/s ({100 * r['frac']:.1f}%) "
              f"{r['synapses_per_s']:.3e} synapses/s", file=sys.stderr, flush=True)
    return {"workload": "stdp_100kx100k_p0.01_rate0.10_f32_in_place", "results": res}


def _percentile(x, q):
    return float(np.percentile(np.asarray(x, np.float64), q)) if len(x) else None


def bench_coba(_lib, torch, steps=1000):
    """C1: 4000-neuron E/I COBA network (3200 E + 800 I, FixedProb p = 0.02,
    we = 0.6, wi = 6.7, input 20 mV, V0 ~ U[-60, -50], dt = 0.1 ms, 1000 steps,
    LifRefLTC, exp_euler).  GPU: the per-step network kernel and the
    multi-step one (one block, 100 steps per launch) on this one network (CSR
    connectivity), HIP-graph captured; the multi-step kernel as one launch
    of 1000 steps; eager per-step launches; plain op-by-op GPU path; CPU: the
    oracle's serial restatement (1 thread)."""
    from brainpy_amd.dyn import EICOBANet
    from brainpy_amd.dyn.batched import BatchedCOBA, combined_csr
    from oracle import orc
    net = EICOBANet(seed=0, mode="atomic")
    csrs = {k: (net.csr[k][1], net.csr[k][2]) for k in ("E2E", "E2I", "I2E", "I2I")}
    spec = dict(n_exc=3200, n_inh=800, seed=0, w_exc=0.6, w_inh=6.7, I_ext=20.0, csr=csrs)
    def graph_run(fused):
        b = BatchedCOBA([spec], V_init=net.V0)
        b.capture(100, fused=fused)
        b.replay()
        torch.cuda.synchronize()
        b.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps // 100):
            b.replay()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, int(b.spike_count.sum()), b.V.cpu().numpy()

    # one launch per step, and one multi-step launch per 100 steps (one
    # block, counts in LDS), both HIP-graph captured
    t_graph, spikes, V_steps = graph_run(False)
    t_graph_multi, spikes_multi, V_multi = graph_run(True)
    assert spikes == spikes_multi and np.array_equal(V_multi, V_steps)
    events = spikes * 80          # every neuron has 64 + 16 targets
    b2 = BatchedCOBA([spec], V_init=net.V0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b2.run(steps, fused=False)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    b3 = BatchedCOBA([spec], V_init=net.V0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b3.run(steps, fused=True)
    torch.cuda.synchronize()
    t_one_launch = time.perf_counter() - t0
    net.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.run(steps, record=False)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    idx, ptr = combined_csr(csrs, 3200, 800)
    t0 = time.perf_counter()
    cpu_spikes, _ = orc.coba_net_run(3200, 800, ptr, idx, 0.6, 6.7, 20.0, net.V0, steps,
                                     (-60.0, -50.0, -60.0, 20.0, 5.0, 1.0), raster=False)
    t_cpu = time.perf_counter() - t0
    r = {"workload": "coba_4000_p0.02_1000steps", "steps": steps, "spikes": spikes, "cpu_spikes": int(cpu_spikes),
         "synaptic_events": events, "gpu_graph_s": t_graph, "gpu_graph_multistep_s": t_graph_multi,
         "gpu_one_launch_s": t_one_launch, "gpu_eager_s": t_eager, "gpu_plain_opbyop_s": t_plain,
         "cpu_1thread_s": t_cpu, "gpu_steps_per_s": steps / t_graph, "gpu_events_per_s": events / t_graph,
         "cpu_events_per_s": events / t_cpu, "mean_rate_hz": spikes / 4000 / (steps * 1e-4)}
    print(f"# coba C1 graph {t_graph*1e3:.2f} ms (multi-step launches {t_graph_multi*1e3:.2f} ms, one launch "
          f"{t_one_launch*1e3:.2f} ms) eager {t_eager*1e3:.2f} ms plain {t_plain*1e3:.2f} ms "
          f"cpu {t_cpu*1e3:.1f} ms spikes {spikes} (cpu {cpu_spikes})", file=sys.stderr, flush=True)
    return r


def bench_small_batch(torch, n_nets=512, steps=1000):
    """A batch of 512 small jitconn networks (N log-uniform in [1k, 5k],
    in-degree 80, FP32 weights), 1000 steps in HIP graphs of 100 steps: one
    launch per step for all networks against one multi-step launch (one
    block per network, counts in LDS) per 100 steps; same spike totals."""
    from brainpy_amd.dyn.batched import BatchedCOBA
    specs = _c4_specs(n_nets, seed=77, lo=1000, hi=5000)
    out = {}
    for fused in (False, True):
        b = BatchedCOBA(specs)
        assert b.fused
        b.capture(100, fused=fused)
        b.replay()
        torch.cuda.synchronize()
        b.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps // 100):
            b.replay()
        torch.cuda.synchronize()
        out[fused] = (time.perf_counter() - t0, int(b.spike_count.sum()), int(b.events.sum()), b.total)
        del b
    (t_step, sp, ev, total), (t_multi, sp2, ev2, _) = out[False], out[True]
    assert sp == sp2 and ev == ev2
    b = BatchedCOBA(specs)
    tune = b.autotune(steps=200)     # tiles and the multi-step launch, from the fresh state (restored)
    picks_fused = bool(b.use_fused)
    del b
    r = {"workload": f"batched_coba_{n_nets}nets_1k-5k_jit80_{steps}steps", "networks": n_nets,
         "total_neurons": total, "spikes": sp, "synaptic_events": ev, "per_step_launch_graph_s": t_step,
         "multi_step_launch_graph_s": t_multi, "speedup": t_step / t_multi,
         "autotune_us_per_step": {str(k): v for k, v in tune.items()}, "autotune_picks_multi_step": picks_fused,
         "us_per_step": {"per_step_launches": t_step / steps * 1e6, "multi_step": t_multi / steps * 1e6}}
    print(f"# small batch {n_nets} nets ({total} neurons): per-step launches {t_step*1e3:.1f} ms, multi-step "
          f"{t_multi*1e3:.1f} ms -> x{t_step/t_multi:.2f}; autotune picks "
          f"{'multi-step' if picks_fused else 'per-step'}", file=sys.stderr, flush=True)
    return r


def _c4_specs(n_nets=256, seed=2024, lo=1000, hi=64000):
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_nets)).astype(np.int64)
    specs = []
    for k, N in enumerate(sizes):
        ne = int(round(N * 0.8))
        specs.append(dict(n_exc=ne, n_inh=int(N) - ne, seed=10_000 + k, w_exc=float(rng.uniform(0.3, 0.9)),
                          w_inh=float(rng.uniform(4, 9)), I_ext=float(rng.uniform(15, 25)), in_degree=80))
    return specs


def _current_error_f64(torch, specs, bf16, steps, stored_weights=False):
    """Per-step synaptic current of the batched FP32 run against an FP64 twin
    fed with the run's own spike counts and the unrounded weights (with
    `stored_weights`, the weights the run stores: arithmetic error only).

    Each step, before the launch, the twin reads the counts the launch will
    consume and the pre-step potentials V; its conductances follow
    g64 = (g64 + count * w64) * exp(-dt / tau) exactly, and the currents are
    I = g_e (E_e - V) + g_i (E_i - V) from the run's g (FP32) and from g64.
    Conditioning on the run's spikes is the only comparison a chaotic network
    allows.  Returns, per network, max over steps of
    max|I32 - I64| / max|I64| (networks with no current yet are skipped)."""
    from brainpy_amd.dyn.batched import BatchedCOBA, bf16_round
    b = BatchedCOBA(specs, bf16=bf16)
    d, T = b.dev, b.total
    nid = torch.repeat_interleave(torch.arange(len(specs), device=d), torch.tensor(b.sizes, device=d))
    rnd = bf16_round if (bf16 and stored_weights) else (lambda x: float(np.float32(x)))
    we = torch.tensor([rnd(s["w_exc"]) for s in specs], dtype=torch.float64, device=d)[nid]
    wi = torch.tensor([rnd(s["w_inh"]) for s in specs], dtype=torch.float64, device=d)[nid]
    P = b.params
    de = math.exp(-float(np.float32(P.dt)) / float(np.float32(P.tau_e)))
    di = math.exp(-float(np.float32(P.dt)) / float(np.float32(P.tau_i)))
    Ee, Ei = float(np.float32(P.E_e)), float(np.float32(P.E_i))
    ge = torch.zeros(T, dtype=torch.float64, device=d)
    gi = torch.zeros(T, dtype=torch.float64, device=d)
    worst = torch.zeros(len(specs), dtype=torch.float64, device=d)
    for _ in range(steps):
        base = (b.i % b.n_buf) * 2 * T
        ce = b.counts[base:base + T].double()
        ci = b.counts[base + T:base + 2 * T].double()
        V = b.V.double()
        b.step()
        ge = (ge + ce * we) * de
        gi = (gi + ci * wi) * di
        I64 = ge * (Ee - V) + gi * (Ei - V)
        I32 = b.g_exc.double() * (Ee - V) + b.g_inh.double() * (Ei - V)
        num = torch.zeros_like(worst).scatter_reduce(0, nid, (I32 - I64).abs(), "amax")
        den = torch.zeros_like(worst).scatter_reduce(0, nid, I64.abs(), "amax")
        worst = torch.maximum(worst, torch.where(den > 0, num / den.clamp_min(1e-300), torch.zeros_like(num)))
    w = worst.cpu().numpy()
    del b
    return {"max": float(w.max()), "mean": float(w.mean()), "steps": steps}


def _plain_service(torch, spec, steps, window=100):
    """Device-bound time of one network on the plain FP32 path: the
    reference's per-network step (four event jitconn projections, g += out,
    two population updates; ~14 launches) captured in a HIP graph of
    `window` steps with the step index on the device -- the counterpart of
    the reference's one compiled scan per simulation -- replayed until
    `steps` steps are done (a shorter tail window is captured for the rest).
    Graph capture is outside the timed region, as compilation is."""
    from brainpy_amd.dyn.batched import PlainCOBA
    p = PlainCOBA(spec)
    k = min(window, steps)
    p.capture(k)
    rest = steps % k
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // k):
        p.replay()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    if rest:
        p.capture(rest)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        p.replay()
        torch.cuda.synchronize()
        t += time.perf_counter() - t0
    return t, int(p.i)


def bench_batched(_lib, torch, steps=2000, n_nets=256, err_steps=400, plain=True):
    """C4: 256 independent E/I networks (4:1), N log-uniform in [1k, 64k],
    per-network seed, we ~ U(0.3, 0.9), wi ~ U(4, 9), input ~ U(15, 25) mV,
    jitconn in-degree 80, BF16 weights, FP32 state, 2000 steps, one launch
    per step for all networks, captured in HIP graphs of 100 steps.
    Baseline: the plain per-network FP32 path over the whole workload (every
    network, every step), each network one HIP graph of 100 steps replayed
    20 times, networks one after the other."""
    from brainpy_amd.dyn.batched import BatchedCOBA, bf16_round

    def timed_batch(bf16):
        b = BatchedCOBA(specs, bf16=bf16)
        tune = b.autotune(steps=300)   # neurons per block, timed on the first 300 steps (state restored)
        b.capture(100)
        b.replay()
        torch.cuda.synchronize()
        b.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps // 100):
            b.replay()
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        res = (t, int(b.events.sum()), int(b.spike_count.sum()), b.total, b.tile, tune)
        del b
        return res

    specs = _c4_specs(n_nets)
    t_b, events, spikes, total_neurons, tile, tune = timed_batch(True)
    print("# batched C4 autotune (us/step by tile): " + ", ".join(f"{k}: {v:.1f}" for k, v in tune.items())
          + f" -> {tile}", file=sys.stderr, flush=True)
    # FP32 homogeneous weights: scalar kernel arguments, same bytes and kernel
    t_b32, events32, spikes32, _, _, _ = timed_batch(False)
    # per-step synaptic current against an FP64 twin (same spikes): BF16-rounded
    # and FP32 homogeneous weights
    cur_err = {"bf16_weights": _current_error_f64(torch, specs, True, err_steps),
               "bf16_weights_vs_stored": _current_error_f64(torch, specs, True, err_steps, stored_weights=True),
               "f32_weights": _current_error_f64(torch, specs, False, err_steps)}
    rel = []
    for sp in specs:
        for key in ("w_exc", "w_inh"):
            w = np.float64(np.float32(sp[key]))
            rel.append(abs(bf16_round(sp[key]) - w) / w)
    r = {"workload": "batched_coba_256nets_1k-64k_jit80_2000steps", "networks": n_nets,
         "total_neurons": total_neurons, "steps": steps,
         "bf16_weights": {"s": t_b, "sims_per_s": n_nets / t_b, "synaptic_events": events, "spikes": spikes,
                          "events_per_s": events / t_b, "us_per_step": t_b / steps * 1e6,
                          "current_rel_err_vs_f64": cur_err["bf16_weights"]},
         "f32_weights": {"s": t_b32, "sims_per_s": n_nets / t_b32, "synaptic_events": events32, "spikes": spikes32,
                         "events_per_s": events32 / t_b32, "us_per_step": t_b32 / steps * 1e6,
                         "current_rel_err_vs_f64": cur_err["f32_weights"]},
         "bf16_arithmetic_err_vs_f64_same_weights": cur_err["bf16_weights_vs_stored"],
         "bf16_weight_rel_err_max": float(max(rel)), "tile": tile,
         "autotune_us_per_step": {str(k): v for k, v in tune.items()}}
    msg = (f"# batched C4 bf16 {t_b:.3f} s ({t_b/steps*1e6:.1f} us/step) {events/t_b:.3e} ev/s, err vs f64 max "
           f"{cur_err['bf16_weights']['max']:.2e} | f32 {t_b32:.3f} s {events32/t_b32:.3e} ev/s, err max "
           f"{cur_err['f32_weights']['max']:.2e}")
    if plain:
        # plain per-network FP32 path, whole workload (its dynamics equal the
        # FP32 batch bit for bit, so its events are events32)
        t_plain, done = 0.0, 0
        for sp in specs:
            t, n = _plain_service(torch, sp, steps)
            t_plain += t
            done += n
        assert done == steps * n_nets
        r["plain_fp32"] = {"s": t_plain, "sims_per_s": n_nets / t_plain, "events_per_s": events32 / t_plain,
                           "how": "every network, 2000 steps each, HIP graphs of 100 steps, one network at a time"}
        r["speedup_bf16_vs_plain_fp32"] = (events / t_b) / (events32 / t_plain)
        r["speedup_f32_vs_plain_fp32"] = t_plain / t_b32
        msg += (f" | plain fp32 (graphs, whole workload) {t_plain:.2f} s {events32/t_plain:.3e} ev/s -> "
                f"x{r['speedup_bf16_vs_plain_fp32']:.1f} (bf16 batch), x{r['speedup_f32_vs_plain_fp32']:.1f} (f32 batch)")
    print(msg, file=sys.stderr, flush=True)
    return r


def bench_served(_lib, torch, n_requests=80, rate=50.0, seed=7, plain=True, err_steps=400, err_sample=16):
    """C5: Poisson arrivals at 50 req/s of COBA requests, N log-uniform in
    [4k, 1M], durations U[100, 5000] steps, jitconn in-degree 80, FP32
    state, admitted against a 200 GB HBM budget (the pools grow with the
    admitted state up to it).  Continuous batching with BF16 and with FP32
    weights, a FIFO queue of one request at a time on the same kernel, and
    the plain per-network FP32 path serving the whole trace one request at a
    time (each request one HIP-graph run, see _plain_service; latencies from
    the arrival times and the measured service times)."""
    from brainpy_amd.dyn.serve import DONE, Request, SimServer
    rng = np.random.Generator(np.random.PCG64(seed))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    sizes = np.exp(rng.uniform(np.log(4000), np.log(1_000_000), n_requests)).astype(np.int64)
    durs = rng.integers(100, 5001, n_requests)
    params = [(int(sizes[k]), 50_000 + k, float(rng.uniform(0.3, 0.9)), float(rng.uniform(4, 9)),
               float(rng.uniform(15, 25)), int(durs[k]), float(arrivals[k])) for k in range(n_requests)]

    def run(max_active, bf16=True):
        srv = SimServer(hbm_budget=200 * (1 << 30), bf16=bf16, max_active=max_active)
        reqs = [Request(k, n, s, we, wi, I, st, arrival=a) for k, (n, s, we, wi, I, st, a) in enumerate(params)]
        done = srv.serve(reqs)
        assert all(r.status == DONE for r in done) and len(done) == n_requests
        lat = [r.completed - r.arrival for r in done]
        ev = sum(r.events for r in done)
        out = {"completed": len(done), "wall_s": srv.wall, "p50_latency_s": _percentile(lat, 50),
               "p99_latency_s": _percentile(lat, 99), "events": ev, "events_per_s": ev / srv.wall,
               "sims_per_s": len(done) / srv.wall, "steps_launched": srv.step_idx,
               "pool_neurons": srv.capacity}
        del srv
        return out, {r.rid: r.events for r in done}

    served, _ = run(None)
    served32, ev32 = run(None, bf16=False)
    fifo, _ = run(1)
    # synaptic current against FP64 on a size-stratified sample of the
    # requests' networks (same kernel, same per-network step)
    order = np.argsort(sizes)
    pick = order[np.linspace(0, n_requests - 1, err_sample).astype(int)]
    sample = []
    for k in pick:
        n_k, seed_k, we, wi, I, _, _ = params[k]
        ne = int(round(n_k * 0.8))
        sample.append(dict(n_exc=ne, n_inh=n_k - ne, seed=seed_k, w_exc=we, w_inh=wi, I_ext=I, in_degree=80))
    served["current_rel_err_vs_f64"] = _current_error_f64(torch, sample, True, err_steps)
    served32["current_rel_err_vs_f64"] = _current_error_f64(torch, sample, False, err_steps)
    r = {"workload": f"served_coba_poisson{rate:g}_{n_requests}req_4k-1M_100-5000steps_jit80",
         "requests": n_requests, "continuous_batching_bf16": served, "continuous_batching_f32": served32,
         "fifo_queue_bf16": fifo,
         "throughput_ratio_vs_fifo": served["events_per_s"] / fifo["events_per_s"] if fifo["events_per_s"] else None,
         "current_error_sample": f"{err_sample} requests spread over the size range, {err_steps} steps"}
    msg = (f"# served C5 batching bf16 {served['events_per_s']:.3e} ev/s p50 {served['p50_latency_s']:.3f} s "
           f"p99 {served['p99_latency_s']:.3f} s err {served['current_rel_err_vs_f64']['max']:.2e} | f32 "
           f"{served32['events_per_s']:.3e} ev/s p99 {served32['p99_latency_s']:.3f} s err "
           f"{served32['current_rel_err_vs_f64']['max']:.2e} | fifo {fifo['events_per_s']:.3e} ev/s p50 "
           f"{fifo['p50_latency_s']:.3f} s p99 {fifo['p99_latency_s']:.3f} s")
    if plain:
        # the plain FP32 path serving the trace in arrival order, one request
        # at a time (its per-request dynamics, hence events, equal the FP32
        # served run's)
        t_free, lat, busy, ev = 0.0, [], 0.0, 0
        for k in np.argsort(arrivals):
            n_k, seed_k, we, wi, I, st, a = params[k]
            ne = int(round(n_k * 0.8))
            svc, _ = _plain_service(torch, dict(n_exc=ne, n_inh=n_k - ne, seed=seed_k, w_exc=we, w_inh=wi,
                                                I_ext=I, in_degree=80), st)
            start = max(a, t_free)
            t_free = start + svc
            busy += svc
            ev += ev32[k]
            lat.append(t_free - a)
        r["plain_fp32_fifo"] = {"busy_s": busy, "events_per_s": ev / max(t_free - float(arrivals.min()), busy),
                                "events_per_busy_s": ev / busy, "p50_latency_s": _percentile(lat, 50),
                                "p99_latency_s": _percentile(lat, 99),
                                "how": "every request of the trace run on the plain path (HIP graphs of 100 "
                                       "steps), one at a time in arrival order; latency = measured service "
                                       "times queued on the arrival times"}
        r["throughput_vs_plain"] = served["events_per_s"] / r["plain_fp32_fifo"]["events_per_busy_s"]
        msg += (f" | plain fp32 fifo {r['plain_fp32_fifo']['events_per_busy_s']:.3e} ev/busy-s p50 "
                f"{r['plain_fp32_fifo']['p50_latency_s']:.2f} s p99 {r['plain_fp32_fifo']['p99_latency_s']:.2f} s"
                f" -> x{r['throughput_vs_plain']:.1f}")
    print(msg, file=sys.stderr, flush=True)
    return r


def bench_batch_ops(_lib, torch, batches=(16, 64), reps=5):
    """Batched event operators (what torch.func.vmap over the events runs)
    against a loop of single calls: jitconn 1M x 1M p=1e-4 row-parallel
    homo (2 % spikes), and the 100k x 100k p=0.01 CSR with heterogeneous
    f32 weights (10 % spikes) in both directions (scatter on the tiled plan
    and stateless, gather)."""
    import brainpy_amd.math as bm
    from brainpy_amd.connect import fixed_prob_csr_device
    dev = _lib.device()
    res = []

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    n = 1_000_000
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    for B in batches:
        E = torch.rand((B, n), generator=g, device=dev) < 0.02
        one = lambda e: bm.jitconn.event_mv_prob_homo(e, 1.0, 1e-4, 42, shape=(n, n))  # noqa: E731
        t_b = timed(lambda: torch.func.vmap(one)(E))
        t_l = timed(lambda: [one(E[b]) for b in range(B)])
        res.append({"op": "jitconn event_mv_prob_homo 1Mx1M p=1e-4 (row-parallel)", "batch": B,
                    "batched_ms": t_b * 1e3, "loop_ms": t_l * 1e3, "speedup": t_l / t_b})
        del E
    m = 100_000
    idx, ptr = fixed_prob_csr_device(m, m, 0.01, seed=2024, device=dev)
    w = torch.empty(idx.numel(), dtype=torch.float32, device=dev).normal_(0.5, 0.1, generator=g)
    for B in batches:
        E = torch.rand((B, m), generator=g, device=dev) < 0.1
        for transpose, stateless, label in ((True, False, "scatter, tiled plan"), (True, True, "scatter, stateless"),
                                            (False, False, "gather")):
            if stateless:
                bm.event.set_plan_cache(enabled=False)
            one = lambda e: bm.event.csrmv(w, idx, ptr, e, shape=(m, m), transpose=transpose)  # noqa: E731
            t_b = timed(lambda: bm.event.csrmv_batch(w, idx, ptr, E, shape=(m, m), transpose=transpose,
                                                     stateless=stateless))
            t_l = timed(lambda: [one(E[b]) for b in range(B)])
            bm.event.set_plan_cache(enabled=True)
            res.append({"op": f"event csrmv 100kx100k p=0.01 heter f32 10% ({label})", "batch": B,
                        "batched_ms": t_b * 1e3, "loop_ms": t_l * 1e3, "speedup": t_l / t_b})
        del E
    for r in res:
        print(f"# batch {r['op']:60s} B={r['batch']:3d} batched {r['batched_ms']:8.2f} ms loop {r['loop_ms']:8.2f} ms"
              f" -> x{r['speedup']:.2f}", file=sys.stderr, flush=True)
    return {"workload": "vmap over event vectors, B = " + "/".join(map(str, batches)), "results": res}


HEADLINE_PROFILE = "event_scatter_100k_p0.01_rate10_heter_f32.md"


def measured_traffic(name=HEADLINE_PROFILE):
    """HBM bytes per launch of the headline kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (copy-calibrated) and the
    summary's own date line, or (None, None)."""
    path = os.path.join(ROOT, "profiles", name)
    traffic, date, kus = None, None, None
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = int(line.split(":")[1])
            elif line.startswith("measured:"):
                date = line.split(":", 1)[1].split()[0]
            elif line.startswith("kernel_mean_us:"):
                kus = float(line.split(":")[1])
    except OSError:
        pass
    return traffic, date, kus


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(indices, indptr, w, evs, events_per_call, seconds, gpu_out):
    """The headline call on the host: the oracle's scatter (the reference CPU
    path's loop, event/_csr_matvec.py:321-340: serial over rows, an f32
    `+=` per synapse) built here with -O3 -march=native (oracle/Makefile
    `native`; the prebuilt -O2 library if the compiler is unavailable), timed
    call by call over about `seconds` at 1 thread (median), then a threaded
    variant (private outputs, summed) at the machine's thread share; the
    calls take the spike vectors `evs` in turn (the first one is checked
    against `gpu_out`)."""
    import ctypes
    import subprocess
    import tempfile
    flags = "gcc -O3 -march=native -ffp-contract=off -fopenmp"
    path = os.path.join(tempfile.mkdtemp(prefix="bpx_cpu_"), "liborc_native.so")
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "native", f"OUT={path}"], check=True,
                       timeout=120, capture_output=True)
    except Exception:
        path = os.path.join(ROOT, "oracle", "liborc.so")
        flags = "gcc -O2 -ffp-contract=off -fopenmp (prebuilt)"
    L = ctypes.CDLL(path)
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    L.orc_event_csrmv_scatter_f32.argtypes = [vp, i64, vp, vp, vp, i64, i64, vp]
    L.orc_event_csrmv_scatter_f32_threads.argtypes = [vp, i64, vp, vp, vp, i64, i64, vp, ctypes.c_int, vp]
    h_idx = np.ascontiguousarray(indices.cpu().numpy())
    h_ptr = np.ascontiguousarray(indptr.cpu().numpy())
    h_w = np.ascontiguousarray(w.cpu().numpy())
    h_evs = [np.ascontiguousarray(ev.cpu().numpy().view(np.uint8)) for ev in evs]
    m, n = h_ptr.size - 1, N
    out = np.empty(n, np.float32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731

    def run(fn, i, *extra):
        fn(P(h_w), h_w.size, P(h_idx), P(h_ptr), P(h_evs[i % len(h_evs)]), m, n, P(out), *extra)

    def timed(fn, budget, *extra):
        run(fn, 0, *extra)   # warm-up
        times = []
        t_end = time.perf_counter() + budget
        while len(times) < 20 or (time.perf_counter() < t_end and len(times) < 5000):
            t0 = time.perf_counter()
            run(fn, len(times) + 1, *extra)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), len(times)

    run(L.orc_event_csrmv_scatter_f32, 0)
    rel = float(np.max(np.abs(gpu_out - out) / np.maximum(np.abs(out), 1.0)))
    med1, calls1 = timed(L.orc_event_csrmv_scatter_f32, seconds)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    scratch = np.empty(threads * n, np.float32)
    medt, callst = timed(L.orc_event_csrmv_scatter_f32_threads, seconds / 3, threads, P(scratch))
    return {"value": events_per_call / med1, "unit": "synaptic events/s", "cores": 1, "kind": "port",
            "sample": f"median of {calls1} calls of the headline call (10% spikes, heter f32, 1e7 events/call, "
                      f"{len(h_evs)} spike vectors in turn), serial scatter of oracle/bpx_oracle.c",
            "cpu": _cpu_model(), "flags": flags,
            "all_threads": {"value": events_per_call / medt, "threads": threads, "calls": callst},
            "gpu_vs_cpu_max_rel_diff": rel}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-sweep", action="store_true", help="time the headline configuration only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the jitconn and network legs")
    ap.add_argument("--only", choices=["jitconn", "batched", "coba", "smallbatch", "served", "coo", "forms", "conn", "stdp",
                                       "bsr", "batchops", "bign"],
                    default=None,
                    help="run only this leg (profiling)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run 3 plain device copies of this many MiB (HBM counter calibration)")
    ap.add_argument("--detail", default=os.path.join(ROOT, "bench_detail.json"),
                    help="file that receives the sweep and the other legs (the last stdout line is the headline)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the library runs on a single MI355X and "
              "has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    import torch
    from brainpy_amd import _lib
    from brainpy_amd.connect import fixed_prob_csr_device
    import brainpy_amd.math as bm

    dev = _lib.device()
    L = _lib.lib()
    if args.only:
        fn = {"jitconn": lambda: bench_jitconn(L, _lib, torch, max(10, args.steps // 10), 3),
              "batched": lambda: bench_batched(_lib, torch),
              "coba": lambda: bench_coba(_lib, torch), "smallbatch": lambda: bench_small_batch(torch),
              "served": lambda: bench_served(_lib, torch),
              "coo": lambda: bench_coo(_lib, torch), "forms": lambda: bench_forms(_lib, torch),
              "conn": lambda: bench_conn(_lib, torch), "stdp": lambda: bench_stdp(_lib, torch),
              "bsr": lambda: bench_bsr(_lib, torch), "batchops": lambda: bench_batch_ops(_lib, torch),
              "bign": lambda: bench_bign(_lib, torch)}[args.only]
        print(json.dumps({"only": args.only, "result": fn()}), flush=True)
        return
    sp = _lib.stream_ptr()
    print(f"# building {M}x{N} CSR, {K_ROW}/row, nnz={M * K_ROW:.3g}", file=sys.stderr, flush=True)
    indices, indptr = fixed_prob_csr_device(M, N, PROB, seed=2024, device=dev)
    nnz = indices.numel()
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    w_f32 = torch.empty(nnz, dtype=torch.float32, device=dev).normal_(0.5, 0.1, generator=gen)
    weights = {
        "homo": torch.ones(1, dtype=torch.float32, device=dev),
        "heter_f32": w_f32,
        "heter_bf16": w_f32.to(torch.bfloat16),
    }
    # NV distinct Bernoulli vectors per rate, taken in turn by the timed
    # calls: the active rows' segments of one call (63 MB at 10 %) fit the
    # 256 MB Infinity Cache, so a vector repeated every call would be timed
    # from that cache; NV vectors (1 GB at 10 %) keep each call's rows in
    # HBM, as a simulation's changing spikes do.  spikes[r] is the first.
    spike_sets = {}
    for r in RATES:
        g = torch.Generator(device=dev)
        g.manual_seed(int(r * 1e6) + 11)
        spike_sets[r] = [torch.rand(M, generator=g, device=dev) < r for _ in range(NV)]
    spikes = {r: v[0] for r, v in spike_sets.items()}
    out = torch.empty(N, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    import ctypes

    def call_public(kind, r, i=0):
        # bm.event.csrmv(data, indices, indptr, events, shape, transpose=True):
        # the reference's operator signature; repeated calls on one matrix
        # run on the cached tiled plan
        return bm.event.csrmv(weights[kind], indices, indptr, spike_sets[r][i], shape=(M, N), transpose=True)

    def call_csr(kind, r, i=0):
        w = weights[kind]
        rc = L.bpx_event_csrmv(sp, _lib.ptr(w), _lib.bpx_dtype(w), w.numel(), _lib.ptr(indices), _lib.ptr(indptr),
                               _lib.BPX_I32, nnz, _lib.ptr(spike_sets[r][i]), _lib.BPX_BOOL, _lib.RULE_POSITIVE, M, N,
                               1, _lib.ptr(out))
        if rc != 0:
            raise RuntimeError(f"bpx_event_csrmv failed ({rc})")

    # the first public call on a matrix is stateless, the second builds the
    # plan (its one-time cost, reported beside the per-call time)
    build_ms = {}
    for kind in ("homo", "heter_f32", "heter_bf16"):
        call_public(kind, HEADLINE[1])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        call_public(kind, HEADLINE[1])
        torch.cuda.synchronize()
        build_ms[kind] = (time.perf_counter() - t0) * 1e3
    info = bm.event.plan_cache_info()
    if info["plans"] != 3:
        raise RuntimeError(f"the public path did not build its plans: {info}")
    global TILES
    TILES = next(p for p, _ in bm.event._PLANS._plans.values() if p is not None).plan.tiles

    def measure(kind, r, steps, warmup, path="public", rotate=True):
        # call i takes spike vector i % nv (nv = 1: the first vector every call)
        call = call_public if path == "public" else call_csr
        kid = _lib.KID["tiled"] if path == "public" else _lib.KID["event_scatter"]
        nv = len(spike_sets[r]) if rotate else 1
        steps = (steps + nv - 1) // nv * nv
        active = sum(int(v.sum().item()) for v in spike_sets[r][:nv]) / nv
        events = active * K_ROW
        for i in range(warmup):
            call(kind, r, i % nv)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            call(kind, r, i % nv)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # device time of the scatter kernel alone, on the launch stream
        _lib.timer_reset()
        _lib.timer_enable(True)
        for i in range(steps):
            call(kind, r, i % nv)
        torch.cuda.synchronize()
        _lib.timer_enable(False)
        ms, launches = _lib.timer_read(kid)
        kern_s = ms / 1e3 / max(launches, 1)
        abytes = algorithmic_bytes(active, kind)
        lbytes = layout_bytes(active, kind, "tiled" if path == "public" else "csr")
        return {
            "path": path, "kind": kind, "rate": r, "spike_vectors": nv, "steps": steps, "active_rows": active,
            "events_per_call": events,
            "ms_per_step": wall / steps * 1e3, "events_per_s": events * steps / wall,
            "kernel_us": kern_s * 1e6, "kernel_launches": launches, "algorithmic_bytes": abytes,
            "kernel_GBs": abytes / kern_s / 1e9, "layout_bytes": lbytes, "layout_GBs": lbytes / kern_s / 1e9,
        }

    if args.copy_mib:
        src = torch.empty(args.copy_mib << 20, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            probe_lib().bpx_probe_copy(sp, _lib.ptr(src), _lib.ptr(dst), src.numel())
        torch.cuda.synchronize()
        del src, dst
    # parity of the public path: integer counts exact against the stateless
    # kernel; heterogeneous weights against an FP64 sum of the same weights
    call_csr("homo", HEADLINE[1])
    ref = out.clone()
    if not torch.equal(call_public("homo", HEADLINE[1]), ref):
        raise RuntimeError("public path (plan) disagrees with the stateless kernel on integer counts")
    ev = spikes[HEADLINE[1]]
    rows = torch.nonzero(ev).flatten()
    starts = indptr[rows].long()
    lens = (indptr[rows + 1] - indptr[rows]).long()
    first = torch.cumsum(lens, 0) - lens
    ent = torch.repeat_interleave(starts - first, lens) + torch.arange(int(lens.sum()), device=dev)
    ref64 = torch.zeros(N, dtype=torch.float64, device=dev)
    ref64.index_add_(0, indices[ent].long(), w_f32[ent].double())
    mag64 = torch.zeros(N, dtype=torch.float64, device=dev)
    mag64.index_add_(0, indices[ent].long(), w_f32[ent].double().abs())
    del ent
    f64_err = {}
    for kind in ("heter_f32", "heter_bf16"):
        y = call_public(kind, HEADLINE[1])
        torch.cuda.synchronize()
        e = (y.double() - ref64).abs() / mag64.clamp_min(1e-30)
        f64_err[kind] = {"max": float(e.max()), "mean": float(e.mean())}
    if f64_err["heter_f32"]["max"] > 1e-5:
        raise RuntimeError(f"public path error against FP64 above 1e-5: {f64_err}")
    detail = {"plan_build_ms": build_ms, "plan_cache": bm.event.plan_cache_info(),
              "current_rel_err_vs_f64": dict(f64_err, against="FP64 sum of the FP32 weights, relative to the "
                                                              "sum of |w| of each output")}
    sweep = []
    if not args.no_sweep:
        for path in ("csr", "public"):
            for kind in ("homo", "heter_f32", "heter_bf16"):
                for r in RATES:
                    res = measure(kind, r, max(20, args.steps // 4), max(5, args.warmup // 4), path)
                    sweep.append(res)
                    print(f"# {path:6s} {kind:10s} rate={r:<6} {res['ms_per_step']*1e3:8.1f} us/call "
                          f"kernel {res['kernel_us']:8.1f} us {res['kernel_GBs']:8.1f} GB/s "
                          f"{res['events_per_s']:.3e} ev/s", file=sys.stderr, flush=True)
    detail["sweep"] = [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()} for s in sweep]
    # graph-amortised: 96 public calls captured in one HIP graph (on the
    # cached plan, the NV spike vectors in turn); two replays are checked
    # against the eager result of the last call
    graph = []
    ng = 6 * NV
    if not args.no_sweep:
        for r in RATES:
            ref_out = call_public(HEADLINE[0], r, (ng - 1) % NV).clone()
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            cs = torch.cuda.Stream()
            cs.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(cs):
                with torch.cuda.graph(g, stream=cs):
                    for i in range(ng):
                        gout = call_public(HEADLINE[0], r, i % NV)
            torch.cuda.current_stream().wait_stream(cs)
            torch.cuda.synchronize()
            reps = 5
            for _ in range(2):
                gout.fill_(-1.0)
                g.replay()
                torch.cuda.synchronize()
                # (the output copies are flushed with FP32 atomics, so the
                # last bits depend on their order)
                if not torch.allclose(gout, ref_out, rtol=1e-5, atol=1e-5):
                    raise RuntimeError(f"graph replay differs from the eager result (rate {r})")
            t0 = time.perf_counter()
            for _ in range(reps):
                g.replay()
            torch.cuda.synchronize()
            per_call = (time.perf_counter() - t0) / (reps * ng)
            active = sum(int(v.sum().item()) for v in spike_sets[r]) / NV
            graph.append({"kind": HEADLINE[0], "rate": r, "us_per_call": per_call * 1e6,
                          "events_per_s": active * K_ROW / per_call})
            print(f"# graph public {HEADLINE[0]} rate={r:<6} {per_call*1e6:8.1f} us/call "
                  f"{active * K_ROW / per_call:.3e} ev/s", file=sys.stderr, flush=True)
            del g, gout
    detail["graph_amortised"] = graph
    head = measure(HEADLINE[0], HEADLINE[1], args.steps, args.warmup, "public")
    plain = measure(HEADLINE[0], HEADLINE[1], args.steps, args.warmup, "csr")
    cached = measure(HEADLINE[0], HEADLINE[1], args.steps, args.warmup, "public", rotate=False)
    print(f"# headline {head}", file=sys.stderr, flush=True)
    print(f"# the first spike vector every call (its rows' segments stay in the Infinity Cache): kernel "
          f"{cached['kernel_us']:.2f} us ({100 * cached['kernel_GBs'] / HBM_PEAK_GBS:.1f}% of HBM peak)",
          file=sys.stderr, flush=True)
    detail["headline_one_vector"] = cached
    print(f"# stateless scatter (first call on a matrix) {plain['ms_per_step']*1e3:.1f} us/call, "
          f"kernel {plain['kernel_us']:.1f} us", file=sys.stderr, flush=True)
    detail["headline"] = head
    detail["stateless_path"] = plain
    high = []
    if not args.no_sweep:
        for r in (0.5, 1.0):
            gr = torch.Generator(device=dev)
            gr.manual_seed(int(r * 1e6) + 11)
            # (half or all of the plan per call: more than the Infinity Cache holds)
            spikes[r] = torch.rand(M, generator=gr, device=dev) < r
            spike_sets[r] = [spikes[r]]
            res_r = measure(HEADLINE[0], r, 50, 5, "public")
            high.append(res_r)
            print(f"# public {HEADLINE[0]} rate={r:<6} {res_r['ms_per_step']*1e3:8.1f} us/call kernel "
                  f"{res_r['kernel_us']:8.1f} us {res_r['kernel_GBs']:7.0f} GB/s "
                  f"({100 * res_r['kernel_GBs'] / HBM_PEAK_GBS:.1f}%) {res_r['events_per_s']:.3e} ev/s",
                  file=sys.stderr, flush=True)
            del spikes[r], spike_sets[r]
    detail["high_rate"] = high
    print(f"# current error vs f64 (10% heter, per output / sum|w|): f32 max {f64_err['heter_f32']['max']:.2e}, "
          f"bf16 max {f64_err['heter_bf16']['max']:.2e}", file=sys.stderr, flush=True)

    extra = {}
    if not args.no_extra:
        legs = (("jitconn", lambda: bench_jitconn(L, _lib, torch, max(10, args.steps // 10), 3, cpu=not args.no_cpu)),
                ("coba", lambda: bench_coba(_lib, torch)), ("small_batch", lambda: bench_small_batch(torch)),
                ("batched", lambda: bench_batched(_lib, torch)), ("served", lambda: bench_served(_lib, torch)),
                ("batch_ops", lambda: bench_batch_ops(_lib, torch)), ("coo", lambda: bench_coo(_lib, torch)),
                ("forms", lambda: bench_forms(_lib, torch)), ("bign", lambda: bench_bign(_lib, torch)),
                ("conn", lambda: bench_conn(_lib, torch)), ("stdp", lambda: bench_stdp(_lib, torch)),
                ("bsr", lambda: bench_bsr(_lib, torch)))
        for name, leg in legs:
            # a leg that fails (its own parity check, an error status) is
            # reported in the detail and on stderr; the headline, measured
            # and checked above, is still printed
            try:
                extra[name] = leg()
            except Exception as e:  # noqa: BLE001
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
                print(f"# leg {name} failed: {type(e).__name__}: {e}", file=sys.stderr, flush=True)
                torch.cuda.synchronize()
    detail["extra"] = extra

    cpu = None
    if not args.no_cpu:
        y = call_public(HEADLINE[0], HEADLINE[1]).cpu().numpy()
        cpu = cpu_baseline(indices, indptr, w_f32, spike_sets[HEADLINE[1]], head["events_per_call"], args.cpu_seconds,
                           y)
        print(f"# cpu {cpu}", file=sys.stderr, flush=True)
        detail["cpu_baseline"] = cpu
    try:
        os.makedirs(os.path.dirname(args.detail), exist_ok=True)
        with open(args.detail, "w") as f:
            json.dump(detail, f, indent=1)
    except OSError as e:
        print(f"# could not write {args.detail}: {e}", file=sys.stderr)

    traffic, traffic_date, traffic_kus = measured_traffic()
    line = {
        "metric": "synaptic events/s",
        "value": head["events_per_s"],
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded FixedProb CSR and Bernoulli spikes, generated on the GPU)",
        "config": {"workload": "event_csrmv_transpose_100kx100k_p0.01_rate0.10_heter_f32", "m": M, "n": N,
                   "nnz": nnz, "spike_rate": HEADLINE[1], "weights": HEADLINE[0], "index_dtype": "int32",
                   "path": "bm.event.csrmv(transpose=True) on its cached tiled plan",
                   "plan_build_ms": round(build_ms[HEADLINE[0]], 3), "events_per_step": head["events_per_call"],
                   "spike_vectors": head["spike_vectors"]},
        "roofline": {"bound": "hbm", "achieved": round(head["kernel_GBs"], 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(head["kernel_GBs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": f"profiles/{HEADLINE_PROFILE} ({traffic_date})",
                     # the profiled kernel's mean duration: a live kernel_us far from it
                     # means the committed counters describe another build
                     "traffic_profile_kernel_us": traffic_kus,
                     "kernel": "k_tiled_scatter", "kernel_us": round(head["kernel_us"], 3),
                     "algorithmic_bytes_per_launch": head["algorithmic_bytes"],
                     # one spike vector every call: its rows' segments stay in the
                     # 256 MB Infinity Cache between calls
                     "one_vector": {"kernel_us": round(cached["kernel_us"], 3),
                                    "frac": round(cached["kernel_GBs"] / HBM_PEAK_GBS, 4)}},
        "cpu_baseline": None if cpu is None else {
            "value": cpu["value"], "unit": cpu["unit"], "cores": cpu["cores"], "kind": cpu["kind"],
            "sample": cpu["sample"], "cpu": cpu["cpu"], "flags": cpu["flags"],
            "all_threads": {"value": cpu["all_threads"]["value"], "threads": cpu["all_threads"]["threads"]}},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
