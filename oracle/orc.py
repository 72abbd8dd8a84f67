"""ctypes binding of the C oracle ``liborc.so`` (test infrastructure only).

Each wrapper takes numpy arrays and returns numpy arrays; the C functions
they call cite the reference loops they restate.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# BPX_ORC_LIB selects another build of the same source (the sanitized one)
_LIB_PATH = os.environ.get("BPX_ORC_LIB") or os.path.join(_HERE, "liborc.so")
_lib = None

# dtype codes (include/bpx.h)
DT_BOOL, DT_F32, DT_F64, DT_I32, DT_I64, DT_BF16, DT_F16, DT_U32, DT_U64, DT_I8, DT_U8, DT_I16, DT_U16 = range(13)
RULE_POSITIVE, RULE_NONZERO = 0, 1
LAYOUT_LANE32, LAYOUT_ROW = 0, 1
DIST = {"homo": 0, "uniform": 1, "normal": 2}

_NP2DT = {
    np.dtype(np.bool_): DT_BOOL, np.dtype(np.float32): DT_F32, np.dtype(np.float64): DT_F64,
    np.dtype(np.int32): DT_I32, np.dtype(np.int64): DT_I64, np.dtype(np.float16): DT_F16,
    np.dtype(np.uint32): DT_U32, np.dtype(np.uint64): DT_U64, np.dtype(np.int8): DT_I8,
    np.dtype(np.uint8): DT_U8, np.dtype(np.int16): DT_I16, np.dtype(np.uint16): DT_U16,
}


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i64, i32, u32, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32, ctypes.c_float
        L.orc_event_info.argtypes = [vp, i32, i64, vp, vp]
        L.orc_event_info_batch.argtypes = [vp, i32, i64, i64, vp, vp]
        L.orc_event_csrmv.argtypes = [vp, i32, i64, vp, i32, vp, i32, vp, i32, i64, i64, i32, i32, vp]
        L.orc_event_csrmv_scatter_f32.argtypes = [vp, i64, vp, vp, vp, i64, i64, vp]
        L.orc_csrmv.argtypes = [vp, i32, i64, vp, i32, vp, i32, vp, i32, i64, i64, i32, vp]
        L.orc_lfsr88_stream.argtypes = [u32, i64, vp]
        L.orc_lfsr88_randint_stream.argtypes = [u32, ctypes.c_int32, ctypes.c_int32, i64, vp]
        L.orc_lfsr88_uniform_stream.argtypes = [u32, f32, f32, i64, vp]
        L.orc_lfsr88_normal_stream.argtypes = [u32, f32, f32, i64, vp]
        L.orc_lfsr88_component_period.argtypes = [i32, u32, i64]
        L.orc_lfsr88_component_period.restype = ctypes.c_int64
        L.orc_jitconn_mv.argtypes = [i32, i32, i32, i32, vp, i32, i64, i64, f32, f32, ctypes.c_int32, u32, vp]
        L.orc_jitconn_enumerate.argtypes = [i32, i32, i32, i64, i64, f32, f32, ctypes.c_int32, u32, vp, vp, vp, i64]
        L.orc_jitconn_enumerate.restype = ctypes.c_int64
        L.bpx_expf_ref.argtypes = [f32]
        L.bpx_expf_ref.restype = f32
        L.orc_expon_coef.argtypes = [f32, f32]
        L.orc_expon_coef.restype = f32
        fp = ctypes.POINTER(ctypes.c_void_p)
        L.orc_lif_coba_step.argtypes = [i64, i32, i32, vp, vp, vp, fp, vp, vp, vp, f32, vp,
                                        f32, f32, f32, f32, f32, f32, f32, f32]
        L.orc_coba_net_run.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, vp, f32, f32, f32, vp, i64, i32,
                                       f32, f32, f32, f32, f32, f32, f32, f32, f32, f32, f32, vp, vp, vp]
        L.orc_coba_net_run.restype = ctypes.c_int64
        d = ctypes.c_double
        L.orc_lif_coba_step_f64.argtypes = [i64, i32, i32, vp, vp, vp, fp, vp, vp, d, vp,
                                            d, d, d, d, d, d, d, d]
        L.orc_csr_on_pre_update.argtypes = [vp, vp, vp, i64, vp, i64, vp, i64, f32, f32, vp]
        L.orc_csc_on_post_update.argtypes = [vp, i64, vp, vp, vp, vp, i64, vp, i64, f32, f32, vp]
        L.orc_dense_on_spike.argtypes = [i32, vp, vp, i64, i64, vp, f32, f32, vp]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def dt_code(a):
    return _NP2DT[np.dtype(a.dtype)]


# --------------------------------------------------------------------------
def event_info(events):
    ev = np.ascontiguousarray(events)
    n = ev.shape[-1]
    if ev.ndim == 1:
        ids = np.empty(n, np.int32)
        num = np.empty(1, np.int32)
        lib().orc_event_info(_p(ev), dt_code(ev), n, _p(ids), _p(num))
    else:
        b = ev.shape[0]
        ids = np.empty((b, n), np.int32)
        num = np.empty(b, np.int32)
        lib().orc_event_info_batch(_p(ev), dt_code(ev), b, n, _p(ids), _p(num))
    return ids, num


def _data_code(data):
    if data.dtype == np.uint16:       # raw bfloat16 bits
        return DT_BF16
    return dt_code(data)


def event_csrmv(data, indices, indptr, events, shape, transpose, rule=RULE_POSITIVE, bf16=False):
    data = np.ascontiguousarray(np.atleast_1d(data))
    indices, indptr, events = (np.ascontiguousarray(x) for x in (indices, indptr, events))
    m, n = shape
    out_dtype = np.float64 if data.dtype == np.float64 else np.float32
    out = np.empty(n if transpose else m, out_dtype)
    dcode = DT_BF16 if bf16 else dt_code(data)
    lib().orc_event_csrmv(_p(data), dcode, data.shape[0], _p(indices), dt_code(indices), _p(indptr),
                          dt_code(indptr), _p(events), dt_code(events), m, n, int(transpose), rule, _p(out))
    return out


def event_csrmv_scatter_f32(data, indices, indptr, events, shape, out=None):
    m, n = shape
    if out is None:
        out = np.empty(n, np.float32)
    lib().orc_event_csrmv_scatter_f32(_p(data), data.shape[0], _p(indices), _p(indptr), _p(events), m, n, _p(out))
    return out


def csrmv(data, indices, indptr, vector, shape, transpose):
    data = np.ascontiguousarray(np.atleast_1d(data))
    indices, indptr, vector = (np.ascontiguousarray(x) for x in (indices, indptr, vector))
    m, n = shape
    out_dtype = np.float64 if data.dtype == np.float64 else np.float32
    out = np.empty(n if transpose else m, out_dtype)
    lib().orc_csrmv(_p(data), dt_code(data), data.shape[0], _p(indices), dt_code(indices), _p(indptr),
                    dt_code(indptr), _p(vector), dt_code(vector), m, n, int(transpose), _p(out))
    return out


# --------------------------------------------------------------------------
def lfsr88_stream(seed, count):
    out = np.empty(count, np.uint32)
    lib().orc_lfsr88_stream(seed & 0xFFFFFFFF, count, _p(out))
    return out


def lfsr88_randint_stream(seed, low, high, count):
    out = np.empty(count, np.int32)
    lib().orc_lfsr88_randint_stream(seed & 0xFFFFFFFF, low, high, count, _p(out))
    return out


def lfsr88_uniform_stream(seed, low, high, count):
    out = np.empty(count, np.float32)
    lib().orc_lfsr88_uniform_stream(seed & 0xFFFFFFFF, low, high, count, _p(out))
    return out


def lfsr88_normal_stream(seed, mu, sigma, count):
    out = np.empty(count, np.float32)
    lib().orc_lfsr88_normal_stream(seed & 0xFFFFFFFF, mu, sigma, count, _p(out))
    return out


def lfsr88_component_period(comp, start, limit):
    return lib().orc_lfsr88_component_period(comp, start, limit)


def jitconn_mv(vec, dist, vmode, outdim_parallel, layout, num_row, num_col, w0, w1, clen, seed):
    """y = M @ vec with M of shape (num_row, num_col) (already shape-swapped)."""
    vec = np.ascontiguousarray(vec)
    out = np.empty(num_row, np.float32)
    rc = lib().orc_jitconn_mv(DIST[dist] if isinstance(dist, str) else dist, vmode, int(outdim_parallel), layout,
                              _p(vec), dt_code(vec), num_row, num_col, np.float32(w0), np.float32(w1),
                              int(clen), seed & 0xFFFFFFFF, _p(out))
    assert rc == 0
    return out


def jitconn_enumerate(dist, outdim_parallel, layout, num_row, num_col, w0, w1, clen, seed):
    L = lib()
    cnt = L.orc_jitconn_enumerate(DIST[dist] if isinstance(dist, str) else dist, int(outdim_parallel), layout,
                                  num_row, num_col, np.float32(w0), np.float32(w1), int(clen), seed & 0xFFFFFFFF,
                                  None, None, None, 0)
    rows = np.empty(cnt, np.int32)
    cols = np.empty(cnt, np.int32)
    ws = np.empty(cnt, np.float32)
    L.orc_jitconn_enumerate(DIST[dist] if isinstance(dist, str) else dist, int(outdim_parallel), layout,
                            num_row, num_col, np.float32(w0), np.float32(w1), int(clen), seed & 0xFFFFFFFF,
                            _p(rows), _p(cols), _p(ws), cnt)
    return rows, cols, ws


def expf(x):
    return np.float32(lib().bpx_expf_ref(np.float32(x)))


def expon_coef(tau, dt):
    return np.float32(lib().orc_expon_coef(np.float32(tau), np.float32(dt)))


def lif_coba_step(V, t_last, spike, gs, tau_g, E_g, c_g, I_ext, params, t, dt, ltc=True, I_vec=None):
    """In-place f32 step of one population.  params = (V_rest, V_th, V_reset, tau, tau_ref, R)."""
    n = V.shape[0]
    ng = len(gs)
    gp = (ctypes.c_void_p * ng)(*[g.ctypes.data for g in gs])
    tau_g = np.asarray(tau_g, np.float32)
    E_g = np.asarray(E_g, np.float32)
    c_g = np.asarray(c_g, np.float32)
    lib().orc_lif_coba_step(n, int(ltc), ng, _p(V), _p(t_last), _p(spike),
                            ctypes.cast(gp, ctypes.POINTER(ctypes.c_void_p)),
                            _p(tau_g), _p(E_g), _p(c_g), np.float32(I_ext), _p(I_vec), *[np.float32(x) for x in params],
                            np.float32(t), np.float32(dt))


def lif_coba_step_f64(V, t_last, spike, gs, tau_g, E_g, I_ext, params, t, dt, ltc=True, I_vec=None):
    n = V.shape[0]
    ng = len(gs)
    gp = (ctypes.c_void_p * ng)(*[g.ctypes.data for g in gs])
    tau_g = np.asarray(tau_g, np.float64)
    E_g = np.asarray(E_g, np.float64)
    lib().orc_lif_coba_step_f64(n, int(ltc), ng, _p(V), _p(t_last), _p(spike),
                                ctypes.cast(gp, ctypes.POINTER(ctypes.c_void_p)),
                                _p(tau_g), _p(E_g), float(I_ext), _p(I_vec), *[float(x) for x in params],
                                float(t), float(dt))


def coba_net_run(n_exc, n_inh, indptr, indices, w_exc, w_inh, I_ext, V0, steps, params, tau_e=5.0, tau_i=10.0,
                 E_e=0.0, E_i=-80.0, dt=0.1, ltc=True, raster=True, delays=None, return_v=False):
    """Serial CPU run of a whole COBA network from its combined CSR; returns
    (total spikes, raster [steps, N] uint8 or None), plus the final
    potentials with `return_v`.  `delays`: steps of delay of the E2E, E2I,
    I2E, I2I projections (None: all 1)."""
    d4 = None if delays is None else np.ascontiguousarray(delays, np.int32)
    N = n_exc + n_inh
    indptr = np.ascontiguousarray(indptr, np.int32)
    indices = np.ascontiguousarray(indices, np.int32)
    V0 = np.ascontiguousarray(V0, np.float32)
    r = np.zeros((steps, N), np.uint8) if raster else None
    v_out = np.empty(N, np.float32) if return_v else None
    tot = lib().orc_coba_net_run(n_exc, n_inh, _p(indptr), _p(indices), np.float32(w_exc), np.float32(w_inh),
                                 np.float32(I_ext), _p(V0), steps, int(ltc), *[np.float32(x) for x in params],
                                 np.float32(tau_e), np.float32(tau_i), np.float32(E_e), np.float32(E_i),
                                 np.float32(dt), _p(r), _p(d4), _p(v_out))
    if return_v:
        return tot, r, v_out
    return tot, r


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def _u8(a):
    return np.ascontiguousarray(np.asarray(a) != 0, np.uint8)


def csr_on_pre_update(w, indices, indptr, spike, trace, w_min=-np.inf, w_max=np.inf):
    w, ind, ptr, sp, tr = _f32(w), _i32(indices), _i32(indptr), _u8(spike), _f32(trace)
    out = np.empty_like(w)
    rc = lib().orc_csr_on_pre_update(_p(w), _p(ind), _p(ptr), w.size, _p(sp), sp.size, _p(tr), tr.size,
                                     np.float32(w_min), np.float32(w_max), _p(out))
    assert rc == 0
    return out


def csc_on_post_update(w, pre_ids, colptr, w_ids, spike, trace, w_min=-np.inf, w_max=np.inf):
    w, pre, cp, wid, sp, tr = _f32(w), _i32(pre_ids), _i32(colptr), _i32(w_ids), _u8(spike), _f32(trace)
    out = np.empty_like(w)
    rc = lib().orc_csc_on_post_update(_p(w), w.size, _p(pre), _p(cp), _p(wid), _p(sp), sp.size, _p(tr), tr.size,
                                      np.float32(w_min), np.float32(w_max), _p(out))
    assert rc == 0
    return out


def dense_on_spike(rows, weight, spike, trace, w_min=-np.inf, w_max=np.inf):
    w, sp, tr = _f32(weight), _u8(spike), _f32(trace)
    out = np.empty_like(w)
    rc = lib().orc_dense_on_spike(int(rows), _p(w), _p(sp), w.shape[0], w.shape[1], _p(tr), np.float32(w_min),
                                  np.float32(w_max), _p(out))
    assert rc == 0
    return out
