// bpx_xla.hip -- XLA custom-call adapters over the C ABI.
//
// The reference binds its GPU operators as legacy XLA custom calls:
// `xla_client.register_custom_call_target(name, capsule, platform)`
// (brainpy/_src/dependency_check.py:75-79) with the GPU calling convention
// `void fn(stream, void** buffers, const char* opaque, size_t opaque_len)`,
// buffers being the operands in order followed by the outputs.  These
// adapters have that signature, take the reference's operand order, and read
// sizes and dtypes from a `bpx_xla_desc` opaque; they launch the same kernels
// as the plain entry points (no second kernel path).  The legacy convention
// has no error channel: a failure is recorded per thread and read with
// bpx_xla_last_status().
#include <string.h>

#include "bpx_common.h"

namespace {

thread_local int t_last_status = BPX_OK;

int set_status(int rc) {
  t_last_status = rc;
  return rc;
}

// the descriptor, copied out of the opaque (which need not be aligned)
bool read_desc(const char* opaque, size_t len, bpx_xla_desc* d) {
  if (!opaque || len != sizeof(bpx_xla_desc)) return false;
  memcpy(d, opaque, sizeof(bpx_xla_desc));
  return d->magic == BPX_XLA_MAGIC;
}

// 4-byte device scalar -> host (the jitconn operands weight/clen/seed are
// one-element device arrays in the reference's lowering)
bool read_scalar(hipStream_t s, const void* p, void* host, size_t bytes) {
  return p && hipMemcpyAsync(host, p, bytes, hipMemcpyDeviceToHost, s) == hipSuccess;
}

}  // namespace

extern "C" {

int bpx_xla_last_status(void) { return t_last_status; }

void bpx_xla_event_csrmv(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
  bpx_xla_desc d;
  if (!buffers || !read_desc(opaque, opaque_len, &d)) {
    set_status(BPX_ERR_ARG);
    return;
  }
  // (data, indices, indptr, events) -> out   (event/_csr_matvec.py:489-500)
  set_status(bpx_event_csrmv(stream, buffers[0], d.data_dtype, d.data_len, buffers[1], buffers[2], d.index_dtype,
                             d.nnz, buffers[3], d.vec_dtype, d.rule, d.m, d.n, d.transpose, buffers[4]));
}

void bpx_xla_csrmv(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
  bpx_xla_desc d;
  if (!buffers || !read_desc(opaque, opaque_len, &d)) {
    set_status(BPX_ERR_ARG);
    return;
  }
  // (data, indices, indptr, vector) -> out   (sparse/_csr_mv.py:328-361)
  set_status(bpx_csrmv(stream, buffers[0], d.data_dtype, d.data_len, buffers[1], buffers[2], d.index_dtype, d.nnz,
                       buffers[3], d.vec_dtype, d.m, d.n, d.transpose, buffers[4]));
}

void bpx_xla_event_info(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
  bpx_xla_desc d;
  if (!buffers || !read_desc(opaque, opaque_len, &d)) {
    set_status(BPX_ERR_ARG);
    return;
  }
  // (events,) -> (ids, num)   (event/_info_collection.py:115-121); the
  // workspace comes as a third output buffer of d.ws_bytes bytes
  set_status(bpx_event_info(stream, buffers[0], d.vec_dtype, d.batch, d.n, (int32_t*)buffers[1],
                            (int32_t*)buffers[2], d.ws_bytes > 0 ? buffers[3] : nullptr, (size_t)d.ws_bytes));
}

void bpx_xla_jitconn_mv(void* stream, void** buffers, const char* opaque, size_t opaque_len) {
  bpx_xla_desc d;
  if (!buffers || !read_desc(opaque, opaque_len, &d)) {
    set_status(BPX_ERR_ARG);
    return;
  }
  // homo: (events, weight, clen, seed) -> out   (jitconn/_event_matvec.py:219-229)
  // uniform / normal: (events, w_low|w_mu, w_high|w_sigma, clen, seed) -> out
  // (jitconn/_event_matvec.py:384-414, :578-608); an optional workspace
  // output of d.ws_bytes bytes follows out
  hipStream_t s = (hipStream_t)stream;
  const int nw = d.dist == BPX_DIST_HOMO ? 1 : 2;
  const bool f64 = d.data_dtype == BPX_F64;
  float wf[2] = {0.f, 0.f};
  double wd[2] = {0.0, 0.0};
  int32_t clen = 0;
  uint32_t seed = 0;
  bool ok = true;
  for (int k = 0; k < nw && ok; ++k)
    ok = f64 ? read_scalar(s, buffers[1 + k], &wd[k], sizeof(double)) : read_scalar(s, buffers[1 + k], &wf[k], sizeof(float));
  ok = ok && read_scalar(s, buffers[1 + nw], &clen, sizeof clen) && read_scalar(s, buffers[2 + nw], &seed, sizeof seed);
  if (!ok || hipStreamSynchronize(s) != hipSuccess) {
    set_status(BPX_ERR_LAUNCH);
    return;
  }
  const float w0 = f64 ? (float)wd[0] : wf[0], w1 = f64 ? (float)wd[1] : wf[1];
  void* out = buffers[3 + nw];
  void* ws = d.ws_bytes > 0 ? buffers[4 + nw] : nullptr;
  set_status(bpx_jitconn_mv(stream, buffers[0], d.vec_dtype, d.event, d.dist, w0, w1, clen, seed, d.m, d.n,
                            d.outdim_parallel, d.layout, (float*)out, ws, (size_t)d.ws_bytes));
}

}  // extern "C"
