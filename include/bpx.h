/*
 * bpx.h -- C ABI of libbpx.so, the gfx950 (MI355X) implementation of
 * BrainPy's event-driven synaptic hot path.
 *
 * Conventions
 *  - Every entry point takes a HIP stream (`void*` = hipStream_t, NULL = the
 *    default stream) and device pointers.  Nothing is allocated and nothing
 *    synchronises the host; inputs are read-only.
 *  - Outputs are fully written by the call (the library zero-fills before
 *    accumulating, on the same stream): callers need not pre-zero them.  The
 *    reference's kernels accumulate with `+=` into outputs that their
 *    launcher must zero (event/_csr_matvec_taichi.py:45,172).
 *  - Return value: BPX_OK (0) or a negative BPX_ERR_* code; bpx_status_string
 *    gives its text.
 *  - Stateless and reentrant; concurrent calls on different streams are safe.
 *
 * Reference paths are relative to the chaoming0625/BrainPy 2.5.0 tree.
 */
#ifndef BPX_H_
#define BPX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BPX_OK 0
#define BPX_ERR_DTYPE (-1)      /* unsupported dtype or dtype combination */
#define BPX_ERR_SHAPE (-2)      /* inconsistent sizes */
#define BPX_ERR_WORKSPACE (-3)  /* workspace missing or too small */
#define BPX_ERR_LAUNCH (-4)     /* a kernel launch failed */
#define BPX_ERR_ARG (-5)        /* invalid scalar argument, or a null array the call would read or write
                                   (an array of length 0 may be null) */

/* dtype codes */
#define BPX_BOOL 0   /* 1 byte, truthy if non-zero */
#define BPX_F32 1
#define BPX_F64 2
#define BPX_I32 3
#define BPX_I64 4
#define BPX_BF16 5   /* weights only; accumulated in f32 */
#define BPX_F16 6    /* weights only; accumulated in f32 */

/* float-event truthiness */
#define BPX_RULE_POSITIVE 0  /* event > 0  (Numba kernels, event/_csr_matvec.py:349) */
#define BPX_RULE_NONZERO 1   /* event != 0 (Taichi kernels, event/_csr_matvec_taichi.py:70) */

/* jitconn weight distributions and layouts */
#define BPX_DIST_HOMO 0
#define BPX_DIST_UNIFORM 1
#define BPX_DIST_NORMAL 2
#define BPX_LAYOUT_LANE32 0  /* Taichi GPU kernels: 32 streams per row/column */
#define BPX_LAYOUT_ROW 1     /* Taichi CPU kernels: one stream per row/column */

const char* bpx_status_string(int status);
const char* bpx_version(void);

/*
 * Spike compaction.  Replaces the `nonzero_*` brainpylib GPU target and the
 * Numba `_event_info` CPU target of bm.event.info
 * (event/_info_collection.py:24-43 wrapper, :78-121 GPU translation,
 * :143-154 CPU loop, batched :54-66).
 * events[batch, n] -> ids[batch, n] (ascending indices of truthy events,
 * then -1), num[batch].  Float events are truthy if != 0.
 */
/* Workspace: bpx_event_info_workspace bytes (tile counts and offsets, then
 * 512 B of event bits per 4096-event tile for bool rows starting 16-byte
 * aligned, which are then read once instead of twice); the minimum accepted
 * is the first part, 8 * batch * ceil(n / 4096) bytes. */
size_t bpx_event_info_workspace(int64_t batch, int64_t n);
int bpx_event_info(void* stream, const void* events, int ev_dtype, int64_t batch, int64_t n,
                   int32_t* ids, int32_t* num, void* workspace, size_t workspace_bytes);

/*
 * bpx_event_info without a workspace and without any allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace lives in a static device array of
 * BPX_EVENT_INFO_SCRATCH_BYTES per GPU, held until the process ends; it
 * holds the full workspace when that fits, else the minimum, and sizes whose
 * minimum exceeds it return BPX_ERR_WORKSPACE (e.g. a batch of 65535 rows of
 * 100,000 events).  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams) race.  Results equal bpx_event_info's.
 */
#define BPX_EVENT_INFO_SCRATCH_BYTES (4u << 20)
int bpx_event_info_noalloc(void* stream, const void* events, int ev_dtype, int64_t batch, int64_t n,
                           int32_t* ids, int32_t* num);

/*
 * Event-driven CSR matvec.  Replaces the brainpylib targets
 * `event_csrmv{_homo,_heter}{_float,_double}{_int,_long}{_bool,}{_transpose,}`
 * (event/_csr_matvec.py:453-500) and the Taichi kernels of csrmv_taichi
 * (event/_csr_matvec_taichi.py:34-312).
 *   transpose=1: out[n] = sum over rows i with a truthy event of row i of A
 *   transpose=0: out[m] = A restricted to columns with a truthy event, row sums
 * data: data_len == 1 (homogeneous) or == nnz; dtype f32/f64/bf16/f16.
 * indices/indptr: int32 or int64 (index_dtype, both the same).
 * events: bool/f32/f64 with ev_rule for float events.
 * out: f64 if data is f64, else f32.
 */
int bpx_event_csrmv(void* stream, const void* data, int data_dtype, int64_t data_len,
                    const void* indices, const void* indptr, int index_dtype, int64_t nnz,
                    const void* events, int ev_dtype, int ev_rule,
                    int64_t m, int64_t n, int transpose, void* out);

/*
 * CSR matvec with a dense vector.  Replaces `cuda_csr_matvec` (cuSPARSE,
 * sparse/_csr_mv.py:197-210) and `csrmv_{homo,heter}_{scalar,vector}`
 * (sparse/_csr_mv.py:257-478), Numba CPU (sparse/_csr_mv.py:126-165) and the
 * Taichi kernels (sparse/_csr_mv_taichi.py:28-155).
 * vector dtype must equal the accumulation type (f32, or f64 for f64 data).
 */
int bpx_csrmv(void* stream, const void* data, int data_dtype, int64_t data_len,
              const void* indices, const void* indptr, int index_dtype, int64_t nnz,
              const void* vector, int vec_dtype, int64_t m, int64_t n, int transpose, void* out);

/*
 * Fixed-order event scatter: the transpose event csrmv computed as a gather
 * over a CSC copy whose row indices ascend within each column, so every
 * output is summed in ascending pre-synaptic order -- bit-identical to the
 * serial CPU scatter of event/_csr_matvec.py:321-340 and independent of
 * scheduling.  csc_data follows the CSC order (data_len 1 = homogeneous).
 * accumulate=1 adds the column sums into out (out[c] = out[c] + sum_c, the
 * order of `g += csrmv(...)`); accumulate=0 overwrites out.
 */
int bpx_event_csrmv_fixed_order(void* stream, const void* csc_data, int data_dtype, int64_t data_len,
                                const int32_t* csc_rows, const int32_t* csc_colptr, int64_t nnz,
                                const void* events, int ev_dtype, int ev_rule,
                                int64_t m, int64_t n, int accumulate, void* out);

/*
 * Accumulating transpose event csrmv: out[n] += events @ A (no zero-fill).
 * No reference counterpart: the reference adds a fresh csrmv result into
 * the conductance (`g += csrmv(...)`, dyn/synapses/abstract_models.py:
 * 133-134); the network step uses this form to add straight into g.
 * Same kernels and arguments as bpx_event_csrmv with transpose=1.
 */
int bpx_event_csrmv_add(void* stream, const void* data, int data_dtype, int64_t data_len,
                        const void* indices, const void* indptr, int index_dtype, int64_t nnz,
                        const void* events, int ev_dtype, int ev_rule, int64_t m, int64_t n, void* out);

/*
 * Tiled CSR plan for the event-driven scatter (transpose event csrmv).
 * A one-time re-layout of a CSR matrix (like a sparse-library SpMV
 * preprocess) after which bpx_event_csrmv_tiled computes exactly
 * bpx_event_csrmv(transpose=1) with LDS accumulation instead of one global
 * atomic per synaptic event: the output is cut into one tile per CU of an
 * XCD (tile_len floats, 16-bit tile-local column indices), each XCD holds
 * one copy of the output in the LDS of its CUs, and rows carry per-tile
 * segment offsets; segments are padded to multiples of 4 entries (index
 * 0xFFFF) so indices and weights stream as 8- and 16-byte vectors.  No reference counterpart (the reference re-reads plain
 * CSR each call, event/_csr_matvec.py:453-500); it serves the same operator.
 * Usage: bpx_tiled_csr_layout (sizes, fills the plan's shape fields) ->
 * caller allocates `bytes` of device memory -> bpx_tiled_csr_build ->
 * bpx_event_csrmv_tiled per call.  Requires nnz < 2^31 and a tile that fits
 * 32 KiB of LDS next to the kernel's row lists (tile_len <= 4096): one tile
 * per CU of an XCD up to n = 4096 * CUs-per-XCD = 131072 on MI355X, then up
 * to 64 tiles (n <= 262144, 16 output copies per call), then up to 1024
 * tiles, a multiple of 8 (n <= 4194304; 8, 4, 2 or 1 copies for up to 128,
 * 256, 512 or 1024 tiles), for which the layout refuses (BPX_ERR_SHAPE)
 * matrices averaging fewer than tiles / 4 synapses per row: their per-row
 * offsets would outweigh the synapses.  Accumulation inside a tile
 * is in 64-bit fixed point (LDS integer atomics; gfx950's LDS float atomics
 * are ~20x slower): fx_shift = min(61 - ceil(log2(max|w| * m)),
 * 50 - ceil(log2 max|w|)) keeps every sum below 2^62 and every term below
 * 2^50, weights with |w| >= 2^(23 - fx_shift) convert exactly (round to
 * nearest even below that), and each copy's tile sum is rounded to f32 once.  The build synchronises the stream
 * once (to read max|w|) and rejects non-finite weights (BPX_ERR_ARG).  Weights: f32, bf16 (rounded to nearest even
 * at build, accumulated in f32) or homogeneous (data_len 1).
 */
#define BPX_TILED_HOMO (-1)
#define BPX_TILED_MAX_TILE 32768
typedef struct {
  int64_t m, n, nnz;
  int32_t tiles;        /* P: output tiles per copy (CUs per XCD, more for n > 131072) */
  int32_t tile_len;     /* T: floats per tile */
  int32_t weight_dtype; /* BPX_F32, BPX_BF16 or BPX_TILED_HOMO */
  int32_t fx_shift;     /* LDS accumulators hold sum(w) * 2^fx_shift as int64 */
  int32_t coarse;       /* 1 when 2^-fx_shift > 2e-6 * min non-zero |w| (weights spanning ~2^40 / m or more):
                           calls then add each weight below 5e5 grid units to its output with an FP32
                           atomic instead of the fixed-point sum (set by build) */
  int32_t reserved;
  int64_t capacity;     /* element slots incl. padding (segments padded to 4) */
  int64_t rowtile_offset, aux_offset, lidx_offset, weights_offset, homo_offset; /* byte offsets in mem */
  void* mem;            /* device memory holding the plan (set by build) */
} bpx_tiled_csr;

int bpx_tiled_csr_layout(int64_t m, int64_t n, int64_t nnz, int weight_dtype, bpx_tiled_csr* plan, size_t* bytes);
int bpx_tiled_csr_build(void* stream, const void* data, int data_dtype, int64_t data_len,
                        const void* indices, const void* indptr, int index_dtype,
                        bpx_tiled_csr* plan, void* mem, size_t mem_bytes);
/* out[n] = events @ A (accumulate=0) or out += events @ A (accumulate=1) */
int bpx_event_csrmv_tiled(void* stream, const bpx_tiled_csr* plan, const void* events, int ev_dtype,
                          int ev_rule, int accumulate, float* out);
/* The same for a batch of event vectors, events [batch][m] -> out
 * [batch][n], in one launch (batch <= 65535); the batching rule of the
 * reference's primitive is a sequential scan over the batch
 * (op_register/utils.py:16-37). */
int bpx_event_csrmv_tiled_batch(void* stream, const bpx_tiled_csr* plan, const void* events, int ev_dtype,
                                int ev_rule, int64_t batch, int accumulate, float* out);
/* Plans of more than 64 tiles: bytes of workspace with which a call lists
 * each vector's active rows once (one pass over the events) and the tile
 * blocks walk that list, instead of every tile block scanning every event
 * of its rows; 0 for smaller plans, which always scan.  Calls that share a
 * workspace must not overlap. */
size_t bpx_tiled_csr_call_workspace(const bpx_tiled_csr* plan, int64_t batch);
/* bpx_event_csrmv_tiled_batch with that workspace (a smaller one, or none,
 * scans as bpx_event_csrmv_tiled_batch does); same results either way. */
int bpx_event_csrmv_tiled_ws(void* stream, const bpx_tiled_csr* plan, const void* events, int ev_dtype, int ev_rule,
                             int64_t batch, int accumulate, float* out, void* workspace, size_t workspace_bytes);
/*
 * The same plan with a dense f32 vector: out (+)= x @ A, the transpose
 * `csrmv` of sparse/_csr_mv.py:29-123 (cuSPARSE transpose SpMV) for a
 * static matrix.  Rows with x != 0 are streamed; each product w * x[row] is
 * rounded to f32 and summed exactly in fixed point (shift lowered by
 * ceil(log2 max|x|) over the finite x, computed on the device: workspace >= 8
 * bytes, or 256 + bpx_tiled_csr_call_workspace(plan, 1) bytes to list the
 * rows with x != 0 once; calls that share a workspace must not overlap).
 * A product below 5e5 units of that grid (where rounding to the grid would
 * cost more than 1e-6 of it), and one that is inf or NaN, is added to its
 * output with an FP32 atomic instead, so tiny terms keep f32 accuracy and
 * non-finite ones reach their outputs as IEEE addition would (an infinity,
 * or NaN for a NaN term or infinities of both signs).  bpx_event_csrmv_tiled
 * does the same for weights below the grid's floor (the plan's coarse flag).
 */
int bpx_csrmv_tiled(void* stream, const bpx_tiled_csr* plan, const void* x, int x_dtype, int accumulate, float* out,
                    void* workspace, size_t workspace_bytes);

/*
 * The same plan read row-wise: out[m] = A @ vec, the transpose=0 gather of
 * bm.sparse.csrmv (dense f32 vec; sparse/_csr_mv.py:29-123), and with
 * event=1 the event gather of bm.event.csrmv (bool events, or f32 events
 * with ev_rule; the weights of rows' truthy columns summed;
 * event/_csr_matvec.py:38-110) -- for events bpx_event_csrmv's bitmask
 * kernel reads whole rows and is the faster of the two.
 * Each block stages one column tile of vec in LDS and sums the tile's
 * segment of each of its rows (w * vec[c] in element order, then a fixed
 * 8-lane tree); the per-tile row partials (workspace of
 * bpx_tiled_csr_gather_workspace bytes = 4 * tiles * m) are added in tile
 * order: deterministic, no atomics, the matrix read once at 2 + w bytes per
 * synapse.  No reference counterpart as a plan (the reference re-reads CSR).
 */
size_t bpx_tiled_csr_gather_workspace(const bpx_tiled_csr* plan);
int bpx_tiled_csr_gather(void* stream, const bpx_tiled_csr* plan, const void* vec, int vec_dtype, int event,
                         int ev_rule, float* out, void* workspace, size_t workspace_bytes);

/*
 * Batched event operators: `batch` event vectors in one launch, in place of
 * the reference's batching rule for these primitives, a sequential scan over
 * the batch (op_register/utils.py:16-37; event/_csr_matvec.py:549-556,
 * jitconn/_event_matvec.py:294-302).  The vectors are packed into per-index
 * bit masks (workspace of bpx_batch_workspace(len) bytes, len = the event
 * length; 32 vectors per pass), then each row's synapses (or each generated
 * connection) are read (generated) once for all vectors.
 *
 * bpx_event_csrmv_batch: events [batch][m] (transpose) or [batch][n], bool
 *   or f32 (ev_rule), data f32 or bf16 (homogeneous when data_len == 1),
 *   out f32 [batch][n] or [batch][m].  transpose=1: one FP32 atomic per
 *   (vector, synapse) of rows active in any vector; transpose=0: gather with
 *   one shuffle reduction per vector.
 * bpx_jitconn_mv_batch: bool events [batch][num_col] -> out [batch][num_row];
 *   row-parallel results are bit-identical to bpx_jitconn_mv on each vector
 *   (finite weights); the scatter form adds with FP32 atomics.
 */
size_t bpx_batch_workspace(int64_t len);
int bpx_event_csrmv_batch(void* stream, const void* data, int data_dtype, int64_t data_len,
                          const void* indices, const void* indptr, int index_dtype, int64_t nnz,
                          const void* events, int ev_dtype, int ev_rule, int64_t batch, int64_t m, int64_t n,
                          int transpose, float* out, void* workspace, size_t workspace_bytes);
int bpx_jitconn_mv_batch(void* stream, const void* events, int ev_dtype, int dist, float w0, float w1,
                         int32_t clen, uint32_t seed, int64_t batch, int64_t num_row, int64_t num_col,
                         int outdim_parallel, int layout, float* out, void* workspace, size_t workspace_bytes);

/*
 * Just-in-time connectivity matvec y = M @ v, M never stored.  Replaces the
 * Taichi kernels of jitconn/_event_matvec_taichi.py:38-1106 (event) and
 * jitconn/_matvec_taichi.py:30-746 (dense vector), and stands in for the
 * brainpylib targets `gpu_jit_{event_,}csrmv_prob_{homo,uniform,normal}_v2_*`
 * (jitconn/_event_matvec.py:198-229, jitconn/_matvec.py:341-378).
 * num_row x num_col is the matrix shape after the transpose swap of
 * jitconn/_matvec_taichi.py:197-207 (num_row = len(out), num_col = len(vec)).
 *   event=1: vec is events (bool or f32); event=0: vec is a dense f32 vector.
 *   dist: BPX_DIST_*; (w0, w1) = weight | (low, high) | (mu, sigma).
 *   clen = ceil(1/p)*2-1; seed as u32.
 * out: f32 [num_row].
 * workspace (optional, bpx_jitconn_workspace bytes = max(num_col / 8,
 * 256 + 4 num_col)): with bool events and outdim_parallel the events are
 * packed into it as a bitmask that every block keeps in LDS; with bool events
 * and outdim_parallel=0 (lane32) it holds the list of active columns, and
 * each block owns a tile of output rows, regenerating every active column's
 * stream up to its tile and summing in LDS (no global atomics; homogeneous
 * weights exact).  With bpx_jitconn_workspace_ex bytes that form runs in two
 * passes instead: every active stream is generated once and its connections
 * listed by row tile, then each tile sums its lists in LDS (homogeneous
 * weights as exact counts; uniform and normal weights in 64-bit fixed point
 * when all draws are finite).  NULL falls back to reading the event bytes
 * (outdim) or to one FP32 atomic per connection (scatter).
 */
size_t bpx_jitconn_workspace(int64_t num_row, int64_t num_col);
/* No counterpart in the reference: the workspace bytes with which
 * bpx_jitconn_mv runs outdim_parallel=0 with bool events in two passes
 * (each active stream generated once, its connections listed by row tile,
 * then each tile summed in LDS); the lists are sized for about 12.5 %
 * active columns, and tiles whose lists overflow regenerate their streams
 * within the call.  Returns bpx_jitconn_workspace's size when the two-pass
 * form does not apply (layouts other than lane32 take neither). */
size_t bpx_jitconn_workspace_ex(int64_t num_row, int64_t num_col, int32_t clen, int dist);
int bpx_jitconn_mv(void* stream, const void* vec, int vec_dtype, int event,
                   int dist, float w0, float w1, int32_t clen, uint32_t seed,
                   int64_t num_row, int64_t num_col, int outdim_parallel, int layout, float* out,
                   void* workspace, size_t workspace_bytes);

/*
 * Fused LIF + Expon + COBA exponential-Euler step of one population (f32).
 * Replaces the per-neuron integrator of LifRefLTC / LifRef
 * (dyn/neurons/lif.py:461-500, 550-556), Expon (dyn/synapses/
 * abstract_models.py:121-134), COBA (dyn/outs/outputs.py:53-54) and
 * ExponentialEuler (integrators/ode/exponential.py:357-362).
 * g: host array of n_g (<= 4) device pointers to conductances [n];
 * tau_g, E_g, c_g: host arrays (c_g = dt*exprel(-dt/tau_g) in f32);
 * I_vec: optional per-neuron external input (NULL -> I_ext).
 */
int bpx_lif_coba_step(void* stream, int64_t n, int ltc, int n_g,
                      float* V, float* t_last, uint8_t* spike, float* const* g,
                      const float* tau_g, const float* E_g, const float* c_g,
                      float I_ext, const float* I_vec,
                      float V_rest, float V_th, float V_reset, float tau, float tau_ref, float R,
                      float t, float dt);

/*
 * The same step with the time read on the device: t = f32(*d_step + step) *
 * dt (the reference's t = i * dt of runners.py:619-643 with the step index
 * i held in device memory), so a captured HIP graph of K steps (step =
 * 0..K-1, then bpx_net_advance(d_step, K)) continues in time on every
 * replay, as the reference's compiled scan carries i on the device.
 */
int bpx_lif_coba_step_dev(void* stream, int64_t n, int ltc, int n_g,
                          float* V, float* t_last, uint8_t* spike, float* const* g,
                          const float* tau_g, const float* E_g, const float* c_g,
                          float I_ext, const float* I_vec,
                          float V_rest, float V_th, float V_reset, float tau, float tau_ref, float R,
                          const int64_t* d_step, int64_t step, float dt);

/*
 * Host helper: the Expon exp_euler factor c = dt * exprel(dt * (-(1/tau)))
 * in f32 with the library's deterministic exp, i.e. the `dt*phi` of
 * integrators/ode/exponential.py:357-362 for dg/dt = -g/tau
 * (dyn/synapses/abstract_models.py:118-119).  Pass it as c_g above.
 */
float bpx_expon_coef(float tau, float dt);

/*
 * Batched E/I COBA network step.  Advances any number of independent
 * networks (LifRefLTC/LifRef neurons, Expon + COBA synapses, homogeneous
 * weights: the EICOBA_PostAlign model of
 * examples/dynamics_simulation/COBA.py:58-109, with optional projection
 * delays) by one step in one launch.
 * No single reference counterpart: it fuses, per step, the four event
 * projections (event csrmv, event/_csr_matvec.py:38-110, or jitconn
 * event_mv_prob_homo with outdim_parallel=False,
 * jitconn/_event_matvec_taichi.py:92-123) and the population updates
 * (dyn/neurons/lif.py:461-500) in the reference order.
 * State pools are indexed by off + local neuron id (E first, then I);
 * counts is [n_buf][2][total] uint32, zero at start (and zero in a network's
 * range when it joins), n_buf >= 2 and above every delay.  A projection with
 * delay d steps reads at step i the spikes of step i - d, as the rotation
 * update of VarDelay does (delay.py:305-349); delays 0 and 1 both read the
 * previous step's spikes.  The step index (count buffer; network time
 * t = (step - t0_step) * dt) is `step`, or *d_step + step when d_step points
 * to a device int64 base -- the form for HIP graphs: capture K launches with
 * step = 0..K-1 followed by bpx_net_advance(d_step, K), and every replay
 * continues in time.
 * tiles: device array of (network, first local neuron) int32 pairs, one per
 * block of tile_size neurons (64, 128, 256, 512 or 1024: the block size of
 * the launch and of its LDS spike list; the host picks it per batch by
 * timing the candidates).  nets: device array of bpx_net_desc prepared with
 * bpx_net_prepare (host).  jitconn projections use seed[k], clen[k] for
 * k = E2E, E2I, I2E, I2I with matrix shape (post, pre); CSR networks give one
 * row per neuron listing local post ids (E targets of E neurons feed the
 * excitatory conductance, targets of I neurons the inhibitory one).
 * spikes/events/spike_count are optional outputs; null state arrays (V,
 * t_last, g_exc, g_inh, counts) with work to do return BPX_ERR_ARG.
 */
#define BPX_NET_JIT 0
#define BPX_NET_CSR 1
typedef struct {
  int64_t off;               /* first pool index of the network */
  int32_t n_exc, n_inh;
  uint32_t seed[4];          /* jitconn seeds: E2E, E2I, I2E, I2I */
  int32_t clen[4];           /* jitconn ceil(1/p)*2-1 per projection */
  uint32_t div[4][4];        /* filled by bpx_net_prepare */
  float w_exc, w_inh;        /* homogeneous weights */
  float I_ext;               /* constant input */
  int32_t t0_step;           /* step index at which the network started (t = (step - t0_step) * dt) */
  int32_t delay[4];          /* delay in steps of E2E, E2I, I2E, I2I (0 or 1: the previous step's spikes) */
  const int32_t* indptr;     /* CSR networks: [N+1] */
  const int32_t* indices;    /* CSR networks: local post ids */
} bpx_net_desc;

typedef struct {
  float V_rest, V_th, V_reset, tau, tau_ref, R;
  float tau_e, tau_i, E_e, E_i;
  float c_e, c_i;            /* bpx_expon_coef(tau_e, dt), bpx_expon_coef(tau_i, dt) */
  float dt;
  int32_t ltc;               /* 1 LifRefLTC, 0 LifRef */
} bpx_net_params;

void bpx_net_prepare(bpx_net_desc* desc);
int bpx_net_step(void* stream, const bpx_net_desc* nets, const int32_t* tiles, int n_tiles, int tile_size,
                 int conn_kind, int64_t total, float* V, float* t_last, float* g_exc, float* g_inh, uint32_t* counts, int n_buf,
                 uint8_t* spikes, unsigned long long* events, unsigned int* spike_count,
                 const bpx_net_params* params, int64_t step, const int64_t* d_step);
int bpx_net_advance(void* stream, int64_t* d_step, int64_t delta);
/* Multi-step form of bpx_net_step (same reference counterpart): advances
 * every network of `nets` (one block each, n_nets of them) n_steps steps in
 * one launch, from step `step` (or *d_step + step), with the same results as
 * n_steps calls of bpx_net_step: spike raster, state, counts, events and
 * spike counts.  For small networks, where the per-step launch dominates:
 * every network has at most max_neurons <= bpx_net_run_max_neurons(n_buf)
 * neurons (5456 at n_buf = 2), and no neuron receives more than
 * max_in_degree <= 65535 synapses from one population (counts are kept as
 * 16-bit halves in LDS); larger values return BPX_ERR_SHAPE.  The
 * descriptors live on the device, so the sizes of the networks are checked
 * there: a network with more than max_neurons neurons is left unchanged and
 * BPX_ERR_SHAPE is written to *status (nullable device int32, left as it is
 * otherwise; the caller zeroes it).  max_in_degree is a precondition: the
 * caller's bound on the synapses one neuron receives from one population
 * (bpx does not recount it).  raster (nullable) receives [n_steps][total]
 * spike flags, spikes (nullable) the flags of the last step. */
int64_t bpx_net_run_max_neurons(int n_buf);
int bpx_net_run(void* stream, const bpx_net_desc* nets, int n_nets, int max_neurons, int max_in_degree,
                int conn_kind, int64_t total, float* V, float* t_last, float* g_exc, float* g_inh, uint32_t* counts,
                int n_buf, uint8_t* raster, uint8_t* spikes, unsigned long long* events, unsigned int* spike_count,
                const bpx_net_params* params, int64_t step, const int64_t* d_step, int64_t n_steps,
                int32_t* status);

/* ---------------------------------------------------------------------------
 * Event-driven STDP weight updates.
 *
 * bpx_csr_on_pre_update replaces csr_on_pre_update (dnn/linear.py:650-676):
 *   out = w; for every pre neuron i with spike[i] != 0 and k in row i:
 *   out[k] = min(max(out[k] + trace[indices[k]], w_min), w_max).
 * bpx_csc_on_post_update replaces csc_on_post_update (dnn/linear.py:679-700)
 * on the CSC view of the same synapses (csr2csc, connect/base.py:645):
 *   for every post neuron j with spike[j] != 0 and k in column j:
 *   out[w_ids[k]] = min(max(out[w_ids[k]] + trace[pre_ids[k]], w_min), w_max).
 * bpx_dense_on_pre / bpx_dense_on_post replace dense_on_pre / dense_on_post
 * (dnn/linear.py:238-275) on a row-major [num_pre, num_post] matrix:
 *   rows of spiking pre neurons += trace[post], or columns of spiking post
 *   neurons += trace[pre], clipped to [w_min, w_max].
 * w, trace and out share w_dtype (F32 or F64); spike is BOOL, F32 or F64
 * (non-zero = spike); indices, indptr (colptr), pre_ids and w_ids share
 * index_dtype (I32 or I64).  out may equal w (in place); otherwise w is
 * copied into out first.  Each weight is written at most once per call, so
 * results do not depend on scheduling.
 * ------------------------------------------------------------------------- */
int bpx_csr_on_pre_update(void* stream, const void* w, int w_dtype, const void* indices, const void* indptr,
                          int index_dtype, int64_t nnz, const void* spike, int spike_dtype, int64_t num_pre,
                          const void* trace, int64_t num_post, double w_min, double w_max, void* out);
int bpx_csc_on_post_update(void* stream, const void* w, int w_dtype, const void* pre_ids, const void* colptr,
                           const void* w_ids, int index_dtype, int64_t nnz, const void* spike, int spike_dtype,
                           int64_t num_post, const void* trace, int64_t num_pre, double w_min, double w_max,
                           void* out);
int bpx_dense_on_pre(void* stream, const void* w, int w_dtype, const void* spike, int spike_dtype, int64_t num_pre,
                     int64_t num_post, const void* trace, double w_min, double w_max, void* out);
int bpx_dense_on_post(void* stream, const void* w, int w_dtype, const void* spike, int spike_dtype, int64_t num_pre,
                      int64_t num_post, const void* trace, double w_min, double w_max, void* out);

/* -------------------------------------------------------------------------
 * COO sparse products.
 *
 * bpx_coomv replaces `coomv` (sparse/_coo_mv.py:17-89; cuSPARSE lowering
 * :122-152, default lowering _coomv_impl :99-105): out (m, or n when
 * transposed) = A @ vector (A^T @ vector), A given by nse (row, col, data)
 * triples in any order (the sorted flags of the reference only choose a
 * cuSPARSE call; the result is the same).  data_len 1 broadcasts one value.
 * Values and vector share one dtype: f32, f64 or i32 (wrapping).  Index
 * dtype i32 or i64.  Output keys outside the shape are dropped and gather
 * indices clamped, as JAX's scatter/gather do.
 *
 * bpx_coo_matmat replaces `seg_matmul` (sparse/_jax_prim.py:97-166) with a
 * two-dimensional dense operand: sparse_left != 0: out (m, k) = S (m, n) @
 * dense (n, k); sparse_left == 0: out (k, n) = dense (k, m) @ S (m, n).
 * Row-major dense and out.
 *
 * bpx_coo_sddmm: out[e] = sum_c a[row[e]*a_stride + c*a_col_stride] *
 * b[col[e]*b_stride + c*b_col_stride] for c < ncols -- the value cotangent
 * of both products (sparse/_coo_mv.py:174-186 for coomv; the JAX gradient
 * of sparse/_jax_prim.py:55-94 for seg_matmul).  Pairs with a row or col
 * outside [0, a_keys) / [0, b_keys) give 0.
 * ------------------------------------------------------------------------- */
int bpx_coomv(void* stream, const void* data, int data_dtype, int64_t data_len, const void* row, const void* col,
              int index_dtype, int64_t nse, const void* vector, int vec_dtype, int64_t m, int64_t n, int transpose,
              void* out);
int bpx_coo_matmat(void* stream, const void* data, int data_dtype, int64_t data_len, const void* row,
                   const void* col, int index_dtype, int64_t nse, const void* dense, int dense_dtype, int64_t m,
                   int64_t n, int64_t k, int sparse_left, void* out);
int bpx_coo_sddmm(void* stream, const void* row, const void* col, int index_dtype, int64_t nse, const void* a,
                  int64_t a_keys, int64_t a_stride, int64_t a_col_stride, const void* b, int64_t b_keys,
                  int64_t b_stride, int64_t b_col_stride, int64_t ncols, int dtype, void* out);

/* -------------------------------------------------------------------------
 * Block-sparse (BSR) products.
 *
 * bpx_bsrmv replaces `cusparse_bcsr_matvec` (sparse/_bsr_mv.py:60-91; GPU
 * lowering :118-146 = cuSPARSE bsrmv): out = A @ vector, A of mb x nb blocks
 * of blocksize x blocksize given by data[nnzb][bs][bs] (row-major blocks, the
 * layout the reference's transpose rule reads, :175-199), block columns
 * indices[nnzb] and block-row pointers indptr[mb + 1]; out has mb * bs
 * entries.  transpose != 0: out = A^T @ vector, nb * bs entries (the
 * reference's transpose rule calls it; its own lowerings do not take it).
 * dtype f32 or f64 for data, vector and out; indices i32 or i64.  Blocks
 * whose column lies outside [0, nb) are skipped.
 *
 * bpx_bsr_sddmm: the value cotangent of bpx_bsrmv -- out[j][p][q] =
 * rowv[i*bs + p] * colv[indices[j]*bs + q] for block j of block row i
 * (pass rowv = cotangent, colv = vector; transposed: rowv = vector,
 * colv = cotangent).
 *
 * bpx_bcsrmm replaces `bcsrmm` (sparse/_bsr_mm.py:210-259; Numba loop
 * :292-307, GPU lowering :344-375): out (m x n, row-major) = A (m x k,
 * row-major) @ B, B (k x n) stored by 32-column blocks: b_ptr[n/32 + 1] over
 * column blocks, b_indices[nblk][2] = (k block, data block),
 * b_data[nblk][32][32] with B[ki*32 + b][ni*32 + a] = b_data[bi][a][b].
 * k and n multiples of 32 (the reference's block size is fixed at 32 x 32).
 * f32 runs on MFMA (A and b_data 16-byte aligned), f64 on FMA.  Blocks with
 * a k block or data block outside range are skipped.
 * ------------------------------------------------------------------------- */
int bpx_bsrmv(void* stream, const void* data, int dtype, const void* indices, const void* indptr, int index_dtype,
              int64_t nnzb, int64_t blocksize, const void* vector, int64_t mb, int64_t nb, int transpose, void* out);
int bpx_bsr_sddmm(void* stream, const void* indices, const void* indptr, int index_dtype, int64_t nnzb,
                  int64_t blocksize, const void* rowv, const void* colv, int dtype, int64_t mb, int64_t nb,
                  void* out);
int bpx_bcsrmm(void* stream, const void* a, int dtype, const void* b_data, const void* b_indices, const void* b_ptr,
               int index_dtype, int64_t nblk, int64_t m, int64_t k, int64_t n, void* out);

/* -------------------------------------------------------------------------
 * FixedProb connectivity on the device.  Replaces FixedProb.build_csr
 * (connect/random_conn.py:85-128, pre_ratio = 1): each of the m rows gets
 * k = int(n * prob) distinct targets drawn uniformly from [0, n) (the
 * reference's rng.choice(n, k, replace=False)), written in ascending order;
 * include_self == 0 (m == n required, as the reference raises otherwise)
 * drops a row's own index when drawn (row length k - 1), which needs a
 * workspace of bpx_fixed_prob_workspace(m) bytes; allow_multi_conn != 0 draws
 * the k targets with replacement, unsorted.  The random stream (SplitMix64
 * of seed, row and draw index) is this library's own: the same seed gives the
 * same matrix on every call, not the reference's matrix.  n <= 1,310,720
 * without replacement (the row bitmap lives in LDS); m * k < 2^31.
 * ------------------------------------------------------------------------- */
size_t bpx_fixed_prob_workspace(int64_t m);
int bpx_fixed_prob_csr(void* stream, int64_t m, int64_t n, int64_t k, uint64_t seed, int include_self,
                       int allow_multi_conn, int32_t* indices, int32_t* indptr, void* workspace,
                       size_t workspace_bytes);

/* -------------------------------------------------------------------------
 * XLA custom-call adapters.
 *
 * The reference registers its GPU operators as legacy XLA custom calls,
 * `xla_client.register_custom_call_target(name, capsule, platform)`
 * (dependency_check.py:75-79), with the GPU convention
 * `void fn(stream, void** buffers, const char* opaque, size_t opaque_len)`:
 * buffers are the operands in the order of the reference's lowering,
 * followed by the outputs.  These adapters have that signature and wrap the
 * entry points above (same kernels).  Their opaque is a bpx_xla_desc (built
 * by the host binding in place of brainpylib's build_double_size_descriptor /
 * build_nonzero_descriptor, which carry only the shape): it adds the dtypes
 * and lengths the kernels need.  The legacy convention has no error
 * channel: the status of the calling thread's last adapter call is
 * bpx_xla_last_status().
 *
 *  bpx_xla_event_csrmv  replaces event_csrmv_{homo,heter}_* targets
 *    (event/_csr_matvec.py:453-500): buffers (data, indices, indptr, events, out).
 *  bpx_xla_csrmv        replaces csrmv_{homo,heter}_vector_* targets
 *    (sparse/_csr_mv.py:328-361): buffers (data, indices, indptr, vector, out).
 *  bpx_xla_event_info   replaces nonzero_* targets (event/_info_collection.py:
 *    78-121): buffers (events, ids, num, workspace) -- the workspace is a third
 *    output of bpx_event_info_workspace(batch, n) bytes in ws_bytes.
 *  bpx_xla_jitconn_mv   replaces gpu_jit_{event_,}csrmv_{atomic_,}prob_* targets
 *    (jitconn/_event_matvec.py:198-229, 384-414, 578-608; jitconn/_matvec.py:
 *    341-378, 535-574, 717-760): buffers (vec, weight, clen, seed, out) for homo, (vec, w_low|w_mu,
 *    w_high|w_sigma, clen, seed, out) otherwise, then an optional bitmask
 *    workspace of ws_bytes (bpx_jitconn_workspace).  weight/clen/seed are
 *    one-element device arrays there; the adapter reads them with one
 *    synchronous copy per call.
 * ------------------------------------------------------------------------- */
#define BPX_XLA_MAGIC 0x31585042u   /* "BPX1" */
typedef struct {
  uint32_t magic;           /* BPX_XLA_MAGIC */
  int32_t data_dtype;       /* csrmv weights; jitconn weight operands (F32 or F64) */
  int32_t index_dtype;      /* csrmv indices / indptr */
  int32_t vec_dtype;        /* events or vector */
  int32_t rule;             /* BPX_RULE_* for float events */
  int32_t transpose;        /* csrmv */
  int32_t dist;             /* jitconn BPX_DIST_* */
  int32_t outdim_parallel;  /* jitconn */
  int32_t layout;           /* jitconn BPX_LAYOUT_* */
  int32_t event;            /* jitconn: 1 event operator, 0 dense-vector operator */
  int64_t m, n;             /* csrmv shape; jitconn (num_row, num_col) after the transpose swap; info n */
  int64_t nnz;              /* csrmv len(indices) */
  int64_t data_len;         /* csrmv 1 (homogeneous) or nnz */
  int64_t batch;            /* info rows */
  int64_t ws_bytes;         /* bytes of the trailing workspace buffer (0: none) */
} bpx_xla_desc;

void bpx_xla_event_csrmv(void* stream, void** buffers, const char* opaque, size_t opaque_len);
void bpx_xla_csrmv(void* stream, void** buffers, const char* opaque, size_t opaque_len);
void bpx_xla_event_info(void* stream, void** buffers, const char* opaque, size_t opaque_len);
void bpx_xla_jitconn_mv(void* stream, void** buffers, const char* opaque, size_t opaque_len);
int bpx_xla_last_status(void);

/*
 * Kernel timer.  While enabled, the dominant kernel of each entry point is
 * launched with HIP start/stop events on the caller's stream; bpx_timer_read
 * synchronises those events and returns the summed device milliseconds and
 * launch count for one kernel id (0 event scatter, 1 event gather,
 * 2 csr scatter, 3 csr gather, 4 jitconn outdim, 5 jitconn scatter,
 * 6 event info, 7 lif step, 8 fixed order, 9 tiled active-row list,
 * 10 network step, 11 tiled event scatter, 12 COO products, 13 BSR
 * products, 14 multi-step network, 15 tiled gather, 16 jitconn scatter
 * list pass; the two-pass jitconn scatter times its list pass under 16
 * and its sum pass under 5).
 */
int bpx_timer_enable(int on);
int bpx_timer_read(int kernel_id, double* total_ms, int64_t* launches);
int bpx_timer_reset(void);

#ifdef __cplusplus
}
#endif

#endif /* BPX_H_ */
