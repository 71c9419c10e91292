// Native substep: the whole of KarmanVortexSolver.substep
// (karman_vortex.py:232-242) issued from C++, with the DFSPH Jacobi loops
// (dimcv_sph.py:376-393, 597-614) converging on the device.
//
// Loop semantics are the reference's: at least one and at most max
// iterations; after iteration k the average residual (FP64 sum / fluid
// count) is compared with eta.  Here the comparison runs in a one-thread
// kernel that raises a device flag; every kernel of later iterations reads
// the flag first and returns, so the host may issue `chunk` iterations
// between synchronisations without changing any result.
#include <algorithm>
#include <vector>

#include "common.h"

namespace dimcv {

int launch_jacobi_pass(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                       const int32_t* done, hipStream_t s, double* partials = nullptr,
                       float offset = 0.0f, const Packed* pk = nullptr);
int launch_pack_xm(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, const Packed& pk,
                   hipStream_t s, bool zero_tiny = true);
int launch_scale_kappa(int which, const dimcv_state* st, const dimcv_params* p, float scale, int32_t fluid_bound,
                       float* kap, hipStream_t s, bool reset_jacobi = false, int32_t* bar_tree = nullptr);
int launch_copy_advect(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, hipStream_t s);
int launch_predict_life(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, hipStream_t s);
int launch_nonpressure(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, int32_t fluid_bound,
                       hipStream_t s, const Packed* pk, float* factor = nullptr, float factor_scale = 1.0f,
                       bool zero_fixed_m = false);
void set_far_timing(hipEvent_t a, hipEvent_t b);
int neighbor_build_launch(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, hipStream_t s);
int launch_fused(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, hipStream_t s,
                 double* partials = nullptr, const Packed* pk = nullptr);
int launch_solve_single(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                        const Packed& pk, double* partials, float offset, double eta, int32_t max_iter,
                        int32_t* bar_tree, hipStream_t s);

// Events for the optional far-field timing (created on first use, kept for
// the life of the process).
static hipEvent_t g_step_ev[2] = {nullptr, nullptr};

__global__ void k_jacobi_reset(int32_t* counters) {
  counters[DIMCV_C_JACOBI_DONE] = counters[DIMCV_C_STEP_ERROR] != 0 ? 1 : 0;
  counters[DIMCV_C_JACOBI_ITERS] = 0;
}

// The build's outcome, before the substep writes anything: an active solid
// keyed past the cell table, a list wider than nbr_kmax (the lists are
// truncated) or no fluid (the reference divides by fluid_particle_num).
// Nonzero makes the substep's state writers return (step_failed).
__global__ void k_step_guard(int32_t* counters, int32_t nbr_kmax, int32_t* tiny) {
  *tiny = 0;  // the tiny-coordinate flag that k_pack_xm raises next
  counters[DIMCV_C_BAR_ABORT] = 0;
  int32_t e = 0;
  if (counters[DIMCV_C_KEY_ERROR] != 0) e |= 1;
  if (counters[DIMCV_C_NBR_MAX] > nbr_kmax) e |= 2;
  if (counters[DIMCV_C_FLUID_NUM] <= 0) e |= 4;
  counters[DIMCV_C_STEP_ERROR] = e;
}

// Host side of the same test, from a copy of the counters.
static int step_status(const int32_t* hc, const dimcv_params* p) {
  if (hc[DIMCV_C_BAR_ABORT] != 0) {
    set_error("dimcv_step: a grid barrier of the device-side Jacobi loop timed out (blocks not all resident)");
    return DIMCV_ERR_HIP;
  }
  if (hc[DIMCV_C_KEY_ERROR] != 0) {
    set_error("%d active solid particles lie past the cell table", hc[DIMCV_C_KEY_ERROR]);
    return DIMCV_ERR_KEY;
  }
  if (hc[DIMCV_C_NBR_MAX] > p->nbr_kmax) {
    set_error("neighbour list width %d exceeds nbr_kmax %d", hc[DIMCV_C_NBR_MAX], p->nbr_kmax);
    return DIMCV_ERR_NBR;
  }
  if (hc[DIMCV_C_FLUID_NUM] <= 0) {
    set_error("dimcv_step: no fluid particles (the reference divides by fluid_particle_num)");
    return DIMCV_ERR_ARG;
  }
  return DIMCV_OK;
}


// which = 0: divergence (iteration, density change, residual offset 0)
// which = 1: pressure (iteration, density_adv, residual offset rho0)
// Iterations are issued in chunks; the first chunk is jacobi_first_chunk()
// (chunk 0: the previous solve's iteration count passed in *iters, at least
// one), later ones `chunk` (or one).
static int jacobi_first_chunk(int which, const dimcv_step_opts* o, int32_t iters) {
  const int32_t max_iter = which == 0 ? o->max_iter_v : o->max_iter_p;
  return o->chunk > 0 ? o->chunk : std::max(1, std::min(iters, max_iter));
}

// The iteration reset and `n` iterations (kernels only).
static int jacobi_issue(int which, const dimcv_state* st, const dimcv_params* p, const dimcv_step_opts* o,
                        int32_t fluid_bound, const Workspace& ws, hipStream_t s, int n, bool reset) {
  const int32_t* done = st->counters + DIMCV_C_JACOBI_DONE;
  const int32_t max_iter = which == 0 ? o->max_iter_v : o->max_iter_p;
  const double eta = which == 0 ? o->eta_v : o->eta_p;
  const float offset = which == 0 ? 0.0f : o->density0;
  const Packed pk{ws.xm, ws.xw, ws.kap, ws.tiny};
  if (reset) hipLaunchKernelGGL(k_jacobi_reset, dim3(1), dim3(1), 0, s, st->counters);
  for (int c = 0; c < n; ++c) {
    // iteration pass; density pass with the residual's block partials;
    // the FP64 reduction and the convergence test in one launch
    int rc = launch_jacobi_pass(which == 0 ? 0 : 2, st, p, fluid_bound, done, s, nullptr, 0.0f, &pk);
    if (!rc) rc = launch_jacobi_pass(which == 0 ? 1 : 3, st, p, fluid_bound, done, s, ws.partials, offset, &pk);
    if (!rc)
      rc = reduce_tail_f64(ws.partials, nblocks(fluid_bound), st->dscal + DIMCV_D_RESIDUAL, st->counters, done, 1,
                           o->fluid_count, eta, max_iter, s);
    if (rc) return rc;
  }
  DIMCV_LAUNCHED("jacobi chunk");
  return DIMCV_OK;
}

// The counter array read back to the host for a convergence check: an
// asynchronous copy into pinned memory and an event the host polls, which
// returns within microseconds of the copy (a blocking synchronisation waits
// for the runtime to wake the thread).  Buffer and event are made once and
// kept for the life of the process.
static int32_t* g_check_host = nullptr;
static hipEvent_t g_check_ev = nullptr;
// The divergence solve's counters of a device-side step (pinned).
static int32_t* g_div_host = nullptr;

static int read_counters(const dimcv_state* st, hipStream_t s, int32_t* out) {
  if (!g_check_host) {
    DIMCV_HIP(hipHostMalloc((void**)&g_check_host, DIMCV_C_COUNT * sizeof(int32_t), hipHostMallocDefault));
    DIMCV_HIP(hipEventCreateWithFlags(&g_check_ev, hipEventDisableTiming));
  }
  DIMCV_HIP(hipMemcpyAsync(g_check_host, st->counters, DIMCV_C_COUNT * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipEventRecord(g_check_ev, s));
  hipError_t e;
  while ((e = hipEventQuery(g_check_ev)) == hipErrorNotReady) {
  }
  DIMCV_HIP(e);
  std::copy(g_check_host, g_check_host + DIMCV_C_COUNT, out);
  return DIMCV_OK;
}

// Host checks of an issued solve until it has converged, issuing further
// chunks.  With `all`, each check reads the whole counter array into it.
static int jacobi_wait(int which, const dimcv_state* st, const dimcv_params* p, const dimcv_step_opts* o,
                       int32_t fluid_bound, const Workspace& ws, hipStream_t s, int32_t* iters_out,
                       int32_t* all = nullptr) {
  int32_t local[DIMCV_C_COUNT];
  int32_t* c = all ? all : local;
  const int chunk = o->chunk > 0 ? o->chunk : 1;
  for (;;) {
    int rc0 = read_counters(st, s, c);
    if (rc0) return rc0;
    if (c[DIMCV_C_JACOBI_DONE]) break;
    int rc = jacobi_issue(which, st, p, o, fluid_bound, ws, s, chunk, false);
    if (rc) return rc;
  }
  *iters_out = c[DIMCV_C_JACOBI_ITERS];
  return DIMCV_OK;
}

// The rest of a solve on the device (dimcv_step with chunk < 0): one
// persistent launch that returns at once when the issued iterations have
// converged.
static int jacobi_device(int which, const dimcv_state* st, const dimcv_params* p, const dimcv_step_opts* o,
                         int32_t fluid_bound, const Workspace& ws, hipStream_t s) {
  const Packed pk{ws.xm, ws.xw, ws.kap, ws.tiny};
  return launch_solve_single(which, st, p, fluid_bound, pk, ws.partials, which == 0 ? 0.0f : o->density0,
                             which == 0 ? o->eta_v : o->eta_p, which == 0 ? o->max_iter_v : o->max_iter_p,
                             ws.bar_tree, s);
}

static int jacobi(int which, const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* o,
                  int32_t fluid_bound, const Workspace& ws, hipStream_t s, int32_t* iters, int32_t* all = nullptr) {
  int rc = jacobi_issue(which, st, p, o, fluid_bound, ws, s, jacobi_first_chunk(which, o, *iters), true);
  if (rc) return rc;
  return jacobi_wait(which, st, p, o, fluid_bound, ws, s, iters, all);
}

}  // namespace dimcv

using namespace dimcv;

#define CHECK(call)        \
  do {                     \
    int rc_ = (call);      \
    if (rc_) return rc_;   \
  } while (0)

// The substep up to the divergence solve's first host check: kernels only
// (no host synchronisation), so it can be captured into a graph.
static int substep_head(const dimcv_state* st, const dimcv_params* p, const dimcv_step_opts* o,
                        int32_t particle_bound, int32_t fluid_bound, hipStream_t s, const Workspace& ws,
                        int n_first) {
  const Packed pk{ws.xm, ws.xw, ws.kap, ws.tiny};
  hipLaunchKernelGGL(k_step_guard, dim3(1), dim3(1), 0, s, st->counters, p->nbr_kmax, ws.tiny);
  CHECK(launch_pack_xm(st, p, particle_bound, pk, s, false));
  // compute_densities + compute_DFSPH_factor + the first
  // compute_density_change of divergence_solve (dimcv_sph.py:376-393) in one pass
  CHECK(launch_fused(3, st, p, fluid_bound, s, nullptr, &pk));
  CHECK(launch_scale_kappa(0, st, p, o->inv_dt, fluid_bound, ws.kap, s, true, ws.bar_tree));  // also resets the solve
  return jacobi_issue(0, st, p, o, fluid_bound, ws, s, n_first, false);
}

// The substep after the divergence solve, in segments that hold kernels
// only (each can be replayed as a graph), around the far field (issued
// directly, between its timing events) and the pressure solve's checks.
static int seg_mid1(const dimcv_state* st, const dimcv_params* p, const dimcv_step_opts* o, int32_t particle_bound,
                    int32_t fluid_bound, hipStream_t s, const Workspace& ws) {
  void* stream = (void*)s;
  const Packed pk{ws.xm, ws.xw, ws.kap, ws.tiny};
  // multiply_time_step (dimcv_sph.py:395-399) inside the non-pressure pass
  // with fixed-M sampling it also zero-fills the non-fluid slots the vorticity
  // pass leaves (launch_fused 7)
  CHECK(launch_nonpressure(st, p, particle_bound, fluid_bound, s, &pk, st->dfsph_factor, o->dt, o->m > 0));
  // dimcv() (dimcv_sph.py:812-819)
  // init() + compute_d_vorticity in one pass; then the KVN (and Bernoulli
  // sample update), which reads only grad_v; then the smoothing pass, which
  // also produces vorticity_vis (compute_vorticity_vis below in the
  // reference) and the kernel density with its kden_sum partials.
  // With fixed-M sampling the KVN is computed inside the vorticity pass, and
  // one zero fill covers the non-fluid slots of both passes.
  if (o->m > 0) {
    CHECK(launch_fused(7, st, p, fluid_bound, s, nullptr, &pk));
  } else {
    CHECK(launch_fused(1, st, p, fluid_bound, s, nullptr, &pk));
    CHECK(dimcv_compute_kvn_sample(st, p, o->u1, o->u2, 1, fluid_bound, stream));
  }
  CHECK(launch_fused(o->m > 0 ? 6 : 4, st, p, fluid_bound, s, ws.partials, &pk));
  CHECK(reduce_tail_f64(ws.partials, nblocks(fluid_bound), st->dscal + DIMCV_D_KDEN_SUM, st->counters, nullptr, 0,
                        0, 0.0, 0, s));
  if (o->m > 0) CHECK(dimcv_cdf_sample(st, p, o->u_cdf, o->m, fluid_bound, stream));
  else CHECK(dimcv_record_sample_idx(st, p, particle_bound, stream));
  return dimcv_vort2vel_near(st, p, fluid_bound, stream);
}

static int seg_far(const dimcv_state* st, const dimcv_params* p, const dimcv_step_opts* o, int32_t fluid_bound,
                   hipStream_t s) {
  if (o->time_far) {
    if (!g_step_ev[0]) {
      DIMCV_HIP(hipEventCreate(&g_step_ev[0]));
      DIMCV_HIP(hipEventCreate(&g_step_ev[1]));
    }
    set_far_timing(g_step_ev[0], g_step_ev[1]);
  }
  int rc_vv = dimcv_vort2vel_far(st, p, fluid_bound, (int32_t)p->n_max, (void*)s);
  set_far_timing(nullptr, nullptr);
  return rc_vv;
}

static int seg_mid2(const dimcv_state* st, const dimcv_params* p, const dimcv_step_opts* o, int32_t particle_bound,
                    int32_t fluid_bound, hipStream_t s, const Workspace& ws, int n_first_p) {
  const Packed pk{ws.xm, ws.xw, ws.kap, ws.tiny};
  // update_particle_life_time + predict_velocity, by slot (the fluid list
  // holds exactly the fluid slots below particle_num)
  CHECK(launch_predict_life(st, p, particle_bound, s));
  // pressure_solve (dimcv_sph.py:597-614)
  CHECK(launch_jacobi_pass(3, st, p, fluid_bound, nullptr, s, nullptr, 0.0f, &pk));
  CHECK(launch_scale_kappa(1, st, p, o->inv_dt2, fluid_bound, ws.kap, s, true, ws.bar_tree));  // also resets the solve
  return jacobi_issue(1, st, p, o, fluid_bound, ws, s, n_first_p, false);
}

static int seg_tail(const dimcv_state* st, const dimcv_params* p, const dimcv_step_opts* o, int32_t particle_bound,
                    hipStream_t s) {
  void* stream = (void*)s;
  CHECK(launch_copy_advect(st, p, particle_bound, s));  // copy_x_temp + advect
  if (o->bounce_back) CHECK(dimcv_enforce_boundary_3d(st, p, particle_bound, stream));
  return DIMCV_OK;
}

// The far field's duration (the pressure solve's check has synchronised past it).
static int far_time(const dimcv_params* p, dimcv_step_opts* o) {
  if (o->time_far && p->far_mode != DIMCV_FAR_EXACT) {
    DIMCV_HIP(hipEventSynchronize(g_step_ev[1]));
    DIMCV_HIP(hipEventElapsedTime(&o->out_far_ms, g_step_ev[0], g_step_ev[1]));
  }
  return DIMCV_OK;
}

// The rest of the substep, issued directly; with `all` the pressure solve's
// last host check leaves the whole counter array there.
static int substep_tail(const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* o, int32_t particle_bound,
                        int32_t fluid_bound, hipStream_t s, const Workspace& ws, int32_t* all) {
  CHECK(jacobi_wait(0, st, p, o, fluid_bound, ws, s, &o->out_kv));
  CHECK(seg_mid1(st, p, o, particle_bound, fluid_bound, s, ws));
  CHECK(seg_far(st, p, o, fluid_bound, s));
  CHECK(seg_mid2(st, p, o, particle_bound, fluid_bound, s, ws, jacobi_first_chunk(1, o, o->out_kp)));
  CHECK(jacobi_wait(1, st, p, o, fluid_bound, ws, s, &o->out_kp, all));
  CHECK(seg_tail(st, p, o, particle_bound, s));
  return far_time(p, o);
}

static int async_idle(const char* who);

extern "C" int dimcv_substep(const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* o,
                             int32_t particle_bound, int32_t fluid_bound, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!st || !p || !o) {
    set_error("dimcv_substep: null argument");
    return DIMCV_ERR_ARG;
  }
  CHECK(async_idle("dimcv_substep"));
  if (o->fluid_count <= 0) {
    set_error("dimcv_substep: no fluid particles (the reference divides by fluid_particle_num)");
    return DIMCV_ERR_ARG;
  }
  if (o->m == 0 && (!o->u1 || !o->u2)) {
    set_error("dimcv_substep: Bernoulli uniforms required");
    return DIMCV_ERR_ARG;
  }
  if (o->m > 0 && !o->u_cdf) {
    set_error("dimcv_substep: inverse-CDF uniforms required");
    return DIMCV_ERR_ARG;
  }
  Workspace ws;
  CHECK(bind_workspace(st, p, &ws, p->far_mode == DIMCV_FAR_EXACT ? 0 : p->n_max));
  CHECK(substep_head(st, p, o, particle_bound, fluid_bound, s, ws, jacobi_first_chunk(0, o, o->out_kv)));
  return substep_tail(st, p, o, particle_bound, fluid_bound, s, ws, nullptr);
}

// The start of a step (grid build, lists, substep up to the divergence
// solve's first host check) as a HIP graph, captured once per distinct set
// of arguments (the two buffer parities of a run) on a private stream and
// replayed on the caller's stream: one launch instead of ~40, whose small
// kernels would otherwise wait for the host to issue them.  If capture is
// unavailable the kernels are issued directly.  The cache holds the last
// eight argument sets (buffer parity x uniform buffers x timing); calls
// from several threads must not overlap.
struct HeadKey {
  std::vector<unsigned char> bytes;
  void add(const void* q, size_t n) {
    const unsigned char* c = (const unsigned char*)q;
    bytes.insert(bytes.end(), c, c + n);
  }
};
struct SegGraph {
  std::vector<unsigned char> key;
  hipGraphExec_t exec;
};
// one cache per captured segment of dimcv_step
enum { SEG_HEAD, SEG_MID1, SEG_MID2, SEG_FINISH, SEG_COUNT };
static std::vector<SegGraph> g_seg_graphs[SEG_COUNT];
static hipStream_t g_capture_stream = nullptr;
static bool g_graphs_off = false;
static bool g_seg_off[SEG_COUNT] = {};

// With `direct_on_error`, a segment whose capture fails (an error inside the
// capture included) is issued directly, now and from then on.
template <class F>
static int run_graph(int seg, F&& issue, const HeadKey& key, hipStream_t s, bool direct_on_error = false) {
  if (g_graphs_off || g_seg_off[seg]) return issue(s);
  std::vector<SegGraph>& cache = g_seg_graphs[seg];
  for (const SegGraph& g : cache)
    if (g.key == key.bytes) {
      DIMCV_HIP(hipGraphLaunch(g.exec, s));
      return DIMCV_OK;
    }
  if (!g_capture_stream) DIMCV_HIP(hipStreamCreateWithFlags(&g_capture_stream, hipStreamNonBlocking));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ok = hipStreamBeginCapture(g_capture_stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
  int rc = DIMCV_OK;
  if (ok) {
    rc = issue(g_capture_stream);
    ok = hipStreamEndCapture(g_capture_stream, &graph) == hipSuccess && graph != nullptr && rc == DIMCV_OK;
    ok = ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  }
  if (graph) (void)hipGraphDestroy(graph);
  if (direct_on_error && (rc || !ok)) {
    (void)hipGetLastError();
    g_seg_off[seg] = true;
    return issue(s);
  }
  if (rc) return rc;
  if (!ok) {  // no capture on this runtime: issue directly from now on
    (void)hipGetLastError();
    g_graphs_off = true;
    return issue(s);
  }
  if (cache.size() >= 8) {
    (void)hipGraphExecDestroy(cache.front().exec);
    cache.erase(cache.begin());
  }
  cache.push_back({key.bytes, exec});
  DIMCV_HIP(hipGraphLaunch(exec, s));
  return DIMCV_OK;
}

// A step whose solves converge on the device (no host checks): both of its
// halves decide it the same way from the same arguments.
static bool device_step(const dimcv_step_opts* o, const dimcv_params* p) { return o->chunk < 0 && !p->dyn_rigid; }

// Arguments, workspace and graph key of one dimcv_step (the same for its
// begin and finish halves).
struct StepPrep {
  Workspace ws;
  HeadKey key;
  int n_first_v, n_first_p;
};

static int step_prepare(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* o,
                        int32_t particle_bound, StepPrep* sp) {
  if (!pre || !st || !p || !o) {
    set_error("dimcv_step: null argument");
    return DIMCV_ERR_ARG;
  }
  if (o->m == 0 && (!o->u1 || !o->u2)) {
    set_error("dimcv_step: Bernoulli uniforms required");
    return DIMCV_ERR_ARG;
  }
  if (o->m > 0 && !o->u_cdf) {
    set_error("dimcv_step: inverse-CDF uniforms required");
    return DIMCV_ERR_ARG;
  }
  if (p->n_max > DIMCV_MAX_SLOTS) {
    set_error("n_max %lld exceeds DIMCV_MAX_SLOTS (%d)", (long long)p->n_max, DIMCV_MAX_SLOTS);
    return DIMCV_ERR_ARG;
  }
  o->fluid_count = 0;  // the residual divisor is read from the device counters
  CHECK(bind_workspace(st, p, &sp->ws, p->far_mode == DIMCV_FAR_EXACT ? 0 : p->n_max));
  sp->n_first_v = jacobi_first_chunk(0, o, o->out_kv);
  sp->n_first_p = jacobi_first_chunk(1, o, o->out_kp);
  // every argument the captured segments depend on
  dimcv_step_opts ko = *o;
  ko.out_kv = ko.out_kp = ko.time_far = 0;
  ko.out_far_ms = 0.0f;
  sp->key.add(pre, sizeof(*pre));
  sp->key.add(st, sizeof(*st));
  sp->key.add(p, sizeof(*p));
  sp->key.add(&ko, sizeof(ko));
  sp->key.add(&particle_bound, sizeof(particle_bound));
  sp->key.add(&sp->n_first_v, sizeof(sp->n_first_v));
  sp->key.add(&sp->n_first_p, sizeof(sp->n_first_p));
  return DIMCV_OK;
}

// The device-side rest of a step (both solves converging on the device):
// the divergence solve's remaining iterations, the copy of its counters to
// g_div_host, the substep to the far field, the far field (between its
// timing events), the pressure solve and the tail, replayed as one graph.
static int finish_device(const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* o, int32_t pb,
                         const StepPrep& sp, hipStream_t s) {
  if (!g_div_host) DIMCV_HIP(hipHostMalloc((void**)&g_div_host, DIMCV_C_COUNT * sizeof(int32_t),
                                           hipHostMallocDefault));
  if (o->time_far && !g_step_ev[0]) {
    DIMCV_HIP(hipEventCreate(&g_step_ev[0]));
    DIMCV_HIP(hipEventCreate(&g_step_ev[1]));
  }
  const Workspace& ws = sp.ws;
  HeadKey key = sp.key;
  key.add(&o->time_far, sizeof(o->time_far));
  return run_graph(SEG_FINISH, [&](hipStream_t hs) -> int {
    CHECK(jacobi_device(0, st, p, o, pb, ws, hs));
    DIMCV_HIP(hipMemcpyAsync(g_div_host, st->counters, DIMCV_C_COUNT * sizeof(int32_t), hipMemcpyDeviceToHost, hs));
    CHECK(seg_mid1(st, p, o, pb, pb, hs, ws));
    CHECK(seg_far(st, p, o, pb, hs));
    CHECK(seg_mid2(st, p, o, pb, pb, hs, ws, sp.n_first_p));
    CHECK(jacobi_device(1, st, p, o, pb, ws, hs));
    return seg_tail(st, p, o, pb, hs);
  }, key, s, true);
}

// The grid build into the sorted buffers of `pre` (= the fields of `st`),
// the lists and the substep up to the divergence solve's first check, all
// bounded by particle_bound; enqueued only (one graph launch).
extern "C" int dimcv_step_begin(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p,
                                dimcv_step_opts* o, int32_t particle_bound, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  StepPrep sp;
  CHECK(step_prepare(pre, st, p, o, particle_bound, &sp));
  const Workspace& ws = sp.ws;
  const int32_t pb = particle_bound;
  const int n_first_v = sp.n_first_v;
  return run_graph(SEG_HEAD, [&](hipStream_t hs) -> int {
    CHECK(dimcv_grid_build(pre, p, pb, nullptr, hs));
    CHECK(neighbor_build_launch(st, p, pb, hs));
    // compute_moving_boundary_volume (karman_vortex.py:223) when the scene has dynamic rigid particles
    if (p->dyn_rigid) CHECK(dimcv_compute_moving_boundary_volume(st, p, pb, hs));
    return substep_head(st, p, o, pb, pb, hs, ws, n_first_v);
  }, sp.key, s);
}

// Asynchronous completion of a device-side step: dimcv_step_finish_async
// enqueues the rest of the step and the copy of the counters into pinned
// memory and returns; dimcv_step_wait, called before the next step's
// _finish (or at any later time), waits for that copy and reports the step
// as dimcv_step_finish would have.  The caller may enqueue the next step's
// _begin in between, so the device does not idle while the host prepares
// the next step.  One asynchronous step may be outstanding.
static hipEvent_t g_async_ev = nullptr;
static bool g_async_pending = false;

static int async_idle(const char* who) {
  if (g_async_pending) {
    set_error("%s: an asynchronous step is outstanding (call dimcv_step_wait first)", who);
    return DIMCV_ERR_ARG;
  }
  return DIMCV_OK;
}

// The rest of the step after dimcv_step_begin with the same arguments; the
// counts come back in host_counters with the pressure solve's last check.
extern "C" int dimcv_step_finish(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p,
                                 dimcv_step_opts* o, int32_t particle_bound, int32_t* host_counters,
                                 void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!host_counters) {
    set_error("dimcv_step: null argument");
    return DIMCV_ERR_ARG;
  }
  CHECK(async_idle("dimcv_step_finish"));
  StepPrep sp;
  CHECK(step_prepare(pre, st, p, o, particle_bound, &sp));
  const Workspace& ws = sp.ws;
  const HeadKey& key = sp.key;
  const int32_t pb = particle_bound;
  const int n_first_p = sp.n_first_p;
  if (device_step(o, p)) {
    // device-side convergence: the rest of the step is enqueued without a
    // host check (each solve's remaining iterations, if any, in a persistent
    // launch); the divergence solve's counters are copied aside before the
    // pressure solve resets them, and the host waits once, at the end.  A
    // failed build runs the step with every writer of x, v, is_sample and
    // life_time returning (the solves start converged).
    CHECK(finish_device(st, p, o, pb, sp, s));
    CHECK(read_counters(st, s, host_counters));
    o->out_kv = g_div_host[DIMCV_C_JACOBI_ITERS];
    o->out_kp = host_counters[DIMCV_C_JACOBI_ITERS];
    CHECK(far_time(p, o));
    return step_status(host_counters, p);
  }
  // the divergence solve's checks also bring the build's outcome: a failed
  // build ends the step here (its solve was done before it began, and no
  // writer of x, v, is_sample or life_time has run)
  CHECK(jacobi_wait(0, st, p, o, pb, ws, s, &o->out_kv, host_counters));
  if (host_counters[DIMCV_C_STEP_ERROR] != 0) return step_status(host_counters, p);
  CHECK(run_graph(SEG_MID1, [&](hipStream_t hs) { return seg_mid1(st, p, o, pb, pb, hs, ws); }, key, s));
  CHECK(seg_far(st, p, o, pb, s));
  CHECK(run_graph(SEG_MID2, [&](hipStream_t hs) { return seg_mid2(st, p, o, pb, pb, hs, ws, n_first_p); }, key,
                  s));
  CHECK(jacobi_wait(1, st, p, o, pb, ws, s, &o->out_kp, host_counters));
  // one or two kernels: a direct launch starts sooner than a graph launch
  CHECK(seg_tail(st, p, o, pb, s));
  CHECK(far_time(p, o));
  return step_status(host_counters, p);
}

extern "C" int dimcv_step_finish_async(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p,
                                       dimcv_step_opts* o, int32_t particle_bound, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  CHECK(async_idle("dimcv_step_finish_async"));
  StepPrep sp;
  CHECK(step_prepare(pre, st, p, o, particle_bound, &sp));
  if (!device_step(o, p)) {
    set_error("dimcv_step_finish_async: only the device-side step (chunk < 0, no dynamic rigid particles)");
    return DIMCV_ERR_ARG;
  }
  const int32_t pb = particle_bound;
  if (!g_check_host) {
    DIMCV_HIP(hipHostMalloc((void**)&g_check_host, DIMCV_C_COUNT * sizeof(int32_t), hipHostMallocDefault));
    DIMCV_HIP(hipEventCreateWithFlags(&g_check_ev, hipEventDisableTiming));
  }
  if (!g_async_ev) DIMCV_HIP(hipEventCreateWithFlags(&g_async_ev, hipEventDisableTiming));
  CHECK(finish_device(st, p, o, pb, sp, s));
  DIMCV_HIP(hipMemcpyAsync(g_check_host, st->counters, DIMCV_C_COUNT * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipEventRecord(g_async_ev, s));
  g_async_pending = true;
  return DIMCV_OK;
}

extern "C" int dimcv_step_wait(const dimcv_params* p, dimcv_step_opts* o, int32_t* host_counters) {
  if (!p || !o || !host_counters) {
    set_error("dimcv_step_wait: null argument");
    return DIMCV_ERR_ARG;
  }
  if (!g_async_pending) {
    set_error("dimcv_step_wait: no asynchronous step is outstanding");
    return DIMCV_ERR_ARG;
  }
  hipError_t e;
  while ((e = hipEventQuery(g_async_ev)) == hipErrorNotReady) {
  }
  g_async_pending = false;
  DIMCV_HIP(e);
  std::copy(g_check_host, g_check_host + DIMCV_C_COUNT, host_counters);
  o->out_kv = g_div_host[DIMCV_C_JACOBI_ITERS];
  o->out_kp = host_counters[DIMCV_C_JACOBI_ITERS];
  CHECK(far_time(p, o));
  return step_status(host_counters, p);
}

extern "C" int dimcv_step(const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* o,
                          int32_t particle_bound, int32_t* host_counters, void* stream) {
  if (!host_counters) {
    set_error("dimcv_step: null argument");
    return DIMCV_ERR_ARG;
  }
  CHECK(dimcv_step_begin(pre, st, p, o, particle_bound, stream));
  return dimcv_step_finish(pre, st, p, o, particle_bound, host_counters, stream);
}
