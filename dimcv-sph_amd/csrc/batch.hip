// Batched scenes: many independent scenes advanced by one set of launches.
//
// Layout.  Slots of scene s are contiguous and in scene order; a scene's cell
// keys are cell_base + its own flattened cell index, so the stable sort keeps
// every scene in the order its solo run would have (inactive slots of all
// scenes follow the last scene, past particle_num, as in a solo run).  The
// fluid list gives scene s the segment [fluid_start, fluid_start + F) with
// fluid_start a multiple of 256 (the rest of the last block holds -1).
// Hence every 256-entry block, every 64-lane wave, every FP64 partial and
// every canonical tree or scan of a scene covers exactly what it covers in a
// solo run, and the per-scene results are bitwise those of the solo run.
//
// Per scene: Jacobi convergence flags (every Jacobi pass reads the flag of
// its block's scene), FP64 reductions of the block partials, the CDF scan
// and draws, and the far-field work (persistent kernel over a work queue).
//
// Each Jacobi solve is one persistent launch (k_batch_solve) that iterates
// until every scene has converged, with the convergence tests on the device,
// so the step needs no host check: it is captured as one HIP graph (grid,
// neighbour lists, substep) and the host reads the counters once at its end.
// Nothing in the step synchronises or allocates.
#include "common.h"

namespace dimcv {

int launch_jacobi_pass(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                       const int32_t* done, hipStream_t s, double* partials = nullptr,
                       float offset = 0.0f, const Packed* pk = nullptr);
int launch_fused(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, hipStream_t s,
                 double* partials = nullptr, const Packed* pk = nullptr);
int launch_block_partials(int which, const dimcv_state* st, const dimcv_params* p, float offset, int32_t bound,
                          double* partials, const int32_t* done, hipStream_t s);
int launch_pack_xm(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, const Packed& pk,
                   hipStream_t s, bool zero_tiny = true);
int launch_scale_kappa(int which, const dimcv_state* st, const dimcv_params* p, float scale, int32_t fluid_bound,
                       float* kap, hipStream_t s, bool reset_jacobi = false, int32_t* bar_tree = nullptr);
int launch_nonpressure(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, int32_t fluid_bound,
                       hipStream_t s, const Packed* pk, float* factor = nullptr, float factor_scale = 1.0f,
                       bool zero_fixed_m = false);
int batch_grid_launch(const dimcv_state* st, const dimcv_params* p, Workspace& ws, int32_t* ulist,
                      int32_t* ucount, hipStream_t s);
int batch_nbr_launch(const dimcv_state* st, const dimcv_params* p, hipStream_t s);
void batch_set_nbr_block(int b);
int batch_cdf_launch(const dimcv_state* st, const dimcv_params* p, Workspace& ws, BatchWs& bw, const double* u,
                     int64_t m_total, hipStream_t s);
int batch_vort2vel_launch(const dimcv_state* st, const dimcv_params* p, Workspace& ws, BatchWs& bw, hipStream_t s);
void batch_set_far_workgroups(int w);
int launch_batch_solve(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                       const Packed& pk, double* partials, float offset, double eta, int32_t max_iter,
                       int32_t* bar, int32_t* bar_tree, hipStream_t s);

static int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

int64_t batch_layout(int64_t n_scenes, int64_t n_max, BatchWs* bw, void* base) {
  char* b = (char*)base;
  int64_t off = 0;
  auto take = [&](int64_t bytes) -> void* {
    void* q = b ? (void*)(b + off) : nullptr;
    off += align256(bytes > 0 ? bytes : 1);
    return q;
  };
  BatchWs w{};
  const int64_t S = n_scenes > 0 ? n_scenes : 1;
  const int64_t n = n_max > 0 ? n_max : 1;
  w.tile_bound = n / DIMCV_SCAN_TILE + S + 1;
  w.ulist = (int32_t*)take(4 * n);
  w.order = (int32_t*)take(4 * S);
  w.item_start = (int32_t*)take(4 * (S + 1));
  w.misc = (int32_t*)take(4 * BM_COUNT);
  w.btot = (double*)take(8 * w.tile_bound);
  w.bpre = (double*)take(8 * w.tile_bound);
  w.bytes = off;
  if (bw) *bw = w;
  return off;
}

int bind_batch(const dimcv_state* st, const dimcv_params* p, BatchWs* bw) {
  if (!st->scenes || st->n_scenes <= 0 || !st->scene_id || !st->scene_id_buffer || !st->scene_counters ||
      !st->scene_dscal || !st->block_scene || !st->batch_workspace) {
    set_error("batched call on a state without a scene table");
    return DIMCV_ERR_ARG;
  }
  if (p->periodic || p->has_vortex_domain || p->has_transit || st->n_emitters > 0) {
    set_error("batched scenes are open boxes without vortex domain, transit area or emitters");
    return DIMCV_ERR_ARG;
  }
  const int64_t need = batch_layout(st->n_scenes, p->n_max, nullptr, nullptr);
  if (st->batch_workspace_bytes < need) {
    set_error("batch workspace too small: %lld < %lld bytes", (long long)st->batch_workspace_bytes,
              (long long)need);
    return DIMCV_ERR_WORKSPACE;
  }
  batch_layout(st->n_scenes, p->n_max, bw, st->batch_workspace);
  return DIMCV_OK;
}

// ---------------------------------------------------------------------------
__global__ void k_batch_begin(int32_t* counters, int32_t* misc) {
  counters[DIMCV_C_KEY_ERROR] = 0;
  counters[DIMCV_C_STEP_ERROR] = 0;
  misc[BM_ERROR] = 0;
}

__device__ __forceinline__ int32_t lower_bound_i32(const int32_t* __restrict__ a, int32_t n, int32_t v) {
  int32_t lo = 0, hi = n;
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Per-scene slot ranges (from the cell table), fluid ranges (in the unpadded
// ascending fluid list), padded fluid segments, CDF tiles, far-field splits,
// partial bases and the far-field work order.  One workgroup.
__global__ void __launch_bounds__(256) k_batch_scene_setup(dimcv_state st, dimcv_params p, BatchWs bw) {
  const int S = st.n_scenes;
  const int32_t U = bw.misc[BM_UCOUNT];
  const int32_t* __restrict__ gpn = st.grid_particles_num;
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const dimcv_scene& sc = st.scenes[s];
    int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
    const int32_t a = sc.cell_base > 0 ? gpn[sc.cell_base - 1] : 0;
    const int32_t b = gpn[sc.cell_base + sc.n_cells - 1];
    const int32_t lo = lower_bound_i32(bw.ulist, U, a);
    const int32_t hi = lower_bound_i32(bw.ulist, U, b);
    c[DIMCV_SC_SLOT_START] = a;
    c[DIMCV_SC_SLOTS] = b - a;
    c[DIMCV_SC_FLUID_SRC] = lo;
    c[DIMCV_SC_FLUID] = hi - lo;
    c[DIMCV_SC_SPLITS] = far_split_count_mc(hi - lo);
  }
  __syncthreads();
  // far-field order: larger per-item cost (the scene's sample count) first,
  // then scene index
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const int64_t cost = st.scenes[s].m;
    int rank = 0;
    for (int t = 0; t < S; ++t) {
      const int64_t cost_t = st.scenes[t].m;
      rank += (cost_t > cost || (cost_t == cost && t < s)) ? 1 : 0;
    }
    bw.order[rank] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  int64_t fs = 0, tiles = 0;
  int err = 0;
  for (int s = 0; s < S; ++s) {
    int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
    const int64_t F = c[DIMCV_SC_FLUID];
    c[DIMCV_SC_FLUID_START] = (int32_t)fs;
    fs += (F + 255) & ~int64_t(255);
    c[DIMCV_SC_TILE_START] = (int32_t)tiles;
    tiles += (F + DIMCV_SCAN_TILE - 1) / DIMCV_SCAN_TILE;
    c[DIMCV_SC_PART_BASE] = 0;  // the batched far field keeps no partial sums
  }
  if (fs > p.n_max) err |= 1;
  int64_t items = 0;
  for (int r = 0; r < S; ++r) {
    const int s = bw.order[r];
    int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
    bw.item_start[r] = (int32_t)items;
    c[DIMCV_SC_ITEM_START] = (int32_t)items;
    items += ((int64_t)c[DIMCV_SC_FLUID] + 255) / 256;
  }
  bw.item_start[S] = (int32_t)items;
  bw.misc[BM_ITEMS] = (int32_t)items;
  bw.misc[BM_QUEUE] = 0;
  bw.misc[BM_TILES] = (int32_t)tiles;
  bw.misc[BM_ERROR] |= err;
  st.counters[DIMCV_C_FLUID_LIST] = (int32_t)(fs < p.n_max ? fs : p.n_max);
}

// Scene of each 256-entry fluid-list block (-1: none).
__global__ void k_batch_block_scene(dimcv_state st, int64_t nb) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int64_t q = b * 256;
  const int s = last_le(st.scene_counters + DIMCV_SC_FLUID_START, st.n_scenes, DIMCV_SC_COUNT, q);
  int r = -1;
  if (s >= 0) {
    const int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
    if (q < (int64_t)c[DIMCV_SC_FLUID_START] + c[DIMCV_SC_FLUID]) r = s;
  }
  st.block_scene[b] = r;
}

// Padded fluid list from the unpadded one.
__global__ void k_batch_place_fluid(dimcv_state st, const int32_t* __restrict__ ulist, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int s = st.block_scene[q >> 8];
  int32_t v = -1;
  if (s >= 0) {
    const int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
    const int64_t lf = q - c[DIMCV_SC_FLUID_START];
    if (lf < c[DIMCV_SC_FLUID]) v = ulist[c[DIMCV_SC_FLUID_SRC] + lf];
  }
  st.fluid_list[q] = v;
}

// Canonical FP64 reduction of each scene's block partials (one workgroup per
// scene) into scene_dscal[s][which].
__global__ void __launch_bounds__(256) k_batch_reduce(dimcv_state st, double* __restrict__ partials, int which) {
  __shared__ double lds[256];
  batch_reduce_scene(st, partials, which, blockIdx.x, lds);
}

// Before a solve: every scene iterates (none when the step failed), no
// iterations counted, the persistent solve's barrier words cleared.
__global__ void k_batch_jacobi_reset(dimcv_state st, int32_t* __restrict__ bar, int32_t* __restrict__ tree) {
  for (int g = threadIdx.x; g < DIMCV_BAR_GROUPS; g += blockDim.x) tree[16 * g] = 0;  // group arrivals
  const int32_t d = step_failed(st) ? 1 : 0;
  for (int s = threadIdx.x; s < st.n_scenes; s += blockDim.x) {
    st.scene_counters[(int64_t)s * DIMCV_SC_COUNT + DIMCV_SC_DONE] = d;
    st.scene_counters[(int64_t)s * DIMCV_SC_COUNT + DIMCV_SC_ITERS] = 0;
  }
  if (threadIdx.x == 0) {
    st.counters[DIMCV_C_JACOBI_DONE] = 0;
    st.counters[DIMCV_C_JACOBI_ITERS] = 0;
    bar[0] = 0;  // arrivals
    bar[2] = 0;  // abort
    bar[3] = 0;  // go flags
    bar[4] = 0;
  }
}

// After the batched list build: a key past a scene's table, a list wider
// than nbr_kmax or a batch capacity error makes the step's state writers
// return (step_failed) and its solves start converged.
__global__ void k_batch_guard(int32_t* counters, const int32_t* misc, int32_t nbr_kmax) {
  int32_t e = 0;
  if (counters[DIMCV_C_KEY_ERROR] != 0) e |= 1;
  if (counters[DIMCV_C_NBR_MAX] > nbr_kmax) e |= 2;
  if (misc[BM_ERROR] != 0) e |= 8;
  counters[DIMCV_C_STEP_ERROR] = e;
}

// ---------------------------------------------------------------------------
// step segments
struct BatchCtx {
  const dimcv_state* pre;  // state before the grid build's swap
  const dimcv_state* st;   // state after it
  const dimcv_params* p;
  const dimcv_step_opts* o;
  Workspace ws;
  BatchWs bw;
  Packed pk;  // the step's packed records (workspace regions)
};

#define CHECK(call)      \
  do {                   \
    int rc_ = (call);    \
    if (rc_) return rc_; \
  } while (0)

static int seg_grid(BatchCtx& c, hipStream_t s) {
  const dimcv_state* st = c.pre;
  const int64_t n = c.p->n_max;
  hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(1), 0, s, st->counters, c.bw.misc);
  CHECK(batch_grid_launch(st, c.p, c.ws, c.bw.ulist, c.bw.misc + BM_UCOUNT, s));
  hipLaunchKernelGGL(k_batch_scene_setup, dim3(1), dim3(256), 0, s, *st, *c.p, c.bw);
  const int64_t nb = n / 256 + 1;
  hipLaunchKernelGGL(k_batch_block_scene, dim3(nblocks(nb)), dim3(256), 0, s, *st, nb);
  hipLaunchKernelGGL(k_batch_place_fluid, dim3(nblocks(n)), dim3(256), 0, s, *st, (const int32_t*)c.bw.ulist, n);
  DIMCV_LAUNCHED("batched grid build");
  return DIMCV_OK;
}

static int seg_div_begin(BatchCtx& c, hipStream_t s) {
  const dimcv_state* st = c.st;
  const int32_t nb = (int32_t)c.p->n_max;
  CHECK(launch_pack_xm(st, c.p, nb, c.pk, s));
  CHECK(launch_fused(3, st, c.p, nb, s, nullptr, &c.pk));
  CHECK(launch_scale_kappa(0, st, c.p, c.o->inv_dt, nb, c.ws.kap, s));
  hipLaunchKernelGGL(k_batch_jacobi_reset, dim3(1), dim3(256), 0, s, *st, c.bw.misc + BM_BAR, c.ws.bar_tree);
  DIMCV_LAUNCHED("k_batch_jacobi_reset");
  return DIMCV_OK;
}

// The divergence (which = 0) or pressure solve of every scene, to
// convergence: the first iteration's two passes (leaving the block residual
// partials), then one persistent launch for the tests and any further
// iterations.
static int seg_solve(BatchCtx& c, int which, hipStream_t s) {
  const dimcv_state* st = c.st;
  const int32_t nb = (int32_t)c.p->n_max;
  const int32_t* done = st->scene_counters + DIMCV_SC_DONE;
  CHECK(launch_jacobi_pass(which == 0 ? 0 : 2, st, c.p, nb, done, s, nullptr, 0.0f, &c.pk));
  CHECK(launch_jacobi_pass(which == 0 ? 1 : 3, st, c.p, nb, done, s, c.ws.partials,
                           which == 0 ? 0.0f : c.o->density0, &c.pk));
  return launch_batch_solve(which, c.st, c.p, (int32_t)c.p->n_max, c.pk, c.ws.partials,
                            which == 0 ? 0.0f : c.o->density0, which == 0 ? c.o->eta_v : c.o->eta_p,
                            which == 0 ? c.o->max_iter_v : c.o->max_iter_p, c.bw.misc + BM_BAR, c.ws.bar_tree,
                            s);
}

static int seg_mid(BatchCtx& c, hipStream_t s) {
  const dimcv_state* st = c.st;
  const dimcv_params* p = c.p;
  const int32_t nb = (int32_t)p->n_max;
  CHECK(dimcv_multiply_time_step(st, p, st->dfsph_factor, c.o->dt, nb, s));
  CHECK(launch_nonpressure(st, p, nb, nb, s, &c.pk));
  CHECK(launch_fused(1, st, p, nb, s, nullptr, &c.pk));
  CHECK(dimcv_compute_kvn_sample(st, p, nullptr, nullptr, 0, nb, s));
  CHECK(launch_fused(4, st, p, nb, s, c.ws.partials, &c.pk));
  hipLaunchKernelGGL(k_batch_reduce, dim3((unsigned)st->n_scenes), dim3(256), 0, s, *st, c.ws.partials,
                     (int)DIMCV_SD_KDEN_SUM);
  DIMCV_LAUNCHED("batched kden");
  CHECK(batch_cdf_launch(st, p, c.ws, c.bw, c.o->u_cdf, c.o->m, s));
  CHECK(batch_vort2vel_launch(st, p, c.ws, c.bw, s));
  CHECK(dimcv_update_particle_life_time(st, p, nb, s));
  CHECK(dimcv_predict_velocity(st, p, nb, s));
  CHECK(launch_jacobi_pass(3, st, p, nb, nullptr, s, nullptr, 0.0f, &c.pk));
  CHECK(launch_scale_kappa(1, st, p, c.o->inv_dt2, nb, c.ws.kap, s));
  hipLaunchKernelGGL(k_batch_jacobi_reset, dim3(1), dim3(256), 0, s, *st, c.bw.misc + BM_BAR, c.ws.bar_tree);
  DIMCV_LAUNCHED("k_batch_jacobi_reset");
  return DIMCV_OK;
}

static int seg_post(BatchCtx& c, hipStream_t s) {
  const int32_t nb = (int32_t)c.p->n_max;
  CHECK(dimcv_copy_x_temp(c.st, c.p, nb, s));
  CHECK(dimcv_advect(c.st, c.p, nb, s));
  return DIMCV_OK;
}

static int make_ctx(BatchCtx& c, const dimcv_state* pre, const dimcv_state* st, const dimcv_params* p,
                    const dimcv_step_opts* o) {
  c.pre = pre;
  c.st = st;
  c.p = p;
  c.o = o;
  CHECK(bind_workspace(st, p, &c.ws, o ? (int64_t)o->m : 0));
  CHECK(bind_batch(st, p, &c.bw));
  c.pk = Packed{c.ws.xm, c.ws.xw, c.ws.kap, c.ws.tiny};
  if (o) {
    if (o->m <= 0 || !o->u_cdf) {
      set_error("batched substep: fixed-M uniforms required (m > 0, u_cdf)");
      return DIMCV_ERR_ARG;
    }
  }
  return DIMCV_OK;
}

static int read_iters(const dimcv_state* st, int32_t* out_iters, int32_t* tmp, hipStream_t s) {
  const int64_t S = st->n_scenes;
  DIMCV_HIP(hipMemcpyAsync(tmp, st->scene_counters, sizeof(int32_t) * DIMCV_SC_COUNT * S, hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipStreamSynchronize(s));
  for (int64_t k = 0; out_iters && k < S; ++k) {
    out_iters[2 * k] = tmp[k * DIMCV_SC_COUNT + DIMCV_SC_ITERS_V];
    out_iters[2 * k + 1] = tmp[k * DIMCV_SC_COUNT + DIMCV_SC_ITERS_P];
  }
  return DIMCV_OK;
}

// hc: the counters; misc: the batch misc words (or null).
static int check_counters(const int32_t* hc, const int32_t* misc, const dimcv_params* p) {
  if (hc && hc[DIMCV_C_KEY_ERROR] != 0) {
    set_error("%d active solid particles lie past their scene's cell table", hc[DIMCV_C_KEY_ERROR]);
    return DIMCV_ERR_KEY;
  }
  if (hc && hc[DIMCV_C_NBR_MAX] > p->nbr_kmax) {
    set_error("neighbour list width %d exceeds nbr_kmax %d", hc[DIMCV_C_NBR_MAX], p->nbr_kmax);
    return DIMCV_ERR_NBR;
  }
  if (misc && misc[BM_ERROR]) {
    set_error("batch capacity exceeded (flags %d: 1 fluid list)", misc[BM_ERROR]);
    return DIMCV_ERR_RANGE;
  }
  if (misc && misc[BM_BAR + 2]) {
    set_error("batched Jacobi solve: a grid barrier wait timed out (blocks not all resident)");
    return DIMCV_ERR_HIP;
  }
  return DIMCV_OK;
}

}  // namespace dimcv

using namespace dimcv;

// ---------------------------------------------------------------------------
// step graphs
struct dimcv_batch_graph {
  hipGraph_t graph[2];  // one step per buffer parity
  hipGraphExec_t exec[2];
  dimcv_state st[2];
  dimcv_params p;
  dimcv_step_opts o;
  int32_t* pinned;  // counters, misc words, scene counters
  int32_t n_scenes;
};

extern "C" {

int64_t dimcv_batch_workspace_bytes(int32_t n_scenes, const int32_t* scene_particles, int64_t n_max) {
  if (n_scenes <= 0 || !scene_particles) return -1;
  return batch_layout(n_scenes, n_max, nullptr, nullptr);
}

int dimcv_batch_set_launch(int32_t far_workgroups, int32_t nbr_block) {
  if (far_workgroups < 0 || far_workgroups > 65535 ||
      !(nbr_block == 0 || nbr_block == 64 || nbr_block == 128 || nbr_block == 256)) {
    set_error("dimcv_batch_set_launch: workgroups %d or block %d out of range", far_workgroups, nbr_block);
    return DIMCV_ERR_ARG;
  }
  batch_set_far_workgroups(far_workgroups);
  batch_set_nbr_block(nbr_block > 0 ? nbr_block : 256);
  return DIMCV_OK;
}

int dimcv_batch_grid_build(const dimcv_state* st, const dimcv_params* p, void* stream) {
  if (!st || !p) {
    set_error("dimcv_batch_grid_build: null argument");
    return DIMCV_ERR_ARG;
  }
  BatchCtx c;
  CHECK(make_ctx(c, st, st, p, nullptr));
  return seg_grid(c, (hipStream_t)stream);
}

int dimcv_batch_neighbor_build(const dimcv_state* st, const dimcv_params* p, void* stream) {
  if (!st || !p) {
    set_error("dimcv_batch_neighbor_build: null argument");
    return DIMCV_ERR_ARG;
  }
  BatchCtx c;
  CHECK(make_ctx(c, st, st, p, nullptr));
  return batch_nbr_launch(st, p, (hipStream_t)stream);
}

int dimcv_batch_status(const dimcv_state* st, const dimcv_params* p, int32_t* host_counters, void* stream) {
  if (!st || !p) {
    set_error("dimcv_batch_status: null argument");
    return DIMCV_ERR_ARG;
  }
  BatchCtx c;
  CHECK(make_ctx(c, st, st, p, nullptr));
  hipStream_t s = (hipStream_t)stream;
  int32_t hc[DIMCV_C_COUNT];
  int32_t misc[BM_COUNT];
  DIMCV_HIP(hipMemcpyAsync(hc, st->counters, sizeof(hc), hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipMemcpyAsync(misc, c.bw.misc, sizeof(misc), hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipStreamSynchronize(s));
  if (host_counters)
    for (int k = 0; k < DIMCV_C_COUNT; ++k) host_counters[k] = hc[k];
  return check_counters(hc, misc, p);
}

int dimcv_batch_substep(const dimcv_state* st, const dimcv_params* p, dimcv_step_opts* o, int32_t* out_iters,
                        void* stream) {
  if (!st || !p || !o) {
    set_error("dimcv_batch_substep: null argument");
    return DIMCV_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  BatchCtx c;
  CHECK(make_ctx(c, st, st, p, o));
  CHECK(seg_div_begin(c, s));
  CHECK(seg_solve(c, 0, s));
  CHECK(seg_mid(c, s));
  CHECK(seg_solve(c, 1, s));
  CHECK(seg_post(c, s));
  int32_t misc[BM_COUNT];
  DIMCV_HIP(hipMemcpyAsync(misc, c.bw.misc, sizeof(misc), hipMemcpyDeviceToHost, s));
  int32_t* tmp = new int32_t[(size_t)st->n_scenes * DIMCV_SC_COUNT];
  int rc = read_iters(st, out_iters, tmp, s);  // synchronises the stream
  delete[] tmp;
  if (rc) return rc;
  return check_counters(nullptr, misc, p);
}

void dimcv_batch_graph_destroy(dimcv_batch_graph* g) {
  if (!g) return;
  for (int a = 0; a < 2; ++a) {
    if (g->exec[a]) (void)hipGraphExecDestroy(g->exec[a]);
    if (g->graph[a]) (void)hipGraphDestroy(g->graph[a]);
  }
  if (g->pinned) (void)hipHostFree(g->pinned);
  delete g;
}

// The whole step of buffer parity par as one graph: grid build, lists, the
// build guard and the substep with both solves.
static int capture(dimcv_batch_graph* g, int par, hipStream_t s) {
  const dimcv_state* pre = &g->st[par];
  const dimcv_state* cur = &g->st[par ^ 1];
  BatchCtx c;
  CHECK(make_ctx(c, pre, cur, &g->p, &g->o));
  DIMCV_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  int rc = seg_grid(c, s);
  if (!rc) rc = batch_nbr_launch(cur, &g->p, s);
  if (!rc) {
    hipLaunchKernelGGL(k_batch_guard, dim3(1), dim3(1), 0, s, cur->counters, (const int32_t*)c.bw.misc,
                       g->p.nbr_kmax);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) rc = hip_fail(le, "k_batch_guard");
  }
  if (!rc) rc = seg_div_begin(c, s);
  if (!rc) rc = seg_solve(c, 0, s);
  if (!rc) rc = seg_mid(c, s);
  if (!rc) rc = seg_solve(c, 1, s);
  if (!rc) rc = seg_post(c, s);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  g->graph[par] = graph;
  DIMCV_HIP(hipGraphInstantiate(&g->exec[par], graph, nullptr, nullptr, 0));
  return DIMCV_OK;
}

int dimcv_batch_graph_create(const dimcv_state* st_even, const dimcv_state* st_odd, const dimcv_params* p,
                             const dimcv_step_opts* opts, void* stream, dimcv_batch_graph** out) {
  if (!st_even || !st_odd || !p || !opts || !out) {
    set_error("dimcv_batch_graph_create: null argument");
    return DIMCV_ERR_ARG;
  }
  if (!stream) {
    set_error("dimcv_batch_graph_create: capture needs a non-null stream");
    return DIMCV_ERR_ARG;
  }
  *out = nullptr;
  dimcv_batch_graph* g = new dimcv_batch_graph();
  g->st[0] = *st_even;
  g->st[1] = *st_odd;
  g->p = *p;
  g->o = *opts;
  g->n_scenes = st_even->n_scenes;
  hipError_t e = hipHostMalloc((void**)&g->pinned, sizeof(int32_t) * (DIMCV_C_COUNT + BM_COUNT + (size_t)DIMCV_SC_COUNT *
                                                                      (size_t)(g->n_scenes > 0 ? g->n_scenes : 1)),
                               hipHostMallocDefault);
  if (e != hipSuccess) {
    g->pinned = nullptr;
    dimcv_batch_graph_destroy(g);
    return hip_fail(e, "hipHostMalloc");
  }
  for (int par = 0; par < 2; ++par) {
    int rc = capture(g, par, (hipStream_t)stream);
    if (rc) {
      dimcv_batch_graph_destroy(g);
      return rc;
    }
  }
  *out = g;
  return DIMCV_OK;
}

// One graph launch, then one readback of the counters, the misc words and
// the scene counters (iteration counts) for the step's status.
int dimcv_batch_graph_step(dimcv_batch_graph* g, int32_t parity, int32_t* out_iters, void* stream) {
  if (!g || (parity != 0 && parity != 1)) {
    set_error("dimcv_batch_graph_step: bad arguments");
    return DIMCV_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  const dimcv_state* cur = &g->st[parity ^ 1];
  BatchWs bw;
  CHECK(bind_batch(cur, &g->p, &bw));
  int32_t* hc = g->pinned;
  int32_t* misc = g->pinned + DIMCV_C_COUNT;
  int32_t* sc = misc + BM_COUNT;
  DIMCV_HIP(hipGraphLaunch(g->exec[parity], s));
  DIMCV_HIP(hipMemcpyAsync(hc, cur->counters, sizeof(int32_t) * DIMCV_C_COUNT, hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipMemcpyAsync(misc, bw.misc, sizeof(int32_t) * BM_COUNT, hipMemcpyDeviceToHost, s));
  CHECK(read_iters(cur, out_iters, sc, s));
  return check_counters(hc, misc, &g->p);
}

}  // extern "C"
