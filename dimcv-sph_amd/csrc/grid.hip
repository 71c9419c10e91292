// Uniform grid: particle insertion, activity, stable cell-key radix sort,
// cell table, permutation gather, fluid list and fixed-radius neighbour lists.
//
// Reference: particle_system.py:316-586 (add_particle, update_activity,
// particle_partition, update_grid_id, PrefixSumExecutor, counting_sort,
// for_all_neighbors), sph_base.py:91-112 (boundary volumes),
// karman_vortex.py:27-36 (init_cylinder).
//
// The reference compacts active slots with an atomic counter and counting-
// sorts by cell with a reversed atomic_sub loop; under serial execution both
// are order preserving, so the slot order after a build is "active slots in
// ascending order, stably sorted by cell".  Here that order is produced by
// one stable LSD radix sort of (cell key, slot) in which inactive slots carry
// a key past the last cell, followed by one gather of every persistent field.
#include "common.h"

namespace dimcv {

// ---------------------------------------------------------------------------
// particle insertion (particle_system.py:316-380): add_particle into slot q
__device__ __forceinline__ void add_one(dimcv_state& st, const dimcv_params& p, int64_t q, int32_t obj,
                                        const float* pos3, const float* vel3, float dens, float pres, int32_t mat,
                                        int32_t dyn, const int32_t* col3) {
  st.object_id[q] = obj;
  for (int d = 0; d < 3; ++d) {
    st.x[3 * q + d] = pos3[d];
    st.x_0[3 * q + d] = pos3[d];
    st.v[3 * q + d] = vel3[d];
    st.color[3 * q + d] = col3[d];
    st.acceleration[3 * q + d] = 0.0f;
  }
  st.density[q] = dens;
  st.m_V[q] = p.m_V0;
  st.m[q] = p.m_V0 * dens;
  st.pressure[q] = pres;
  st.material[q] = mat;
  st.is_dynamic[q] = dyn != 0 ? 1 : 0;
  st.is_active[q] = 1;
  st.is_sample[q] = 0;
  st.life_time[q] = 0.0f;
}

__global__ void k_add_particles(dimcv_state st, dimcv_params p, int32_t obj, int32_t n,
                                const float* __restrict__ pos, const float* __restrict__ vel,
                                const float* __restrict__ dens, const float* __restrict__ pres,
                                const int32_t* __restrict__ mat, const int32_t* __restrict__ dyn,
                                const int32_t* __restrict__ col) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int64_t pn = st.counters[DIMCV_C_PARTICLE_NUM];
  int64_t q = pn + k;
  if (q >= p.n_max) return;
  add_one(st, p, q, obj, pos + 3 * k, vel + 3 * k, dens[k], pres[k], mat[k], dyn[k], col + 3 * k);
  if (mat[k] == 1) atomicAdd(&st.counters[DIMCV_C_FLUID_NUM], 1);
}

// emit_particle (dimcv_sph.py:616-637) for one emitter: add_cube of its
// lattice (kept on the device) at particle_num, velocity, density and colour
// shared by the layer, fluid, dynamic, pressure 0.
__global__ void k_emit(dimcv_state st, dimcv_params p, dimcv_emitter e) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= e.n) return;
  const int64_t q = (int64_t)st.counters[DIMCV_C_PARTICLE_NUM] + k;
  if (q >= p.n_max) return;  // dropped beyond capacity (particle_system.py:361)
  add_one(st, p, q, e.object_id, e.pos + 3 * k, e.velocity, e.density, 0.0f, 1, 1, e.color);
}

// The layer's counts: particle_num and fluid_particle_num by the slots
// written (:378-380, :330-331), the emitter's particleNum by the whole layer
// (dimcv_sph.py:626-629 counts compute_cube_particle_num, drops included).
__global__ void k_emit_finish(dimcv_state st, int64_t n_max, int32_t n, int32_t obj) {
  const int64_t pn = st.counters[DIMCV_C_PARTICLE_NUM];
  const int64_t got = pn + n < n_max ? pn + n : (pn < n_max ? n_max : pn);
  st.counters[DIMCV_C_PARTICLE_NUM] = (int32_t)got;
  st.counters[DIMCV_C_FLUID_NUM] += (int32_t)(got - pn);
  if (obj >= 0 && obj < st.n_emitters) st.num_particles_each_emitter[obj] += n;
}

__global__ void k_add_finish(int32_t* counters, int64_t n_max, int32_t n, int32_t* dropped) {
  int64_t pn = counters[DIMCV_C_PARTICLE_NUM];
  int64_t want = pn + n;
  int64_t got = want < n_max ? want : n_max;
  counters[DIMCV_C_PARTICLE_NUM] = (int32_t)got;
  *dropped = (int32_t)(want - got);
}

// karman_vortex.py:27-36
__global__ void k_init_cylinder(dimcv_state st, int32_t bound, int32_t obj, float cx, float cy, float radius) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bound || i >= st.counters[DIMCV_C_PARTICLE_NUM]) return;
  if (st.object_id[i] != obj) return;
  float dx = st.x[3 * i] - cx;
  float dy = st.x[3 * i + 1] - cy;
  if (sqrtf(dx * dx + dy * dy) > radius) {
    st.x[3 * i] = 0.0f;
    st.x[3 * i + 1] = 0.0f;
    st.x[3 * i + 2] = 0.0f;
    st.is_active[i] = 0;
  }
}

// ---------------------------------------------------------------------------
// pos_to_index / flatten (particle_system.py:382-402): f32 division by the
// f32 grid size, truncation toward zero, row-major flattening.
// In a periodic box (positions wrapped into [0, L)) the index is clamped to
// the last cell, which f32 rounding of x / h can otherwise overshoot.
__device__ __forceinline__ void cell_of(const dimcv_params& p, v3 x, int c[3]) {
  c[0] = (int)(x.x / p.grid_size);
  c[1] = (int)(x.y / p.grid_size);
  c[2] = (int)(x.z / p.grid_size);
  if (p.periodic)
    for (int d = 0; d < 3; ++d) c[d] = min(max(c[d], 0), p.grid_num[d] - 1);
}

// update_activity (:413-427) fused with update_grid_id's key computation
// (:490-498).  Inactive slots get key n_cells (sorted past every cell).
__global__ void k_activity_keys(dimcv_state st, dimcv_params p, int64_t n_in,
                                uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_in) return;
  const uint32_t sentinel = (uint32_t)p.n_cells;
  vals[i] = (uint32_t)i;
  if (i >= st.counters[DIMCV_C_PARTICLE_NUM]) {
    keys[i] = sentinel;
    return;
  }
  int32_t mat = st.material[i];
  v3 xi = ld3(st.x, i);
  uint8_t active = st.is_active[i];
  if (mat == 0) {
    active = 1;
  } else if (mat == 1 && p.periodic) {
    // periodic box: wrap instead of deactivating (build-defined variant)
    float q[3] = {xi.x, xi.y, xi.z};
    for (int j = 0; j < 3; ++j) {
      if (q[j] < p.domain_start[j]) q[j] += p.box_size[j];
      else if (q[j] >= p.domain_end[j]) q[j] -= p.box_size[j];
    }
    xi = mk(q[0], q[1], q[2]);
    st3(st.x, i, xi);
  } else if (mat == 1) {
    float q[3] = {xi.x, xi.y, xi.z};
    for (int j = 0; j < 3; ++j) {
      if (q[j] < p.domain_start[j] || q[j] > p.domain_end[j]) {
        active = 0;
        atomicSub(&st.counters[DIMCV_C_FLUID_NUM], 1);
        int32_t o = st.object_id[i];
        if (st.n_emitters > 0 && o >= 0 && o < st.n_emitters)
          atomicSub(&st.num_particles_each_emitter[o], 1);
        break;
      }
    }
  }
  // the flat index as the reference forms it: a cell coordinate one past the
  // grid in y or z (a particle exactly on an upper face) lands in the next
  // row's cell there too
  int c[3];
  cell_of(p, xi, c);
  const int64_t g = ((int64_t)c[0] * p.grid_num[1] + c[1]) * p.grid_num[2] + c[2];
  const bool in_table = g >= 0 && g < p.n_cells;
  if (active && !in_table && mat == 1 && !p.periodic) {
    // an active fluid particle whose index lies past the table (exactly on
    // the domain's upper x face, or the last row's upper y/z face): the
    // reference indexes out of bounds there; it leaves the domain as one
    // past the face does
    active = 0;
    atomicSub(&st.counters[DIMCV_C_FLUID_NUM], 1);
    int32_t o = st.object_id[i];
    if (st.n_emitters > 0 && o >= 0 && o < st.n_emitters) atomicSub(&st.num_particles_each_emitter[o], 1);
  }
  st.is_active[i] = active;
  if (!active) {
    keys[i] = sentinel;
    return;
  }
  if (!in_table) {  // a solid outside the table
    atomicAdd(&st.counters[DIMCV_C_KEY_ERROR], 1);
    keys[i] = sentinel;
    return;
  }
  keys[i] = (uint32_t)g;
}

// particle_num = fluid + solid (:467-468), never above the slots scanned.
__global__ void k_set_particle_num(int32_t* counters, int32_t n_in) {
  int32_t pn = counters[DIMCV_C_FLUID_NUM] + counters[DIMCV_C_SOLID_NUM];
  counters[DIMCV_C_PREV_NUM] = counters[DIMCV_C_PARTICLE_NUM];
  counters[DIMCV_C_PARTICLE_NUM] = pn < n_in ? pn : n_in;
  counters[DIMCV_C_NBR_MAX] = 0;
  counters[DIMCV_C_STEP_ERROR] = 0;
}

// ---------------------------------------------------------------------------
// stable LSD radix sort, 8-bit digits, 4096 keys per workgroup
// Digits of BITS bits (8 or 9: 9 when that saves a pass over the key bits).
template <int BITS>
__global__ void __launch_bounds__(256) k_radix_hist(const uint32_t* __restrict__ keys, int64_t n,
                                                    int shift, uint32_t* __restrict__ hist, int nblk) {
  constexpr int R = 1 << BITS;
  constexpr uint32_t MASK = R - 1;
  __shared__ uint32_t h[R];
  for (int d = threadIdx.x; d < R; d += 256) h[d] = 0;
  __syncthreads();
  int64_t base = (int64_t)blockIdx.x * DIMCV_SORT_TILE;
#pragma unroll 4
  for (int k = 0; k < DIMCV_SORT_TILE / 256; ++k) {
    int64_t i = base + k * 256 + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & MASK], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += 256) hist[(int64_t)d * nblk + blockIdx.x] = h[d];
}

// Each workgroup walks its tile in 16 strips of 256 keys (index order).  In a
// strip, every wave ranks its lanes among same-digit lanes with BITS ballots;
// wave counts per digit go to LDS so a key's output position is
//   base(digit, block) + keys of that digit in earlier strips
//   + same-digit keys in lower waves + same-digit lanes below it,
// which is exactly its stable rank.
template <int BITS>
__global__ void __launch_bounds__(256) k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, const uint32_t* __restrict__ offs, int nblk) {
  constexpr int R = 1 << BITS;
  constexpr uint32_t MASK = R - 1;
  __shared__ uint32_t cnt[R];
  __shared__ uint32_t wc[4][R];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int d = t; d < R; d += 256) {
    cnt[d] = offs[(int64_t)d * nblk + blockIdx.x];
    wc[0][d] = wc[1][d] = wc[2][d] = wc[3][d] = 0;
  }
  __syncthreads();
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int64_t base = (int64_t)blockIdx.x * DIMCV_SORT_TILE;
  for (int k = 0; k < DIMCV_SORT_TILE / 256; ++k) {
    int64_t i = base + k * 256 + t;
    bool valid = i < n;
    uint32_t key = valid ? kin[i] : 0u;
    uint32_t val = valid ? vin[i] : 0u;
    uint32_t d = (key >> shift) & MASK;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = cnt[d] + rank;
      for (int ww = 0; ww < w; ++ww) pos += wc[ww][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    for (int dd = t; dd < R; dd += 256) {
      cnt[dd] += wc[0][dd] + wc[1][dd] + wc[2][dd] + wc[3][dd];
      wc[0][dd] = wc[1][dd] = wc[2][dd] = wc[3][dd] = 0;
    }
    __syncthreads();
  }
}

// Cell table: grid_particles_num[c] = number of sorted keys <= c (the
// reference's inclusive prefix sum of per-cell counts); the counting sort's
// remaining per-cell counter is zero for every cell afterwards.
__global__ void k_cell_end(const uint32_t* __restrict__ keys, int64_t n, int32_t n_cells,
                           int32_t* __restrict__ gpn, int32_t* __restrict__ gpn_temp) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (keys[mid] <= (uint32_t)c) lo = mid + 1;
    else hi = mid;
  }
  gpn[c] = (int32_t)lo;
  gpn_temp[c] = 0;
}

// One gather of every sorted field into its buffer (particle_partition +
// counting_sort copies, :435-453 and :513-532).
__global__ void k_gather(dimcv_state st, int64_t n_in, const uint32_t* __restrict__ keys,
                         const uint32_t* __restrict__ perm) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_in || s >= st.counters[DIMCV_C_PARTICLE_NUM]) return;
  int64_t i = perm[s];
  st.grid_ids_buffer[s] = (int32_t)keys[s];
  st.object_id_buffer[s] = st.object_id[i];
  st3(st.x_buffer, s, ld3(st.x, i));
  st3(st.x_0_buffer, s, ld3(st.x_0, i));
  st3(st.v_buffer, s, ld3(st.v, i));
  st.m_V_buffer[s] = st.m_V[i];
  st.m_buffer[s] = st.m[i];
  st.density_buffer[s] = st.density[i];
  st.pressure_buffer[s] = st.pressure[i];
  st.material_buffer[s] = st.material[i];
  st.color_buffer[3 * s] = st.color[3 * i];
  st.color_buffer[3 * s + 1] = st.color[3 * i + 1];
  st.color_buffer[3 * s + 2] = st.color[3 * i + 2];
  st.is_dynamic_buffer[s] = st.is_dynamic[i];
  st.dfsph_factor_buffer[s] = st.dfsph_factor[i];
  st.density_adv_buffer[s] = st.density_adv[i];
  st.is_active_buffer[s] = 1;
  st.is_sample_buffer[s] = st.is_sample[i];
  st.life_time_buffer[s] = st.life_time[i];
  if (st.scene_id) st.scene_id_buffer[s] = st.scene_id[i];
  if (st.particle_id) st.particle_id_buffer[s] = st.particle_id[i];
}

// ---------------------------------------------------------------------------
// ordered compaction: out = { i < n : pred(i) } ascending, count -> *total
struct PredFluid {
  const int32_t* material;
  const int32_t* n;
  __device__ bool operator()(int64_t i) const { return i < *n && material[i] == 1; }
};
struct PredSample {
  const uint8_t* is_sample;
  const int32_t* n;
  __device__ bool operator()(int64_t i) const { return i < *n && is_sample[i] == 1; }
};

template <class Pred>
__global__ void __launch_bounds__(256) k_compact_count(Pred pred, int64_t n, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t wsum[4];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  bool f = i < n && pred(i);
  uint64_t b = __ballot(f);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

template <class Pred>
__global__ void __launch_bounds__(256) k_compact_write(Pred pred, int64_t n, const uint32_t* __restrict__ off,
                                                       int32_t* __restrict__ out) {
  __shared__ uint32_t wsum[4];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  bool f = i < n && pred(i);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t b = __ballot(f);
  if (lane == 0) wsum[w] = (uint32_t)__popcll(b);
  __syncthreads();
  uint32_t pre = off[blockIdx.x];
  for (int ww = 0; ww < w; ++ww) pre += wsum[ww];
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (f) out[pre + (uint32_t)__popcll(b & lt)] = (int32_t)i;
}

template <class Pred>
static int compact(Pred pred, int64_t n, int32_t* out, int32_t* total, Workspace& ws, hipStream_t s) {
  unsigned nb = nblocks(n);
  if (nb == 0) {
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, total);
    DIMCV_LAUNCHED("k_zero_i32");
    return DIMCV_OK;
  }
  hipLaunchKernelGGL(k_compact_count<Pred>, dim3(nb), dim3(256), 0, s, pred, n, ws.compact_cnt);
  int rc = exclusive_scan_u32(ws.compact_cnt, ws.compact_cnt, nb, ws.scan_tmp, (uint32_t*)total, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_compact_write<Pred>, dim3(nb), dim3(256), 0, s, pred, n, ws.compact_cnt, out);
  DIMCV_LAUNCHED("compact");
  return DIMCV_OK;
}

int compact_samples(const dimcv_state* st, int64_t n, Workspace& ws, hipStream_t s) {
  PredSample pr{st->is_sample, st->counters + DIMCV_C_PARTICLE_NUM};
  return compact(pr, n, st->sample_idx, st->counters + DIMCV_C_NUM_SAMPLES, ws, s);
}

// ---------------------------------------------------------------------------
// for_all_neighbors (particle_system.py:562-586)
template <class F>
__device__ __forceinline__ void scan_cells(const dimcv_params& p, const float* __restrict__ x,
                                           const int32_t* __restrict__ gpn, int64_t i, v3 xi, F&& f) {
  int cc[3];
  cell_of(p, xi, cc);
  const int nx = p.grid_num[0], ny = p.grid_num[1], nz = p.grid_num[2];
  for (int o0 = -1; o0 <= 1; ++o0) {
    int g0 = cc[0] + o0;
    if (p.periodic) g0 = (g0 + nx) % nx;
    else if (g0 < 0 || g0 >= nx) continue;
    for (int o1 = -1; o1 <= 1; ++o1) {
      int g1 = cc[1] + o1;
      if (p.periodic) g1 = (g1 + ny) % ny;
      else if (g1 < 0 || g1 >= ny) continue;
      for (int o2 = -1; o2 <= 1; ++o2) {
        int g2 = cc[2] + o2;
        if (p.periodic) g2 = (g2 + nz) % nz;
        else if (g2 < 0 || g2 >= nz) continue;
        int fl = g0 * ny * nz + g1 * nz + g2;
        int beg = gpn[fl > 0 ? fl - 1 : 0];
        int end = gpn[fl];
        for (int j = beg; j < end; ++j) {
          if (j == i) continue;
          v3 xj = ld3(x, j);
          if (vnorm(rel(p, xi, xj)) < p.support_radius) f(j);
        }
      }
    }
  }
}

// Open grid: the three cells (c0+o0, c1+o1, c2-1..c2+1) are consecutive in the
// flattened order, so the 27 cell ranges collapse into 9 contiguous slot runs
// [gpn[max(0, first-1)], gpn[last]) visited in the same order; the run form
// keeps cell 0 empty exactly as the per-cell rule does.  The distance test
// sqrt_rn(r.r) < h is evaluated as r.r < nbr_r2_limit, which is equivalent
// for every f32 r.r (sqrt is monotone and correctly rounded).
// Candidates [beg, end) of one run, SCAN_U at a time: their positions are
// loaded (indices clamped into the run) before any is tested, so a lane
// waits for one memory latency per SCAN_U candidates rather than per
// candidate.  f(j, accepted) is called for every candidate in slot order
// (the emitter below records the accepted ones without a branch per
// candidate).  Entries carry the neighbour's fluid flag in their top bit;
// it is read from material when an entry is written to the list, so the
// candidate loads are the 12-byte positions alone.
#define SCAN_U 8
#define NBR_STAGE 31  // staged list rows per lane (the neighbour build below)
#define NBR_FLUID_BIT 0x80000000u
__device__ __forceinline__ uint32_t nbr_entry(const int32_t* __restrict__ material, uint32_t j) {
  return j | (material[j] == 1 ? NBR_FLUID_BIT : 0u);
}

// The acceptance test (j < end, r.r < limit, j != i) is formed from sign
// bits in vector registers, not as compares into lane masks: (j - end),
// (bits(r.r) - bits(limit)) (non-negative floats order as their bit
// patterns) and (d | -d) with d = j - i each have the sign bit set exactly
// when their condition holds, and none of the differences leaves 32 bits.
// Compares would each write a scalar mask that the next instruction waits
// for.
// Tests candidates j .. j + SCAN_U - 1 of a run ending at end (positions
// xj) and hands each to the emitter f in slot order.
template <class F>
__device__ __forceinline__ void test_trip(int j, int end, const v3* xj, int64_t i, v3 xi, uint32_t lim_bits,
                                          F&& f) {
  const uint32_t ui = (uint32_t)i, uend = (uint32_t)end;
  uint32_t ok[SCAN_U];
#pragma unroll
  for (int u = 0; u < SCAN_U; ++u) {
    const v3 r = vsub(xi, xj[u]);
    const float r2 = r.x * r.x + r.y * r.y + r.z * r.z;
    const uint32_t jj = (uint32_t)(j + u), d = jj - ui;
    ok[u] = ((jj - uend) & (__float_as_uint(r2) - lim_bits) & (d | (0u - d))) >> 31;
  }
  // While every lane of the wave has SCAN_U free stage rows left, the trip
  // writes row cnt directly (no clamp to the dump row, no direct list
  // write): one wave-wide test per trip instead of a branch per candidate.
  if (!__any(f.cnt > NBR_STAGE - SCAN_U)) {
#pragma unroll
    for (int u = 0; u < SCAN_U; ++u) f.staged(j + u, ok[u]);
  } else {
#pragma unroll
    for (int u = 0; u < SCAN_U; ++u) f(j + u, ok[u]);
  }
}

template <class F>
__device__ __forceinline__ void scan_run(const float* __restrict__ x, int beg, int end, int64_t i, v3 xi, float lim,
                                         F&& f) {
  for (int j = beg; j < end; j += SCAN_U) {
    v3 xj[SCAN_U];
#pragma unroll
    for (int u = 0; u < SCAN_U; ++u) xj[u] = ld3s(x, (uint32_t)min(j + u, end - 1));
    test_trip(j, end, xj, i, xi, __float_as_uint(lim), f);
  }
}

template <class F>
__device__ __forceinline__ void scan_runs(const dimcv_params& p, const float* __restrict__ x,
                                          const int32_t* __restrict__ gpn, int64_t i, v3 xi, F&& f) {
  int cc[3];
  cell_of(p, xi, cc);
  const int nx = p.grid_num[0], ny = p.grid_num[1], nz = p.grid_num[2];
  const int zlo = max(cc[2] - 1, 0), zhi = min(cc[2] + 1, nz - 1);
  const float lim = p.nbr_r2_limit;
  for (int o0 = -1; o0 <= 1; ++o0) {
    int g0 = cc[0] + o0;
    if (g0 < 0 || g0 >= nx) continue;
    for (int o1 = -1; o1 <= 1; ++o1) {
      int g1 = cc[1] + o1;
      if (g1 < 0 || g1 >= ny) continue;
      int fl = (g0 * ny + g1) * nz;
      int beg = gpn[max(fl + zlo - 1, 0)];
      int end = gpn[fl + zhi];
      scan_run(x, beg, end, i, xi, lim, f);
    }
  }
}

// Neighbour lists are k-major (nbr[k * n_max + f]) so that the passes read
// 64 consecutive entries per wave and row.  Lanes find their k-th neighbour
// at different times, so writing entries as found scatters 4-byte stores
// over many rows and partially written lines leave L2 before they fill.  The
// first NBR_STAGE entries of each lane are therefore staged in LDS and
// written row by row by the whole wave (one coalesced 256-byte row store);
// entries past NBR_STAGE are written directly.
//
// Every candidate is stored to the stage, accepted or not, at row cnt (which
// only an accepted candidate advances) or, from cnt = NBR_STAGE on, at the
// dump row NBR_STAGE: a rejected candidate's word is overwritten by the next
// accepted one or lies past the list, so recording costs a few vector
// instructions per candidate and no branch (only the rare direct write of an
// entry past the stage branches).

struct NbrEmitter {
  uint32_t (*stage)[64];  // the wave's NBR_STAGE + 1 rows
  uint32_t* __restrict__ nbr;
  int64_t stride;
  int64_t f;
  int kmax, lane, cnt;
  const int32_t* __restrict__ material;
  // ok is 0 or 1.  staged(): only while cnt + SCAN_U <= NBR_STAGE.  The
  // stage holds slots; nbr_flush adds the fluid flags.
  __device__ __forceinline__ void staged(int j, uint32_t ok) {
    stage[cnt][lane] = (uint32_t)j;
    cnt += (int)ok;
  }
  __device__ __forceinline__ void operator()(int j, uint32_t ok) {
    stage[min(cnt, NBR_STAGE)][lane] = (uint32_t)j;
    const uint32_t direct = ok & ((uint32_t)(NBR_STAGE - 1 - cnt) >> 31) & ((uint32_t)(cnt - kmax) >> 31);
    if (direct) nbr[(int64_t)cnt * stride + f] = nbr_entry(material, (uint32_t)j);
    cnt += (int)ok;
  }
};

// Raises counters[DIMCV_C_NBR_MAX] to the wave's longest list: one atomic per
// wave (its lane 0) instead of one per lane on the same word.  Called by every
// lane of the wave (cnt = 0 for lanes without an entry).
__device__ __forceinline__ void nbr_wave_max(dimcv_state& st, int lane, int cnt) {
  int m = cnt;
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  if (lane == 0 && m > 0) atomicMax(&st.counters[DIMCV_C_NBR_MAX], m);
}

__device__ __forceinline__ void nbr_flush(dimcv_state& st, const dimcv_params& p, uint32_t (*stage)[64], int lane,
                                          int64_t f, int cnt) {
  nbr_wave_max(st, lane, cnt);
  const int c = min(cnt, min(NBR_STAGE, p.nbr_kmax));
  int cw = c;
  for (int o = 32; o > 0; o >>= 1) cw = max(cw, __shfl_xor(cw, o, 64));
  const int64_t stride = p.n_max;
  for (int k = 0; k < cw; ++k)
    if (k < c) st.nbr[(int64_t)k * stride + f] = nbr_entry(st.material, stage[k][lane]);
}

__global__ void __launch_bounds__(256) k_nbr_build(dimcv_state st, dimcv_params p, int32_t bound) {
  __shared__ uint32_t stage[4][NBR_STAGE + 1][64];
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  NbrEmitter emit{stage[w], st.nbr, p.n_max, f, p.nbr_kmax, lane, 0, st.material};
  if (f < bound && f < st.counters[DIMCV_C_FLUID_LIST]) {
    const int64_t i = st.fluid_list[f];
    const v3 xi = ld3(st.x, i);
    if (p.periodic)
      scan_cells(p, st.x, st.grid_particles_num, i, xi,
                 [&](int j) { emit(j, 1u); });
    else scan_runs(p, st.x, st.grid_particles_num, i, xi, emit);
    st.nbr_count[f] = emit.cnt;
  }
  nbr_flush(st, p, stage[w], lane, f, emit.cnt);
}

// dimcv_sph.py:536-549: the pressure update's reaction on a dynamic rigid
// neighbour j of fluid particle i, acceleration[j] += -vel_change inv_dt
// rho_i / rho_j with vel_change = -dt k_i m_V,j rho0 gradW(x_i - x_j) for
// |k_i| > m_eps.  Gathered per rigid particle from its fluid neighbours in
// ascending slot order, the order of the reference's serial adds, onto the
// value the acceleration holds; no list holds cell 0, but a fluid particle in
// cell 0 still visits j, so cell 0's slots (the lowest) are scanned first when
// j's cell borders it.  A rigid particle in cell 0 is visited by nobody.
__global__ void k_rigid_reaction(dimcv_state st, dimcv_params p, int32_t bound, const int32_t* done) {
  if (jacobi_gated(st, done)) return;
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= bound || j >= st.counters[DIMCV_C_PARTICLE_NUM]) return;
  if (st.material[j] != 0 || !st.is_dynamic[j]) return;
  const v3 xj = ld3(st.x, j);
  int cc[3];
  cell_of(p, xj, cc);
  const int nx = p.grid_num[0], ny = p.grid_num[1], nz = p.grid_num[2];
  if (((int64_t)cc[0] * ny + cc[1]) * nz + cc[2] == 0) return;
  const float mvj = st.m_V[j], rho_j = st.density[j];
  v3 acc = ld3(st.acceleration, j);
  auto visit = [&](int i) {
    if (st.material[i] != 1) return;
    const float k_i = (st.density_adv[i] - p.density0) * st.dfsph_factor[i];
    if (fabsf(k_i) > p.m_eps) {
      const v3 g = smul(mvj * p.density0, cubic_dW(p, rel(p, ld3(st.x, i), xj)));
      const v3 vc = smul(-p.dt * 1.0f * k_i, g);
      const float rho_i = st.density[i];
      acc = vadd(acc, mk(((-vc.x) * p.inv_dt) * rho_i / rho_j, ((-vc.y) * p.inv_dt) * rho_i / rho_j,
                         ((-vc.z) * p.inv_dt) * rho_i / rho_j));
    }
  };
  bool near0 = true;
  const int lim[3] = {nx, ny, nz};
  for (int d = 0; d < 3; ++d) near0 = near0 && (cc[d] <= 1 || (p.periodic && cc[d] == lim[d] - 1));
  if (near0) {
    const int32_t end0 = st.grid_particles_num[0];
    for (int32_t i = 0; i < end0; ++i)
      if (i != j && vnorm(rel(p, xj, ld3(st.x, i))) < p.support_radius) visit(i);
  }
  scan_cells(p, st.x, st.grid_particles_num, j, xj, visit);
  st3(st.acceleration, j, acc);
}

int launch_rigid_reaction(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, const int32_t* done,
                          hipStream_t s) {
  if (!p->dyn_rigid || particle_bound <= 0) return DIMCV_OK;
  hipLaunchKernelGGL(k_rigid_reaction, dim3(nblocks(particle_bound)), dim3(256), 0, s, *st, *p, particle_bound, done);
  DIMCV_LAUNCHED("k_rigid_reaction");
  return DIMCV_OK;
}

// compute_com / solve_constraints (sph_base.py:183-222) of one dynamic rigid
// object: FP64 moments over its particles, per block with block_tree_sum256
// into part[b * 9 + c], then summed over blocks in block order (one block).
// mode 0: (sum m, sum m x), m = m_V0 density (f32); mode 1: sum m p q^T
// (row-major), p = x - cm, q = x_0 - rest_cm.
__global__ void __launch_bounds__(256) k_rigid_moments(dimcv_state st, dimcv_params p, int32_t bound,
                                                        int32_t obj, int mode, v3 cm, v3 rest,
                                                        double* __restrict__ part) {
  __shared__ double lds[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (i < bound && i < st.counters[DIMCV_C_PARTICLE_NUM] && st.material[i] == 0 && st.is_dynamic[i] &&
      st.object_id[i] == obj) {
    const float m = p.m_V0 * st.density[i];
    const v3 x = ld3(st.x, i);
    if (mode == 0) {
      c[0] = m;
      c[1] = (double)m * x.x;
      c[2] = (double)m * x.y;
      c[3] = (double)m * x.z;
    } else {
      const v3 pp = vsub(x, cm), q = vsub(ld3(st.x_0, i), rest);
      const float a[3] = {pp.x, pp.y, pp.z}, b[3] = {q.x, q.y, q.z};
      for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) c[3 * r + k] = (double)m * a[r] * b[k];
    }
  }
  const int nc = mode == 0 ? 4 : 9;
  for (int k = 0; k < nc; ++k) {
    const double sum = block_tree_sum256(c[k], lds);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * 9 + k] = sum;
  }
}

__global__ void __launch_bounds__(256) k_rigid_moments_sum(const double* __restrict__ part, int64_t nb,
                                                            double* __restrict__ out) {
  __shared__ double lds[256];
  for (int k = 0; k < 9; ++k) {
    double acc = 0.0;
    for (int64_t b0 = 0; b0 < nb; b0 += 256) {
      const int64_t b = b0 + threadIdx.x;
      const double s = block_tree_sum256(b < nb ? part[b * 9 + k] : 0.0, lds);
      acc += s;  // thread 0's running total over groups, in block order
    }
    if (threadIdx.x == 0) out[k] = acc;
  }
}

// sph_base.py:215-220: x += (goal - x) * 1.0, goal = cm + R (x_0 - rest_cm)
// (each component of R q summed over k = 0, 1, 2 in f32).
struct RigidT {
  float cm[3], R[9], rest[3];
};
__global__ void k_rigid_apply(dimcv_state st, int32_t bound, int32_t obj, RigidT t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bound || i >= st.counters[DIMCV_C_PARTICLE_NUM]) return;
  if (st.material[i] != 0 || !st.is_dynamic[i] || st.object_id[i] != obj) return;
  const v3 q = vsub(ld3(st.x_0, i), mk(t.rest[0], t.rest[1], t.rest[2]));
  const v3 rq = mk(t.R[0] * q.x + t.R[1] * q.y + t.R[2] * q.z, t.R[3] * q.x + t.R[4] * q.y + t.R[5] * q.z,
                   t.R[6] * q.x + t.R[7] * q.y + t.R[8] * q.z);
  const v3 goal = vadd(mk(t.cm[0], t.cm[1], t.cm[2]), rq);
  const v3 x = ld3(st.x, i);
  const v3 corr = smul(1.0f, vsub(goal, x));
  st3(st.x, i, vadd(x, corr));
}

// sph_base.py:91-112: Akinci boundary volume over solid neighbours.
__global__ void k_boundary_volume(dimcv_state st, dimcv_params p, int32_t bound, int dynamic_sel) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bound || i >= st.counters[DIMCV_C_PARTICLE_NUM]) return;
  if (st.material[i] != 0 || (st.is_dynamic[i] != 0) != (dynamic_sel != 0)) return;
  v3 xi = ld3(st.x, i);
  float delta = cubic_W(p, 0.0f);
  scan_cells(p, st.x, st.grid_particles_num, i, xi, [&](int j) {
    if (st.material[j] == 0) delta += cubic_W(p, vnorm(rel(p, xi, ld3(st.x, j))));
  });
  st.m_V[i] = 1.0f / delta * 3.0f;
}

// Stable radix sort of (key, slot) over the key bits in use (sentinel =
// n_cells), the cell table and the gather into the *_buffer fields.  Keys
// and slots are in ws.keys[0] / ws.vals[0].
int sort_cells_gather(const dimcv_state* st, const dimcv_params* p, Workspace& ws, int64_t n_in, hipStream_t s) {
  // key bits in use (the sentinel n_cells included); 9-bit digits where
  // they need fewer passes than 8-bit ones (18 bits: 2 passes, not 3)
  const int nbits = 32 - __builtin_clz((unsigned)p->n_cells);
  const int bits = (nbits + 8) / 9 < (nbits + 7) / 8 ? 9 : 8;
  const int passes = (nbits + bits - 1) / bits;
  int cur = 0;
  if (n_in > 0) {
    int nblk = (int)((n_in + DIMCV_SORT_TILE - 1) / DIMCV_SORT_TILE);
    for (int ps = 0; ps < passes; ++ps) {
      int shift = bits * ps;
      if (bits == 9)
        hipLaunchKernelGGL(k_radix_hist<9>, dim3(nblk), dim3(256), 0, s, ws.keys[cur], n_in, shift, ws.radix_hist,
                           nblk);
      else
        hipLaunchKernelGGL(k_radix_hist<8>, dim3(nblk), dim3(256), 0, s, ws.keys[cur], n_in, shift, ws.radix_hist,
                           nblk);
      int rc = exclusive_scan_u32(ws.radix_hist, ws.radix_hist, (int64_t)(1 << bits) * nblk, ws.scan_tmp, nullptr,
                                  s);
      if (rc) return rc;
      if (bits == 9)
        hipLaunchKernelGGL(k_radix_scatter<9>, dim3(nblk), dim3(256), 0, s, ws.keys[cur], ws.vals[cur],
                           ws.keys[cur ^ 1], ws.vals[cur ^ 1], n_in, shift, ws.radix_hist, nblk);
      else
        hipLaunchKernelGGL(k_radix_scatter<8>, dim3(nblk), dim3(256), 0, s, ws.keys[cur], ws.vals[cur],
                           ws.keys[cur ^ 1], ws.vals[cur ^ 1], n_in, shift, ws.radix_hist, nblk);
      DIMCV_LAUNCHED("k_radix_scatter");
      cur ^= 1;
    }
  }
  hipLaunchKernelGGL(k_cell_end, dim3(nblocks(p->n_cells)), dim3(256), 0, s, ws.keys[cur], n_in,
                     p->n_cells, st->grid_particles_num, st->grid_particles_num_temp);
  if (n_in > 0)
    hipLaunchKernelGGL(k_gather, dim3(nblocks(n_in)), dim3(256), 0, s, *st, n_in, ws.keys[cur],
                       ws.vals[cur]);
  DIMCV_LAUNCHED("k_gather");
  return DIMCV_OK;
}

// ---------------------------------------------------------------------------
// Batched scenes: keys with per-scene boxes, grids and cell bases; neighbour
// lists over each scene's own cells.

// update_activity + update_grid_id of the slot's scene (open box, fluid
// only).  Inactive slots get the sentinel key (past every scene's cells).
__global__ void k_batch_keys(dimcv_state st, dimcv_params p, int64_t n_in, uint32_t* __restrict__ keys,
                             uint32_t* __restrict__ vals) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_in) return;
  const uint32_t sentinel = (uint32_t)p.n_cells;
  vals[i] = (uint32_t)i;
  if (i >= st.counters[DIMCV_C_PARTICLE_NUM]) {
    keys[i] = sentinel;
    return;
  }
  const dimcv_scene& sc = st.scenes[st.scene_id[i]];
  const int32_t mat = st.material[i];
  const v3 xi = ld3(st.x, i);
  uint8_t active = st.is_active[i];
  if (mat == 0) {
    active = 1;
  } else if (mat == 1) {
    // batched scenes have no emitters (SceneBatch rejects them), so unlike
    // k_activity_keys no per-emitter count is decremented here
    const float q[3] = {xi.x, xi.y, xi.z};
    for (int j = 0; j < 3; ++j) {
      if (q[j] < sc.domain_start[j] || q[j] > sc.domain_end[j]) {
        active = 0;
        atomicSub(&st.counters[DIMCV_C_FLUID_NUM], 1);
        break;
      }
    }
  }
  // the scene's flat index, with the solo step's rule at the upper faces
  // (k_activity_keys), so a batched scene keys as a solo one
  int c[3];
  c[0] = (int)(xi.x / p.grid_size);
  c[1] = (int)(xi.y / p.grid_size);
  c[2] = (int)(xi.z / p.grid_size);
  const int64_t g = ((int64_t)c[0] * sc.grid_num[1] + c[1]) * sc.grid_num[2] + c[2];
  const bool in_table = g >= 0 && g < (int64_t)sc.n_cells;
  if (active && !in_table && mat == 1) {
    active = 0;
    atomicSub(&st.counters[DIMCV_C_FLUID_NUM], 1);
  }
  st.is_active[i] = active;
  if (!active) {
    keys[i] = sentinel;
    return;
  }
  if (!in_table) {
    atomicAdd(&st.counters[DIMCV_C_KEY_ERROR], 1);
    keys[i] = sentinel;
    return;
  }
  keys[i] = (uint32_t)(sc.cell_base + g);
}

// Neighbour lists of fluid-list entry q over the 9 z-runs of its scene's
// grid; the scene's own cell 0 stays invisible (its run starts at the end of
// that cell), as in a solo build.  Entries are staged as in k_nbr_build
// (workgroups of at most 256 threads).
__global__ void __launch_bounds__(256) k_batch_nbr_build(dimcv_state st, dimcv_params p, int64_t bound) {
  __shared__ uint32_t stage[4][NBR_STAGE + 1][64];
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  NbrEmitter emit{stage[w], st.nbr, p.n_max, f, p.nbr_kmax, lane, 0, st.material};
  const int64_t i = (f < bound && f < st.counters[DIMCV_C_FLUID_LIST]) ? st.fluid_list[f] : -2;
  if (i == -1) st.nbr_count[f] = 0;
  if (i >= 0) {
    const dimcv_scene& sc = st.scenes[st.scene_id[i]];
    const v3 xi = ld3(st.x, i);
    int cc[3];
    cc[0] = (int)(xi.x / p.grid_size);
    cc[1] = (int)(xi.y / p.grid_size);
    cc[2] = (int)(xi.z / p.grid_size);
    const int nx = sc.grid_num[0], ny = sc.grid_num[1], nz = sc.grid_num[2];
    const int32_t* __restrict__ gpn = st.grid_particles_num + sc.cell_base;
    const int zlo = max(cc[2] - 1, 0), zhi = min(cc[2] + 1, nz - 1);
    const float lim = p.nbr_r2_limit;
    for (int o0 = -1; o0 <= 1; ++o0) {
      const int g0 = cc[0] + o0;
      if (g0 < 0 || g0 >= nx) continue;
      for (int o1 = -1; o1 <= 1; ++o1) {
        const int g1 = cc[1] + o1;
        if (g1 < 0 || g1 >= ny) continue;
        const int fl = (g0 * ny + g1) * nz;
        const int lo = fl + zlo - 1;
        const int beg = lo >= 0 ? gpn[lo] : gpn[0];
        const int end = gpn[fl + zhi];
        scan_run(st.x, beg, end, i, xi, lim, emit);
      }
    }
    st.nbr_count[f] = emit.cnt;
  }
  nbr_flush(st, p, stage[w], lane, f, emit.cnt);
}

static int g_nbr_block = 256;

void batch_set_nbr_block(int b) { g_nbr_block = b; }

// Activity + keys, sort, cell table, gather (state before the swap) and the
// unpadded fluid list (ascending slots) into ulist / *ucount.
int batch_grid_launch(const dimcv_state* st, const dimcv_params* p, Workspace& ws, int32_t* ulist,
                      int32_t* ucount, hipStream_t s) {
  const int64_t n_in = p->n_max;
  hipLaunchKernelGGL(k_batch_keys, dim3(nblocks(n_in)), dim3(256), 0, s, *st, *p, n_in, ws.keys[0], ws.vals[0]);
  hipLaunchKernelGGL(k_set_particle_num, dim3(1), dim3(1), 0, s, st->counters, (int32_t)n_in);
  DIMCV_LAUNCHED("k_batch_keys");
  int rc = sort_cells_gather(st, p, ws, n_in, s);
  if (rc) return rc;
  PredFluid pf{st->material_buffer, st->counters + DIMCV_C_PARTICLE_NUM};
  return compact(pf, n_in, ulist, ucount, ws, s);
}

// Neighbour lists of the swapped state (padded fluid list in place).
int batch_nbr_launch(const dimcv_state* st, const dimcv_params* p, hipStream_t s) {
  const int b = g_nbr_block;
  hipLaunchKernelGGL(k_batch_nbr_build, dim3(nblocks(p->n_max, b)), dim3(b), 0, s, *st, *p, (int64_t)p->n_max);
  DIMCV_LAUNCHED("k_batch_nbr_build");
  return DIMCV_OK;
}

// The list build without the width check (the caller reads DIMCV_C_NBR_MAX).
int neighbor_build_launch(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, hipStream_t s) {
  if (p->n_max > DIMCV_MAX_SLOTS) {
    set_error("n_max %lld exceeds DIMCV_MAX_SLOTS (%d)", (long long)p->n_max, DIMCV_MAX_SLOTS);
    return DIMCV_ERR_ARG;
  }
  if (fluid_bound > 0)
    hipLaunchKernelGGL(k_nbr_build, dim3(nblocks(fluid_bound)), dim3(256), 0, s, *st, *p, fluid_bound);
  DIMCV_LAUNCHED("k_nbr_build");
  return DIMCV_OK;
}

}  // namespace dimcv

using namespace dimcv;

extern "C" {

int dimcv_add_particles(const dimcv_state* st, const dimcv_params* p, int32_t object_id, int32_t n,
                        const float* pos, const float* vel, const float* density,
                        const float* pressure, const int32_t* material, const int32_t* is_dynamic,
                        const int32_t* color, int32_t* dropped, void* stream) {
  if (!st || !p || n < 0) {
    set_error("dimcv_add_particles: bad arguments");
    return DIMCV_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  if (n > 0)
    hipLaunchKernelGGL(k_add_particles, dim3(nblocks(n)), dim3(256), 0, s, *st, *p, object_id, n, pos,
                       vel, density, pressure, material, is_dynamic, color);
  int32_t* dev_drop = (int32_t*)ws.compact_cnt;
  hipLaunchKernelGGL(k_add_finish, dim3(1), dim3(1), 0, s, st->counters, p->n_max, n, dev_drop);
  DIMCV_LAUNCHED("k_add_particles");
  if (dropped) {
    DIMCV_HIP(hipMemcpyAsync(dropped, dev_drop, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    DIMCV_HIP(hipStreamSynchronize(s));
  }
  return DIMCV_OK;
}

int dimcv_emit(const dimcv_state* st, const dimcv_params* p, const dimcv_emitter* emitters, int32_t n_emitters,
               void* stream) {
  if (!st || !p || (n_emitters > 0 && !emitters) || n_emitters < 0) {
    set_error("dimcv_emit: bad arguments");
    return DIMCV_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  for (int32_t e = 0; e < n_emitters; ++e) {
    const dimcv_emitter& em = emitters[e];
    if (em.n < 0 || (em.n > 0 && !em.pos)) {
      set_error("dimcv_emit: emitter %d has no lattice", e);
      return DIMCV_ERR_ARG;
    }
    if (em.n > 0) hipLaunchKernelGGL(k_emit, dim3(nblocks(em.n)), dim3(256), 0, s, *st, *p, em);
    hipLaunchKernelGGL(k_emit_finish, dim3(1), dim3(1), 0, s, *st, p->n_max, em.n, em.object_id);
  }
  DIMCV_LAUNCHED("k_emit");
  return DIMCV_OK;
}

int dimcv_init_cylinder(const dimcv_state* st, const dimcv_params* p, int32_t object_id, float cx,
                        float cy, float radius, void* stream) {
  (void)p;
  int32_t bound = (int32_t)p->n_max;
  hipLaunchKernelGGL(k_init_cylinder, dim3(nblocks(bound)), dim3(256), 0, (hipStream_t)stream, *st,
                     bound, object_id, cx, cy, radius);
  DIMCV_LAUNCHED("k_init_cylinder");
  return DIMCV_OK;
}

int dimcv_grid_build(const dimcv_state* st, const dimcv_params* p, int32_t prev_particle_num,
                     int32_t* host_counters, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  if (prev_particle_num < 0 || prev_particle_num > p->n_max || p->n_cells <= 0) {
    set_error("dimcv_grid_build: particle bound %d outside [0, %lld]", prev_particle_num,
              (long long)p->n_max);
    return DIMCV_ERR_ARG;
  }
  if (p->periodic && (p->grid_num[0] < 3 || p->grid_num[1] < 3 || p->grid_num[2] < 3)) {
    set_error("periodic boxes need at least 3 cells per axis");
    return DIMCV_ERR_ARG;
  }
  const int64_t n_in = prev_particle_num;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, st->counters + DIMCV_C_KEY_ERROR);
  if (n_in > 0)
    hipLaunchKernelGGL(k_activity_keys, dim3(nblocks(n_in)), dim3(256), 0, s, *st, *p, n_in, ws.keys[0],
                       ws.vals[0]);
  hipLaunchKernelGGL(k_set_particle_num, dim3(1), dim3(1), 0, s, st->counters, (int32_t)n_in);
  DIMCV_LAUNCHED("k_activity_keys");

  rc = sort_cells_gather(st, p, ws, n_in, s);
  if (rc) return rc;
  PredFluid pf{st->material_buffer, st->counters + DIMCV_C_PARTICLE_NUM};
  rc = compact(pf, n_in, st->fluid_list, st->counters + DIMCV_C_FLUID_LIST, ws, s);
  if (rc) return rc;
  if (host_counters) {
    DIMCV_HIP(hipMemcpyAsync(host_counters, st->counters, sizeof(int32_t) * DIMCV_C_COUNT,
                             hipMemcpyDeviceToHost, s));
    DIMCV_HIP(hipStreamSynchronize(s));
    if (host_counters[DIMCV_C_KEY_ERROR] != 0) {
      set_error("%d active solid particles lie past the cell table", host_counters[DIMCV_C_KEY_ERROR]);
      return DIMCV_ERR_KEY;
    }
  }
  return DIMCV_OK;
}


int dimcv_neighbor_build(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                         void* stream) {
  hipStream_t s = (hipStream_t)stream;
  int rc = dimcv::neighbor_build_launch(st, p, fluid_bound, s);
  if (rc) return rc;
  int32_t mx = 0;
  DIMCV_HIP(hipMemcpyAsync(&mx, st->counters + DIMCV_C_NBR_MAX, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipStreamSynchronize(s));
  if (mx > p->nbr_kmax) {
    set_error("neighbour list width %d exceeds nbr_kmax %d", mx, p->nbr_kmax);
    return DIMCV_ERR_NBR;
  }
  return DIMCV_OK;
}

int dimcv_compute_static_boundary_volume(const dimcv_state* st, const dimcv_params* p,
                                         int32_t particle_bound, void* stream) {
  if (particle_bound > 0)
    hipLaunchKernelGGL(k_boundary_volume, dim3(nblocks(particle_bound)), dim3(256), 0, (hipStream_t)stream,
                       *st, *p, particle_bound, 0);
  DIMCV_LAUNCHED("k_boundary_volume");
  return DIMCV_OK;
}

int dimcv_compute_moving_boundary_volume(const dimcv_state* st, const dimcv_params* p,
                                         int32_t particle_bound, void* stream) {
  if (particle_bound > 0)
    hipLaunchKernelGGL(k_boundary_volume, dim3(nblocks(particle_bound)), dim3(256), 0, (hipStream_t)stream,
                       *st, *p, particle_bound, 1);
  DIMCV_LAUNCHED("k_boundary_volume");
  return DIMCV_OK;
}

int dimcv_rigid_moments(const dimcv_state* st, const dimcv_params* p, int32_t object_id, int32_t mode,
                        const float* cm, const float* rest_cm, int32_t particle_bound, double* out, void* stream) {
  if (!st || !p || !out || (mode != 0 && mode != 1) || (mode == 1 && (!cm || !rest_cm))) {
    set_error("dimcv_rigid_moments: bad arguments");
    return DIMCV_ERR_ARG;
  }
  for (int k = 0; k < 9; ++k) out[k] = 0.0;
  if (particle_bound <= 0) return DIMCV_OK;
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  const int64_t nb = nblocks(particle_bound);
  // block partials (9 doubles per block) and the result, in the far-field
  // partial region (at least min(16 n_max, 2^21) x 32 bytes)
  if ((nb * 9 + 9) * 8 > ws.far_partial_cap * 32) {
    set_error("dimcv_rigid_moments: workspace too small");
    return DIMCV_ERR_WORKSPACE;
  }
  double* part = (double*)ws.far_partials;
  double* dout = part + nb * 9;
  const v3 c = mode == 1 ? v3{cm[0], cm[1], cm[2]} : v3{0.0f, 0.0f, 0.0f};
  const v3 r = mode == 1 ? v3{rest_cm[0], rest_cm[1], rest_cm[2]} : v3{0.0f, 0.0f, 0.0f};
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_rigid_moments, dim3((unsigned)nb), dim3(256), 0, s, *st, *p, particle_bound, object_id, mode,
                     c, r, part);
  hipLaunchKernelGGL(k_rigid_moments_sum, dim3(1), dim3(256), 0, s, (const double*)part, nb, dout);
  DIMCV_LAUNCHED("k_rigid_moments");
  DIMCV_HIP(hipMemcpyAsync(out, dout, 9 * sizeof(double), hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipStreamSynchronize(s));
  return DIMCV_OK;
}

int dimcv_rigid_apply(const dimcv_state* st, const dimcv_params* p, int32_t object_id, const float* T,
                      int32_t particle_bound, void* stream) {
  if (!st || !p || !T) {
    set_error("dimcv_rigid_apply: null argument");
    return DIMCV_ERR_ARG;
  }
  if (particle_bound <= 0) return DIMCV_OK;
  RigidT t;
  for (int k = 0; k < 3; ++k) t.cm[k] = T[k];
  for (int k = 0; k < 9; ++k) t.R[k] = T[3 + k];
  for (int k = 0; k < 3; ++k) t.rest[k] = T[12 + k];
  hipLaunchKernelGGL(k_rigid_apply, dim3(nblocks(particle_bound)), dim3(256), 0, (hipStream_t)stream, *st,
                     particle_bound, object_id, t);
  DIMCV_LAUNCHED("k_rigid_apply");
  return DIMCV_OK;
}

}  // extern "C"
