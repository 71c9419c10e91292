// SPH / DFSPH / DIMCV neighbour passes over the per-step neighbour lists.
//
// One thread per fluid-list entry f (fluid slot i = fluid_list[f]); the
// neighbours are read from the k-major list nbr[k * n_max + f] in the order
// for_all_neighbors visits them (particle_system.py:562-586), so every
// accumulation below is the reference's serial sum, term by term, in the
// reference's expression order (dimcv_sph.py, sph_base.py).
#include <algorithm>
#include <mutex>

#include "common.h"

namespace dimcv {

#define FLUID_ENTRY() FLUID_ENTRY_LB(xcd_block())
// The same for logical block lb of 256 fluid-list entries (persistent loops).
#define FLUID_ENTRY_LB(lb)                                                  \
  const int64_t f = (lb) * 256 + threadIdx.x;                              \
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;          \
  const int64_t i = st.fluid_list[f];                                      \
  if (i < 0) return;                                                       \
  const int nn = min(st.nbr_count[f], p.nbr_kmax);                         \
  const NbrList nl(st.nbr, f);                            \
  const int64_t ns = p.n_max

// (x, m_V) of neighbour slot j: with XM from the step's packed 16-byte
// records (one aligned load), otherwise from the separate fields.
template <bool XM>
__device__ __forceinline__ float4 nb_xmv(const dimcv_state& st, const float4* __restrict__ xm, uint32_t j) {
  if constexpr (XM) return *(const float4*)((const char*)xm + j * 16u);
  const v3 a = ld3s(st.x, j);
  return make_float4(a.x, a.y, a.z, ld1s(st.m_V, j));
}
__device__ __forceinline__ v3 xyz(const float4& q) { return mk(q.x, q.y, q.z); }

// (x, w) of neighbour entry e (slot j), w = m for fluid and m_V rho0 for
// solids, the mass factor of the pressure passes: with XM from the packed
// records.
template <bool XM>
__device__ __forceinline__ float4 nb_xw(const dimcv_state& st, const dimcv_params& p, const float4* __restrict__ xw,
                                        uint32_t j, uint32_t e) {
  if constexpr (XM) return *(const float4*)((const char*)xw + j * 16u);
  const v3 a = ld3s(st.x, j);
  return make_float4(a.x, a.y, a.z, nb_fluid(e) ? ld1s(st.m, j) : ld1s(st.m_V, j) * p.density0);
}


// Zero-fill of slots that a "field.fill(0) + fluid writes" kernel leaves at
// zero: every slot >= particle_num or not fluid.
// Up to five per-slot arrays (null = unused) and their floats per slot.
struct ZeroSet {
  float* a[5];
  int c[5];
};

__global__ void k_zero_nonfluid(const int32_t* __restrict__ material, const int32_t* __restrict__ counters,
                                int64_t n_max, ZeroSet z) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_max) return;
  if (s < counters[DIMCV_C_PARTICLE_NUM] && material[s] == 1) return;
#pragma unroll
  for (int q = 0; q < 5; ++q)
    if (z.a[q])
      for (int k = 0; k < z.c[q]; ++k) z.a[q][s * z.c[q] + k] = 0.0f;
}

static int zero_nonfluid(const dimcv_state* st, const dimcv_params* p, hipStream_t s, const ZeroSet& z) {
  hipLaunchKernelGGL(k_zero_nonfluid, dim3(nblocks(p->n_max)), dim3(256), 0, s, st->material, st->counters,
                     p->n_max, z);
  DIMCV_LAUNCHED("k_zero_nonfluid");
  return DIMCV_OK;
}
static int zero_nonfluid(const dimcv_state* st, const dimcv_params* p, hipStream_t s, float* a, int ca,
                         float* b = nullptr, int cb = 0, float* c = nullptr, int cc = 0) {
  return zero_nonfluid(st, p, s, ZeroSet{{a, b, c, nullptr, nullptr}, {ca, cb, cc, 0, 0}});
}

// dimcv_sph.py:126-148
__global__ void __launch_bounds__(256) k_densities(dimcv_state st, dimcv_params p, int32_t bound) {
  FLUID_ENTRY();
  v3 xi = ld3(st.x, i);
  float dens = st.m_V[i] * cubic_W(p, 0.0f);
  float den = 0.0f;
  FOR_NBR(k, e, j) { den += st.m_V[j] * cubic_W(p, vnorm(rel(p, xi, ld3(st.x, j)))); }
  dens += den;
  dens *= p.density0;
  st.density[i] = dens;
}

// compute_densities (dimcv_sph.py:126-148) and compute_DFSPH_factor (:86-124)
// in one neighbour loop: both read only geometry and masses, and each sum is
// accumulated exactly as in its own kernel.  CHANGE also accumulates the
// first compute_density_change of the divergence solve (:337-374), which
// reads only x and v and the same kernel gradients.
template <bool CHANGE, bool XM>
__global__ void __launch_bounds__(256) k_density_factor(dimcv_state st, dimcv_params p, int32_t bound,
                                                        Packed pk) {
  FLUID_ENTRY();
  const bool sd = XM && *pk.tiny == 0;  // shared-reciprocal divisions in cubic_dW
  v3 xi = ld3(st.x, i);
  const v3 vi = CHANGE ? ld3(st.v, i) : mk(0.0f, 0.0f, 0.0f);
  float dens = st.m_V[i] * cubic_W(p, 0.0f);
  float den = 0.0f;
  float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
  float dadv = 0.0f;
  struct G { float4 xmv; v3 v; float m; };
  nbr_loop(nl, ns, nn,
           [&](uint32_t j) {
             return G{nb_xmv<XM>(st, pk.xm, j), CHANGE ? ld3s(st.v, j) : mk(0.0f, 0.0f, 0.0f), ld1s(st.m, j)};
           },
           [&](int, uint32_t e, const G& g) {
             const float mV = g.xmv.w;
             v3 r = rel(p, xi, xyz(g.xmv));
             den += mV * cubic_W(p, vnorm(r));
             v3 gw = cubic_dW(p, r, sd);
             if (CHANGE) dadv += mV * vdot(vsub(vi, g.v), gw);
             if (nb_fluid(e)) {
               v3 gg = smul(g.m, gw);
               r3 += vdot(gg, gg);
               r0 += gg.x;
               r1 += gg.y;
               r2 += gg.z;
             } else {
               v3 gg = smul(mV * p.density0, gw);
               r0 += gg.x;
               r1 += gg.y;
               r2 += gg.z;
             }
           });
  dens += den;
  dens *= p.density0;
  st.density[i] = dens;
  float sum = r3;
  v3 gp = mk(r0, r1, r2);
  sum += vdot(gp, gp);
  float factor = 0.0f;
  if (sum > 1e-6f) factor = 1.0f / sum;
  st.dfsph_factor[i] = factor;
  if (CHANGE) {
    float d = fmaxf(dadv, 0.0f);
    if (st.nbr_count[f] < 20) d = 0.0f;
    st.density_adv[i] = d;
  }
}

// dimcv_sph.py:86-124
__global__ void __launch_bounds__(256) k_dfsph_factor(dimcv_state st, dimcv_params p, int32_t bound) {
  FLUID_ENTRY();
  v3 xi = ld3(st.x, i);
  float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
  FOR_NBR(k, e, j) {
    v3 gw = cubic_dW(p, rel(p, xi, ld3(st.x, j)));
    if (nb_fluid(e)) {
      v3 g = smul(st.m[j], gw);
      r3 += vdot(g, g);
      r0 += g.x;
      r1 += g.y;
      r2 += g.z;
    } else {
      v3 g = smul(st.m_V[j] * p.density0, gw);
      r0 += g.x;
      r1 += g.y;
      r2 += g.z;
    }
  }
  float sum = r3;
  v3 gp = mk(r0, r1, r2);
  sum += vdot(gp, gp);
  float factor = 0.0f;
  if (sum > 1e-6f) factor = 1.0f / sum;
  st.dfsph_factor[i] = factor;
}

// Residual term of compute_density_error (dimcv_sph.py:407-419) of slot i
// whose density_adv is da, and the block's canonical FP64 partial of it
// (as k_residual forms them); every thread of the block calls it.
__device__ __forceinline__ double residual_term(const dimcv_params& p, const dimcv_state& st, int64_t i, float da,
                                                float offset) {
  const float t = !p.has_transit ? da - offset : transit_weight(p, st.x[3 * i]) * (da - offset);
  return (double)t;
}
__device__ __forceinline__ void residual_partial(double v, double* __restrict__ partials, double* lds) {
  const double sum = block_tree_sum256(v, lds);
  if (threadIdx.x == 0) partials[xcd_block()] = sum;
}

// dimcv_sph.py:337-374 for the entries of logical block lb; returns the
// entry's divergence-residual term (offset 0) of the value just written.
template <bool XM>
__device__ __forceinline__ double density_change_body(dimcv_state& st, const dimcv_params& p, int32_t bound,
                                                      const Packed& pk, int64_t lb) {
  double res = 0.0;
  [&] {
    FLUID_ENTRY_LB(lb);
    const bool sd = XM && *pk.tiny == 0;  // shared-reciprocal divisions in cubic_dW
    v3 xi = ld3(st.x, i), vi = ld3(st.v, i);
    float dadv = 0.0f;
    struct G { float4 xmv; v3 v; };
    nbr_loop<8>(nl, ns, nn, [&](uint32_t j) { return G{nb_xmv<XM>(st, pk.xm, j), ld3s(st.v, j)}; },
                [&](int, uint32_t, const G& g) {
                  dadv += g.xmv.w * vdot(vsub(vi, g.v), cubic_dW(p, rel(p, xi, xyz(g.xmv)), sd));
                });
    float d = fmaxf(dadv, 0.0f);
    if (st.nbr_count[f] < 20) d = 0.0f;
    st.density_adv[i] = d;
    if (XM) pk.kap[i] = d * st.dfsph_factor[i] * st.density[i];  // the next iteration's kappa_i
    res = residual_term(p, st, i, d, 0.0f);
  }();
  return res;
}

// With `partials`, also the block partial of the divergence residual.
template <bool XM>
__global__ void __launch_bounds__(256) k_density_change(dimcv_state st, dimcv_params p, int32_t bound,
                                                  const int32_t* __restrict__ done,
                                                  double* __restrict__ partials, Packed pk) {
  __shared__ double lds[256];
  if (jacobi_gated(st, done)) return;  // uniform per block
  const double res = density_change_body<XM>(st, p, bound, pk, xcd_block());
  if (partials) residual_partial(res, partials, lds);
}

// (x, m_V) of every slot as one 16-byte record, for the neighbour gathers
// of a step (positions and volumes do not change until advection).
// Also raises *tiny when a coordinate is nonzero and below 2^-76 in
// magnitude (the condition under which cubic_dW must not share reciprocals).
__device__ __forceinline__ bool tiny_coord(float c) { return c != 0.0f && fabsf(c) < 0x1p-76f; }

__global__ void k_pack_xm(dimcv_state st, dimcv_params p, int32_t bound, float4* __restrict__ xm,
                          float4* __restrict__ xw, int32_t* __restrict__ tiny) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= bound || s >= st.counters[DIMCV_C_PARTICLE_NUM]) return;
  const float x = st.x[3 * s], y = st.x[3 * s + 1], z = st.x[3 * s + 2], mV = st.m_V[s];
  xm[s] = make_float4(x, y, z, mV);
  xw[s] = make_float4(x, y, z, st.material[s] == 1 ? st.m[s] : mV * p.density0);
  if (tiny_coord(x) || tiny_coord(y) || tiny_coord(z)) atomicOr(tiny, 1);
}

// zero_tiny false: the caller has cleared *pk.tiny already (k_step_guard).
int launch_pack_xm(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, const Packed& pk,
                   hipStream_t s, bool zero_tiny) {
  if (zero_tiny) hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, s, (int32_t*)pk.tiny);
  if (particle_bound <= 0) return DIMCV_OK;
  hipLaunchKernelGGL(k_pack_xm, dim3(nblocks(particle_bound)), dim3(256), 0, s, *st, *p, particle_bound,
                     (float4*)pk.xm, (float4*)pk.xw, (int32_t*)pk.tiny);
  DIMCV_LAUNCHED("k_pack_xm");
  return DIMCV_OK;
}

// multiply_time_step (dimcv_sph.py:395-399) of the DFSPH factor at the start
// of a Jacobi solve, also writing kappa_i of its first iteration:
// b alpha rho (which 0, divergence) or (da - rho0) alpha (which 1, pressure).
// With reset, it also clears the Jacobi counters (done flag, iteration count)
// for the solve that follows it.
__global__ void k_scale_kappa(dimcv_state st, dimcv_params p, float scale, int which, int32_t bound,
                              float* __restrict__ kap, int reset, int32_t* __restrict__ tree) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (reset && f == 0) {
    // a failed build leaves the solve done before its first iteration
    st.counters[DIMCV_C_JACOBI_DONE] = step_failed(st) ? 1 : 0;
    st.counters[DIMCV_C_JACOBI_ITERS] = 0;
    st.counters[DIMCV_C_BAR_ARRIVE] = 0;  // barrier of a persistent solve (k_solve_single)
  }
  // the group arrival words of that barrier (the launch has >= 256 threads)
  if (reset && tree && f < DIMCV_BAR_GROUPS) tree[16 * f] = 0;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  float a = st.dfsph_factor[i];
  a *= scale;
  st.dfsph_factor[i] = a;
  kap[i] = which == 0 ? st.density_adv[i] * a * st.density[i] : (st.density_adv[i] - p.density0) * a;
}

int launch_scale_kappa(int which, const dimcv_state* st, const dimcv_params* p, float scale, int32_t fluid_bound,
                       float* kap, hipStream_t s, bool reset_jacobi, int32_t* bar_tree) {
  // with reset_jacobi, at least one block runs (its first thread clears the counters)
  if (fluid_bound <= 0 && !reset_jacobi) return DIMCV_OK;
  hipLaunchKernelGGL(k_scale_kappa, dim3(fluid_bound > 0 ? nblocks(fluid_bound) : 1), dim3(256), 0, s, *st, *p,
                     scale, which, fluid_bound, kap, reset_jacobi ? 1 : 0, reset_jacobi ? bar_tree : nullptr);
  DIMCV_LAUNCHED("k_scale_kappa");
  return DIMCV_OK;
}

// dimcv_sph.py:395-399
__global__ void k_multiply(dimcv_state st, float* field, float scale, int32_t bound) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  field[i] *= scale;
}

// dimcv_sph.py:407-419: per-block canonical FP64 tree over fluid-list order.
__global__ void __launch_bounds__(256) k_residual(dimcv_state st, dimcv_params p, float offset, int32_t bound,
                                                  double* __restrict__ partials, const int32_t* __restrict__ done) {
  __shared__ double lds[256];
  if (jacobi_gated(st, done)) return;  // uniform per block
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (f < bound && f < st.counters[DIMCV_C_FLUID_LIST]) {
    const int64_t i = st.fluid_list[f];
    if (i >= 0) {
      float t;
      if (!p.has_transit) t = st.density_adv[i] - offset;
      else t = transit_weight(p, st.x[3 * i]) * (st.density_adv[i] - offset);
      v = (double)t;
    }
  }
  double s = block_tree_sum256(v, lds);
  if (threadIdx.x == 0) partials[xcd_block()] = s;
}

// dimcv_sph.py:421-468.  kappa_j = b_j alpha_j rho_j is gathered first (with
// XM from the array the previous density change wrote); positions and mass
// factors only for the pairs that pass |kappa_i + kappa_j| > eps.
template <bool XM>
__device__ __forceinline__ void div_iter_body(dimcv_state& st, const dimcv_params& p, int32_t bound,
                                              const Packed& pk, int64_t lb) {
  FLUID_ENTRY_LB(lb);
  const bool sd = XM && *pk.tiny == 0;  // shared-reciprocal divisions in cubic_dW
  const float dt = p.dt;
  v3 xi = ld3(st.x, i);
  const float k_i = XM ? pk.kap[i] : st.density_adv[i] * st.dfsph_factor[i] * st.density[i];
  v3 dv = mk(0.0f, 0.0f, 0.0f);
  nbr_loop2<8>(
      nl, ns, nn,
      [&](uint32_t j) {
        if constexpr (XM) return ld1s(pk.kap, j);
        return ld1s(st.density_adv, j) * ld1s(st.dfsph_factor, j) * ld1s(st.density, j);
      },
      [&](uint32_t e, float k_j) { return fabsf(nb_fluid(e) ? k_i + k_j : k_i) > p.m_eps; },
      [&](uint32_t j, uint32_t e) { return nb_xw<XM>(st, p, pk.xw, j, e); },
      [&](int, uint32_t e, float k_j, const float4& q) {
        const v3 gg = smul(q.w, cubic_dW(p, rel(p, xi, xyz(q)), sd));
        if (nb_fluid(e)) {
          float k_sum = k_i + k_j;
          dv = vsub(dv, smul(dt * k_sum, gg));
        } else {
          v3 vc = smul(-dt * 1.0f * k_i, gg);
          dv = vadd(dv, vc);
        }
      });
  v3 vi = ld3(st.v, i);
  if (!p.has_transit) vi = vadd(vi, dv);
  else vi = vadd(vi, smul(transit_weight(p, xi.x), dv));
  st3(st.v, i, vi);
}

template <bool XM>
__global__ void __launch_bounds__(256) k_div_iter(dimcv_state st, dimcv_params p, int32_t bound,
                                                  const int32_t* __restrict__ done, Packed pk) {
  if (jacobi_gated(st, done)) return;
  div_iter_body<XM>(st, p, bound, pk, xcd_block());
}

// dimcv_sph.py:150-208, fluid particles (surface tension sigma = 0 and the
// boundary viscosity 0 contribute exact zeros and are not evaluated).
// With factor, also multiply_time_step (dimcv_sph.py:395-399) of that field
// (the DFSPH factor, which this pass does not read): factor[i] *= factor_scale.
template <bool XM>
__global__ void __launch_bounds__(256) k_nonpressure_fluid(dimcv_state st, dimcv_params p, int32_t bound,
                                                           Packed pk, float* __restrict__ factor,
                                                           float factor_scale) {
  FLUID_ENTRY();
  const bool sd = XM && *pk.tiny == 0;  // shared-reciprocal divisions in cubic_dW
  const float visc_coef = st.block_scene ? st.scenes[block_scene_of(st)].visc_coef : p.visc_coef;
  v3 xi = ld3(st.x, i), vi = ld3(st.v, i);
  v3 dv = mk(p.gravity[0], p.gravity[1], p.gravity[2]);
  struct G { float4 xmv; v3 v; };
  nbr_loop(nl, ns, nn, [&](uint32_t j) { return G{nb_xmv<XM>(st, pk.xm, j), ld3s(st.v, j)}; },
           [&](int, uint32_t e, const G& g) {
             if (nb_fluid(e)) {
               v3 r = rel(p, xi, xyz(g.xmv));
               float v_xy = vdot(vsub(vi, g.v), r);
               float rn = vnorm(r);
               v3 fv = smul(visc_coef * g.xmv.w * v_xy / (rn * rn + p.visc_h2), cubic_dW(p, r, sd));
               dv = vadd(dv, fv);
             }
           });
  st3(st.acceleration, i, dv);
  if (factor) factor[i] *= factor_scale;
}

// dimcv_sph.py:197-204 for non-fluid slots: static rigid -> 0, otherwise g.
// One value and one store, chosen by selects: written as two stores under
// if/else, the gfx950 build stored the static branch's zero for dynamic
// rigid slots as well (the join overwrote the gravity registers).
// With zero_n > 0 it also does k_zero_nonfluid's fill of z for slots below
// zero_n (every slot but the fluid ones below particle_num).
__global__ void k_nonpressure_other(dimcv_state st, dimcv_params p, int32_t bound, ZeroSet z, int64_t zero_n) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < zero_n && !(s < st.counters[DIMCV_C_PARTICLE_NUM] && st.material[s] == 1)) {
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (z.a[q])
        for (int k = 0; k < z.c[q]; ++k) z.a[q][s * z.c[q] + k] = 0.0f;
  }
  if (s >= bound || s >= st.counters[DIMCV_C_PARTICLE_NUM]) return;
  const int32_t mat = st.material[s];
  if (mat == 1) return;
  const bool stat = mat == 0 && st.is_dynamic[s] == 0;
  st3(st.acceleration, s, mk(stat ? 0.0f : p.gravity[0], stat ? 0.0f : p.gravity[1], stat ? 0.0f : p.gravity[2]));
}

// dimcv_sph.py:220-226
__global__ void k_predict(dimcv_state st, dimcv_params p, int32_t bound) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= bound || s >= st.counters[DIMCV_C_PARTICLE_NUM] || step_failed(st)) return;
  if (st.is_dynamic[s] && st.material[s] == 1)
    st3(st.v, s, vadd(ld3(st.v, s), smul(p.dt, ld3(st.acceleration, s))));
}

// dimcv_sph.py:210-218
__global__ void k_advect(dimcv_state st, dimcv_params p, int32_t bound) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= bound || s >= st.counters[DIMCV_C_PARTICLE_NUM] || step_failed(st)) return;
  if (!st.is_dynamic[s]) return;
  v3 v = ld3(st.v, s);
  if (st.material[s] == 0) {
    v = vadd(v, smul(p.dt, ld3(st.acceleration, s)));
    st3(st.v, s, v);
  }
  st3(st.x, s, vadd(ld3(st.x, s), smul(p.dt, v)));
}

// dimcv_sph.py:470-493 for the entries of logical block lb; returns the
// entry's pressure-residual term (offset: the caller's density0).
template <bool XM>
__device__ __forceinline__ double density_adv_body(dimcv_state& st, const dimcv_params& p, int32_t bound,
                                                   float offset, const Packed& pk, int64_t lb) {
  double res = 0.0;
  [&] {
    FLUID_ENTRY_LB(lb);
    const bool sd = XM && *pk.tiny == 0;  // shared-reciprocal divisions in cubic_dW
    v3 xi = ld3(st.x, i), vi = ld3(st.v, i);
    float delta = 0.0f;
    // w = m (fluid) or m_V rho0 (solid), as the reference weights the two
    // sums; without the packed records w = m and solids read m_V below
    struct G { float4 xw; v3 v; };
    nbr_loop<8>(nl, ns, nn,
                [&](uint32_t j) {
                  if constexpr (XM) return G{nb_xw<true>(st, p, pk.xw, j, 0u), ld3s(st.v, j)};
                  const v3 a = ld3s(st.x, j);
                  return G{make_float4(a.x, a.y, a.z, ld1s(st.m, j)), ld3s(st.v, j)};
                },
                [&](int, uint32_t e, const G& g) {
                  float dd = vdot(vsub(vi, g.v), cubic_dW(p, rel(p, xi, xyz(g.xw)), sd));
                  if (XM || nb_fluid(e)) delta += g.xw.w * dd;
                  else delta += ld1s(st.m_V, nb_slot(e)) * p.density0 * dd;
                });
    float da = st.density[i] + p.dt * delta;
    const float d = fmaxf(da, p.density0);
    st.density_adv[i] = d;
    if (XM) pk.kap[i] = (d - p.density0) * st.dfsph_factor[i];  // the next iteration's kappa_i
    res = residual_term(p, st, i, d, offset);
  }();
  return res;
}

// With `partials`, also the block partial of the pressure residual.
template <bool XM>
__global__ void __launch_bounds__(256) k_density_adv(dimcv_state st, dimcv_params p, int32_t bound,
                                                  const int32_t* __restrict__ done,
                                                  double* __restrict__ partials, float offset,
                                                  Packed pk) {
  __shared__ double lds[256];
  if (jacobi_gated(st, done)) return;  // uniform per block
  const double res = density_adv_body<XM>(st, p, bound, offset, pk, xcd_block());
  if (partials) residual_partial(res, partials, lds);
}

// dimcv_sph.py:495-549 (v_i updated in place, neighbour by neighbour);
// kappa_j = (da_j - rho0) alpha_j first, as in k_div_iter.
template <bool XM>
__device__ __forceinline__ void pressure_iter_body(dimcv_state& st, const dimcv_params& p, int32_t bound,
                                                   const Packed& pk, int64_t lb) {
  FLUID_ENTRY_LB(lb);
  const bool sd = XM && *pk.tiny == 0;  // shared-reciprocal divisions in cubic_dW
  const float dt = p.dt;
  v3 xi = ld3(st.x, i);
  const float k_i = XM ? pk.kap[i] : (st.density_adv[i] - p.density0) * st.dfsph_factor[i];
  v3 vi = ld3(st.v, i);
  const float w = p.has_transit ? transit_weight(p, xi.x) : 1.0f;
  nbr_loop2<8>(
      nl, ns, nn,
      [&](uint32_t j) {
        if constexpr (XM) return ld1s(pk.kap, j);
        return (ld1s(st.density_adv, j) - p.density0) * ld1s(st.dfsph_factor, j);
      },
      [&](uint32_t e, float k_j) { return fabsf(nb_fluid(e) ? k_i + k_j : k_i) > p.m_eps; },
      [&](uint32_t j, uint32_t e) { return nb_xw<XM>(st, p, pk.xw, j, e); },
      [&](int, uint32_t e, float k_j, const float4& q) {
        const v3 gg = smul(q.w, cubic_dW(p, rel(p, xi, xyz(q)), sd));
        if (nb_fluid(e)) {
          float k_sum = k_i + k_j;
          if (!p.has_transit) vi = vsub(vi, smul(dt * k_sum, gg));
          else vi = vsub(vi, smul(w * dt * k_sum, gg));
        } else {
          v3 vc = smul(-dt * 1.0f * k_i, gg);
          if (!p.has_transit) vi = vadd(vi, vc);
          else vi = vadd(vi, smul(w, vc));
        }
      });
  st3(st.v, i, vi);
}

template <bool XM>
__global__ void __launch_bounds__(256) k_pressure_iter(dimcv_state st, dimcv_params p, int32_t bound,
                                                       const int32_t* __restrict__ done,
                                                       Packed pk) {
  if (jacobi_gated(st, done)) return;
  pressure_iter_body<XM>(st, p, bound, pk, xcd_block());
}

// A batched Jacobi solve to convergence in one persistent launch (which 0:
// divergence, dimcv_sph.py:376-393; 1: pressure, :597-614), after its first
// iteration's two passes, which run as ordinary full-occupancy launches and
// leave the block residual partials.  The launch tests every scene (its
// canonical FP64 residual and convergence test) and, while a scene iterates
// on, runs further iterations: the velocity update (k_div_iter /
// k_pressure_iter), the density change (k_density_change / k_density_adv)
// with its block partials, and the tests again.  Grid barriers separate the
// phases, and the blocks leave together when no scene iterates on.  Blocks
// stride over the logical 256-entry blocks; a block of a converged scene is
// skipped as the launched passes skip it, so every scene's arithmetic and
// iteration count are those of separate launches.
// bar: BM_BAR of the batch misc words (arrivals, generation, abort, go flags
// for even and odd iterations; cleared by k_batch_jacobi_reset); tree: the
// barrier's group words (grid_sync_tree; arrivals cleared by the same reset).
template <int WHICH>
__global__ void __launch_bounds__(256) k_batch_solve(dimcv_state st, dimcv_params p, int32_t bound, Packed pk,
                                                     double* __restrict__ partials, float offset, double eta,
                                                     int32_t max_iter, int32_t* __restrict__ bar,
                                                     int32_t* __restrict__ tree) {
  __shared__ double lds[256];
  __shared__ int go;
  int32_t gen = 0, ggen = 0;
  grid_sync_tree_init(bar, tree, gen, ggen);
  const int32_t* done = st.scene_counters + DIMCV_SC_DONE;
  int32_t* flag = bar + 3;
  // XCD-aware: block b runs on XCD b % 8 (gridDim.x is a multiple of 8) and
  // strides over that XCD's contiguous eighth of the logical blocks
  const int64_t nlb = ((int64_t)bound + 255) / 256;
  const int64_t xcd = blockIdx.x & 7u, per = gridDim.x >> 3;
  const int64_t lb0 = nlb * xcd / 8 + (blockIdx.x >> 3), lb1 = nlb * (xcd + 1) / 8;
  for (int it = 0;; ++it) {
    if (it > 0) {  // iteration it + 1 (the first ran before the launch)
      for (int64_t lb = lb0; lb < lb1; lb += per) {
        if (jacobi_gated_lb(st, done, lb)) continue;  // uniform per block
        if (WHICH == 0) div_iter_body<true>(st, p, bound, pk, lb);
        else pressure_iter_body<true>(st, p, bound, pk, lb);
      }
      if (!grid_sync_tree(bar, tree, gen, ggen)) return;
      for (int64_t lb = lb0; lb < lb1; lb += per) {
        if (jacobi_gated_lb(st, done, lb)) continue;
        const double res = WHICH == 0 ? density_change_body<true>(st, p, bound, pk, lb)
                                      : density_adv_body<true>(st, p, bound, offset, pk, lb);
        const double sum = block_tree_sum256(res, lds);
        if (threadIdx.x == 0) partials[lb] = sum;
      }
      if (!grid_sync_tree(bar, tree, gen, ggen)) return;
    }
    for (int sc = blockIdx.x; sc < st.n_scenes; sc += gridDim.x) {
      if (done[(int64_t)sc * DIMCV_SC_COUNT]) continue;  // uniform
      batch_reduce_scene(st, partials, DIMCV_SD_RESIDUAL, sc, lds);
      if (threadIdx.x == 0 && batch_scene_check(st, eta, max_iter, WHICH, sc)) atomicOr(&flag[it & 1], 1);
    }
    // the other flag was last read before this iteration's barriers
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicExch(&flag[(it + 1) & 1], 0);
    if (!grid_sync_tree(bar, tree, gen, ggen)) return;
    if (threadIdx.x == 0) go = __hip_atomic_load(&flag[it & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (!go) break;
  }
}

// Resident blocks of k_batch_solve<which> on the current device: the
// occupancy bound, rounded down to a multiple of the 8 XCDs.
static int solve_capacity(int which) {
  static std::mutex mu;
  static int cap[64][2];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  std::lock_guard<std::mutex> lock(mu);
  if (cap[dev][which] == 0) {
    int per_cu = 0, cus = 0;
    const void* f = which == 0 ? (const void*)k_batch_solve<0> : (const void*)k_batch_solve<1>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 256, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    cap[dev][which] = (per_cu * cus) & ~7;
  }
  return cap[dev][which];
}

int launch_batch_solve(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                       const Packed& pk, double* partials, float offset, double eta, int32_t max_iter,
                       int32_t* bar, int32_t* bar_tree, hipStream_t s) {
  const int cap = solve_capacity(which);
  if (cap <= 0) {
    set_error("persistent Jacobi solve: no resident block capacity on this device");
    return DIMCV_ERR_HIP;
  }
  // a multiple of 8 blocks (idle blocks of a small batch only take part in
  // the barriers)
  const int64_t nlb = ((int64_t)fluid_bound + 255) / 256;
  const unsigned g = (unsigned)std::max<int64_t>(8, std::min<int64_t>((nlb + 7) & ~int64_t(7), cap));
  if (g > DIMCV_BAR_GROUP * DIMCV_BAR_GROUPS) {
    set_error("persistent Jacobi solve: %u blocks exceed the barrier's %d groups", g, DIMCV_BAR_GROUPS);
    return DIMCV_ERR_HIP;
  }
  if (which == 0)
    hipLaunchKernelGGL(k_batch_solve<0>, dim3(g), dim3(256), 0, s, *st, *p, fluid_bound, pk, partials, offset, eta,
                       max_iter, bar, bar_tree);
  else
    hipLaunchKernelGGL(k_batch_solve<1>, dim3(g), dim3(256), 0, s, *st, *p, fluid_bound, pk, partials, offset, eta,
                       max_iter, bar, bar_tree);
  DIMCV_LAUNCHED("k_batch_solve");
  return DIMCV_OK;
}

// The single-scene Jacobi solve's remaining iterations in one persistent
// launch (the device-side form of dimcv_step's convergence checks): the
// first iterations ran as ordinary launches and left the solve's state in
// counters[DIMCV_C_JACOBI_DONE / _ITERS]; if the solve has converged every
// block returns at once, otherwise the launch iterates: the velocity update,
// a grid barrier, the density pass with its block partials, a grid barrier,
// then every block forms the residual's canonical tree over all the block
// partials itself (the tree of k_reduce_tail, so the same value in every
// block) and applies the test, so all blocks leave the loop together with no
// third barrier; block 0 records the residual and the counters.  Same
// arithmetic and order as the launched passes.
// bar = counters + DIMCV_C_BAR_ARRIVE (arrivals, generation, abort); tree:
// the barrier's group words (grid_sync_tree).
//
// The canonical tree of a[0, n) (levels of 256 consecutive values, as
// k_reduce_tail and k_reduce_levels form it), by one block: each group of 256
// level-1 sums is kept in l1 and summed into l2; n <= 2^24.
__device__ double block_canonical_sum(const double* __restrict__ a, int64_t n, double* lds, double* l1, double* l2) {
  const int t = threadIdx.x;
  if (n <= 0) return 0.0;
  const int64_t G = (n + 255) / 256;  // level-1 groups
  const int64_t C = (G + 255) / 256;  // level-2 groups
  for (int64_t c = 0; c < C; ++c) {
    const int64_t q1 = G - c * 256 < 256 ? G - c * 256 : 256;
    for (int64_t q = 0; q < q1; ++q) {
      const int64_t k = (c * 256 + q) * 256 + t;
      const double sum = block_tree_sum256(k < n ? a[k] : 0.0, lds);
      if (t == 0) l1[q] = sum;
    }
    __syncthreads();
    if (G == 1) return l1[0];
    const double s2 = block_tree_sum256(t < q1 ? l1[t] : 0.0, lds);
    if (t == 0) l2[c] = s2;
    __syncthreads();
  }
  if (C == 1) return l2[0];
  return block_tree_sum256(t < C ? l2[t] : 0.0, lds);
}

template <int WHICH>
__global__ void __launch_bounds__(256) k_solve_single(dimcv_state st, dimcv_params p, int32_t bound, Packed pk,
                                                      double* __restrict__ partials, float offset, double eta,
                                                      int32_t max_iter, int32_t* __restrict__ tree) {
  __shared__ double lds[256], l1[256], l2[256];
  __shared__ int go, it0, fluid;
  int32_t* bar = st.counters + DIMCV_C_BAR_ARRIVE;
  if (threadIdx.x == 0) {
    go = st.counters[DIMCV_C_JACOBI_DONE] == 0;
    it0 = st.counters[DIMCV_C_JACOBI_ITERS];
    fluid = st.counters[DIMCV_C_FLUID_NUM];
  }
  __syncthreads();
  if (!go) return;  // uniform: converged before the launch
  int32_t gen = 0, ggen = 0;
  grid_sync_tree_init(bar, tree, gen, ggen);
  int32_t it = it0;
  const int64_t nlb = ((int64_t)bound + 255) / 256;
  const int64_t xcd = blockIdx.x & 7u, per = gridDim.x >> 3;
  const int64_t lb0 = nlb * xcd / 8 + (blockIdx.x >> 3), lb1 = nlb * (xcd + 1) / 8;
  for (;;) {
    for (int64_t lb = lb0; lb < lb1; lb += per) {
      if (WHICH == 0) div_iter_body<true>(st, p, bound, pk, lb);
      else pressure_iter_body<true>(st, p, bound, pk, lb);
    }
    if (!grid_sync_tree(bar, tree, gen, ggen)) return;
    for (int64_t lb = lb0; lb < lb1; lb += per) {
      const double res = WHICH == 0 ? density_change_body<true>(st, p, bound, pk, lb)
                                    : density_adv_body<true>(st, p, bound, offset, pk, lb);
      const double sum = block_tree_sum256(res, lds);
      if (threadIdx.x == 0) partials[lb] = sum;
    }
    if (!grid_sync_tree(bar, tree, gen, ggen)) return;
    const double r = block_canonical_sum(partials, nlb, lds, l1, l2);
    ++it;
    const bool done = r / (double)fluid <= eta || it >= max_iter;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st.dscal[DIMCV_D_RESIDUAL] = r;
      st.counters[DIMCV_C_JACOBI_ITERS] = it;
      if (done) st.counters[DIMCV_C_JACOBI_DONE] = 1;
    }
    if (done) break;  // the same decision in every block
    // the next velocity update overwrites nothing this reduction reads
    // (partials are rewritten only after the next barrier)
  }
}

static int solve_single_capacity(int which) {
  static std::mutex mu;
  static int cap[64][2];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  std::lock_guard<std::mutex> lock(mu);
  if (cap[dev][which] == 0) {
    int per_cu = 0, cus = 0;
    const void* f = which == 0 ? (const void*)k_solve_single<0> : (const void*)k_solve_single<1>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, 256, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    cap[dev][which] = (per_cu * cus) & ~7;
  }
  return cap[dev][which];
}

int launch_solve_single(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                        const Packed& pk, double* partials, float offset, double eta, int32_t max_iter,
                        int32_t* bar_tree, hipStream_t s) {
  const int cap = solve_single_capacity(which);
  if (cap <= 0) {
    set_error("persistent Jacobi solve: no resident block capacity on this device");
    return DIMCV_ERR_HIP;
  }
  const int64_t nlb = ((int64_t)fluid_bound + 255) / 256;
  const unsigned g = (unsigned)std::max<int64_t>(8, std::min<int64_t>((nlb + 7) & ~int64_t(7), cap));
  if (g > DIMCV_BAR_GROUP * DIMCV_BAR_GROUPS) {
    set_error("persistent Jacobi solve: %u blocks exceed the barrier's %d groups", g, DIMCV_BAR_GROUPS);
    return DIMCV_ERR_HIP;
  }
  if (which == 0)
    hipLaunchKernelGGL(k_solve_single<0>, dim3(g), dim3(256), 0, s, *st, *p, fluid_bound, pk, partials, offset, eta,
                       max_iter, bar_tree);
  else
    hipLaunchKernelGGL(k_solve_single<1>, dim3(g), dim3(256), 0, s, *st, *p, fluid_bound, pk, partials, offset, eta,
                       max_iter, bar_tree);
  DIMCV_LAUNCHED("k_solve_single");
  return DIMCV_OK;
}

// dimcv_sph.py:797-801
__global__ void k_copy_x_temp(dimcv_state st, int64_t n_max) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_max) return;
  v3 q = s < st.counters[DIMCV_C_PARTICLE_NUM] ? ld3(st.x, s) : mk(0.0f, 0.0f, 0.0f);
  st3(st.x_temp, s, q);
}

// k_copy_x_temp then k_advect of each slot in one launch over n_max slots
// (advection for slots below `bound`, as k_advect's launch).
__global__ void k_copy_advect(dimcv_state st, dimcv_params p, int64_t n_max, int32_t bound) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_max) return;
  const int32_t pn = st.counters[DIMCV_C_PARTICLE_NUM];
  const v3 x = s < pn ? ld3(st.x, s) : mk(0.0f, 0.0f, 0.0f);
  st3(st.x_temp, s, x);
  if (s >= bound || s >= pn || !st.is_dynamic[s] || step_failed(st)) return;
  v3 v = ld3(st.v, s);
  if (st.material[s] == 0) {
    v = vadd(v, smul(p.dt, ld3(st.acceleration, s)));
    st3(st.v, s, v);
  }
  st3(st.x, s, vadd(x, smul(p.dt, v)));
}

// k_life_time and k_predict of each slot in one launch: the fluid list holds
// exactly the fluid slots below particle_num, in slot order, so the life-time
// update runs by slot on the same particles.
__global__ void k_predict_life(dimcv_state st, dimcv_params p, int32_t bound) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= bound || s >= st.counters[DIMCV_C_PARTICLE_NUM] || step_failed(st)) return;
  if (st.material[s] != 1) return;
  if (in_vortex_domain(p, ld3(st.x, s))) st.life_time[s] += p.dt;
  if (st.is_dynamic[s]) st3(st.v, s, vadd(ld3(st.v, s), smul(p.dt, ld3(st.acceleration, s))));
}

int launch_copy_advect(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, hipStream_t s) {
  if (p->n_max > 0)
    hipLaunchKernelGGL(k_copy_advect, dim3(nblocks(p->n_max)), dim3(256), 0, s, *st, *p, p->n_max,
                       particle_bound);
  DIMCV_LAUNCHED("k_copy_advect");
  return DIMCV_OK;
}

int launch_predict_life(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, hipStream_t s) {
  if (particle_bound > 0)
    hipLaunchKernelGGL(k_predict_life, dim3(nblocks(particle_bound)), dim3(256), 0, s, *st, *p, particle_bound);
  DIMCV_LAUNCHED("k_predict_life");
  return DIMCV_OK;
}

// dimcv_sph.py:778-795
__global__ void __launch_bounds__(256) k_vorticity_vis(dimcv_state st, dimcv_params p, int32_t bound) {
  FLUID_ENTRY();
  v3 xi = ld3(st.x, i);
  float w = st.m_V[i] * cubic_W(p, 0.0f);
  v3 vort = smul(w, ld3(st.vorticity, i));
  FOR_NBR(k, e, j) {
    float wj = st.m_V[j] * cubic_W(p, vnorm(rel(p, xi, ld3(st.x, j))));
    vort = vadd(vort, smul(wj, ld3(st.vorticity, j)));
    w += wj;
  }
  st3(st.vorticity_vis, i, vdivs(vort, w + 1e-6f));
}

// sph_base.py:118-123,149-179 for fluid particles
__global__ void k_bounce(dimcv_state st, dimcv_params p, int32_t bound) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= bound || s >= st.counters[DIMCV_C_PARTICLE_NUM] || step_failed(st)) return;
  if (st.material[s] != 1 || !st.is_dynamic[s]) return;
  v3 pos = ld3(st.x, s);
  float q[3] = {pos.x, pos.y, pos.z};
  float nrm[3] = {0.0f, 0.0f, 0.0f};
  for (int d = 0; d < 3; ++d) {
    if (q[d] > p.bounce_hi[d]) {
      nrm[d] += 1.0f;
      st.x[3 * s + d] = p.bounce_hi[d];
    }
    if (q[d] <= p.bounce_padding) {
      nrm[d] += -1.0f;
      st.x[3 * s + d] = p.bounce_padding;
    }
  }
  v3 cn = mk(nrm[0], nrm[1], nrm[2]);
  float len = vnorm(cn);
  if (len > 1e-6f) {
    v3 vec = vdivs(cn, len);
    v3 v = ld3(st.v, s);
    const float c_f = 0.5f;
    st3(st.v, s, vsub(v, smul((1.0f + c_f) * vdot(v, vec), vec)));
  }
}

// ---- DIMCV vorticity passes -------------------------------------------------

// dimcv_sph.py:264-271
__global__ void k_v_star(dimcv_state st, dimcv_params p, int32_t bound) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  st3(st.v_star, i, vadd(ld3(st.v, i), vmuls(ld3(st.acceleration, i), p.dt)));
}

// dimcv_sph.py:228-248,284-293 and 250-262,273-282; MODE 1 = vorticity,
// 2 = grad_v, 3 = both (+ v*) in one neighbour pass.
// MODE 7 also forms d_vort = vorticity - curl v* (:551-583) in the same loop:
// v*_j = v_j + a_j dt is formed on the fly exactly as compute_v_star does.
// Kinematic vorticity number |W| / (|S| + 1e-8) of a velocity gradient g
// (row-major 3x3), W and S its antisymmetric and symmetric parts times two
// (dimcv_sph.py:295-310).
__device__ __forceinline__ float kvn_of(const float* g) {
  float Wn = 0.0f, Sn = 0.0f;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      float w = g[3 * r + c] - g[3 * c + r];
      float s = g[3 * r + c] + g[3 * c + r];
      if (r == 0 && c == 0) {
        Wn = w * w;
        Sn = s * s;
      } else {
        Wn = Wn + w * w;
        Sn = Sn + s * s;
      }
    }
  return sqrtf(Wn) / (sqrtf(Sn) + 1e-8f);
}

// MODE bits: 1 vorticity, 2 grad v, 4 d_vort from v* (v* written when MODE
// >= 3), 8 the kinematic vorticity number from the pass's own grad v (the
// KVN pass of fixed-M sampling, which reads nothing else that changes).
template <int MODE, bool XM>
__global__ void __launch_bounds__(256) k_vort_gradv(dimcv_state st, dimcv_params p, int32_t bound,
                                                    Packed pk) {
  FLUID_ENTRY();
  const bool sd = XM && *pk.tiny == 0;  // shared-reciprocal divisions in cubic_dW
  v3 xi = ld3(st.x, i), vi = ld3(st.v, i);
  v3 vort = mk(0.0f, 0.0f, 0.0f);
  v3 curl = mk(0.0f, 0.0f, 0.0f);
  v3 vsi = mk(0.0f, 0.0f, 0.0f);
  if (MODE & 4) vsi = vadd(vi, vmuls(ld3(st.acceleration, i), p.dt));
  float G[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  struct Gt { float4 xmv; v3 v, a; };
  nbr_loop(nl, ns, nn,
           [&](uint32_t j) {
             return Gt{nb_xmv<XM>(st, pk.xm, j), ld3s(st.v, j),
                       (MODE & 4) ? ld3s(st.acceleration, j) : mk(0.0f, 0.0f, 0.0f)};
           },
           [&](int, uint32_t e, const Gt& g) {
             if (!nb_fluid(e)) return;
             v3 vj = g.v;
             v3 kd = cubic_dW(p, rel(p, xi, xyz(g.xmv)), sd);
             v3 vji = vsub(vj, vi);
             float mv = g.xmv.w;
             if (MODE & 1) vort = vadd(vort, smul(mv, vcross(vji, kd)));
             if (MODE & 2) {
               v3 a = smul(mv, vji);
               G[0] = G[0] + a.x * kd.x; G[1] = G[1] + a.x * kd.y; G[2] = G[2] + a.x * kd.z;
               G[3] = G[3] + a.y * kd.x; G[4] = G[4] + a.y * kd.y; G[5] = G[5] + a.y * kd.z;
               G[6] = G[6] + a.z * kd.x; G[7] = G[7] + a.z * kd.y; G[8] = G[8] + a.z * kd.z;
             }
             if (MODE & 4) {
               v3 vsj = vadd(vj, vmuls(g.a, p.dt));
               curl = vadd(curl, smul(mv, vcross(vsub(vsj, vsi), kd)));
             }
           });
  if (MODE & 1) st3(st.vorticity, i, vort);
  if (MODE & 2)
    for (int q = 0; q < 9; ++q) st.grad_v[9 * i + q] = G[q];
  if (MODE >= 3) st3(st.v_star, i, vadd(vi, vmuls(ld3(st.acceleration, i), p.dt)));
  if (MODE & 4) st3(st.d_vort, i, vsub(vort, curl));
  if ((MODE & 8) && in_vortex_domain(p, xi)) st.kinematic_vort_num[i] = kvn_of(G);
}

// dimcv_sph.py:551-583
__global__ void __launch_bounds__(256) k_d_vorticity(dimcv_state st, dimcv_params p, int32_t bound) {
  FLUID_ENTRY();
  v3 xi = ld3(st.x, i), vsi = ld3(st.v_star, i);
  v3 curl = mk(0.0f, 0.0f, 0.0f);
  FOR_NBR(k, e, j) {
    if (!nb_fluid(e)) continue;
    curl = vadd(curl, smul(st.m_V[j], vcross(vsub(ld3(st.v_star, j), vsi), cubic_dW(p, rel(p, xi, ld3(st.x, j))))));
  }
  st3(st.d_vort, i, vsub(ld3(st.vorticity, i), curl));
}

// dimcv_sph.py:660-677
__global__ void __launch_bounds__(256) k_d_vort_smoothed(dimcv_state st, dimcv_params p, int32_t bound) {
  FLUID_ENTRY();
  v3 xi = ld3(st.x, i);
  if (!in_vortex_domain(p, xi)) return;
  v3 acc = vmuls(smul(st.m_V[i], ld3(st.d_vort, i)), cubic_W(p, 0.0f));
  FOR_NBR(k, e, j) {
    if (!nb_fluid(e)) continue;
    acc = vadd(acc, vmuls(smul(st.m_V[j], ld3(st.d_vort, j)), cubic_W(p, vnorm(rel(p, xi, ld3(st.x, j))))));
  }
  st3(st.d_vort_smoothed, i, acc);
}

// compute_all_d_vort_smoothed (dimcv_sph.py:660-677) and compute_vorticity_vis
// (:778-795) in one neighbour loop.  Both read only geometry, m_V, d_vort and
// vorticity, none of which changes between the vorticity passes and the end
// of the substep, so the visualisation field is the one the reference
// computes after the pressure solve.
__global__ void __launch_bounds__(256) k_smooth_vis(dimcv_state st, dimcv_params p, int32_t bound) {
  FLUID_ENTRY();
  v3 xi = ld3(st.x, i);
  const bool dom = in_vortex_domain(p, xi);
  const float w0 = cubic_W(p, 0.0f);
  v3 acc = vmuls(smul(st.m_V[i], ld3(st.d_vort, i)), w0);
  float w = st.m_V[i] * w0;
  v3 vort = smul(w, ld3(st.vorticity, i));
  FOR_NBR(k, e, j) {
    float Wr = cubic_W(p, vnorm(rel(p, xi, ld3(st.x, j))));
    float mv = st.m_V[j];
    if (dom && nb_fluid(e)) acc = vadd(acc, vmuls(smul(mv, ld3(st.d_vort, j)), Wr));
    float wj = mv * Wr;
    vort = vadd(vort, smul(wj, ld3(st.vorticity, j)));
    w += wj;
  }
  if (dom) st3(st.d_vort_smoothed, i, acc);
  st3(st.vorticity_vis, i, vdivs(vort, w + 1e-6f));
}

// dimcv_sph.py:295-324 with host-fed uniforms u1/u2 per slot.
__global__ void k_kvn_sample(dimcv_state st, dimcv_params p, const float* __restrict__ u1,
                             const float* __restrict__ u2, int do_bernoulli, int32_t bound) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0 || !in_vortex_domain(p, ld3(st.x, i))) return;
  const float* G = st.grad_v + 9 * i;
  float g[9];
  for (int q = 0; q < 9; ++q) g[q] = G[q];
  const float kvn = kvn_of(g);
  float t0 = fminf(fmaxf(st.life_time[i], 0.0f), 1.0f);
  float rr = 0.018f * exp_f32(p.neg_gen_a * t0) + 0.004f;
  float produce = 1.0f - 1.0f / pow_f32(kvn + 1.0f, rr);
  float del = exp_f32(p.neg_del_b * kvn);
  st.kinematic_vort_num[i] = kvn;
  if (!do_bernoulli || step_failed(st)) return;
  uint8_t smp = st.is_sample[i];
  if (smp == 0 && u1[i] < produce) st.is_sample[i] = 1;
  else if (smp == 1 && kvn < 2.0f && u2[i] < del) st.is_sample[i] = 0;
}

__device__ __forceinline__ float epan(const dimcv_params& p, v3 r) {
  float ret = 1.0f;
  float q0 = fabsf(r.x) * p.inv_support_radius;
  float q1 = fabsf(r.y) * p.inv_support_radius;
  float q2 = fabsf(r.z) * p.inv_support_radius;
  if (q0 < 1.0f) ret *= 0.75f * (1.0f - q0 * q0); else ret *= 0.0f;
  if (q1 < 1.0f) ret *= 0.75f * (1.0f - q1 * q1); else ret *= 0.0f;
  if (q2 < 1.0f) ret *= 0.75f * (1.0f - q2 * q2); else ret *= 0.0f;
  return ret;
}

// k_smooth_vis and compute_kernel_density (dimcv_sph.py:721-751) in one
// neighbour loop, with the block's canonical FP64 partial of kden_sum.  KVN
// must be current (it depends only on grad_v, so computing it before the
// smoothing changes no value).
template <bool XM>
__global__ void __launch_bounds__(256) k_smooth_vis_kden(dimcv_state st, dimcv_params p, int32_t bound,
                                                         double* __restrict__ partials,
                                                         Packed pk) {
  __shared__ double lds[256];
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  const int64_t i = (f < bound && f < st.counters[DIMCV_C_FLUID_LIST]) ? st.fluid_list[f] : -1;
  if (i >= 0) {
    const int nn = min(st.nbr_count[f], p.nbr_kmax);
    const NbrList nl(st.nbr, f);
    const int64_t ns = p.n_max;
    const v3 xi = ld3(st.x, i);
    const bool dom = in_vortex_domain(p, xi);
    const float w0 = cubic_W(p, 0.0f);
    v3 acc = vmuls(smul(st.m_V[i], ld3(st.d_vort, i)), w0);
    float w = st.m_V[i] * w0;
    v3 vort = smul(w, ld3(st.vorticity, i));
    float kden = st.kinematic_vort_num[i] * epan(p, mk(0.0f, 0.0f, 0.0f));
    struct G { float4 xmv; v3 dv, vo; float kvn; };
    nbr_loop<8>(nl, ns, nn,
             [&](uint32_t j) {
               return G{nb_xmv<XM>(st, pk.xm, j), ld3s(st.d_vort, j), ld3s(st.vorticity, j),
                        ld1s(st.kinematic_vort_num, j)};
             },
             [&](int, uint32_t e, const G& g) {
               const v3 r = rel(p, xi, xyz(g.xmv));
               float Wr = cubic_W(p, vnorm(r));
               float mv = g.xmv.w;
               if (dom && nb_fluid(e)) {
                 acc = vadd(acc, vmuls(smul(mv, g.dv), Wr));
                 kden += g.kvn * epan(p, r);
               }
               float wj = mv * Wr;
               vort = vadd(vort, smul(wj, g.vo));
               w += wj;
             });
    if (dom) {
      st3(st.d_vort_smoothed, i, acc);
      st.kernel_density[i] = kden;
      contrib = (double)kden;
    }
    st3(st.vorticity_vis, i, vdivs(vort, w + 1e-6f));
  }
  const double sum = block_tree_sum256(contrib, lds);
  if (threadIdx.x == 0) partials[xcd_block()] = sum;
}

// dimcv_sph.py:721-751; kden_sum partials in canonical tree order.
__global__ void __launch_bounds__(256) k_kernel_density(dimcv_state st, dimcv_params p, int32_t bound,
                                                        double* __restrict__ partials) {
  __shared__ double lds[256];
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (f < bound && f < st.counters[DIMCV_C_FLUID_LIST]) {
    const int64_t i = st.fluid_list[f];
    if (i >= 0 && in_vortex_domain(p, ld3(st.x, i))) {
      const v3 xi = ld3(st.x, i);
      const int nn = min(st.nbr_count[f], p.nbr_kmax);
      const NbrList nl(st.nbr, f);
      const int64_t ns = p.n_max;
      float kden = st.kinematic_vort_num[i] * epan(p, mk(0.0f, 0.0f, 0.0f));
      FOR_NBR(k, e, j) {
        if (!nb_fluid(e)) continue;
        kden += st.kinematic_vort_num[j] * epan(p, rel(p, xi, ld3(st.x, j)));
      }
      st.kernel_density[i] = kden;
      contrib = (double)kden;
    }
  }
  double s = block_tree_sum256(contrib, lds);
  if (threadIdx.x == 0) partials[xcd_block()] = s;
}

// dimcv_sph.py:753-760
__global__ void k_life_time(dimcv_state st, dimcv_params p, int32_t bound) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST] || step_failed(st)) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0 || !in_vortex_domain(p, ld3(st.x, i))) return;
  st.life_time[i] += p.dt;
}

int compact_samples(const dimcv_state* st, int64_t n, Workspace& ws, hipStream_t s);

// Jacobi pieces gated by a device flag (iterations after convergence are
// skipped on the device), used by the native substep.
// which 1 and 3 also write the residual's block partials when `partials`
// is given (offset: that of the pressure residual).
int launch_jacobi_pass(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                       const int32_t* done, hipStream_t s, double* partials, float offset, const Packed* pk) {
  if (fluid_bound <= 0) return DIMCV_OK;
  const dim3 g(nblocks(fluid_bound)), b(256);
  const Packed q = pk ? *pk : Packed{nullptr, nullptr, nullptr, nullptr};
  switch (which) {
    case 0:
      if (pk) hipLaunchKernelGGL(k_div_iter<true>, g, b, 0, s, *st, *p, fluid_bound, done, q);
      else hipLaunchKernelGGL(k_div_iter<false>, g, b, 0, s, *st, *p, fluid_bound, done, q);
      break;
    case 1:
      if (pk) hipLaunchKernelGGL(k_density_change<true>, g, b, 0, s, *st, *p, fluid_bound, done, partials, q);
      else hipLaunchKernelGGL(k_density_change<false>, g, b, 0, s, *st, *p, fluid_bound, done, partials, q);
      break;
    case 2:
      if (pk) hipLaunchKernelGGL(k_pressure_iter<true>, g, b, 0, s, *st, *p, fluid_bound, done, q);
      else hipLaunchKernelGGL(k_pressure_iter<false>, g, b, 0, s, *st, *p, fluid_bound, done, q);
      // its reaction on dynamic rigid particles, before density_adv moves on
      launch_rigid_reaction(st, p, (int32_t)p->n_max, done, s);
      break;
    default:
      if (pk) hipLaunchKernelGGL(k_density_adv<true>, g, b, 0, s, *st, *p, fluid_bound, done, partials, offset, q);
      else hipLaunchKernelGGL(k_density_adv<false>, g, b, 0, s, *st, *p, fluid_bound, done, partials, offset, q);
      break;
  }
  DIMCV_LAUNCHED("jacobi pass");
  return DIMCV_OK;
}

// compute_non_pressure_forces (dimcv_sph.py:150-208): non-fluid slots, then
// the fluid pass (gathering the packed records when pk is given).
// zero_fixed_m: also the non-fluid zero fill of the fixed-M vorticity pass
// (launch_fused 7 then skips it).
int launch_nonpressure(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, int32_t fluid_bound,
                       hipStream_t s, const Packed* pk, float* factor, float factor_scale, bool zero_fixed_m) {
  const Packed q = pk ? *pk : Packed{nullptr, nullptr, nullptr, nullptr};
  const ZeroSet z = zero_fixed_m ? ZeroSet{{st->v_star, st->vorticity, st->grad_v, st->d_vort, st->vorticity_vis},
                                           {3, 3, 9, 3, 3}}
                                 : ZeroSet{{nullptr, nullptr, nullptr, nullptr, nullptr}, {0, 0, 0, 0, 0}};
  const int64_t zero_n = zero_fixed_m ? p->n_max : 0;
  const int64_t other_n = std::max<int64_t>(particle_bound, zero_n);
  if (other_n > 0)
    hipLaunchKernelGGL(k_nonpressure_other, dim3(nblocks(other_n)), dim3(256), 0, s, *st, *p, particle_bound, z,
                       zero_n);
  DIMCV_LAUNCHED("k_nonpressure_other");
  if (fluid_bound > 0) {
    const dim3 g(nblocks(fluid_bound)), b(256);
    if (pk) hipLaunchKernelGGL(k_nonpressure_fluid<true>, g, b, 0, s, *st, *p, fluid_bound, q, factor, factor_scale);
    else hipLaunchKernelGGL(k_nonpressure_fluid<false>, g, b, 0, s, *st, *p, fluid_bound, q, factor, factor_scale);
  }
  DIMCV_LAUNCHED("k_nonpressure_fluid");
  return DIMCV_OK;
}

// which: 0 density + factor; 3 density + factor + first density change; 1
// v*, vorticity, grad v, d_vort; 5 as 1 with the fixed-M KVN, also zeroing
// vorticity_vis's non-fluid slots; 7 as 5 after launch_nonpressure with
// zero_fixed_m (which did the zero fill); 4 smoothing + AKDE with kden
// partials; 6 as 4 after a 5 or 7 (no zero fill); 2 smoothing alone.
int launch_fused(int which, const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, hipStream_t s,
                 double* partials, const Packed* pk) {
  const Packed q = pk ? *pk : Packed{nullptr, nullptr, nullptr, nullptr};
  if (which == 1) {  // v*, vorticity, grad_v, d_vort: zero the non-fluid slots first
    int rc = zero_nonfluid(st, p, s, st->v_star, 3, st->vorticity, 3, st->grad_v, 9);
    if (!rc) rc = zero_nonfluid(st, p, s, st->d_vort, 3);
    if (rc) return rc;
  } else if (which == 7) {  // as 5, its zero fill done by launch_nonpressure (zero_fixed_m)
  } else if (which == 5) {  // as 1, and vorticity_vis for a following which 6, in one launch
    int rc = zero_nonfluid(st, p, s, ZeroSet{{st->v_star, st->vorticity, st->grad_v, st->d_vort, st->vorticity_vis},
                                             {3, 3, 9, 3, 3}});
    if (rc) return rc;
  } else if (which == 2 || which == 4) {
    int rc = zero_nonfluid(st, p, s, st->vorticity_vis, 3);
    if (rc) return rc;
  }
  if ((which == 4 || which == 6) && !partials) {
    set_error("launch_fused: kden partials required");
    return DIMCV_ERR_ARG;
  }
  if (fluid_bound <= 0) return DIMCV_OK;
  dim3 g(nblocks(fluid_bound)), b(256);
  switch (which) {
    case 0:
      if (pk) hipLaunchKernelGGL((k_density_factor<false, true>), g, b, 0, s, *st, *p, fluid_bound, q);
      else hipLaunchKernelGGL((k_density_factor<false, false>), g, b, 0, s, *st, *p, fluid_bound, q);
      break;
    case 1:
      if (pk) hipLaunchKernelGGL((k_vort_gradv<7, true>), g, b, 0, s, *st, *p, fluid_bound, q);
      else hipLaunchKernelGGL((k_vort_gradv<7, false>), g, b, 0, s, *st, *p, fluid_bound, q);
      break;
    case 3:
      if (pk) hipLaunchKernelGGL((k_density_factor<true, true>), g, b, 0, s, *st, *p, fluid_bound, q);
      else hipLaunchKernelGGL((k_density_factor<true, false>), g, b, 0, s, *st, *p, fluid_bound, q);
      break;
    case 5:
    case 7:
      if (pk) hipLaunchKernelGGL((k_vort_gradv<15, true>), g, b, 0, s, *st, *p, fluid_bound, q);
      else hipLaunchKernelGGL((k_vort_gradv<15, false>), g, b, 0, s, *st, *p, fluid_bound, q);
      break;
    case 4:
    case 6:
      if (pk) hipLaunchKernelGGL(k_smooth_vis_kden<true>, g, b, 0, s, *st, *p, fluid_bound, partials, q);
      else hipLaunchKernelGGL(k_smooth_vis_kden<false>, g, b, 0, s, *st, *p, fluid_bound, partials, q);
      break;
    default: hipLaunchKernelGGL(k_smooth_vis, g, b, 0, s, *st, *p, fluid_bound); break;
  }
  DIMCV_LAUNCHED("fused pass");
  return DIMCV_OK;
}

// Per-block FP64 partials of the residual (which = 0, with offset) or of
// kden (which = 1), 256 fluid-list entries per block; the batched path
// reduces them per scene.
int launch_block_partials(int which, const dimcv_state* st, const dimcv_params* p, float offset, int32_t bound,
                          double* partials, const int32_t* done, hipStream_t s) {
  const unsigned nb = nblocks(bound);
  if (nb == 0) return DIMCV_OK;
  if (which == 0)
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, *st, *p, offset, bound, partials, done);
  else
    hipLaunchKernelGGL(k_kernel_density, dim3(nb), dim3(256), 0, s, *st, *p, bound, partials);
  DIMCV_LAUNCHED("block partials");
  return DIMCV_OK;
}

}  // namespace dimcv

using namespace dimcv;

// single-mode vorticity passes of the per-pass entry points
constexpr auto k_vort_gradv_1 = k_vort_gradv<1, false>;
constexpr auto k_vort_gradv_2 = k_vort_gradv<2, false>;
constexpr auto k_vort_gradv_3 = k_vort_gradv<3, false>;

#define LAUNCH_FLUID(kern, ...)                                                                 \
  do {                                                                                          \
    if (fluid_bound > 0)                                                                        \
      hipLaunchKernelGGL(kern, dim3(nblocks(fluid_bound)), dim3(256), 0, (hipStream_t)stream, \
                         __VA_ARGS__);                                                          \
    DIMCV_LAUNCHED(#kern);                                                                      \
  } while (0)

#define LAUNCH_SLOTS(kern, bound, ...)                                                          \
  do {                                                                                          \
    if ((bound) > 0)                                                                            \
      hipLaunchKernelGGL(kern, dim3(nblocks(bound)), dim3(256), 0, (hipStream_t)stream,        \
                         __VA_ARGS__);                                                          \
    DIMCV_LAUNCHED(#kern);                                                                      \
  } while (0)

extern "C" {

int dimcv_compute_densities(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  LAUNCH_FLUID(k_densities, *st, *p, fluid_bound);
  return DIMCV_OK;
}

int dimcv_compute_DFSPH_factor(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  LAUNCH_FLUID(k_dfsph_factor, *st, *p, fluid_bound);
  return DIMCV_OK;
}

int dimcv_compute_density_change(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  LAUNCH_FLUID(k_density_change<false>, *st, *p, fluid_bound, nullptr, nullptr, Packed{});
  return DIMCV_OK;
}

int dimcv_multiply_time_step(const dimcv_state* st, const dimcv_params* p, float* field, float scale,
                             int32_t fluid_bound, void* stream) {
  (void)p;
  LAUNCH_FLUID(k_multiply, *st, field, scale, fluid_bound);
  return DIMCV_OK;
}

int dimcv_compute_density_error(const dimcv_state* st, const dimcv_params* p, float offset,
                                int32_t fluid_bound, double* host_sum, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  unsigned nb = nblocks(fluid_bound);
  if (nb > 0) {
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, s, *st, *p, offset, fluid_bound, ws.partials, nullptr);
    DIMCV_LAUNCHED("k_residual");
  }
  rc = reduce_partials_f64(ws.partials, nb, st->dscal + DIMCV_D_RESIDUAL, stream, nullptr);
  if (rc) return rc;
  if (host_sum) {
    DIMCV_HIP(hipMemcpyAsync(host_sum, st->dscal + DIMCV_D_RESIDUAL, sizeof(double), hipMemcpyDeviceToHost, s));
    DIMCV_HIP(hipStreamSynchronize(s));
  }
  return DIMCV_OK;
}

int dimcv_divergence_solver_iteration(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  LAUNCH_FLUID(k_div_iter<false>, *st, *p, fluid_bound, nullptr, Packed{});
  return DIMCV_OK;
}

int dimcv_compute_non_pressure_forces(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound,
                                      int32_t fluid_bound, void* stream) {
  return dimcv::launch_nonpressure(st, p, particle_bound, fluid_bound, (hipStream_t)stream, nullptr, nullptr, 1.0f,
                                   false);
}

int dimcv_predict_velocity(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream) {
  LAUNCH_SLOTS(k_predict, particle_bound, *st, *p, particle_bound);
  return DIMCV_OK;
}

int dimcv_advect(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream) {
  LAUNCH_SLOTS(k_advect, particle_bound, *st, *p, particle_bound);
  return DIMCV_OK;
}

int dimcv_compute_density_adv(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  LAUNCH_FLUID(k_density_adv<false>, *st, *p, fluid_bound, nullptr, nullptr, 0.0f, Packed{});
  return DIMCV_OK;
}

int dimcv_pressure_solve_iteration(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  LAUNCH_FLUID(k_pressure_iter<false>, *st, *p, fluid_bound, nullptr, Packed{});
  return dimcv::launch_rigid_reaction(st, p, (int32_t)p->n_max, nullptr, (hipStream_t)stream);
}

int dimcv_copy_x_temp(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream) {
  (void)particle_bound;
  LAUNCH_SLOTS(k_copy_x_temp, p->n_max, *st, p->n_max);
  return DIMCV_OK;
}

int dimcv_compute_vorticity_vis(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  int rc = zero_nonfluid(st, p, (hipStream_t)stream, st->vorticity_vis, 3);
  if (rc) return rc;
  LAUNCH_FLUID(k_vorticity_vis, *st, *p, fluid_bound);
  return DIMCV_OK;
}

int dimcv_enforce_boundary_3d(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream) {
  LAUNCH_SLOTS(k_bounce, particle_bound, *st, *p, particle_bound);
  return DIMCV_OK;
}

int dimcv_compute_v_star(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  int rc = zero_nonfluid(st, p, (hipStream_t)stream, st->v_star, 3);
  if (rc) return rc;
  LAUNCH_FLUID(k_v_star, *st, *p, fluid_bound);
  return DIMCV_OK;
}

int dimcv_compute_vorticity(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  int rc = zero_nonfluid(st, p, (hipStream_t)stream, st->vorticity, 3);
  if (rc) return rc;
  LAUNCH_FLUID(k_vort_gradv_1, *st, *p, fluid_bound, Packed{});
  return DIMCV_OK;
}

int dimcv_compute_grad_v(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  int rc = zero_nonfluid(st, p, (hipStream_t)stream, st->grad_v, 9);
  if (rc) return rc;
  LAUNCH_FLUID(k_vort_gradv_2, *st, *p, fluid_bound, Packed{});
  return DIMCV_OK;
}

int dimcv_dimcv_init(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  int rc = zero_nonfluid(st, p, (hipStream_t)stream, st->v_star, 3, st->vorticity, 3, st->grad_v, 9);
  if (rc) return rc;
  LAUNCH_FLUID(k_vort_gradv_3, *st, *p, fluid_bound, Packed{});
  return DIMCV_OK;
}

int dimcv_compute_d_vorticity(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  int rc = zero_nonfluid(st, p, (hipStream_t)stream, st->d_vort, 3);
  if (rc) return rc;
  LAUNCH_FLUID(k_d_vorticity, *st, *p, fluid_bound);
  return DIMCV_OK;
}

int dimcv_compute_all_d_vort_smoothed(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                                      void* stream) {
  LAUNCH_FLUID(k_d_vort_smoothed, *st, *p, fluid_bound);
  return DIMCV_OK;
}

int dimcv_compute_kvn_sample(const dimcv_state* st, const dimcv_params* p, const float* u1, const float* u2,
                             int32_t do_bernoulli, int32_t fluid_bound, void* stream) {
  if (do_bernoulli && (!u1 || !u2)) {
    set_error("dimcv_compute_kvn_sample: uniforms required for the Bernoulli update");
    return DIMCV_ERR_ARG;
  }
  LAUNCH_FLUID(k_kvn_sample, *st, *p, u1, u2, (int)do_bernoulli, fluid_bound);
  return DIMCV_OK;
}

int dimcv_record_sample_idx(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  if (p->n_max > 0) {
    hipLaunchKernelGGL(k_zero_i32n, dim3(((unsigned)p->n_max + 255u) / 256u), dim3(256), 0, s, st->sample_idx,
                       (int)p->n_max);
    DIMCV_LAUNCHED("k_zero_i32n");
  }
  return compact_samples(st, particle_bound, ws, s);
}

int dimcv_compute_kernel_density(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  unsigned nb = nblocks(fluid_bound);
  if (nb > 0) {
    hipLaunchKernelGGL(k_kernel_density, dim3(nb), dim3(256), 0, (hipStream_t)stream, *st, *p, fluid_bound,
                       ws.partials);
    DIMCV_LAUNCHED("k_kernel_density");
  }
  return reduce_partials_f64(ws.partials, nb, st->dscal + DIMCV_D_KDEN_SUM, stream, nullptr);
}

int dimcv_update_particle_life_time(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound,
                                    void* stream) {
  LAUNCH_FLUID(k_life_time, *st, *p, fluid_bound);
  return DIMCV_OK;
}

}  // extern "C"
