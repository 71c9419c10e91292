// Velocity recovery from vorticity (vort2vel, dimcv_sph.py:649-719), the
// inverse-CDF importance sampler, the direct O(N^2) Biot-Savart comparator and
// the vortex-ring initial condition.
//
// Far field.  For target i the reference sums, over samples s with
// |x_i - x_s| > h,
//     K(|r|) (r_xy x w_s) S_i / (kden_s + 1e-8),    K(r) = 1 / r^2 for r > h,
// then scales by 0.1592 V / n_i (n_i = accepted samples).  The three modes:
//   EXACT: the reference expression, IEEE div/sqrt, per pair (bit-exact);
//   FAST : c_s = w_s / (kden_s + 1e-8) folded once per sample, S_i factored
//          out of the sum, K from a hardware reciprocal, FMA accumulation.
//          The acceptance test |r| > h is r^2 > T with T the largest f32 whose
//          correctly rounded sqrt is <= h, and r^2 formed without FMA, so the
//          accepted set (and n_i) equals the reference's exactly.
//   KAHAN: FAST with Kahan-compensated f32 accumulators.
// Samples are read with wave-uniform addresses (scalar loads): every lane of a
// wave walks the same sample at the same time; no LDS staging is needed.
#include <mutex>

#include "common.h"

namespace dimcv {

// ---- near field + S_i (one thread per fluid-list entry) --------------------
__global__ void __launch_bounds__(256) k_vv_near(dimcv_state st, dimcv_params p, int32_t bound,
                                                 float* __restrict__ near3, float* __restrict__ scale_s) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  v3 xi = ld3(st.x, i);
  if (!in_vortex_domain(p, xi)) return;
  const int nn = min(st.nbr_count[f], p.nbr_kmax);
  const NbrList nl(st.nbr, f);
  const int64_t ns = p.n_max;
  v3 vv = mk(0.0f, 0.0f, 0.0f);
  float kns = st.kernel_density[i];
  struct G { v3 x, w; float kd; };
  nbr_loop<8>(nl, ns, nn,
           [&](uint32_t j) { return G{ld3s(st.x, j), ld3s(st.d_vort_smoothed, j), ld1s(st.kernel_density, j)}; },
           [&](int, uint32_t e, const G& g) {
             if (!nb_fluid(e)) return;
             v3 r2 = rel(p, xi, g.x);
             v3 r = mk(r2.x, r2.y, 0.0f);
             float rn = vnorm(r);
             const float ex = exp_f32(-rn * p.inv_eps);  // both exponentials of the expression
             float A = (1.0f - ex) / (rn * (rn + p.eps * ex));
             vv = vadd(vv, smul(A, vcross(r, g.w)));
             kns += g.kd;
           });
  vv = vmuls(vv, p.vp_factor);
  st3(near3, f, vv);
  scale_s[f] = (float)(kden_sum_of(st) - (double)kns);
}

// Final combination (dimcv_sph.py:711-719).
__device__ __forceinline__ void vv_finish(dimcv_state& st, const dimcv_params& p, int64_t i, v3 near, v3 vs,
                                          int added) {
  if (step_failed(st)) return;
  if (added == 0) vs = mk(0.0f, 0.0f, 0.0f);
  else vs = vmuls(vs, p.vp_factor / (float)added);
  v3 vvort = vadd(vs, near);
  if (st.v_vort) st3(st.v_vort, i, vvort);
  v3 vi = ld3(st.v, i);
  if (vnorm(vvort) < vnorm(vi)) st3(st.v, i, vadd(vi, vvort));
}

// EXACT far field: the reference expression per pair.
__global__ void __launch_bounds__(256) k_vv_far_exact(dimcv_state st, dimcv_params p, int32_t bound,
                                                      const float* __restrict__ near3,
                                                      const float* __restrict__ scale_s) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  v3 xi = ld3(st.x, i);
  if (!in_vortex_domain(p, xi)) return;
  const float S = scale_s[f];
  int M;
  const int32_t* __restrict__ sidx = st.sample_idx;
  if (st.block_scene) {  // batched: the scene's own samples
    const int sc = block_scene_of(st);
    M = st.scene_counters[(int64_t)sc * DIMCV_SC_COUNT + DIMCV_SC_NUM_SAMPLES];
    sidx += st.scenes[sc].sample_base;
  } else {
    M = st.counters[DIMCV_C_NUM_SAMPLES];
  }
  v3 vs = mk(0.0f, 0.0f, 0.0f);
  int added = 0;
  for (int js = 0; js < M; ++js) {
    const int64_t pj = sidx[js];
    v3 r2 = rel(p, xi, ld3(st.x, pj));
    float rn = vnorm(r2);
    if (rn > p.support_radius) {
      v3 r = mk(r2.x, r2.y, 0.0f);
      const float ex = exp_f32(-rn * p.inv_eps);  // both exponentials of the expression
      float A = (1.0f - ex) / (rn * (rn + p.eps * ex));
      v3 t = smul(A, vcross(r, ld3(st.d_vort_smoothed, pj)));
      t = vmuls(t, S);
      t = vdivs(t, st.kernel_density[pj] + 1e-8f);
      vs = vadd(vs, t);
      added += 1;
    }
  }
  vv_finish(st, p, i, ld3(near3, f), vs, added);
}

// Pack far-field sources: (x, y, z, 0), (c_x, c_y, c_z, 0) with
// c = w / (kden + 1e-8) (weighted = 1) or c = w (weighted = 0).
__global__ void k_pack_sources(dimcv_state st, const int32_t* __restrict__ idx, const int32_t* __restrict__ count,
                               int64_t cap, int weighted, float4* __restrict__ out) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cap || k >= *count) return;
  int64_t s = idx[k];
  v3 w = ld3(st.d_vort_smoothed, s);
  if (weighted) {
    float den = st.kernel_density[s] + 1e-8f;
    w = mk(w.x / den, w.y / den, w.z / den);
  }
  out[2 * k] = make_float4(st.x[3 * s], st.x[3 * s + 1], st.x[3 * s + 2], 0.0f);
  out[2 * k + 1] = make_float4(w.x, w.y, w.z, 0.0f);
}

// FAST / KAHAN far field: sum_s 1/r^2 (r_xy x c_s) over accepted samples s.
//
// Common path (no lane of the wave near the cut-off): r^2 with FMAs, one
// reciprocal shared by two samples (1/a = b / (ab)), four FMAs into the
// accumulators, no masking.  A wave takes the exact path for a sample pair
// when any lane has r^2 <= T_hi = T (1 + 2^-20), a margin far wider than the
// few-ulp difference between the FMA and the reference r^2: there every lane
// forms r^2 in the reference order, accepts iff r^2 > T (<=> sqrt_rn(r^2) >
// h), and counts rejections, so the accepted set and n_i are exact.
template <bool KAHAN>
__device__ __forceinline__ void acc_add(float& s, float& comp, float term) {
  if (KAHAN) {
    float y = term - comp, t = s + y;
    comp = (t - s) - y;
    s = t;
  } else {
    s += term;
  }
}

template <bool PERIODIC>
__device__ __forceinline__ void far_delta(v3 xi, float4 a, v3 L, float& dx, float& dy, float& dz) {
  dx = xi.x - a.x;
  dy = xi.y - a.y;
  dz = xi.z - a.z;
  if (PERIODIC) {
    const float hx = 0.5f * L.x, hy = 0.5f * L.y, hz = 0.5f * L.z;
    dx = dx > hx ? dx - L.x : (dx < -hx ? dx + L.x : dx);
    dy = dy > hy ? dy - L.y : (dy < -hy ? dy + L.y : dy);
    dz = dz > hz ? dz - L.z : (dz < -hz ? dz + L.z : dz);
  }
}

template <bool KAHAN>
__device__ __forceinline__ void far_accumulate(float dx, float dy, float inv, float4 c, float* A, float* C) {
  const float t1 = dy * inv, t2 = dx * inv;
  if (KAHAN) {
    acc_add<true>(A[0], C[0], t1 * c.z);
    acc_add<true>(A[1], C[1], -(t2 * c.z));
    acc_add<true>(A[2], C[2], fmaf(t2, c.y, -(t1 * c.x)));
  } else {
    A[0] = fmaf(t1, c.z, A[0]);
    A[1] = fmaf(-t2, c.z, A[1]);
    A[2] = fmaf(t2, c.y, A[2]);
    A[2] = fmaf(-t1, c.x, A[2]);
  }
}

template <bool KAHAN, bool PERIODIC>
__device__ __forceinline__ void far_exact_sample(v3 xi, float4 a, float4 c, float T, v3 L, float* A, float* C,
                                                 int& rej) {
  float dx, dy, dz;
  far_delta<PERIODIC>(xi, a, L, dx, dy, dz);
  float r2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
  bool ok = r2 > T;
  float inv = ok ? __builtin_amdgcn_rcpf(r2) : 0.0f;
  rej += ok ? 0 : 1;
  far_accumulate<KAHAN>(dx, dy, inv, c, A, C);
}

// One pair of samples.  Both paths only produce the two reciprocals (0 for a
// rejected sample);
// the accumulation after them is shared, so the accumulators stay in place
// (no register shuffling between the paths).
template <bool KAHAN, bool PERIODIC>
__device__ __forceinline__ void far_pair(float4 a0, float4 c0, float4 a1, float4 c1, v3 xi, float T, float T_hi,
                                         v3 L, float* A, float* C, int& rej) {
  float dx0, dy0, dz0, dx1, dy1, dz1;
  far_delta<PERIODIC>(xi, a0, L, dx0, dy0, dz0);
  far_delta<PERIODIC>(xi, a1, L, dx1, dy1, dz1);
  const float r20 = fmaf(dz0, dz0, fmaf(dy0, dy0, dx0 * dx0));
  const float r21 = fmaf(dz1, dz1, fmaf(dy1, dy1, dx1 * dx1));
  float inv0, inv1;
  int nrej = 0;
  if (__builtin_expect(__any(fminf(r20, r21) <= T_hi), 0)) {
    // reference order r^2 and the exact acceptance test, per sample
    const float e0 = __fadd_rn(__fadd_rn(__fmul_rn(dx0, dx0), __fmul_rn(dy0, dy0)), __fmul_rn(dz0, dz0));
    const float e1 = __fadd_rn(__fadd_rn(__fmul_rn(dx1, dx1), __fmul_rn(dy1, dy1)), __fmul_rn(dz1, dz1));
    const bool ok0 = e0 > T, ok1 = e1 > T;
    inv0 = ok0 ? __builtin_amdgcn_rcpf(e0) : 0.0f;
    inv1 = ok1 ? __builtin_amdgcn_rcpf(e1) : 0.0f;
    nrej = (ok0 ? 0 : 1) + (ok1 ? 0 : 1);
  } else {
    const float q = __builtin_amdgcn_rcpf(r20 * r21);
    inv0 = r21 * q;
    inv1 = r20 * q;
  }
  far_accumulate<KAHAN>(dx0, dy0, inv0, c0, A, C);
  far_accumulate<KAHAN>(dx1, dy1, inv1, c1, A, C);
  rej += nrej;
}

// Four samples: the common path's reciprocals (shared by samples (0, 1) and
// by (2, 3)), then one wave-wide cut-off test for the four (a three-way and a
// two-way minimum) that replaces them with far_pair's exact path when any lane
// is near the cut-off; accumulated in sample order.  Rejections are counted
// only on the exact path.
template <bool KAHAN, bool PERIODIC>
__device__ __forceinline__ void far_quad(float4 a0, float4 c0, float4 a1, float4 c1, float4 a2, float4 c2,
                                         float4 a3, float4 c3, v3 xi, float T, float T_hi, v3 L, float* A,
                                         float* C, int& rej) {
  float dx0, dy0, dz0, dx1, dy1, dz1, dx2, dy2, dz2, dx3, dy3, dz3;
  far_delta<PERIODIC>(xi, a0, L, dx0, dy0, dz0);
  far_delta<PERIODIC>(xi, a1, L, dx1, dy1, dz1);
  far_delta<PERIODIC>(xi, a2, L, dx2, dy2, dz2);
  far_delta<PERIODIC>(xi, a3, L, dx3, dy3, dz3);
  const float r20 = fmaf(dz0, dz0, fmaf(dy0, dy0, dx0 * dx0));
  const float r21 = fmaf(dz1, dz1, fmaf(dy1, dy1, dx1 * dx1));
  const float r22 = fmaf(dz2, dz2, fmaf(dy2, dy2, dx2 * dx2));
  const float r23 = fmaf(dz3, dz3, fmaf(dy3, dy3, dx3 * dx3));
  // common path first (overwritten below when any lane is near the cut-off)
  const float q01 = __builtin_amdgcn_rcpf(r20 * r21);
  const float q23 = __builtin_amdgcn_rcpf(r22 * r23);
  float inv0 = r21 * q01, inv1 = r20 * q01, inv2 = r23 * q23, inv3 = r22 * q23;
  if (__builtin_expect(__any(fminf(fminf(r20, r21), fminf(r22, r23)) <= T_hi), 0)) {
    // reference order r^2 and the exact acceptance test, per sample
    const float e0 = __fadd_rn(__fadd_rn(__fmul_rn(dx0, dx0), __fmul_rn(dy0, dy0)), __fmul_rn(dz0, dz0));
    const float e1 = __fadd_rn(__fadd_rn(__fmul_rn(dx1, dx1), __fmul_rn(dy1, dy1)), __fmul_rn(dz1, dz1));
    const float e2 = __fadd_rn(__fadd_rn(__fmul_rn(dx2, dx2), __fmul_rn(dy2, dy2)), __fmul_rn(dz2, dz2));
    const float e3 = __fadd_rn(__fadd_rn(__fmul_rn(dx3, dx3), __fmul_rn(dy3, dy3)), __fmul_rn(dz3, dz3));
    const bool ok0 = e0 > T, ok1 = e1 > T, ok2 = e2 > T, ok3 = e3 > T;
    inv0 = ok0 ? __builtin_amdgcn_rcpf(e0) : 0.0f;
    inv1 = ok1 ? __builtin_amdgcn_rcpf(e1) : 0.0f;
    inv2 = ok2 ? __builtin_amdgcn_rcpf(e2) : 0.0f;
    inv3 = ok3 ? __builtin_amdgcn_rcpf(e3) : 0.0f;
    rej += (ok0 ? 0 : 1) + (ok1 ? 0 : 1) + (ok2 ? 0 : 1) + (ok3 ? 0 : 1);
  }
  far_accumulate<KAHAN>(dx0, dy0, inv0, c0, A, C);
  far_accumulate<KAHAN>(dx1, dy1, inv1, c1, A, C);
  far_accumulate<KAHAN>(dx2, dy2, inv2, c2, A, C);
  far_accumulate<KAHAN>(dx3, dy3, inv3, c3, A, C);
}

// Sum over sources [j0, j1) four samples per trip (far_quad; a remainder in
// pairs, then one): the next four samples' data are loaded (scalar loads)
// before the current four are evaluated.
template <bool KAHAN, bool PERIODIC>
__device__ __forceinline__ void far_sum_range(const float4* __restrict__ src, int j0, int j1, v3 xi, float T, v3 L,
                                              float* A, float* C, int& rej) {
  const float T_hi = T * (1.0f + 9.5367431640625e-07f);
  int js = j0;
  if (js + 3 < j1) {
    float4 a0 = src[2 * js], c0 = src[2 * js + 1], a1 = src[2 * js + 2], c1 = src[2 * js + 3];
    float4 a2 = src[2 * js + 4], c2 = src[2 * js + 5], a3 = src[2 * js + 6], c3 = src[2 * js + 7];
    for (; js + 3 < j1; js += 4) {
      const int nx = js + 7 < j1 ? js + 4 : js;  // in-range prefetch index
      const float4 na0 = src[2 * nx], nc0 = src[2 * nx + 1], na1 = src[2 * nx + 2], nc1 = src[2 * nx + 3];
      const float4 na2 = src[2 * nx + 4], nc2 = src[2 * nx + 5], na3 = src[2 * nx + 6], nc3 = src[2 * nx + 7];
      far_quad<KAHAN, PERIODIC>(a0, c0, a1, c1, a2, c2, a3, c3, xi, T, T_hi, L, A, C, rej);
      a0 = na0; c0 = nc0; a1 = na1; c1 = nc1;
      a2 = na2; c2 = nc2; a3 = na3; c3 = nc3;
    }
  }
  for (; js + 1 < j1; js += 2)
    far_pair<KAHAN, PERIODIC>(src[2 * js], src[2 * js + 1], src[2 * js + 2], src[2 * js + 3], xi, T, T_hi, L, A, C,
                              rej);
  if (js < j1) far_exact_sample<KAHAN, PERIODIC>(xi, src[2 * js], src[2 * js + 1], T, L, A, C, rej);
}

template <bool KAHAN, bool PERIODIC>
__device__ __forceinline__ void far_sum(const float4* __restrict__ src, int M, v3 xi, float T, v3 L, v3& acc,
                                        int& added) {
  float A[3] = {0.0f, 0.0f, 0.0f}, C[3] = {0.0f, 0.0f, 0.0f};
  int rej = 0;
  far_sum_range<KAHAN, PERIODIC>(src, 0, M, xi, T, L, A, C, rej);
  acc = mk(A[0], A[1], A[2]);
  added = M - rej;
}

// Split form: workgroup (b, s) sums sources [s*chunk, (s+1)*chunk) for the
// targets of block b and stores the partial: (A, rejections) as one float4
// (naive sums), or (A, compensation, rejections) as a FarPartial (Kahan).
// Split 0 marks a target outside the vortex domain (rejections -1), so that
// k_far_combine does not read its position again.  k_far_combine adds the
// partials in split order.
template <bool KAHAN>
__device__ __forceinline__ void far_put_partial(void* part, int64_t idx, const float* A, const float* C, int rej) {
  if (KAHAN) {
    FarPartial& o = ((FarPartial*)part)[idx];
    for (int d = 0; d < 3; ++d) {
      o.a[d] = A[d];
      o.c[d] = C[d];
    }
    o.rej = rej;
  } else {
    ((float4*)part)[idx] = make_float4(A[0], A[1], A[2], __int_as_float(rej));
  }
}

template <bool KAHAN, bool PERIODIC>
__global__ void __launch_bounds__(256) k_far_split(dimcv_state st, dimcv_params p, int32_t bound,
                                                   const float4* __restrict__ src, const int32_t* __restrict__ msrc,
                                                   int cap, int splits, void* __restrict__ part) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  v3 xi = ld3(st.x, i);
  const int sidx = blockIdx.y;
  float A[3] = {0.0f, 0.0f, 0.0f}, C[3] = {0.0f, 0.0f, 0.0f};
  if (!in_vortex_domain(p, xi)) {
    if (sidx == 0) far_put_partial<KAHAN>(part, f, A, C, -1);
    return;
  }
  const int M = min(*msrc, cap);
  const int chunk = (M + splits - 1) / splits;
  const int j0 = min(sidx * chunk, M), j1 = min(j0 + chunk, M);
  int rej = 0;
  far_sum_range<KAHAN, PERIODIC>(src, j0, j1, xi, p.far_r2_max_rejected,
                                 mk(p.box_size[0], p.box_size[1], p.box_size[2]), A, C, rej);
  far_put_partial<KAHAN>(part, (int64_t)sidx * bound + f, A, C, rej);
}

// Combine the partials: naive sums in split order, or Kahan sums carrying
// each partial's compensation.  mode 0: MC far field (scale by S_i, gate into
// v); mode 1: direct sum into v_vort.
template <bool KAHAN>
__global__ void k_far_combine(dimcv_state st, dimcv_params p, int32_t bound, const int32_t* __restrict__ msrc, int cap,
                              int splits, const void* __restrict__ part, const float* __restrict__ near3,
                              const float* __restrict__ scale_s, int mode) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  float A[3] = {0.0f, 0.0f, 0.0f}, C[3] = {0.0f, 0.0f, 0.0f};
  int rej = 0;
  for (int s2 = 0; s2 < splits; ++s2) {
    const int64_t idx = (int64_t)s2 * bound + f;
    float a[3], c[3] = {0.0f, 0.0f, 0.0f};
    int r;
    if (KAHAN) {
      const FarPartial& q = ((const FarPartial*)part)[idx];
      for (int d = 0; d < 3; ++d) {
        a[d] = q.a[d];
        c[d] = q.c[d];
      }
      r = q.rej;
    } else {
      const float4 q = ((const float4*)part)[idx];
      a[0] = q.x;
      a[1] = q.y;
      a[2] = q.z;
      r = __float_as_int(q.w);
    }
    if (r < 0) return;  // outside the vortex domain (split 0's mark)
    for (int d = 0; d < 3; ++d) {
      if (KAHAN) {
        acc_add<true>(A[d], C[d], a[d]);
        acc_add<true>(A[d], C[d], -c[d]);
      } else {
        A[d] += a[d];
      }
    }
    rej += r;
  }
  const int M = min(*msrc, cap);
  v3 acc = mk(A[0], A[1], A[2]);
  if (mode == 0) vv_finish(st, p, i, ld3(near3, f), vmuls(acc, scale_s[f]), M - rej);
  else st3(st.v_vort, i, vadd(ld3(near3, f), vmuls(acc, p.vp_factor)));
}

template <bool KAHAN, bool PERIODIC>
__global__ void __launch_bounds__(256) k_vv_far_fast(dimcv_state st, dimcv_params p, int32_t bound,
                                                     const float* __restrict__ near3,
                                                     const float* __restrict__ scale_s,
                                                     const float4* __restrict__ src, int cap) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  v3 xi = ld3(st.x, i);
  if (!in_vortex_domain(p, xi)) return;
  const int M = min(st.counters[DIMCV_C_NUM_SAMPLES], cap);
  v3 acc;
  int added;
  far_sum<KAHAN, PERIODIC>(src, M, xi, p.far_r2_max_rejected, mk(p.box_size[0], p.box_size[1], p.box_size[2]),
                           acc, added);
  const float S = scale_s[f];
  vv_finish(st, p, i, ld3(near3, f), vmuls(acc, S), added);
}

// ---- direct O(N^2) Biot-Savart ---------------------------------------------
template <bool KAHAN, bool PERIODIC>
__global__ void __launch_bounds__(256) k_direct_f32(dimcv_state st, dimcv_params p, int32_t bound,
                                                    const float* __restrict__ near3,
                                                    const float4* __restrict__ src,
                                                    const int32_t* __restrict__ nsrc) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  v3 xi = ld3(st.x, i);
  if (!in_vortex_domain(p, xi)) return;
  v3 acc;
  int added;
  far_sum<KAHAN, PERIODIC>(src, *nsrc, xi, p.far_r2_max_rejected, mk(p.box_size[0], p.box_size[1], p.box_size[2]),
                           acc, added);
  st3(st.v_vort, i, vadd(ld3(near3, f), vmuls(acc, p.vp_factor)));
}

// minimum image of an FP64 separation (periodic box)
__device__ __forceinline__ void wrap_d(const dimcv_params& p, double& rx, double& ry, double& rz) {
  double L0 = p.box_size[0], L1 = p.box_size[1], L2 = p.box_size[2];
  if (rx > 0.5 * L0) rx -= L0; else if (rx < -0.5 * L0) rx += L0;
  if (ry > 0.5 * L1) ry -= L1; else if (ry < -0.5 * L1) ry += L1;
  if (rz > 0.5 * L2) rz -= L2; else if (rz < -0.5 * L2) rz += L2;
}

// FP64 direct sum (near part with the 2-D norm, far part 1/|r|^2, same
// acceptance sets as the f32 paths).
__global__ void __launch_bounds__(256) k_direct_f64(dimcv_state st, dimcv_params p, int32_t bound,
                                                    const float4* __restrict__ src,
                                                    const int32_t* __restrict__ nsrc, double* __restrict__ out,
                                                    double vp_factor_d) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  v3 xi = ld3(st.x, i);
  if (!in_vortex_domain(p, xi)) return;
  const double X = xi.x, Y = xi.y, Z = xi.z;
  double nx = 0.0, ny = 0.0, nz = 0.0;
  const int nn = min(st.nbr_count[f], p.nbr_kmax);
  for (int k = 0; k < nn; ++k) {
    uint32_t e = st.nbr[(int64_t)k * p.n_max + f];
    if (!nb_fluid(e)) continue;
    int64_t j = nb_slot(e);
    double rx = X - (double)st.x[3 * j], ry = Y - (double)st.x[3 * j + 1];
    if (p.periodic) {
      double rz = 0.0;
      wrap_d(p, rx, ry, rz);
    }
    double rn = sqrt(rx * rx + ry * ry);
    double ex = exp(-rn / 1e-6);
    double A = (1.0 - ex) / (rn * (rn + 1e-6 * ex));
    const float* w = st.d_vort_smoothed + 3 * j;
    nx += A * (ry * (double)w[2]);
    ny += A * (-rx * (double)w[2]);
    nz += A * (rx * (double)w[1] - ry * (double)w[0]);
  }
  const int M = *nsrc;
  const float T = p.far_r2_max_rejected;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int js = 0; js < M; ++js) {
    const float4 a = src[2 * js];
    const float4 c = src[2 * js + 1];
    v3 rf = rel(p, xi, mk(a.x, a.y, a.z));
    float r2f = __fadd_rn(__fadd_rn(__fmul_rn(rf.x, rf.x), __fmul_rn(rf.y, rf.y)), __fmul_rn(rf.z, rf.z));
    if (!(r2f > T)) continue;
    double rx = X - (double)a.x, ry = Y - (double)a.y, rz = Z - (double)a.z;
    if (p.periodic) wrap_d(p, rx, ry, rz);
    double inv = 1.0 / (rx * rx + ry * ry + rz * rz);
    fx += inv * (ry * (double)c.z);
    fy += inv * (-rx * (double)c.z);
    fz += inv * (rx * (double)c.y - ry * (double)c.x);
  }
  out[3 * i] = vp_factor_d * (nx + fx);
  out[3 * i + 1] = vp_factor_d * (ny + fy);
  out[3 * i + 2] = vp_factor_d * (nz + fz);
}

// ---- canonical FP64 inclusive scan (block-local part) -----------------------
// 256 threads x 16 consecutive values: l0 serial per thread, l1 serial per 16
// threads over thread totals, L2 serial over the 16 group totals;
// L1[t] = L2[g-1] + l1[t], L[i] = L1[t-1] + l0[i].
// y(k) gives value k of the tile.
template <class Y>
__device__ void block_local_scan_f64(Y&& y, double* __restrict__ L, int64_t m, double* tt, double* l1s,
                                     double* L2s, double* L1s) {
  const int t = threadIdx.x;
  const int nthreads = (int)((m + 15) / 16);
  double l0[16];
  double acc = 0.0;
  if (t < nthreads) {
    int64_t s = (int64_t)t * 16;
    int64_t e = min(s + 16, m);
    acc = y(s);
    l0[0] = acc;
    for (int64_t k = s + 1; k < e; ++k) {
      acc = acc + y(k);
      l0[k - s] = acc;
    }
    tt[t] = acc;
  }
  __syncthreads();
  const int ngroups = (nthreads + 15) / 16;
  if (t < ngroups) {
    int s = t * 16, e = min(s + 16, nthreads);
    double a = tt[s];
    l1s[s] = a;
    for (int k = s + 1; k < e; ++k) {
      a = a + tt[k];
      l1s[k] = a;
    }
    L2s[t] = a;  // group total, scanned serially below
  }
  __syncthreads();
  if (t == 0) {
    double a = L2s[0];
    for (int g = 1; g < ngroups; ++g) {
      a = a + L2s[g];
      L2s[g] = a;
    }
  }
  __syncthreads();
  if (t < nthreads) {
    int g = t / 16;
    L1s[t] = g == 0 ? l1s[t] : L2s[g - 1] + l1s[t];
  }
  __syncthreads();
  if (t < nthreads) {
    int64_t s = (int64_t)t * 16;
    int64_t e = min(s + 16, m);
    for (int64_t k = s; k < e; ++k) L[k] = t == 0 ? l0[k - s] : L1s[t - 1] + l0[k - s];
  }
  __syncthreads();
}

__device__ void block_local_scan_f64(const double* __restrict__ y, double* __restrict__ L, int64_t m,
                                     double* tt, double* l1s, double* L2s, double* L1s) {
  block_local_scan_f64([&](int64_t k) { return y[k]; }, L, m, tt, l1s, L2s, L1s);
}

// weights: in-domain fluid kden (fluid-list order), 0 otherwise
__device__ __forceinline__ double cdf_weight(const dimcv_state& st, const dimcv_params& p, int64_t f) {
  double v = 0.0;
  if (f < st.counters[DIMCV_C_FLUID_LIST]) {
    const int64_t i = st.fluid_list[f];
    if (i >= 0 && in_vortex_domain(p, ld3(st.x, i))) v = (double)st.kernel_density[i];
  }
  return v;
}

__global__ void k_cdf_weights(dimcv_state st, dimcv_params p, int32_t bound, double* __restrict__ w) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound) return;
  w[f] = cdf_weight(st, p, f);
}

// Canonical scan tiles of the CDF weights (cdf_weight), each weight formed
// where the tile scan reads it.
__global__ void __launch_bounds__(256) k_weights_scan_blocks(dimcv_state st, dimcv_params p, double* __restrict__ L,
                                                             int64_t cap, double* __restrict__ btot) {
  __shared__ double tt[256], l1s[256], L2s[16], L1s[256];
  const int64_t n = min((int64_t)st.counters[DIMCV_C_FLUID_LIST], cap);
  const int64_t s = (int64_t)blockIdx.x * DIMCV_SCAN_TILE;
  if (s >= n) return;  // uniform per block
  const int64_t m = min((int64_t)DIMCV_SCAN_TILE, n - s);
  block_local_scan_f64([&](int64_t k) { return cdf_weight(st, p, s + k); }, L + s, m, tt, l1s, L2s, L1s);
  if (threadIdx.x == 0) btot[blockIdx.x] = L[s + m - 1];
}

// Scan of the block totals (at most DIMCV_MAX_SCAN_BLOCKS of them: one block-
// local scan), then out = P[b-1] + L for b > 0.
__global__ void __launch_bounds__(256) k_scan_totals(const double* __restrict__ btot, double* __restrict__ bpre,
                                                     const int32_t* __restrict__ nptr, int64_t cap) {
  __shared__ double tt[256], l1s[256], L2s[16], L1s[256];
  const int64_t n = min((int64_t)*nptr, cap);
  const int64_t nb = (n + DIMCV_SCAN_TILE - 1) / DIMCV_SCAN_TILE;
  if (nb <= 0) return;
  block_local_scan_f64(btot, bpre, nb, tt, l1s, L2s, L1s);
}

__global__ void k_scan_add(double* __restrict__ L, const double* __restrict__ bpre, const int32_t* __restrict__ nptr,
                           int64_t cap, double* __restrict__ total) {
  const int64_t n = min((int64_t)*nptr, cap);
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t b = i / DIMCV_SCAN_TILE;
  if (b > 0) L[i] = bpre[b - 1] + L[i];
  if (i == n - 1) *total = L[i];
}

__global__ void k_cdf_draw(dimcv_state st, const double* __restrict__ cdf, int64_t cdf_cap,
                           const double* __restrict__ u, int32_t m, int64_t zero_cap) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t F = min((int64_t)st.counters[DIMCV_C_FLUID_LIST], cdf_cap);
  const double total = F > 0 ? cdf[F - 1] : 0.0;
  const bool ok = F > 0 && total > 0.0;
  if (k == 0) st.counters[DIMCV_C_NUM_SAMPLES] = ok ? m : 0;
  if (k >= zero_cap) return;
  if (k >= m || !ok) {
    st.sample_idx[k] = 0;
    return;
  }
  const double t = u[k] * total;
  int64_t lo = 0, hi = F;  // first cdf > t
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (cdf[mid] > t) hi = mid;
    else lo = mid + 1;
  }
  if (lo >= F) {  // first cdf >= total
    int64_t a = 0, b = F;
    while (a < b) {
      int64_t mid = (a + b) >> 1;
      if (cdf[mid] >= total) b = mid;
      else a = mid + 1;
    }
    lo = a;
  }
  st.sample_idx[k] = st.fluid_list[lo];
}

// ---- vortex rings ------------------------------------------------------------
__global__ void k_ring_velocity(int64_t npts, const double* __restrict__ pts, double inv_s2, int64_t n,
                                const float* __restrict__ x, float* __restrict__ vel) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double X0 = x[3 * i], X1 = x[3 * i + 1], X2 = x[3 * i + 2];
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  for (int64_t k = 0; k < npts; ++k) {
    const double* q = pts + 7 * k;
    double rx = X0 - q[0], ry = X1 - q[1], rz = X2 - q[2];
    double r2 = rx * rx + ry * ry + rz * rz;
    if (r2 <= 0.0) continue;
    double core = 1.0 - exp(-r2 * inv_s2);
    double f = q[6] * core / (r2 * sqrt(r2));
    u0 += f * (q[4] * rz - q[5] * ry);
    u1 += f * (q[5] * rx - q[3] * rz);
    u2 += f * (q[3] * ry - q[4] * rx);
  }
  vel[3 * i] = (float)u0;
  vel[3 * i + 1] = (float)u1;
  vel[3 * i + 2] = (float)u2;
}

struct PredDomainFluid {
  const int32_t* material;
  const float* x;
  const int32_t* n;
  dimcv_params p;
  __device__ bool operator()(int64_t i) const {
    return i < *n && material[i] == 1 && in_vortex_domain(p, ld3(x, i));
  }
};

template <class Pred>
__global__ void __launch_bounds__(256) k_compact_count2(Pred pred, int64_t n, uint32_t* __restrict__ cnt) {
  __shared__ uint32_t wsum[4];
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  bool f = i < n && pred(i);
  uint64_t b = __ballot(f);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) cnt[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

template <class Pred>
__global__ void __launch_bounds__(256) k_compact_write2(Pred pred, int64_t n, const uint32_t* __restrict__ off,
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

// Source splits so that a small target set still fills the GPU
// (far_split_count).
static int far_splits(int64_t targets, const Workspace& ws) {
  const int s = far_split_count(targets);
  // never binds: far_partial_cap >= min(16 n_max, 2^21) > s * targets
  return (int64_t)s * targets > ws.far_partial_cap ? 1 : s;
}

// far_split_count_mc as far as the partials fit the workspace.
static int far_splits_mc(int64_t targets, const Workspace& ws) {
  int64_t s = far_split_count_mc(targets);
  const int64_t fit = targets > 0 ? ws.far_partial_cap / targets : 1;
  if (s > fit) s = fit;
  return s < 1 ? 1 : (int)s;
}

static int launch_far_split(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, const float4* src,
                            const int32_t* msrc, int cap, int splits, bool kahan, Workspace& ws, int mode,
                            hipStream_t s) {
  void* part = ws.far_partials;
  dim3 g(nblocks(fluid_bound), splits);
  auto kern = kahan ? (p->periodic ? k_far_split<true, true> : k_far_split<true, false>)
                    : (p->periodic ? k_far_split<false, true> : k_far_split<false, false>);
  hipLaunchKernelGGL(kern, g, dim3(256), 0, s, *st, *p, fluid_bound, src, msrc, cap, splits, part);
  auto comb = kahan ? k_far_combine<true> : k_far_combine<false>;
  hipLaunchKernelGGL(comb, dim3(nblocks(fluid_bound)), dim3(256), 0, s, *st, *p, fluid_bound, msrc, cap, splits,
                     (const void*)part, ws.near3, ws.scale_s, mode);
  DIMCV_LAUNCHED("far field (split)");
  return DIMCV_OK;
}

// ---- batched scenes ------------------------------------------------------------
__device__ __forceinline__ int scene_of_sample(const dimcv_state& st, int64_t g) {
  int lo = 0, hi = st.n_scenes;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (st.scenes[mid].sample_base <= g) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

// Canonical scan tiles of every scene's CDF weights: tile t of scene s covers
// entries [4096 t, 4096 t + 4096) of the scene's segment, as in a solo scan.
__global__ void __launch_bounds__(256) k_batch_scan_tiles(dimcv_state st, const double* __restrict__ w,
                                                          double* __restrict__ L, double* __restrict__ btot,
                                                          const int32_t* __restrict__ misc) {
  __shared__ double tt[256], l1s[256], L2s[16], L1s[256];
  const int64_t T = blockIdx.x;
  if (T >= misc[BM_TILES]) return;  // uniform per block
  const int s = last_le(st.scene_counters + DIMCV_SC_TILE_START, st.n_scenes, DIMCV_SC_COUNT, T);
  const int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
  const int64_t t = T - c[DIMCV_SC_TILE_START];
  const int64_t start = c[DIMCV_SC_FLUID_START] + t * DIMCV_SCAN_TILE;
  const int64_t m = min((int64_t)DIMCV_SCAN_TILE, (int64_t)c[DIMCV_SC_FLUID] - t * DIMCV_SCAN_TILE);
  block_local_scan_f64(w + start, L + start, m, tt, l1s, L2s, L1s);
  if (threadIdx.x == 0) btot[T] = L[start + m - 1];
}

// Scan of each scene's tile totals (one workgroup per scene).
__global__ void __launch_bounds__(256) k_batch_scan_totals(dimcv_state st, const double* __restrict__ btot,
                                                           double* __restrict__ bpre) {
  __shared__ double tt[256], l1s[256], L2s[16], L1s[256];
  const int32_t* c = st.scene_counters + (int64_t)blockIdx.x * DIMCV_SC_COUNT;
  const int64_t nt = ((int64_t)c[DIMCV_SC_FLUID] + DIMCV_SCAN_TILE - 1) / DIMCV_SCAN_TILE;
  if (nt <= 0 || nt > DIMCV_MAX_SCAN_BLOCKS) return;  // scenes are limited to 4096^2 particles
  block_local_scan_f64(btot + c[DIMCV_SC_TILE_START], bpre + c[DIMCV_SC_TILE_START], nt, tt, l1s, L2s, L1s);
}

__global__ void k_batch_scan_add(dimcv_state st, double* __restrict__ L, const double* __restrict__ bpre,
                                 int64_t bound) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= bound || q >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int s = st.block_scene[q >> 8];
  if (s < 0) return;
  const int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
  const int64_t lf = q - c[DIMCV_SC_FLUID_START];
  if (lf >= c[DIMCV_SC_FLUID]) return;
  const int64_t t = lf / DIMCV_SCAN_TILE;
  if (t > 0) L[q] = bpre[c[DIMCV_SC_TILE_START] + t - 1] + L[q];
  if (lf == c[DIMCV_SC_FLUID] - 1) st.scene_dscal[(int64_t)s * DIMCV_SD_COUNT + DIMCV_SD_CDF_TOTAL] = L[q];
}

// Inverse-CDF draws of every scene (sample g of the batch is draw
// g - sample_base of its scene), then the packed far-field source of the
// draw (k_pack_sources, weighted).
__global__ void k_batch_cdf_draw(dimcv_state st, const double* __restrict__ L, const double* __restrict__ u,
                                 int64_t m_total, float4* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= m_total) return;
  const int s = scene_of_sample(st, g);
  if (s < 0) return;
  const dimcv_scene& sc = st.scenes[s];
  const int64_t k = g - sc.sample_base;
  if (k >= sc.m) return;
  int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
  const int64_t F = c[DIMCV_SC_FLUID];
  const double* cdf = L + c[DIMCV_SC_FLUID_START];
  const double total = F > 0 ? cdf[F - 1] : 0.0;
  const bool ok = F > 0 && total > 0.0;
  if (k == 0) c[DIMCV_SC_NUM_SAMPLES] = ok ? sc.m : 0;
  if (!ok) {
    st.sample_idx[g] = c[DIMCV_SC_SLOT_START];
    return;
  }
  const double t = u[g] * total;
  int64_t lo = 0, hi = F;  // first cdf > t
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (cdf[mid] > t) hi = mid;
    else lo = mid + 1;
  }
  if (lo >= F) {  // first cdf >= total
    int64_t a = 0, b = F;
    while (a < b) {
      int64_t mid = (a + b) >> 1;
      if (cdf[mid] >= total) b = mid;
      else a = mid + 1;
    }
    lo = a;
  }
  const int64_t slot = st.fluid_list[c[DIMCV_SC_FLUID_START] + lo];
  st.sample_idx[g] = (int32_t)slot;
  v3 w = ld3(st.d_vort_smoothed, slot);
  const float den = st.kernel_density[slot] + 1e-8f;
  w = mk(w.x / den, w.y / den, w.z / den);
  out[2 * g] = make_float4(st.x[3 * slot], st.x[3 * slot + 1], st.x[3 * slot + 2], 0.0f);
  out[2 * g + 1] = make_float4(w.x, w.y, w.z, 0.0f);
}

// Persistent far field over a work queue.  Work item q = (scene, 256-target
// block) in the order set up by the scene set-up (largest sample count
// first); a workgroup takes the next item until none is left.  A target's
// sum is the one of the scene run alone: k_vv_far_fast (splits = 1), or
// k_far_split + k_far_combine (splits > 1), whose source ranges are summed
// here one after another in the same 64-lane waves and added in split order
// in registers.  A batch fills the GPU with its scenes' target blocks, so it
// needs no partial sums and no combine pass.
template <bool KAHAN>
__global__ void __launch_bounds__(256) k_batch_far(dimcv_state st, dimcv_params p, const float* __restrict__ near3,
                                                   const float* __restrict__ scale_s, const float4* __restrict__ src,
                                                   int32_t* __restrict__ misc, const int32_t* __restrict__ order,
                                                   const int32_t* __restrict__ item_start) {
  __shared__ int q_sh;
  const int total = __builtin_amdgcn_readfirstlane(misc[BM_ITEMS]);
  const float T = p.far_r2_max_rejected;
  const v3 L0 = mk(0.0f, 0.0f, 0.0f);
  // The queue grab and its barriers must stay in uniform control flow: the
  // per-target work below is an if-region (no continue / return), so every
  // wave reaches the next grab together.
  for (;;) {
    if (threadIdx.x == 0) q_sh = atomicAdd(&misc[BM_QUEUE], 1);
    __syncthreads();
    const int q = __builtin_amdgcn_readfirstlane(q_sh);
    __syncthreads();
    if (q >= total) break;
    const int r = last_le(item_start, st.n_scenes, 1, q);
    const int s = order[r];
    const int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
    const int tb = q - item_start[r];
    const int splits = c[DIMCV_SC_SPLITS];
    const int F = c[DIMCV_SC_FLUID];
    const int lf = tb * 256 + (int)threadIdx.x;
    if (lf < F) {
      const int64_t f = (int64_t)c[DIMCV_SC_FLUID_START] + lf;
      const int64_t i = st.fluid_list[f];
      const v3 xi = ld3(st.x, i);
      if (in_vortex_domain(p, xi)) {
        const int M = c[DIMCV_SC_NUM_SAMPLES];
        const float4* sp = src + 2 * (int64_t)st.scenes[s].sample_base;
        if (splits == 1) {
          v3 acc;
          int added;
          far_sum<KAHAN, false>(sp, M, xi, T, L0, acc, added);
          vv_finish(st, p, i, ld3(near3, f), vmuls(acc, scale_s[f]), added);
        } else {
          // the solo run's source splits one after another, each summed
          // from fresh accumulators, then added in split order exactly as
          // k_far_combine adds the stored partials
          const int chunk = (M + splits - 1) / splits;
          float A[3] = {0.0f, 0.0f, 0.0f}, C[3] = {0.0f, 0.0f, 0.0f};
          int rej_all = 0;
          for (int k = 0; k < splits; ++k) {
            const int j0 = min(k * chunk, M), j1 = min(j0 + chunk, M);
            float a[3] = {0.0f, 0.0f, 0.0f}, cc[3] = {0.0f, 0.0f, 0.0f};
            int rej = 0;
            far_sum_range<KAHAN, false>(sp, j0, j1, xi, T, L0, a, cc, rej);
            for (int d = 0; d < 3; ++d) {
              if (KAHAN) {
                acc_add<true>(A[d], C[d], a[d]);
                acc_add<true>(A[d], C[d], -cc[d]);
              } else {
                A[d] += a[d];
              }
            }
            rej_all += rej;
          }
          vv_finish(st, p, i, ld3(near3, f), vmuls(mk(A[0], A[1], A[2]), scale_s[f]), M - rej_all);
        }
      }
    }
  }
}

static int g_far_wgs = 2048;
void batch_set_far_workgroups(int w) { g_far_wgs = w > 0 ? w : 2048; }

// CDF of every scene, draws and packed sources.
int batch_cdf_launch(const dimcv_state* st, const dimcv_params* p, Workspace& ws, BatchWs& bw, const double* u,
                     int64_t m_total, hipStream_t s) {
  const int64_t n = p->n_max;
  hipLaunchKernelGGL(k_cdf_weights, dim3(nblocks(n)), dim3(256), 0, s, *st, *p, (int32_t)n, ws.cdf_weights);
  hipLaunchKernelGGL(k_batch_scan_tiles, dim3((unsigned)bw.tile_bound), dim3(256), 0, s, *st, ws.cdf_weights,
                     ws.cdf, bw.btot, bw.misc);
  hipLaunchKernelGGL(k_batch_scan_totals, dim3((unsigned)st->n_scenes), dim3(256), 0, s, *st, bw.btot, bw.bpre);
  hipLaunchKernelGGL(k_batch_scan_add, dim3(nblocks(n)), dim3(256), 0, s, *st, ws.cdf, bw.bpre, n);
  if (m_total > 0)
    hipLaunchKernelGGL(k_batch_cdf_draw, dim3(nblocks(m_total)), dim3(256), 0, s, *st, ws.cdf, u, m_total,
                       ws.samples);
  DIMCV_LAUNCHED("batched CDF");
  return DIMCV_OK;
}

// vort2vel of every scene: near field + S_i, then the far field.
int batch_vort2vel_launch(const dimcv_state* st, const dimcv_params* p, Workspace& ws, BatchWs& bw, hipStream_t s) {
  const int64_t n = p->n_max;
  hipLaunchKernelGGL(k_vv_near, dim3(nblocks(n)), dim3(256), 0, s, *st, *p, (int32_t)n, ws.near3, ws.scale_s);
  if (p->far_mode == DIMCV_FAR_EXACT) {
    hipLaunchKernelGGL(k_vv_far_exact, dim3(nblocks(n)), dim3(256), 0, s, *st, *p, (int32_t)n, ws.near3, ws.scale_s);
  } else {
    const bool kahan = p->far_mode == DIMCV_FAR_KAHAN;
    auto far = kahan ? k_batch_far<true> : k_batch_far<false>;
    hipLaunchKernelGGL(far, dim3((unsigned)g_far_wgs), dim3(256), 0, s, *st, *p, ws.near3, ws.scale_s, ws.samples,
                       bw.misc, bw.order, bw.item_start);
  }
  DIMCV_LAUNCHED("batched vort2vel");
  return DIMCV_OK;
}

// ---- inverse-CDF sampling (B3) ---------------------------------------------
static int cdf_check(const dimcv_params* p, int32_t m, int32_t fluid_bound, const char* who) {
  if (m < 0 || m > p->n_max) {
    set_error("%s: %d samples outside [0, n_max]", who, m);
    return DIMCV_ERR_RANGE;
  }
  if (fluid_bound < 0 || (int64_t)fluid_bound > (int64_t)DIMCV_SCAN_TILE * DIMCV_MAX_SCAN_BLOCKS ||
      fluid_bound > p->n_max) {
    set_error("%s: fluid bound %d outside [0, min(n_max, scan limit)]", who, fluid_bound);
    return DIMCV_ERR_RANGE;
  }
  return DIMCV_OK;
}

// Weights, canonical FP64 scan and draws into caller-chosen scratch of at
// least fluid_bound doubles (weights, cdf) and fluid_bound / 4096 + 16
// (block totals, prefixes); every kernel reads at most fluid_bound entries.
static int cdf_launch(const dimcv_state* st, const dimcv_params* p, const double* u, int32_t m,
                      int32_t fluid_bound, double* weights, double* cdf, double* btot, double* bpre,
                      hipStream_t s) {
  const int32_t* nF = st->counters + DIMCV_C_FLUID_LIST;
  const int64_t cap = fluid_bound;
  if (fluid_bound > 0) {
    (void)weights;
    unsigned nb = nblocks(fluid_bound, DIMCV_SCAN_TILE);
    hipLaunchKernelGGL(k_weights_scan_blocks, dim3(nb), dim3(256), 0, s, *st, *p, cdf, cap, btot);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(256), 0, s, btot, bpre, nF, cap);
    hipLaunchKernelGGL(k_scan_add, dim3(nblocks(fluid_bound)), dim3(256), 0, s, cdf, bpre, nF, cap,
                       st->dscal + DIMCV_D_CDF_TOTAL);
    DIMCV_LAUNCHED("canonical scan");
  }
  const int64_t zero_cap = p->n_max;
  hipLaunchKernelGGL(k_cdf_draw, dim3(nblocks(zero_cap)), dim3(256), 0, s, *st, cdf, cap, u, m, zero_cap);
  DIMCV_LAUNCHED("k_cdf_draw");
  return DIMCV_OK;
}

// Scratch of dimcv_cdf_sample_noalloc, in place of a workspace: FP64
// weights and CDF for up to kScratchFluid fluid particles plus the scan's
// block totals and prefixes (16 MiB and 4 KiB).  Every GPU the library runs on holds
// one copy, at its own address, until the process ends; host code uses the
// address hipGetSymbolAddress returns for the current device (the array's
// own name is not a device address on the host).  All calls on one GPU share
// it: calls that overlap on one GPU race.  Nothing in it needs clearing: each
// call writes every entry it reads.
constexpr int64_t kScratchFluid = DIMCV_NOALLOC_MAX_FLUID;
constexpr int64_t kScratchBlocks = kScratchFluid / DIMCV_SCAN_TILE + 16;
constexpr size_t kScratchBytes = 8 * (size_t)(2 * kScratchFluid + 2 * kScratchBlocks);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The scratch address on the current device, looked up once per device.
static int scratch_base(unsigned char** out) {
  int dev = -1;
  DIMCV_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("dimcv_cdf_sample_noalloc: device %d beyond the %d the library tracks", dev, kMaxDevices);
    return DIMCV_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* q = nullptr;
    hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || q == nullptr) {
      set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
      return DIMCV_ERR_HIP;
    }
    g_scratch_ptr[dev] = (unsigned char*)q;
  }
  *out = g_scratch_ptr[dev];
  return DIMCV_OK;
}

}  // namespace dimcv

using namespace dimcv;

extern "C" {

int dimcv_cdf_sample(const dimcv_state* st, const dimcv_params* p, const double* u, int32_t m,
                     int32_t fluid_bound, void* stream) {
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  rc = cdf_check(p, m, fluid_bound, "dimcv_cdf_sample");
  if (rc) return rc;
  return cdf_launch(st, p, u, m, fluid_bound, ws.cdf_weights, ws.cdf, ws.block_tot, ws.block_pre,
                    (hipStream_t)stream);
}

int dimcv_cdf_sample_noalloc(const dimcv_state* st, const dimcv_params* p, const double* u, int32_t m,
                             int32_t fluid_bound, void* stream) {
  if (!st || !p) {
    set_error("dimcv_cdf_sample_noalloc: null state or params");
    return DIMCV_ERR_ARG;
  }
  int rc = cdf_check(p, m, fluid_bound, "dimcv_cdf_sample_noalloc");
  if (rc) return rc;
  if ((int64_t)fluid_bound > kScratchFluid) {
    set_error("dimcv_cdf_sample_noalloc: %d fluid particles exceed the scratch maximum %lld "
              "(use dimcv_cdf_sample with a workspace)", fluid_bound, (long long)kScratchFluid);
    return DIMCV_ERR_RANGE;
  }
  unsigned char* base = nullptr;
  rc = scratch_base(&base);
  if (rc) return rc;
  double* weights = (double*)base;
  double* cdf = weights + kScratchFluid;
  double* btot = cdf + kScratchFluid;
  double* bpre = btot + kScratchBlocks;
  return cdf_launch(st, p, u, m, fluid_bound, weights, cdf, btot, bpre, (hipStream_t)stream);
}

int dimcv_vort2vel_near(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  if (fluid_bound <= 0) return DIMCV_OK;
  hipLaunchKernelGGL(k_vv_near, dim3(nblocks(fluid_bound)), dim3(256), 0, s, *st, *p, fluid_bound, ws.near3,
                     ws.scale_s);
  DIMCV_LAUNCHED("k_vv_near");
  return DIMCV_OK;
}

}  // extern "C"

namespace dimcv {
// Optional HIP events recorded around the far-field kernel launch (used by
// dimcv_substep's timing option; null when not timing).
static hipEvent_t g_far_ev[2] = {nullptr, nullptr};
// A timing event on s; inside a stream capture it becomes an event-record
// node of the graph, recorded at each replay.
static hipError_t record_far_event(hipEvent_t ev, hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive)
    return hipEventRecordWithFlags(ev, s, hipEventRecordExternal);
  return hipEventRecord(ev, s);
}
void set_far_timing(hipEvent_t a, hipEvent_t b) {
  g_far_ev[0] = a;
  g_far_ev[1] = b;
}
}  // namespace dimcv

extern "C" {

int dimcv_vort2vel_far(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, int32_t max_samples,
                       void* stream) {
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, p->far_mode == DIMCV_FAR_EXACT ? 0 : max_samples);
  if (rc) return rc;
  if (fluid_bound <= 0) return DIMCV_OK;
  if (p->far_mode != DIMCV_FAR_EXACT && max_samples < 1) {
    set_error("dimcv_vort2vel: max_samples must bound num_samples (use n_max)");
    return DIMCV_ERR_ARG;
  }
  unsigned nb = nblocks(fluid_bound);
  if (p->far_mode == DIMCV_FAR_EXACT) {
    hipLaunchKernelGGL(k_vv_far_exact, dim3(nb), dim3(256), 0, s, *st, *p, fluid_bound, ws.near3, ws.scale_s);
  } else {
    hipLaunchKernelGGL(k_pack_sources, dim3(nblocks(max_samples)), dim3(256), 0, s, *st, st->sample_idx,
                       st->counters + DIMCV_C_NUM_SAMPLES, (int64_t)max_samples, 1, ws.samples);
    const int splits = far_splits_mc(fluid_bound, ws);
    const bool kahan = p->far_mode == DIMCV_FAR_KAHAN;
    if (g_far_ev[0]) DIMCV_HIP(record_far_event(g_far_ev[0], s));
    if (splits == 1) {
      auto kern = kahan ? (p->periodic ? k_vv_far_fast<true, true> : k_vv_far_fast<true, false>)
                        : (p->periodic ? k_vv_far_fast<false, true> : k_vv_far_fast<false, false>);
      hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, s, *st, *p, fluid_bound, ws.near3, ws.scale_s, ws.samples,
                         (int)max_samples);
    } else {
      rc = launch_far_split(st, p, fluid_bound, ws.samples, st->counters + DIMCV_C_NUM_SAMPLES, (int)max_samples,
                            splits, kahan, ws, 0, s);
      if (rc) return rc;
    }
    if (g_far_ev[1]) DIMCV_HIP(record_far_event(g_far_ev[1], s));
  }
  DIMCV_LAUNCHED("vort2vel far field");
  return DIMCV_OK;
}

int32_t dimcv_far_split_count(int64_t targets) { return targets < 0 ? -1 : far_split_count_mc(targets); }

int dimcv_vort2vel(const dimcv_state* st, const dimcv_params* p, int32_t fluid_bound, int32_t max_samples,
                   void* stream) {
  int rc = dimcv_vort2vel_near(st, p, fluid_bound, stream);
  if (rc) return rc;
  return dimcv_vort2vel_far(st, p, fluid_bound, max_samples, stream);
}

int dimcv_direct_bs(const dimcv_state* st, const dimcv_params* p, int32_t precision, double* out_f64,
                    int32_t fluid_bound, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  int rc = bind_workspace(st, p, &ws, p->n_max);
  if (rc) return rc;
  if (precision < 0 || precision > 2 || (precision == 2 && !out_f64) || (precision < 2 && !st->v_vort)) {
    set_error("dimcv_direct_bs: bad precision/output");
    return DIMCV_ERR_ARG;
  }
  if (fluid_bound <= 0) return DIMCV_OK;
  // sources: in-domain fluid slots, ascending
  int64_t n = p->n_max;
  unsigned nb256 = nblocks(n);
  int32_t* nsrc = st->counters + DIMCV_C_SAMPLE_SLOTS;
  int32_t* list = (int32_t*)ws.keys[0];
  PredDomainFluid pr{st->material, st->x, st->counters + DIMCV_C_PARTICLE_NUM, *p};
  hipLaunchKernelGGL(k_compact_count2<PredDomainFluid>, dim3(nb256), dim3(256), 0, s, pr, n, ws.compact_cnt);
  rc = exclusive_scan_u32(ws.compact_cnt, ws.compact_cnt, nb256, ws.scan_tmp, (uint32_t*)nsrc, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_compact_write2<PredDomainFluid>, dim3(nb256), dim3(256), 0, s, pr, n, ws.compact_cnt, list);
  hipLaunchKernelGGL(k_pack_sources, dim3(nb256), dim3(256), 0, s, *st, list, nsrc, n, 0, ws.samples);
  DIMCV_LAUNCHED("direct sources");
  unsigned nb = nblocks(fluid_bound);
  if (precision == 2) {
    hipLaunchKernelGGL(k_direct_f64, dim3(nb), dim3(256), 0, s, *st, *p, fluid_bound, ws.samples, nsrc, out_f64,
                       p->vp_factor_f64);
  } else {
    hipLaunchKernelGGL(k_vv_near, dim3(nb), dim3(256), 0, s, *st, *p, fluid_bound, ws.near3, ws.scale_s);
    const int splits = far_splits(fluid_bound, ws);
    if (splits == 1) {
      auto kern = precision == 1 ? (p->periodic ? k_direct_f32<true, true> : k_direct_f32<true, false>)
                                 : (p->periodic ? k_direct_f32<false, true> : k_direct_f32<false, false>);
      hipLaunchKernelGGL(kern, dim3(nb), dim3(256), 0, s, *st, *p, fluid_bound, ws.near3, ws.samples, nsrc);
    } else {
      rc = launch_far_split(st, p, fluid_bound, ws.samples, nsrc, (int)p->n_max, splits, precision == 1, ws, 1, s);
      if (rc) return rc;
    }
  }
  DIMCV_LAUNCHED("direct Biot-Savart");
  return DIMCV_OK;
}

int dimcv_ring_velocity(int64_t npts, const double* pts, double inv_sigma2, int64_t n, const float* x, float* vel,
                        void* stream) {
  if (n <= 0) return DIMCV_OK;
  hipLaunchKernelGGL(k_ring_velocity, dim3(nblocks(n)), dim3(256), 0, (hipStream_t)stream, npts, pts, inv_sigma2,
                     n, x, vel);
  DIMCV_LAUNCHED("k_ring_velocity");
  return DIMCV_OK;
}

}  // extern "C"
