// The segment solver's O(N^2) regularised Biot-Savart sums
// (segment_solver.py:378-418 _compute_endpoint_velocity_bs and :468-521
// _compute_velocity_at_factor) and its RK4 endpoint update (:458-466,
// :523-542), on the direct-sum pattern of the particle comparator: a prologue
// packs each source segment once (centre, vortex vector, active flag) into a
// table of two float4; every lane of a wave then walks the same source at the
// same time (wave-uniform scalar loads, no LDS staging), one target segment
// (both endpoints) per lane.
//
// For target i, over sources j != i that are active, in ascending j:
//   u(q) = u_inf + sum_j inv4pi (w_j x r) / (|r|^2 + R^2)^(3/2) + 1e-12),
//   r = q - c_j, c_j = 0.5 (x-_j + x+_j),
//   w_j = (gamma_j L_j) (d_j / L_j) with L_j = |d_j| + 1e-8 (endpoint form) or
//   w_j = gamma_j d_j (query form, q = x + (factor dt) k), d_j = x+_j - x-_j.
// Every f32 operation is the reference expression's, one IEEE operation each
// (no contraction), so the sums are bitwise the serial ones.
#include "common.h"

namespace dimcv {

__global__ void k_seg_sources(dimcv_segments sg, int form, float4* __restrict__ src) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= sg.n) return;
  const v3 xm = ld3(sg.x_minus, j), xp = ld3(sg.x_plus, j);
  const v3 c = smul(0.5f, vadd(xm, xp));
  const v3 d = vsub(xp, xm);
  const float g = sg.gamma[j];
  v3 w;
  if (form == 0) {
    const float L = vnorm(d) + 1e-8f;
    const v3 t = vdivs(d, L);
    w = smul(g * L, t);
  } else {
    w = smul(g, d);
  }
  src[2 * j] = make_float4(c.x, c.y, c.z, sg.active[j] == 1 ? 1.0f : 0.0f);
  src[2 * j + 1] = make_float4(w.x, w.y, w.z, 0.0f);
}

// inv4pi (w x r) / (r.r + R^2) r2^(1/2) + 1e-12), reference operation order
__device__ __forceinline__ v3 seg_term(v3 w, v3 r, float R2, float inv4pi) {
  const float r2 = vdot(r, r) + R2;
  const float den = r2 * sqrtf(r2) + 1e-12f;
  return vdivs(smul(inv4pi, vcross(w, r)), den);
}

// LANES = 1: one lane per target segment, both endpoints; LANES = 2: lanes
// (2i, 2i + 1) take the minus and plus endpoints of segment i (twice the
// waves, for pools too small to fill the GPU).  Sources in ascending j.
template <int LANES>
__global__ void __launch_bounds__(256) k_seg_bs(dimcv_segments sg, const float4* __restrict__ src, float qscale,
                                                const float* __restrict__ k_minus, const float* __restrict__ k_plus,
                                                float* __restrict__ out_minus, float* __restrict__ out_plus) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t i = LANES == 2 ? t >> 1 : t;
  if (i >= sg.n || sg.active[i] != 1) return;
  const bool do_m = LANES == 1 || (t & 1) == 0, do_p = LANES == 1 || (t & 1) == 1;
  v3 qm = do_m ? ld3(sg.x_minus, i) : mk(0.0f, 0.0f, 0.0f);
  v3 qp = do_p ? ld3(sg.x_plus, i) : mk(0.0f, 0.0f, 0.0f);
  if (k_minus) {
    if (do_m) qm = vadd(qm, smul(qscale, ld3(k_minus, i)));
    if (do_p) qp = vadd(qp, smul(qscale, ld3(k_plus, i)));
  }
  const float R2 = sg.reg_radius * sg.reg_radius;
  const float inv4pi = sg.inv4pi;
  v3 um = mk(sg.u_inf[0], sg.u_inf[1], sg.u_inf[2]), up = um;
  const int32_t n = sg.n;
  for (int32_t j = 0; j < n; ++j) {
    const float4 a = src[2 * j];  // wave-uniform address
    if (a.w == 0.0f) continue;     // inactive source: uniform
    const float4 b = src[2 * j + 1];
    if (j == i) continue;
    const v3 c = mk(a.x, a.y, a.z), w = mk(b.x, b.y, b.z);
    if (LANES == 1) {
      um = vadd(um, seg_term(w, vsub(qm, c), R2, inv4pi));
      up = vadd(up, seg_term(w, vsub(qp, c), R2, inv4pi));
    } else {
      // the lane's endpoint: q is qm or qp, the other one unused
      const v3 q = do_m ? qm : qp;
      um = vadd(um, seg_term(w, vsub(q, c), R2, inv4pi));
    }
  }
  if (LANES == 1) {
    st3(out_minus, i, um);
    st3(out_plus, i, up);
  } else {
    st3(do_m ? out_minus : out_plus, i, um);
  }
}

// segment_solver.py:523-542: x += (dt/6)(k1 + 2 k2 + 2 k3 + k4), gamma *= decay, age += dt
__global__ void k_seg_rk4_update(dimcv_segments sg, float dt6, const float* __restrict__ k1m,
                                 const float* __restrict__ k2m, const float* __restrict__ k3m,
                                 const float* __restrict__ k4m, const float* __restrict__ k1p,
                                 const float* __restrict__ k2p, const float* __restrict__ k3p,
                                 const float* __restrict__ k4p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sg.n || sg.active[i] != 1) return;
  const v3 sm = vadd(vadd(vadd(ld3(k1m, i), smul(2.0f, ld3(k2m, i))), smul(2.0f, ld3(k3m, i))), ld3(k4m, i));
  const v3 sp = vadd(vadd(vadd(ld3(k1p, i), smul(2.0f, ld3(k2p, i))), smul(2.0f, ld3(k3p, i))), ld3(k4p, i));
  st3(sg.x_minus, i, vadd(ld3(sg.x_minus, i), smul(dt6, sm)));
  st3(sg.x_plus, i, vadd(ld3(sg.x_plus, i), smul(dt6, sp)));
  sg.gamma[i] *= sg.gamma_decay;
  sg.age[i] += sg.dt;
}

// segment_solver.py:458-466: k1 = v for active segments
__global__ void k_seg_copy(int32_t n, const int32_t* __restrict__ active, const float* __restrict__ a,
                           float* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || active[i] != 1) return;
  st3(b, i, ld3(a, i));
}

static int seg_check(const dimcv_segments* sg) {
  if (!sg || sg->n < 0 || (sg->n > 0 && (!sg->x_minus || !sg->x_plus || !sg->gamma || !sg->active || !sg->sources))) {
    set_error("dimcv_segment: bad segment arguments");
    return DIMCV_ERR_ARG;
  }
  return DIMCV_OK;
}

static void seg_velocity(const dimcv_segments* sg, int form, float qscale, const float* km, const float* kp,
                         float* om, float* op, hipStream_t s) {
  float4* src = (float4*)sg->sources;
  hipLaunchKernelGGL(k_seg_sources, dim3(nblocks(sg->n)), dim3(256), 0, s, *sg, form, src);
  // two lanes per segment while one per segment gives fewer than four waves
  // per SIMD (4096 waves: 262,144 segments)
  if (sg->n < (1 << 18))
    hipLaunchKernelGGL(k_seg_bs<2>, dim3(nblocks(2 * (int64_t)sg->n)), dim3(256), 0, s, *sg, (const float4*)src,
                       qscale, km, kp, om, op);
  else
    hipLaunchKernelGGL(k_seg_bs<1>, dim3(nblocks(sg->n)), dim3(256), 0, s, *sg, (const float4*)src, qscale, km, kp,
                       om, op);
}

}  // namespace dimcv

using namespace dimcv;

extern "C" {

int dimcv_segment_velocity(const dimcv_segments* sg, int32_t form, float factor, const float* k_minus,
                           const float* k_plus, float* out_minus, float* out_plus, void* stream) {
  int rc = seg_check(sg);
  if (rc) return rc;
  if (form != 0 && form != 1) {
    set_error("dimcv_segment_velocity: form %d is not 0 (endpoints) or 1 (query points)", form);
    return DIMCV_ERR_ARG;
  }
  if (form == 1 && (!k_minus || !k_plus)) {
    set_error("dimcv_segment_velocity: the query form needs k_minus and k_plus");
    return DIMCV_ERR_ARG;
  }
  if (!out_minus || !out_plus) {
    set_error("dimcv_segment_velocity: null output");
    return DIMCV_ERR_ARG;
  }
  if (sg->n == 0) return DIMCV_OK;
  // the query offset factor * dt * k (segment_solver.py:494-495): f32(factor * dt) first
  seg_velocity(sg, form, factor * sg->dt, form == 1 ? k_minus : nullptr, form == 1 ? k_plus : nullptr, out_minus,
               out_plus, (hipStream_t)stream);
  DIMCV_LAUNCHED("k_seg_bs");
  return DIMCV_OK;
}

int dimcv_segment_rk4(const dimcv_segments* sg, const float* v_minus, const float* v_plus, float* k, void* stream) {
  int rc = seg_check(sg);
  if (rc) return rc;
  if (!v_minus || !v_plus || !k) {
    set_error("dimcv_segment_rk4: null argument");
    return DIMCV_ERR_ARG;
  }
  if (sg->n == 0) return DIMCV_OK;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n3 = 3 * (int64_t)sg->n;
  float* km[4] = {k, k + n3, k + 2 * n3, k + 3 * n3};
  float* kp[4] = {k + 4 * n3, k + 5 * n3, k + 6 * n3, k + 7 * n3};
  const unsigned nb = nblocks(sg->n);
  hipLaunchKernelGGL(k_seg_copy, dim3(nb), dim3(256), 0, s, sg->n, sg->active, v_minus, km[0]);
  hipLaunchKernelGGL(k_seg_copy, dim3(nb), dim3(256), 0, s, sg->n, sg->active, v_plus, kp[0]);
  const float fac[3] = {0.5f, 0.5f, 1.0f};
  for (int st = 0; st < 3; ++st)
    seg_velocity(sg, 1, fac[st] * sg->dt, km[st], kp[st], km[st + 1], kp[st + 1], s);
  hipLaunchKernelGGL(k_seg_rk4_update, dim3(nb), dim3(256), 0, s, *sg, sg->dt6, km[0], km[1], km[2], km[3], kp[0],
                     kp[1], kp[2], kp[3]);
  DIMCV_LAUNCHED("segment RK4");
  return DIMCV_OK;
}

}  // extern "C"
