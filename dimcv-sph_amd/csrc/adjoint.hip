// Hand-written adjoints (vector-Jacobian products) of the step's
// kernel-gradient passes and of MC-VVPE, for differentiation through a
// rollout with the discrete structure of each step frozen: neighbour lists,
// importance samples, Jacobi iteration counts and every gate of the forward
// (the DFSPH factor threshold, max(rho_dot, 0) with the 20-neighbour rule,
// max(rho*, rho0), |kappa_i + kappa_j| > eps, the far-field cut-off |r| > h
// and the |v_vort| < |v| update) keep the values they had in the forward.
// Within that structure each pass is smooth and its VJP is exact.
//
// One thread per fluid-list entry i walks i's neighbour list (the forward's
// visiting order); terms of the pair (i, j) that belong to j are added with
// float atomics.  Every pass accumulates (+=) into its adjoint outputs.
// Pair terms depend on x through r = x_i - x_j: r_bar goes to x_i and its
// negative to x_j.
//
// Kernel derivatives: grad W(r) = phi(|r|) r/|r| with phi = dW/ds
// (sph_base.py:46-68), so the Hessian-vector product is
//   H g = phi'(s) (rh.g) rh + (phi(s)/s) (g - (rh.g) rh);
// the product Epanechnikov kernel (dimcv_sph.py:721-736) differentiates
// factor by factor; A(s) = (1 - E)/(s (s + eps E)), E = exp(-s/eps), is the
// kernel of vort2vel (dimcv_sph.py:652-657, 698-706).
#include <algorithm>

#include "common.h"

namespace dimcv {

__device__ __forceinline__ void atomic_add3(float* a, int64_t i, v3 q) {
  atomicAdd(a + 3 * i, q.x);
  atomicAdd(a + 3 * i + 1, q.y);
  atomicAdd(a + 3 * i + 2, q.z);
}
__device__ __forceinline__ void acc3(v3& a, v3 q) { a = vadd(a, q); }
__device__ __forceinline__ v3 vneg(v3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ v3 lds3(const float* __restrict__ a, int64_t i) {
  return a ? ld3(a, i) : mk(0.0f, 0.0f, 0.0f);
}

// phi(s) and phi'(s) of the cubic spline gradient magnitude.
__device__ __forceinline__ bool dw_parts(const dimcv_params& p, float s, float& phi, float& dphi) {
  const float q = s * p.inv_support_radius;
  if (!(s > 1e-5f && q <= 1.0f)) return false;
  const float ih = p.inv_support_radius;
  if (q <= 0.5f) {
    phi = p.kernel_dk * q * (3.0f * q - 2.0f) * ih;
    dphi = p.kernel_dk * (6.0f * q - 2.0f) * ih * ih;
  } else {
    const float f = 1.0f - q;
    phi = -p.kernel_dk * f * f * ih;
    dphi = 2.0f * p.kernel_dk * f * ih * ih;
  }
  return true;
}

// H(r) g for grad W.
__device__ __forceinline__ v3 hess_w(const dimcv_params& p, v3 r, v3 g) {
  const float s = vnorm(r);
  float phi, dphi;
  if (!dw_parts(p, s, phi, dphi)) return mk(0.0f, 0.0f, 0.0f);
  const v3 rh = vdivs(r, s);
  const float rg = vdot(rh, g);
  const float c = phi / s;
  return mk(dphi * rg * rh.x + c * (g.x - rg * rh.x), dphi * rg * rh.y + c * (g.y - rg * rh.y),
            dphi * rg * rh.z + c * (g.z - rg * rh.z));
}

// Epanechnikov product kernel and its gradient.
__device__ __forceinline__ float epan_f(const dimcv_params& p, v3 r, v3* grad) {
  const float q[3] = {fabsf(r.x) * p.inv_support_radius, fabsf(r.y) * p.inv_support_radius,
                      fabsf(r.z) * p.inv_support_radius};
  const float rc[3] = {r.x, r.y, r.z};
  float e[3], de[3];
  for (int d = 0; d < 3; ++d) {
    if (q[d] < 1.0f) {
      e[d] = 0.75f * (1.0f - q[d] * q[d]);
      de[d] = -1.5f * rc[d] * p.inv_support_radius * p.inv_support_radius;
    } else {
      e[d] = 0.0f;
      de[d] = 0.0f;
    }
  }
  if (grad) *grad = mk(de[0] * e[1] * e[2], e[0] * de[1] * e[2], e[0] * e[1] * de[2]);
  return e[0] * e[1] * e[2];
}

// A(s) and A'(s).
__device__ __forceinline__ float a_kernel(const dimcv_params& p, float s, float& dA) {
  const float E = expf(-s * p.inv_eps);
  const float N = 1.0f - E, D = s * (s + p.eps * E);
  const float dN = E * p.inv_eps, dD = 2.0f * s + p.eps * E - s * E;
  dA = (dN * D - N * dD) / (D * D);
  return N / D;
}

#define ADJ_ENTRY()                                                    \
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;   \
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;     \
  const int64_t i = st.fluid_list[f];                                  \
  if (i < 0) return;                                                   \
  const int nn = min(st.nbr_count[f], p.nbr_kmax);                     \
  const uint32_t* __restrict__ nl = st.nbr + f;                        \
  const int64_t ns = p.n_max;                                          \
  const v3 xi = ld3(st.x, i)

#define FOR_FLUID_NBR(k, j) \
  FOR_NBR(k, e_##k, j)         \
  if (nb_fluid(e_##k))

// rho_i = rho0 (mV_i W(0) + sum_j mV_j W(|r_ij|))            (dimcv_sph.py:126-148)
// in: g_in = rho_bar.  out: x_bar.
__global__ void k_adj_density(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  const float rb = a.g_in[i];
  if (rb == 0.0f) return;
  v3 xb = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const v3 r = vsub(xi, ld3(st.x, j));
    const v3 rbar = smul(p.density0 * st.m_V[j] * rb, cubic_dW(p, r));
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomic_add3(a.x_bar, i, xb);
}

// alpha_i = 1 / (sum_j |m_j gW_ij|^2 + |sum_j m_j gW_ij|^2), 0 below 1e-6
// (dimcv_sph.py:86-124).  in: f0 = alpha (forward), g_in = alpha_bar.
// out: x_bar.
__global__ void k_adj_factor(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  const float al = a.f0[i];
  const float ab = a.g_in[i];
  if (al == 0.0f || ab == 0.0f) return;
  const float sbar = -ab * al * al;
  v3 G = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) { acc3(G, smul(st.m[j], cubic_dW(p, vsub(xi, ld3(st.x, j))))); }
  v3 xb = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const v3 r = vsub(xi, ld3(st.x, j));
    const v3 g = cubic_dW(p, r);
    const float mj = st.m[j];
    const v3 gbar = vadd(smul(2.0f * mj * mj * sbar, g), smul(2.0f * mj * sbar, G));
    const v3 rbar = hess_w(p, r, gbar);
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomic_add3(a.x_bar, i, xb);
}

// mode 0: b_i = [b_i > 0] sum_j mV_j (v_i - v_j).gW_ij   (dimcv_sph.py:337-374)
//   in: f0 = v, f1 = b (forward, gives the gate), g_in = b_bar.
// mode 1: da_i = max(rho_i + dt sum_j m_j (v_i - v_j).gW_ij, rho0) (:470-493)
//   in: f0 = v, f1 = da (forward), g_in = da_bar.  out also b0 = rho_bar.
// out: x_bar, b1 = v_bar.
__global__ void k_adj_div(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  const float gb = a.g_in[i];
  bool on;
  float ub;
  if (a.mode == 0) {
    on = a.f1[i] > 0.0f;
    ub = gb;
  } else {
    on = a.f1[i] > p.density0;
    ub = p.dt * gb;
    if (on && a.b0) atomicAdd(a.b0 + i, gb);
  }
  if (!on || ub == 0.0f) return;
  const v3 vi = ld3(a.f0, i);
  v3 xb = mk(0.0f, 0.0f, 0.0f), vb = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const v3 r = vsub(xi, ld3(st.x, j));
    const v3 g = cubic_dW(p, r);
    const float w = (a.mode == 0 ? st.m_V[j] : st.m[j]) * ub;
    const v3 d = vsub(vi, ld3(a.f0, j));
    acc3(vb, smul(w, g));
    atomic_add3(a.b1, j, smul(-w, g));
    const v3 rbar = hess_w(p, r, smul(w, d));
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomic_add3(a.b1, i, vb);
  atomic_add3(a.x_bar, i, xb);
}

// Jacobi update v_i -= sum_{j: |k_i + k_j| > eps} dt (k_i + k_j) m_j gW_ij
// (dimcv_sph.py:421-468 / 495-549) with
//   mode 0: k = b * alpha * rho   (f0 = b, f1 = alpha, f2 = rho)
//   mode 1: k = (da - rho0) * alpha (f0 = da, f1 = alpha)
// in: g_in = P_bar (the adjoint of the subtracted sum, i.e. -v_bar).
// out: x_bar, b0 = kappa_bar.
__device__ __forceinline__ float kappa_of(const dimcv_params& p, const dimcv_adj_args& a, int64_t i) {
  return a.mode == 0 ? a.f0[i] * a.f1[i] * a.f2[i] : (a.f0[i] - p.density0) * a.f1[i];
}

__global__ void k_adj_jacobi(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  const v3 pb = ld3(a.g_in, i);
  if (pb.x == 0.0f && pb.y == 0.0f && pb.z == 0.0f) return;
  const float ki = kappa_of(p, a, i);
  float kb = 0.0f;
  v3 xb = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const float ksum = ki + kappa_of(p, a, j);
    if (!(fabsf(ksum) > p.m_eps)) continue;
    const v3 r = vsub(xi, ld3(st.x, j));
    const v3 g = cubic_dW(p, r);
    const float mj = st.m[j];
    const float s = p.dt * mj * vdot(g, pb);
    kb += s;
    atomicAdd(a.b0 + j, s);
    const v3 rbar = hess_w(p, r, smul(p.dt * ksum * mj, pb));
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomicAdd(a.b0 + i, kb);
  atomic_add3(a.x_bar, i, xb);
}

// a_i = g + sum_j c mV_j ((v_i - v_j).r) / (|r|^2 + 0.01 h^2) gW_ij  (dimcv_sph.py:150-208)
// in: f0 = v, g_in = a_bar.  out: x_bar, b1 = v_bar.
__global__ void k_adj_nonpressure(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  const v3 ab = ld3(a.g_in, i);
  if (ab.x == 0.0f && ab.y == 0.0f && ab.z == 0.0f) return;
  const v3 vi = ld3(a.f0, i);
  v3 xb = mk(0.0f, 0.0f, 0.0f), vb = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const v3 r = vsub(xi, ld3(st.x, j));
    const v3 g = cubic_dW(p, r);
    const v3 d = vsub(vi, ld3(a.f0, j));
    const float c = p.visc_coef * st.m_V[j];
    const float D = vdot(r, r) + p.visc_h2;
    const float dr = vdot(d, r);
    const float beta = c * dr / D;
    const float bb = vdot(ab, g);
    const v3 dv = smul(bb * c / D, r);
    acc3(vb, dv);
    atomic_add3(a.b1, j, vneg(dv));
    const v3 rbar = vadd(smul(bb * c / D, vsub(d, smul(2.0f * dr / D, r))), hess_w(p, r, smul(beta, ab)));
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomic_add3(a.b1, i, vb);
  atomic_add3(a.x_bar, i, xb);
}

// C_i = sum_j mV_j (u_j - u_i) x gW_ij   (vorticity, dimcv_sph.py:228-248; the
// curl of v* in compute_d_vorticity, :551-583)
// in: f0 = u, g_in = C_bar.  out: x_bar, b1 = u_bar.
__global__ void k_adj_curl(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  const v3 cb = ld3(a.g_in, i);
  if (cb.x == 0.0f && cb.y == 0.0f && cb.z == 0.0f) return;
  const v3 ui = ld3(a.f0, i);
  v3 xb = mk(0.0f, 0.0f, 0.0f), ubi = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const v3 r = vsub(xi, ld3(st.x, j));
    const v3 g = cubic_dW(p, r);
    const float mv = st.m_V[j];
    const v3 d = vsub(ld3(a.f0, j), ui);
    const v3 du = smul(mv, vcross(g, cb));
    atomic_add3(a.b1, j, du);
    ubi = vsub(ubi, du);
    const v3 rbar = hess_w(p, r, smul(mv, vcross(cb, d)));
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomic_add3(a.b1, i, ubi);
  atomic_add3(a.x_bar, i, xb);
}

// G_i[r][c] = sum_j mV_j (v_j - v_i)_r gW_ij_c   (dimcv_sph.py:250-262)
// in: f0 = v, g_in = G_bar (9 per particle).  out: x_bar, b1 = v_bar.
__global__ void k_adj_grad_v(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  float Gb[9];
  bool any = false;
  for (int q = 0; q < 9; ++q) {
    Gb[q] = a.g_in[9 * i + q];
    any |= Gb[q] != 0.0f;
  }
  if (!any) return;
  const v3 vi = ld3(a.f0, i);
  v3 xb = mk(0.0f, 0.0f, 0.0f), vbi = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const v3 r = vsub(xi, ld3(st.x, j));
    const v3 g = cubic_dW(p, r);
    const float mv = st.m_V[j];
    const v3 d = vsub(ld3(a.f0, j), vi);
    const float gc[3] = {g.x, g.y, g.z}, ar[3] = {mv * d.x, mv * d.y, mv * d.z};
    float abar[3], gbar[3];
    for (int rr = 0; rr < 3; ++rr) abar[rr] = Gb[3 * rr] * gc[0] + Gb[3 * rr + 1] * gc[1] + Gb[3 * rr + 2] * gc[2];
    for (int c = 0; c < 3; ++c) gbar[c] = Gb[c] * ar[0] + Gb[3 + c] * ar[1] + Gb[6 + c] * ar[2];
    const v3 dv = mk(mv * abar[0], mv * abar[1], mv * abar[2]);
    atomic_add3(a.b1, j, dv);
    vbi = vsub(vbi, dv);
    const v3 rbar = hess_w(p, r, mk(gbar[0], gbar[1], gbar[2]));
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomic_add3(a.b1, i, vbi);
  atomic_add3(a.x_bar, i, xb);
}

// w_i = mV_i d_i W(0) + sum_j mV_j d_j W(|r_ij|)   (dimcv_sph.py:660-677)
// in: f0 = d_vort, g_in = w_bar.  out: x_bar, b1 = d_vort_bar.
__global__ void k_adj_smooth(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  const v3 wb = ld3(a.g_in, i);
  if (wb.x == 0.0f && wb.y == 0.0f && wb.z == 0.0f) return;
  atomic_add3(a.b1, i, smul(st.m_V[i] * cubic_W(p, 0.0f), wb));
  v3 xb = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const v3 r = vsub(xi, ld3(st.x, j));
    const float mv = st.m_V[j];
    atomic_add3(a.b1, j, smul(mv * cubic_W(p, vnorm(r)), wb));
    const v3 rbar = smul(mv * vdot(ld3(a.f0, j), wb), cubic_dW(p, r));
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomic_add3(a.x_bar, i, xb);
}

// kvn = sqrt(Wn) / (sqrt(Sn) + 1e-8) of grad_v (dimcv_sph.py:295-312),
// elementwise.  in: f0 = grad_v (9), g_in = kvn_bar.  out: b1 = grad_v_bar (9).
__global__ void k_adj_kvn(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  const float kb = a.g_in[i];
  if (kb == 0.0f) return;
  const float* G = a.f0 + 9 * i;
  float Wn = 0.0f, Sn = 0.0f;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const float w = G[3 * r + c] - G[3 * c + r], s = G[3 * r + c] + G[3 * c + r];
      Wn += w * w;
      Sn += s * s;
    }
  const float sw = sqrtf(Wn), ss = sqrtf(Sn), den = ss + 1e-8f;
  const float dW = sw > 0.0f ? 1.0f / (2.0f * sw * den) : 0.0f;
  const float dS = ss > 0.0f ? -sw / (den * den * 2.0f * ss) : 0.0f;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      const float w = G[3 * r + c] - G[3 * c + r], s = G[3 * r + c] + G[3 * c + r];
      a.b1[9 * i + 3 * r + c] += kb * (4.0f * dW * w + 4.0f * dS * s);
    }
}

// kden_i = kvn_i E(0) + sum_j kvn_j E(r_ij)   (dimcv_sph.py:721-751)
// in: f0 = kvn, g_in = kden_bar.  out: x_bar, b1 = kvn_bar.
__global__ void k_adj_kden(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  const float kb = a.g_in[i];
  if (kb == 0.0f) return;
  atomicAdd(a.b1 + i, epan_f(p, mk(0.0f, 0.0f, 0.0f), nullptr) * kb);
  v3 xb = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    const v3 r = vsub(xi, ld3(st.x, j));
    v3 ge;
    const float E = epan_f(p, r, &ge);
    atomicAdd(a.b1 + j, E * kb);
    const v3 rbar = smul(a.f0[j] * kb, ge);
    acc3(xb, rbar);
    atomic_add3(a.x_bar, j, vneg(rbar));
  }
  atomic_add3(a.x_bar, i, xb);
}

// The update gate of vort2vel: 1 where |v_vort| < |v| (dimcv_sph.py:717-719).
// in: f0 = v_vort, f1 = v (before the update).  out: b2 = gate (1.0 / 0.0).
__global__ void k_adj_gate(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  const int64_t f = xcd_block() * blockDim.x + threadIdx.x;
  if (f >= bound || f >= st.counters[DIMCV_C_FLUID_LIST]) return;
  const int64_t i = st.fluid_list[f];
  if (i < 0) return;
  a.b2[i] = (in_vortex_domain(p, ld3(st.x, i)) && vnorm(ld3(a.f0, i)) < vnorm(ld3(a.f1, i))) ? 1.0f : 0.0f;
}

// Near field and S_i of vort2vel (dimcv_sph.py:649-658, 687-692):
//   vv_i = fac sum_j A(|r_xy|) (r_xy x w_j),  S_i = kden_sum - kden_i - sum_j kden_j.
// in: f0 = w (d_vort_smoothed), g_in = u_bar (gate * v_bar), f1 = S_bar (from
// the far-field target pass).  out: x_bar, b0 = w_bar, b1 = kden_bar.
__global__ void k_adj_near(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_ENTRY();
  if (!in_vortex_domain(p, xi)) return;
  const v3 ub = smul(p.vp_factor, ld3(a.g_in, i));
  const float sb = a.f1 ? a.f1[i] : 0.0f;
  if (sb != 0.0f) atomicAdd(a.b1 + i, -sb);
  v3 xb = mk(0.0f, 0.0f, 0.0f);
  FOR_FLUID_NBR(k, j) {
    if (sb != 0.0f) atomicAdd(a.b1 + j, -sb);
    const v3 r3 = vsub(xi, ld3(st.x, j));
    const v3 r = mk(r3.x, r3.y, 0.0f);
    const float s = vnorm(r);
    float dA;
    const float A = a_kernel(p, s, dA);
    const v3 wj = ld3(a.f0, j);
    atomic_add3(a.b0, j, smul(A, vcross(ub, r)));
    const float tu = vdot(vcross(r, wj), ub);
    v3 rb = smul(A, vcross(wj, ub));
    if (s > 0.0f) rb = vadd(rb, smul(tu * dA / s, r));
    rb.z = 0.0f;
    acc3(xb, rb);
    atomic_add3(a.x_bar, j, vneg(rb));
  }
  atomic_add3(a.x_bar, i, xb);
}

// Far field of vort2vel (dimcv_sph.py:693-716) for target i over the
// accepted samples k (reference-order r^2 > T):
//   v_vort_i += (vp_factor / added_i) S_i sum_k A(|r|) c_k (r_xy x w_k),
// r = x_i - x_k, c_k = 1 / (kden_k + 1e-8), A(s) = (1 - E) / (s (s + eps E)),
// E = exp(-s / eps).  Every accepted pair has s > h >> eps, where E is 0 in
// f32 (eps = 1e-6: s / eps > 4e4), so A = 1 / r^2 and dA/ds / s = -2 A^2.
// The samples are read from a packed table (scalar loads, one entry per
// sample for the whole wave): (x, c) and (w, slot).
__global__ void k_adj_pack_samples(dimcv_state st, dimcv_adj_args a, int32_t cap, float4* __restrict__ tab) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int M = min(st.counters[DIMCV_C_NUM_SAMPLES], cap);
  if (q >= M) return;
  const int64_t sk = st.sample_idx[q];
  tab[2 * q] = make_float4(st.x[3 * sk], st.x[3 * sk + 1], st.x[3 * sk + 2], 1.0f / (a.f1[sk] + 1e-8f));
  tab[2 * q + 1] = make_float4(a.f0[3 * sk], a.f0[3 * sk + 1], a.f0[3 * sk + 2], __int_as_float((int)sk));
}

// Target side.  in: f0 = w, f1 = kden, g_in = u_bar (v_vort adjoint),
// dscalar = kden_sum.  out: x_bar, b0 += S_bar, b2 = the target's weights
// (vp_factor / added) S u_bar (0 for targets without samples).  One loop
// counts the accepted samples, the second sums with u_bar and scales once.
__global__ void k_adj_far_targets(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound, int32_t cap,
                                  const float4* __restrict__ tab) {
  ADJ_ENTRY();
  st3(a.b2, i, mk(0.0f, 0.0f, 0.0f));
  if (!in_vortex_domain(p, xi)) return;
  const v3 ub = ld3(a.g_in, i);
  float kns = a.f1[i];
  FOR_FLUID_NBR(k, j) { kns += a.f1[j]; }
  const float S = (float)(a.dscalar - (double)kns);
  const int M = min(st.counters[DIMCV_C_NUM_SAMPLES], cap);
  const float T = p.far_r2_max_rejected;
  auto r2_of = [&](const float4& s4, v3& r) {
    r = mk(xi.x - s4.x, xi.y - s4.y, xi.z - s4.z);
    return __fadd_rn(__fadd_rn(__fmul_rn(r.x, r.x), __fmul_rn(r.y, r.y)), __fmul_rn(r.z, r.z));
  };
  int added = 0;
#pragma unroll 4
  for (int q = 0; q < M; ++q) {
    v3 r;
    added += r2_of(tab[2 * q], r) > T ? 1 : 0;
  }
  if (added == 0) return;
  const float coef = p.vp_factor / (float)added;
  const float sc = coef * S;
  st3(a.b2, i, smul(sc, ub));
  v3 sum = mk(0.0f, 0.0f, 0.0f), xbu = mk(0.0f, 0.0f, 0.0f);
#pragma unroll 2
  for (int q = 0; q < M; ++q) {
    const float4 s4 = tab[2 * q], w4 = tab[2 * q + 1];
    v3 r;
    const float r2 = r2_of(s4, r);
    if (!(r2 > T)) continue;
    const float A = __builtin_amdgcn_rcpf(r2);
    const float c = s4.w;
    const v3 t = mk(r.y * w4.z, -(r.x * w4.z), r.x * w4.y - r.y * w4.x);  // r_xy x w_k
    sum = vadd(sum, smul(A * c, t));
    // d/dx_i of A (r_xy x w_k).u: A (w_k x u)_xy + (t.u) (dA/ds / s) r
    const float kk = -2.0f * A * A * vdot(t, ub);
    const float cx = w4.y * ub.z - w4.z * ub.y, cy = w4.z * ub.x - w4.x * ub.z;
    xbu = vadd(xbu, smul(c, mk(A * cx + kk * r.x, A * cy + kk * r.y, kk * r.z)));
  }
  a.b0[i] += coef * vdot(ub, sum);
  atomic_add3(a.x_bar, i, smul(sc, xbu));
}

// Targets packed for the source side: (x_i, -) and (weights_i, -) per fluid
// entry; entries without weight are all zero (skipped).
__global__ void k_adj_pack_targets(dimcv_state st, dimcv_adj_args a, int32_t bound, float4* __restrict__ tab) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nf = min((int64_t)bound, (int64_t)st.counters[DIMCV_C_FLUID_LIST]);
  if (f >= nf) return;
  const int64_t i = st.fluid_list[f];
  float4 x4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), w4 = x4;
  if (i >= 0) {
    const v3 w = ld3(a.f2, i);
    if (w.x != 0.0f || w.y != 0.0f || w.z != 0.0f) {
      x4 = make_float4(st.x[3 * i], st.x[3 * i + 1], st.x[3 * i + 2], 0.0f);
      w4 = make_float4(w.x, w.y, w.z, 1.0f);
    }
  }
  tab[2 * f] = x4;
  tab[2 * f + 1] = w4;
}

// Source side: for sample k the sums over targets of w_bar_k, kden_bar_k
// (through c_k) and x_bar_k.  One lane per sample, targets read from their
// packed table (scalar loads) in blockIdx.y chunks; one set of atomics per
// lane.  in: f0 = w, f1 = kden, f2 = per-target weights (b2 of the target
// pass), mode = sample bound.  out: x_bar, b0 = w_bar, b1 = kden_bar.
__global__ void k_adj_far_sources(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound,
                                  const float4* __restrict__ stab, const float4* __restrict__ ttab, int32_t cap) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int M = min(st.counters[DIMCV_C_NUM_SAMPLES], cap);
  const int64_t nf = min((int64_t)bound, (int64_t)st.counters[DIMCV_C_FLUID_LIST]);
  const int64_t chunk = (nf + gridDim.y - 1) / gridDim.y;
  const int64_t f0 = (int64_t)blockIdx.y * chunk, f1 = min(f0 + chunk, nf);
  if (q >= M) return;
  const float4 s4 = stab[2 * q], k4 = stab[2 * q + 1];
  const v3 xs = mk(s4.x, s4.y, s4.z), wk = mk(k4.x, k4.y, k4.z);
  const float c = s4.w;
  const int64_t sk = __float_as_int(k4.w);
  const float T = p.far_r2_max_rejected;
  v3 wbar = mk(0.0f, 0.0f, 0.0f), xb = mk(0.0f, 0.0f, 0.0f);
  float cbar = 0.0f;
#pragma unroll 2
  for (int64_t f = f0; f < f1; ++f) {
    const float4 x4 = ttab[2 * f], w4 = ttab[2 * f + 1];
    if (w4.w == 0.0f) continue;  // uniform: target without weight
    const v3 r = mk(x4.x - xs.x, x4.y - xs.y, x4.z - xs.z);
    const float r2 = __fadd_rn(__fadd_rn(__fmul_rn(r.x, r.x), __fmul_rn(r.y, r.y)), __fmul_rn(r.z, r.z));
    if (!(r2 > T)) continue;
    const float A = __builtin_amdgcn_rcpf(r2);
    const v3 t = mk(r.y * wk.z, -(r.x * wk.z), r.x * wk.y - r.y * wk.x);  // r_xy x w_k
    const float tw = t.x * w4.x + t.y * w4.y + t.z * w4.z;
    // w x r_xy
    wbar = vadd(wbar, smul(A * c, mk(-(w4.z * r.y), w4.z * r.x, w4.x * r.y - w4.y * r.x)));
    cbar += A * tw;
    const float kk = -2.0f * A * A * tw;
    const float cx = wk.y * w4.z - wk.z * w4.y, cy = wk.z * w4.x - wk.x * w4.z;
    xb = vsub(xb, smul(c, mk(A * cx + kk * r.x, A * cy + kk * r.y, kk * r.z)));
  }
  atomic_add3(a.b0, sk, wbar);
  atomicAdd(a.b1 + sk, -c * c * cbar);
  atomic_add3(a.x_bar, sk, xb);
}

}  // namespace dimcv

using namespace dimcv;

extern "C" int dimcv_adj_pass(const dimcv_state* st, const dimcv_params* p, int32_t pass, const dimcv_adj_args* a,
                              int32_t fluid_bound, void* stream) {
  if (!st || !p || !a) {
    set_error("dimcv_adj_pass: null argument");
    return DIMCV_ERR_ARG;
  }
  if (p->periodic || st->block_scene) {
    set_error("dimcv_adj_pass: open single-scene systems only");
    return DIMCV_ERR_ARG;
  }
  if (fluid_bound <= 0) return DIMCV_OK;
  hipStream_t s = (hipStream_t)stream;
  const dim3 g(nblocks(fluid_bound)), b(256);
  switch (pass) {
    case DIMCV_ADJ_DENSITY: hipLaunchKernelGGL(k_adj_density, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_FACTOR: hipLaunchKernelGGL(k_adj_factor, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_DIV: hipLaunchKernelGGL(k_adj_div, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_JACOBI: hipLaunchKernelGGL(k_adj_jacobi, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_NONPRESSURE: hipLaunchKernelGGL(k_adj_nonpressure, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_CURL: hipLaunchKernelGGL(k_adj_curl, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_GRAD_V: hipLaunchKernelGGL(k_adj_grad_v, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_SMOOTH: hipLaunchKernelGGL(k_adj_smooth, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_KVN: hipLaunchKernelGGL(k_adj_kvn, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_KDEN: hipLaunchKernelGGL(k_adj_kden, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_GATE: hipLaunchKernelGGL(k_adj_gate, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_NEAR: hipLaunchKernelGGL(k_adj_near, g, b, 0, s, *st, *p, *a, fluid_bound); break;
    case DIMCV_ADJ_FAR_TARGETS:
    case DIMCV_ADJ_FAR_SOURCES: {
      // a->mode bounds the sample count.  Sample table: the far-field
      // partials region; target table: the sample region (n_max entries).
      if (a->mode <= 0) break;
      Workspace ws;
      int rc = bind_workspace(st, p, &ws, 0);
      if (rc) return rc;
      if ((int64_t)a->mode > ws.far_partial_cap || (int64_t)fluid_bound > ws.max_samples) {
        set_error("dimcv_adj_pass: far-field tables exceed the workspace (%d samples, %d targets)", a->mode,
                  fluid_bound);
        return DIMCV_ERR_WORKSPACE;
      }
      float4* stab = (float4*)ws.far_partials;
      hipLaunchKernelGGL(k_adj_pack_samples, dim3(nblocks(a->mode)), b, 0, s, *st, *a, a->mode, stab);
      if (pass == DIMCV_ADJ_FAR_TARGETS) {
        hipLaunchKernelGGL(k_adj_far_targets, g, b, 0, s, *st, *p, *a, fluid_bound, a->mode, (const float4*)stab);
      } else {
        hipLaunchKernelGGL(k_adj_pack_targets, g, b, 0, s, *st, *a, fluid_bound, ws.samples);
        // targets split into chunks of ~256 over blockIdx.y
        const unsigned ych = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (fluid_bound + 255) / 256));
        hipLaunchKernelGGL(k_adj_far_sources, dim3(nblocks(a->mode), ych), b, 0, s, *st, *p, *a, fluid_bound,
                           (const float4*)stab, (const float4*)ws.samples, a->mode);
      }
      break;
    }
    default:
      set_error("dimcv_adj_pass: unknown pass %d", pass);
      return DIMCV_ERR_ARG;
  }
  DIMCV_LAUNCHED("dimcv_adj_pass");
  return DIMCV_OK;
}
