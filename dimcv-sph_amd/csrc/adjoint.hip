// Hand-written adjoints (vector-Jacobian products) of the step's
// kernel-gradient passes and of MC-VVPE, for differentiation through a
// rollout with the discrete structure of each step frozen: neighbour lists,
// importance samples, Jacobi iteration counts and every gate of the forward
// (the DFSPH factor threshold, max(rho_dot, 0) with the 20-neighbour rule,
// max(rho*, rho0), |kappa_i + kappa_j| > eps, the far-field cut-off |r| > h
// and the |v_vort| < |v| update) keep the values they had in the forward.
// Within that structure each pass is smooth and its VJP is exact.
//
// One thread per fluid-list entry i walks i's neighbour list.  Pair terms
// depend on x through r = x_i - x_j: r_bar goes to x_i and its negative to
// x_j.  The neighbour passes gather: i's thread evaluates each pair from both
// ends (its own term and the one the pair sends to i), so only particles no
// list contains scatter.  Every pass accumulates (+=) into its adjoint
// outputs with float atomics (one per particle and component in the
// neighbour passes).
//
// Kernel derivatives: grad W(r) = phi(|r|) r/|r| with phi = dW/ds
// (sph_base.py:46-68), so the Hessian-vector product is
//   H g = phi'(s) (rh.g) rh + (phi(s)/s) (g - (rh.g) rh);
// the product Epanechnikov kernel (dimcv_sph.py:721-736) differentiates
// factor by factor; A(s) = (1 - E)/(s (s + eps E)), E = exp(-s/eps), is the
// kernel of vort2vel (dimcv_sph.py:652-657, 698-706).
#include <algorithm>

// The adjoint reproduces no reference bits (its sums are reordered and
// atomic anyway), so a * b + c may contract to one fused multiply-add in this
// file, the shared vector helpers included; the forward passes, compiled in
// their own files, keep every operation separate.  Tests that must repeat a
// forward decision bit for bit (the far-field r^2 > T) use __fmul_rn /
// __fadd_rn, which never contract.
#pragma clang fp contract(fast)

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
  const NbrList nl(st.nbr, f);                        \
  const int64_t ns = p.n_max;                                          \
  const v3 xi = ld3(st.x, i)

#define FOR_FLUID_NBR(k, j) \
  FOR_NBR(k, e_##k, j)         \
  if (nb_fluid(e_##k))

// Neighbour-pass VJPs in gather form.  The thread of particle i adds, for
// each fluid neighbour j in i's list, both i's own term of the pair (i, j)
// and the term that the pair (j, i) of j's list sends to i, so no pair term
// is scattered.  The list test (|x_i - x_j| < h, j != i) is symmetric: j's
// list holds i exactly when i's list holds j, except for the particles of
// cell 0 (slots [0, gpn[0]) in cell order), which no list contains.  Such a
// particle (vis = false) receives nothing and adds its terms for j to j's
// outputs with atomics.  Pair identities used, with r = x_i - x_j and
// g = grad W(r): grad W(-r) = -g, H(-r) = H(r), W, A and E are even, grad E
// is odd.  Each thread adds its sums into the outputs with one atomic add
// per component (a cell-0 particle's terms may reach any particle).
#define ADJ_GATHER_ENTRY() \
  ADJ_ENTRY();             \
  const bool vis = i >= st.grid_particles_num[0]

__device__ __forceinline__ v3 zero3() { return mk(0.0f, 0.0f, 0.0f); }

// rho_i = rho0 (mV_i W(0) + sum_j mV_j W(|r_ij|))            (dimcv_sph.py:126-148)
// in: g_in = rho_bar.  out: x_bar += rho0 sum_j (mV_j rho_bar_i + mV_i rho_bar_j) gW_ij.
__global__ void k_adj_density(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  const float ci = p.density0 * a.g_in[i];
  const float cr = vis ? p.density0 * st.m_V[i] : 0.0f;
  struct G {
    v3 x;
    float mv, rb;
  };
  v3 xb = zero3();
  nbr_loop<8>(
      nl, ns, nn, [&](uint32_t j) { return G{ld3s(st.x, j), ld1s(st.m_V, j), ld1s(a.g_in, j)}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const v3 g = cubic_dW(p, vsub(xi, q.x));
        acc3(xb, smul(ci * q.mv + cr * q.rb, g));
        if (!vis) atomic_add3(a.x_bar, nb_slot(e), smul(-(ci * q.mv), g));
      });
  atomic_add3(a.x_bar, i, xb);
}

// alpha_i = 1 / (sum_j |m_j gW_ij|^2 + |sum_j m_j gW_ij|^2), 0 below 1e-6
// (dimcv_sph.py:86-124).  in: f0 = alpha (forward), g_in = alpha_bar.
// out: x_bar.  With s_i = -alpha_bar_i alpha_i^2 and G_i = sum_j m_j gW_ij,
// the pair (i, j) gives gW_ij the adjoint 2 m_j s_i (m_j gW_ij + G_i).  A
// first pass tabulates G in slot order.
__device__ __forceinline__ float factor_s(const dimcv_adj_args& a, int64_t i) {
  const float al = a.f0[i];
  return al == 0.0f ? 0.0f : -a.g_in[i] * al * al;
}

__global__ void k_adj_factor_sum(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound,
                                 float* __restrict__ Gt) {
  ADJ_ENTRY();
  v3 G = zero3();
  FOR_FLUID_NBR(k, j) { acc3(G, smul(st.m[j], cubic_dW(p, vsub(xi, ld3(st.x, j))))); }
  st3(Gt, i, G);
}

__global__ void k_adj_factor(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound,
                             const float* __restrict__ Gt) {
  ADJ_GATHER_ENTRY();
  const float si = factor_s(a, i), mi = st.m[i];
  const v3 Gi = ld3(Gt, i);
  struct G {
    v3 x, G;
    float m, s;
  };
  v3 xb = zero3();
  nbr_loop<4>(
      nl, ns, nn, [&](uint32_t j) { return G{ld3s(st.x, j), ld3s(Gt, j), ld1s(st.m, j), factor_s(a, j)}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const v3 r = vsub(xi, q.x);
        const v3 g = cubic_dW(p, r);
        // own: 2 m_j s_i (m_j g + G_i); pair (j, i), gW_ji = -g: 2 m_i s_j (G_j - m_i g)
        const v3 own = smul(2.0f * q.m * si, vadd(smul(q.m, g), Gi));
        const v3 rec = smul(2.0f * mi * (vis ? q.s : 0.0f), vsub(q.G, smul(mi, g)));
        acc3(xb, hess_w(p, r, vsub(own, rec)));
        if (!vis) atomic_add3(a.x_bar, nb_slot(e), vneg(hess_w(p, r, own)));
      });
  atomic_add3(a.x_bar, i, xb);
}

// mode 0: b_i = [b_i > 0] sum_j mV_j (v_i - v_j).gW_ij   (dimcv_sph.py:337-374)
//   in: f0 = v, f1 = b (forward, gives the gate), g_in = b_bar.
// mode 1: da_i = max(rho_i + dt sum_j m_j (v_i - v_j).gW_ij, rho0) (:470-493)
//   in: f0 = v, f1 = da (forward), g_in = da_bar.  out also b0 = rho_bar.
// out: x_bar, b1 = v_bar.  With u_i the gated upstream (b_bar_i, or
// dt da_bar_i) and c = mV (mode 0) or m (mode 1), the pair weight is
// w = c_j u_i + c_i u_j: v_bar_i += w gW_ij, x_bar_i += H_ij w (v_i - v_j).
__device__ __forceinline__ float div_u(const dimcv_params& p, const dimcv_adj_args& a, int64_t i) {
  if (a.mode == 0) return a.f1[i] > 0.0f ? a.g_in[i] : 0.0f;
  return a.f1[i] > p.density0 ? p.dt * a.g_in[i] : 0.0f;
}

__global__ void k_adj_div(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  const float* __restrict__ cm = a.mode == 0 ? st.m_V : st.m;
  if (a.mode == 1 && a.b0 && a.f1[i] > p.density0) atomicAdd(a.b0 + i, a.g_in[i]);
  const float ui = div_u(p, a, i), ci = vis ? cm[i] : 0.0f;
  const v3 vi = ld3(a.f0, i);
  struct G {
    v3 x, v;
    float c, u;
  };
  v3 xb = zero3(), vb = zero3();
  nbr_loop<4>(
      nl, ns, nn, [&](uint32_t j) { return G{ld3s(st.x, j), ld3s(a.f0, j), ld1s(cm, j), div_u(p, a, j)}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const v3 r = vsub(xi, q.x);
        const v3 g = cubic_dW(p, r);
        const v3 d = vsub(vi, q.v);
        const float w = q.c * ui + ci * q.u;
        acc3(vb, smul(w, g));
        acc3(xb, hess_w(p, r, smul(w, d)));
        if (!vis && ui != 0.0f) {
          const float wo = q.c * ui;
          const int64_t j = nb_slot(e);
          atomic_add3(a.b1, j, smul(-wo, g));
          atomic_add3(a.x_bar, j, vneg(hess_w(p, r, smul(wo, d))));
        }
      });
  atomic_add3(a.b1, i, vb);
  atomic_add3(a.x_bar, i, xb);
}

// Jacobi update v_i -= sum_{j: |k_i + k_j| > eps} dt (k_i + k_j) m_j gW_ij
// (dimcv_sph.py:421-468 / 495-549) with
//   mode 0: k = b * alpha * rho   (f0 = b, f1 = alpha, f2 = rho)
//   mode 1: k = (da - rho0) * alpha (f0 = da, f1 = alpha)
// in: g_in = P_bar (the adjoint of the subtracted sum, i.e. -v_bar).
// out: x_bar, b0 = kappa_bar.  Per kept pair:
//   kappa_bar_i += dt gW_ij.(m_j P_bar_i - m_i P_bar_j),
//   x_bar_i += H_ij dt (k_i + k_j) (m_j P_bar_i - m_i P_bar_j).
__device__ __forceinline__ float kappa_of(const dimcv_params& p, const dimcv_adj_args& a, int64_t i) {
  return a.mode == 0 ? a.f0[i] * a.f1[i] * a.f2[i] : (a.f0[i] - p.density0) * a.f1[i];
}

__global__ void k_adj_jacobi(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  const v3 pbi = ld3(a.g_in, i);
  const float ki = kappa_of(p, a, i), mi = vis ? st.m[i] : 0.0f;
  struct G {
    v3 x, pb;
    float m, k;
  };
  float kb = 0.0f;
  v3 xb = zero3();
  nbr_loop<4>(
      nl, ns, nn, [&](uint32_t j) { return G{ld3s(st.x, j), ld3s(a.g_in, j), ld1s(st.m, j), kappa_of(p, a, j)}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const float ksum = ki + q.k;
        if (!(fabsf(ksum) > p.m_eps)) return;
        const v3 r = vsub(xi, q.x);
        const v3 g = cubic_dW(p, r);
        const float so = p.dt * q.m * vdot(g, pbi);
        kb += so - p.dt * mi * vdot(g, q.pb);
        acc3(xb, hess_w(p, r, smul(p.dt * ksum, vsub(smul(q.m, pbi), smul(mi, q.pb)))));
        if (!vis) {
          const int64_t j = nb_slot(e);
          atomicAdd(a.b0 + j, so);
          atomic_add3(a.x_bar, j, vneg(hess_w(p, r, smul(p.dt * ksum * q.m, pbi))));
        }
      });
  atomicAdd(a.b0 + i, kb);
  atomic_add3(a.x_bar, i, xb);
}

// a_i = g + sum_j c mV_j ((v_i - v_j).r) / (|r|^2 + 0.01 h^2) gW_ij  (dimcv_sph.py:150-208)
// in: f0 = v, g_in = a_bar.  out: x_bar, b1 = v_bar.  With
// Q = c (mV_j a_bar_i - mV_i a_bar_j), D = |r|^2 + 0.01 h^2, d = v_i - v_j and
// t = Q.gW_ij / D: v_bar_i += t r, x_bar_i += t (d - 2 (d.r) r / D) + H_ij (d.r / D) Q.
__global__ void k_adj_nonpressure(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  const v3 abi = ld3(a.g_in, i);
  const v3 vi = ld3(a.f0, i);
  const float ci = vis ? p.visc_coef * st.m_V[i] : 0.0f;
  struct G {
    v3 x, v, ab;
    float mv;
  };
  v3 xb = zero3(), vb = zero3();
  nbr_loop<4>(
      nl, ns, nn, [&](uint32_t j) { return G{ld3s(st.x, j), ld3s(a.f0, j), ld3s(a.g_in, j), ld1s(st.m_V, j)}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const v3 r = vsub(xi, q.x);
        const v3 g = cubic_dW(p, r);
        const v3 d = vsub(vi, q.v);
        const float cj = p.visc_coef * q.mv;
        const float D = vdot(r, r) + p.visc_h2;
        const float dr = vdot(d, r);
        const v3 dd = vsub(d, smul(2.0f * dr / D, r));
        const v3 Q = vsub(smul(cj, abi), smul(ci, q.ab));
        const float t = vdot(Q, g) / D;
        acc3(vb, smul(t, r));
        acc3(xb, vadd(smul(t, dd), hess_w(p, r, smul(dr / D, Q))));
        if (!vis) {
          const int64_t j = nb_slot(e);
          const float to = cj * vdot(abi, g) / D;
          atomic_add3(a.b1, j, smul(-to, r));
          atomic_add3(a.x_bar, j, vneg(vadd(smul(to, dd), hess_w(p, r, smul(cj * dr / D, abi)))));
        }
      });
  atomic_add3(a.b1, i, vb);
  atomic_add3(a.x_bar, i, xb);
}

// C_i = sum_j mV_j (u_j - u_i) x gW_ij   (vorticity, dimcv_sph.py:228-248; the
// curl of v* in compute_d_vorticity, :551-583)
// in: f0 = u, g_in = C_bar.  out: x_bar, b1 = u_bar.  With
// Q = mV_j C_bar_i + mV_i C_bar_j: u_bar_i -= gW_ij x Q, x_bar_i += H_ij (Q x (u_j - u_i)).
__global__ void k_adj_curl(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  const v3 cbi = ld3(a.g_in, i);
  const v3 ui = ld3(a.f0, i);
  const float mvi = vis ? st.m_V[i] : 0.0f;
  struct G {
    v3 x, u, cb;
    float mv;
  };
  v3 xb = zero3(), ub = zero3();
  nbr_loop<4>(
      nl, ns, nn, [&](uint32_t j) { return G{ld3s(st.x, j), ld3s(a.f0, j), ld3s(a.g_in, j), ld1s(st.m_V, j)}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const v3 r = vsub(xi, q.x);
        const v3 g = cubic_dW(p, r);
        const v3 d = vsub(q.u, ui);
        const v3 Q = vadd(smul(q.mv, cbi), smul(mvi, q.cb));
        ub = vsub(ub, vcross(g, Q));
        acc3(xb, hess_w(p, r, vcross(Q, d)));
        if (!vis) {
          const int64_t j = nb_slot(e);
          atomic_add3(a.b1, j, smul(q.mv, vcross(g, cbi)));
          atomic_add3(a.x_bar, j, vneg(hess_w(p, r, smul(q.mv, vcross(cbi, d)))));
        }
      });
  atomic_add3(a.b1, i, ub);
  atomic_add3(a.x_bar, i, xb);
}

// M g and M^T d for a row-major 3x3 M.
__device__ __forceinline__ v3 mat_vec(const float* M, v3 g) {
  return mk(M[0] * g.x + M[1] * g.y + M[2] * g.z, M[3] * g.x + M[4] * g.y + M[5] * g.z,
            M[6] * g.x + M[7] * g.y + M[8] * g.z);
}
__device__ __forceinline__ v3 mat_tvec(const float* M, v3 d) {
  return mk(M[0] * d.x + M[3] * d.y + M[6] * d.z, M[1] * d.x + M[4] * d.y + M[7] * d.z,
            M[2] * d.x + M[5] * d.y + M[8] * d.z);
}

// G_i[r][c] = sum_j mV_j (v_j - v_i)_r gW_ij_c   (dimcv_sph.py:250-262)
// in: f0 = v, g_in = G_bar (9 per particle).  out: x_bar, b1 = v_bar.  With
// M = mV_j G_bar_i + mV_i G_bar_j: v_bar_i -= M gW_ij, x_bar_i += H_ij (M^T (v_j - v_i)).
__global__ void k_adj_grad_v(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  float Gi[9];
#pragma unroll
  for (int u = 0; u < 9; ++u) Gi[u] = a.g_in[9 * i + u];
  const v3 vi = ld3(a.f0, i);
  const float mvi = vis ? st.m_V[i] : 0.0f;
  struct G {
    v3 x, v;
    float mv;
    float gb[9];
  };
  v3 xb = zero3(), vb = zero3();
  nbr_loop<4>(
      nl, ns, nn,
      [&](uint32_t j) {
        G q;
        q.x = ld3s(st.x, j);
        q.v = ld3s(a.f0, j);
        q.mv = ld1s(st.m_V, j);
#pragma unroll
        for (int u = 0; u < 9; ++u) q.gb[u] = a.g_in[9 * j + u];
        return q;
      },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const v3 r = vsub(xi, q.x);
        const v3 g = cubic_dW(p, r);
        const v3 d = vsub(q.v, vi);
        float M[9];
#pragma unroll
        for (int u = 0; u < 9; ++u) M[u] = q.mv * Gi[u] + mvi * q.gb[u];
        vb = vsub(vb, mat_vec(M, g));
        acc3(xb, hess_w(p, r, mat_tvec(M, d)));
        if (!vis) {
          const int64_t j = nb_slot(e);
          atomic_add3(a.b1, j, smul(q.mv, mat_vec(Gi, g)));
          atomic_add3(a.x_bar, j, vneg(hess_w(p, r, smul(q.mv, mat_tvec(Gi, d)))));
        }
      });
  atomic_add3(a.b1, i, vb);
  atomic_add3(a.x_bar, i, xb);
}

// w_i = mV_i d_i W(0) + sum_j mV_j d_j W(|r_ij|)   (dimcv_sph.py:660-677)
// in: f0 = d_vort, g_in = w_bar.  out: x_bar, b1 = d_vort_bar:
//   d_bar_i += mV_i (W(0) w_bar_i + sum_j W(|r_ij|) w_bar_j),
//   x_bar_i += sum_j (mV_j d_j.w_bar_i + mV_i d_i.w_bar_j) gW_ij.
__global__ void k_adj_smooth(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  const v3 wbi = ld3(a.g_in, i);
  const v3 di = ld3(a.f0, i);
  const float mvi = st.m_V[i];
  struct G {
    v3 x, d, wb;
    float mv;
  };
  v3 xb = zero3(), ws = smul(cubic_W(p, 0.0f), wbi);
  nbr_loop<4>(
      nl, ns, nn, [&](uint32_t j) { return G{ld3s(st.x, j), ld3s(a.f0, j), ld3s(a.g_in, j), ld1s(st.m_V, j)}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const v3 r = vsub(xi, q.x);
        const v3 g = cubic_dW(p, r);
        const float W = cubic_W(p, vnorm(r));
        const float own = q.mv * vdot(q.d, wbi);
        if (vis) acc3(ws, smul(W, q.wb));
        acc3(xb, smul(own + (vis ? mvi * vdot(di, q.wb) : 0.0f), g));
        if (!vis) {
          const int64_t j = nb_slot(e);
          atomic_add3(a.b1, j, smul(q.mv * W, wbi));
          atomic_add3(a.x_bar, j, smul(-own, g));
        }
      });
  atomic_add3(a.b1, i, smul(mvi, ws));
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
// in: f0 = kvn, g_in = kden_bar.  out: x_bar, b1 = kvn_bar:
//   kvn_bar_i += E(0) kden_bar_i + sum_j E(r_ij) kden_bar_j,
//   x_bar_i += sum_j (kvn_j kden_bar_i + kvn_i kden_bar_j) grad E(r_ij).
__global__ void k_adj_kden(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  const float kbi = a.g_in[i];
  const float kvi = vis ? a.f0[i] : 0.0f;
  struct G {
    v3 x;
    float kv, kb;
  };
  float bs = epan_f(p, zero3(), nullptr) * kbi;
  v3 xb = zero3();
  nbr_loop<8>(
      nl, ns, nn, [&](uint32_t j) { return G{ld3s(st.x, j), ld1s(a.f0, j), ld1s(a.g_in, j)}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        v3 ge;
        const float E = epan_f(p, vsub(xi, q.x), &ge);
        if (vis) bs += E * q.kb;
        acc3(xb, smul(q.kv * kbi + kvi * q.kb, ge));
        if (!vis) {
          const int64_t j = nb_slot(e);
          atomicAdd(a.b1 + j, E * kbi);
          atomic_add3(a.x_bar, j, smul(-(q.kv * kbi), ge));
        }
      });
  atomicAdd(a.b1 + i, bs);
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
// Targets outside the vortex domain carry no upstream.  With u_i = fac u_bar_i
// the pair (i, j) sends w_bar_j += A (u_i x r_xy), kden_bar_j -= S_bar_i and
// x_bar_i += R(w_j, u_i) (x_bar_j -= the same), where
// R(w, u) = A (w x u)_xy + ((r_xy x w).u) (A'(s) / s) r_xy, s = |r_xy|.
__device__ __forceinline__ v3 near_rbar(v3 r, float s, float A, float dA, v3 w, v3 u) {
  v3 rb = smul(A, vcross(w, u));
  if (s > 0.0f) rb = vadd(rb, smul(vdot(vcross(r, w), u) * dA / s, r));
  rb.z = 0.0f;
  return rb;
}

__global__ void k_adj_near(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound) {
  ADJ_GATHER_ENTRY();
  const bool dom = in_vortex_domain(p, xi);
  const v3 ubi = dom ? smul(p.vp_factor, ld3(a.g_in, i)) : zero3();
  const float sbi = (dom && a.f1) ? a.f1[i] : 0.0f;
  const v3 wi = ld3(a.f0, i);
  struct G {
    v3 x, w, ub;
    float sb;
  };
  float kb = -sbi;
  v3 xb = zero3(), wb = zero3();
  nbr_loop<4>(
      nl, ns, nn,
      [&](uint32_t j) { return G{ld3s(st.x, j), ld3s(a.f0, j), ld3s(a.g_in, j), a.f1 ? ld1s(a.f1, j) : 0.0f}; },
      [&](int, uint32_t e, const G& q) {
        if (!nb_fluid(e)) return;
        const v3 r3 = vsub(xi, q.x);
        const v3 r = mk(r3.x, r3.y, 0.0f);
        const float s = vnorm(r);
        float dA;
        const float A = a_kernel(p, s, dA);
        const v3 ro = near_rbar(r, s, A, dA, q.w, ubi);
        v3 rb = ro;
        if (vis && in_vortex_domain(p, q.x)) {  // pair (j, i): r_xy -> -r_xy
          const v3 uj = smul(p.vp_factor, q.ub);
          kb -= q.sb;
          wb = vsub(wb, smul(A, vcross(uj, r)));
          rb = vsub(rb, near_rbar(r, s, A, dA, wi, uj));
        }
        acc3(xb, rb);
        if (!vis && dom) {
          const int64_t j = nb_slot(e);
          if (sbi != 0.0f) atomicAdd(a.b1 + j, -sbi);
          atomic_add3(a.b0, j, smul(A, vcross(ubi, r)));
          atomic_add3(a.x_bar, j, vneg(ro));
        }
      });
  if (kb != 0.0f) atomicAdd(a.b1 + i, kb);
  atomic_add3(a.b0, i, wb);
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
// (vp_factor / added) S u_bar (0 for targets without samples).
// Per target the pass needs the accepted-sample count and two sums over the
// accepted samples, unscaled: sum A c (r_xy x w) and the gradient sum
// c d/dx_i [A (r_xy x w).u_bar]; all are linear, so the sample range can be
// split over blockIdx.y (partials in split order, added by the finish pass).
// d/dx_i of A (r_xy x w_k).u is A (w_k x u)_xy + (t.u) (dA/ds / s) r with
// t = r_xy x w_k; the first term is summed as W = sum A c w_k and crossed
// with u once per thread (far_t_close).
struct FarT {
  float n;
  v3 sum, xbu, W;
};

// One sample (s4 = (x, c), w4 = (w, slot)) for target xi, u2 = -2 u_bar; a
// rejected sample (r^2 <= T, the forward's test) contributes 0.
__device__ __forceinline__ void far_t_pair(FarT& acc, v3 xi, v3 u2, float T, const float4& s4, const float4& w4) {
  const v3 r = mk(xi.x - s4.x, xi.y - s4.y, xi.z - s4.z);
  const float r2 = __fadd_rn(__fadd_rn(__fmul_rn(r.x, r.x), __fmul_rn(r.y, r.y)), __fmul_rn(r.z, r.z));
  const bool ok = r2 > T;
  const float A = ok ? __builtin_amdgcn_rcpf(r2) : 0.0f;
  acc.n += ok ? 1.0f : 0.0f;
  const float Ac = A * s4.w;
  const v3 w = mk(w4.x, w4.y, w4.z);
  const v3 t = mk(r.y * w.z, -(r.x * w.z), r.x * w.y - r.y * w.x);  // r_xy x w_k
  acc.sum = vadd(acc.sum, smul(Ac, t));
  acc.W = vadd(acc.W, smul(Ac, w));
  acc.xbu = vadd(acc.xbu, smul(Ac * A * vdot(t, u2), r));  // c (dA/ds / s) (t.u) r, dA/ds / s = -2 A^2
}

// Adds (W x u)_xy to the gradient sum (u = u_bar, not doubled).
__device__ __forceinline__ void far_t_close(FarT& acc, v3 ub) {
  acc.xbu.x += acc.W.y * ub.z - acc.W.z * ub.y;
  acc.xbu.y += acc.W.z * ub.x - acc.W.x * ub.z;
  acc.W = zero3();
}

// Samples [q0, q1) of the packed table, four per trip, the next four loaded
// (scalar loads) while the current four are summed.
__device__ __forceinline__ void far_t_range(FarT& acc, const float4* __restrict__ tab, int q0, int q1, v3 xi, v3 u2,
                                            float T) {
  int q = q0;
  if (q + 3 < q1) {
    float4 s0 = tab[2 * q], w0 = tab[2 * q + 1], s1 = tab[2 * q + 2], w1 = tab[2 * q + 3];
    float4 s2 = tab[2 * q + 4], w2 = tab[2 * q + 5], s3 = tab[2 * q + 6], w3 = tab[2 * q + 7];
    for (; q + 3 < q1; q += 4) {
      const int nx = q + 7 < q1 ? q + 4 : q;  // in-range prefetch index
      const float4 ns0 = tab[2 * nx], nw0 = tab[2 * nx + 1], ns1 = tab[2 * nx + 2], nw1 = tab[2 * nx + 3];
      const float4 ns2 = tab[2 * nx + 4], nw2 = tab[2 * nx + 5], ns3 = tab[2 * nx + 6], nw3 = tab[2 * nx + 7];
      far_t_pair(acc, xi, u2, T, s0, w0);
      far_t_pair(acc, xi, u2, T, s1, w1);
      far_t_pair(acc, xi, u2, T, s2, w2);
      far_t_pair(acc, xi, u2, T, s3, w3);
      s0 = ns0; w0 = nw0; s1 = ns1; w1 = nw1;
      s2 = ns2; w2 = nw2; s3 = ns3; w3 = nw3;
    }
  }
  for (; q < q1; ++q) far_t_pair(acc, xi, u2, T, tab[2 * q], tab[2 * q + 1]);
}

// Scaling of target i's sums: S_i = kden_sum - kden_i - sum_j kden_j (the
// forward's order), coef = vp_factor / added.
__device__ __forceinline__ void far_t_finish(const dimcv_params& p, const dimcv_adj_args& a, int64_t i,
                                             const NbrList& nl, int64_t ns, int nn, v3 ub,
                                             const FarT& acc) {
  if (acc.n == 0.0f) return;
  float kns = a.f1[i];
  nbr_loop<8>(
      nl, ns, nn, [&](uint32_t j) { return ld1s(a.f1, j); },
      [&](int, uint32_t e, float kj) {
        if (nb_fluid(e)) kns += kj;
      });
  const float S = (float)(a.dscalar - (double)kns);
  const float coef = p.vp_factor / acc.n;
  const float sc = coef * S;
  st3(a.b2, i, smul(sc, ub));
  a.b0[i] += coef * vdot(ub, acc.sum);
  atomic_add3(a.x_bar, i, smul(sc, acc.xbu));
}

// Split y = blockIdx.y of `splits` sums samples [y chunk, (y + 1) chunk);
// with one split it finishes the target itself, otherwise it stores its
// partial (sum, n) (xbu, 0) at part[2 (y bound + f)].
__global__ void __launch_bounds__(256) k_adj_far_targets(dimcv_state st, dimcv_params p, dimcv_adj_args a,
                                                         int32_t bound, int32_t cap, const float4* __restrict__ tab,
                                                         int32_t splits, float4* __restrict__ part) {
  ADJ_ENTRY();
  if (splits == 1) st3(a.b2, i, zero3());
  if (!in_vortex_domain(p, xi)) return;
  const v3 ub = ld3(a.g_in, i);
  const int M = min(st.counters[DIMCV_C_NUM_SAMPLES], cap);
  const int chunk = (M + splits - 1) / splits;
  const int q0 = min(M, (int)blockIdx.y * chunk), q1 = min(M, q0 + chunk);
  FarT acc{0.0f, zero3(), zero3(), zero3()};
  far_t_range(acc, tab, q0, q1, xi, smul(-2.0f, ub), p.far_r2_max_rejected);
  far_t_close(acc, ub);
  if (splits == 1) {
    far_t_finish(p, a, i, nl, ns, nn, ub, acc);
  } else {
    const int64_t o = 2 * ((int64_t)blockIdx.y * bound + f);
    part[o] = make_float4(acc.sum.x, acc.sum.y, acc.sum.z, acc.n);
    part[o + 1] = make_float4(acc.xbu.x, acc.xbu.y, acc.xbu.z, 0.0f);
  }
}

__global__ void k_adj_far_targets_finish(dimcv_state st, dimcv_params p, dimcv_adj_args a, int32_t bound,
                                         int32_t splits, const float4* __restrict__ part) {
  ADJ_ENTRY();
  st3(a.b2, i, zero3());
  if (!in_vortex_domain(p, xi)) return;
  FarT acc{0.0f, zero3(), zero3(), zero3()};
  for (int y = 0; y < splits; ++y) {
    const int64_t o = 2 * ((int64_t)y * bound + f);
    const float4 u = part[o], v = part[o + 1];
    acc.n += u.w;
    acc.sum = vadd(acc.sum, mk(u.x, u.y, u.z));
    acc.xbu = vadd(acc.xbu, mk(v.x, v.y, v.z));
  }
  far_t_finish(p, a, i, nl, ns, nn, ld3(a.g_in, i), acc);
}

// Targets packed for the source side: (x_i, -) and (weights_i, flag) per
// fluid entry; entries without weight are all zero.
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
// (through c_k) and x_bar_k.  One lane per sample; targets are read from
// their packed table (scalar loads, the next four loaded while four are
// summed) in blockIdx.y chunks.  A target without weight contributes 0;
// a group of four such targets is skipped.  Each (chunk, sample) stores its
// partial (w_bar, c sum) (x_bar, 0) at part[2 (y cap + q)]; the finish pass
// adds them in chunk order.  in: f0 = w, f1 = kden, f2 = per-target weights
// (b2 of the target pass), mode = sample bound.  out: x_bar, b0 = w_bar,
// b1 = kden_bar.
// Per lane, unscaled by c_k: wbar = sum A (w x r_xy), cbar = sum A t.w,
// K = sum A^2 (t.w) r and Wt = sum A w (w = the target's weights,
// t = r_xy x w_k); far_s_close forms the lane's (w_bar, x_bar) from them.
struct FarS {
  v3 wbar, K, Wt;
  float cbar;
};

__device__ __forceinline__ void far_s_pair(FarS& acc, v3 xs, v3 wk, float T, const float4& x4, const float4& w4) {
  const v3 r = mk(x4.x - xs.x, x4.y - xs.y, x4.z - xs.z);
  const float r2 = __fadd_rn(__fadd_rn(__fmul_rn(r.x, r.x), __fmul_rn(r.y, r.y)), __fmul_rn(r.z, r.z));
  const float A = r2 > T ? __builtin_amdgcn_rcpf(r2) : 0.0f;
  const v3 w = mk(w4.x, w4.y, w4.z);
  const v3 t = mk(r.y * wk.z, -(r.x * wk.z), r.x * wk.y - r.y * wk.x);  // r_xy x w_k
  const float tw = vdot(t, w);
  acc.wbar = vadd(acc.wbar, smul(A, mk(-(w.z * r.y), w.z * r.x, w.x * r.y - w.y * r.x)));  // w x r_xy
  acc.cbar += A * tw;
  acc.K = vadd(acc.K, smul(A * A * tw, r));
  acc.Wt = vadd(acc.Wt, smul(A, w));
}

// x_bar_k = -c sum [A (w_k x w)_xy + (dA/ds / s) (t.w) r], dA/ds / s = -2 A^2.
__device__ __forceinline__ void far_s_close(const FarS& acc, v3 wk, float c, float4& u, float4& v) {
  const float xx = wk.y * acc.Wt.z - wk.z * acc.Wt.y - 2.0f * acc.K.x;
  const float xy = wk.z * acc.Wt.x - wk.x * acc.Wt.z - 2.0f * acc.K.y;
  u = make_float4(c * acc.wbar.x, c * acc.wbar.y, c * acc.wbar.z, acc.cbar);
  v = make_float4(-c * xx, -c * xy, 2.0f * c * acc.K.z, 0.0f);
}

__global__ void __launch_bounds__(256) k_adj_far_sources(dimcv_state st, dimcv_params p, dimcv_adj_args a,
                                                         int32_t bound, const float4* __restrict__ stab,
                                                         const float4* __restrict__ ttab, int32_t cap,
                                                         float4* __restrict__ part) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int M = min(st.counters[DIMCV_C_NUM_SAMPLES], cap);
  const int64_t nf = min((int64_t)bound, (int64_t)st.counters[DIMCV_C_FLUID_LIST]);
  const int64_t chunk = (nf + gridDim.y - 1) / gridDim.y;
  const int64_t f0 = min(nf, (int64_t)blockIdx.y * chunk), f1 = min(f0 + chunk, nf);
  if (q >= M) return;
  const float4 s4 = stab[2 * q], k4 = stab[2 * q + 1];
  const v3 xs = mk(s4.x, s4.y, s4.z), wk = mk(k4.x, k4.y, k4.z);
  const float c = s4.w;
  const float T = p.far_r2_max_rejected;
  FarS acc{zero3(), zero3(), zero3(), 0.0f};
  int64_t t = f0;
  if (t + 3 < f1) {
    float4 x0 = ttab[2 * t], w0 = ttab[2 * t + 1], x1 = ttab[2 * t + 2], w1 = ttab[2 * t + 3];
    float4 x2 = ttab[2 * t + 4], w2 = ttab[2 * t + 5], x3 = ttab[2 * t + 6], w3 = ttab[2 * t + 7];
    for (; t + 3 < f1; t += 4) {
      const int64_t nx = t + 7 < f1 ? t + 4 : t;  // in-range prefetch index
      const float4 nx0 = ttab[2 * nx], nw0 = ttab[2 * nx + 1], nx1 = ttab[2 * nx + 2], nw1 = ttab[2 * nx + 3];
      const float4 nx2 = ttab[2 * nx + 4], nw2 = ttab[2 * nx + 5], nx3 = ttab[2 * nx + 6], nw3 = ttab[2 * nx + 7];
      if (w0.w + w1.w + w2.w + w3.w != 0.0f) {  // uniform
        far_s_pair(acc, xs, wk, T, x0, w0);
        far_s_pair(acc, xs, wk, T, x1, w1);
        far_s_pair(acc, xs, wk, T, x2, w2);
        far_s_pair(acc, xs, wk, T, x3, w3);
      }
      x0 = nx0; w0 = nw0; x1 = nx1; w1 = nw1;
      x2 = nx2; w2 = nw2; x3 = nx3; w3 = nw3;
    }
  }
  for (; t < f1; ++t) far_s_pair(acc, xs, wk, T, ttab[2 * t], ttab[2 * t + 1]);
  const int64_t o = 2 * ((int64_t)blockIdx.y * cap + q);
  far_s_close(acc, wk, c, part[o], part[o + 1]);
}

// Sum of the chunk partials: block of 1024 threads per 64 samples, wave w
// adding chunks w, w + 16, ...; the 16 wave sums are added in wave order.
// Several samples may be the same particle, hence the atomics.
__global__ void __launch_bounds__(1024) k_adj_far_sources_finish(dimcv_state st, dimcv_adj_args a, int32_t cap,
                                                                 int32_t chunks, const float4* __restrict__ stab,
                                                                 const float4* __restrict__ part) {
  __shared__ float red[16][7][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int q = blockIdx.x * 64 + lane;
  const int M = min(st.counters[DIMCV_C_NUM_SAMPLES], cap);
  float s[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (q < M) {
#pragma unroll 4
    for (int y = w; y < chunks; y += 16) {
      const int64_t o = 2 * ((int64_t)y * cap + q);
      const float4 u = part[o], v = part[o + 1];
      s[0] += u.x;
      s[1] += u.y;
      s[2] += u.z;
      s[3] += u.w;
      s[4] += v.x;
      s[5] += v.y;
      s[6] += v.z;
    }
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) red[w][k][lane] = s[k];
  __syncthreads();
  if (w != 0 || q >= M) return;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    float t = red[0][k][lane];
    for (int u = 1; u < 16; ++u) t += red[u][k][lane];
    s[k] = t;
  }
  const float c = stab[2 * q].w;
  const int64_t sk = __float_as_int(stab[2 * q + 1].w);
  atomic_add3(a.b0, sk, mk(s[0], s[1], s[2]));
  atomicAdd(a.b1 + sk, -c * c * s[3]);
  atomic_add3(a.x_bar, sk, mk(s[4], s[5], s[6]));
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
    case DIMCV_ADJ_FACTOR: {
      // G_i = sum_j m_j gW_ij in slot order, in the workspace's near-field region
      Workspace ws;
      int rc = bind_workspace(st, p, &ws, 0);
      if (rc) return rc;
      hipLaunchKernelGGL(k_adj_factor_sum, g, b, 0, s, *st, *p, *a, fluid_bound, ws.near3);
      hipLaunchKernelGGL(k_adj_factor, g, b, 0, s, *st, *p, *a, fluid_bound, (const float*)ws.near3);
      break;
    }
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
      // a->mode bounds the sample count M.  The far-field partials region
      // holds the sample table (M entries of 32 bytes) followed by the
      // split partials (32 bytes each); the target table of the source side
      // goes to the sample region (n_max entries).
      if (a->mode <= 0) break;
      Workspace ws;
      int rc = bind_workspace(st, p, &ws, 0);
      if (rc) return rc;
      const int64_t M = a->mode, room = ws.far_partial_cap - M;
      if (room < M || (int64_t)fluid_bound > ws.max_samples) {
        set_error("dimcv_adj_pass: far-field tables exceed the workspace (%d samples, %d targets)", a->mode,
                  fluid_bound);
        return DIMCV_ERR_WORKSPACE;
      }
      float4* stab = (float4*)ws.far_partials;
      float4* part = stab + 2 * M;
      hipLaunchKernelGGL(k_adj_pack_samples, dim3(nblocks(a->mode)), b, 0, s, *st, *a, a->mode, stab);
      if (pass == DIMCV_ADJ_FAR_TARGETS) {
        // enough splits for ~8192 waves, at least 64 samples each
        const int64_t waves = (fluid_bound + 63) / 64;
        const int64_t sp = std::max<int64_t>(
            1, std::min<int64_t>({16, (8192 + waves - 1) / waves, M / 64, room / fluid_bound}));
        hipLaunchKernelGGL(k_adj_far_targets, dim3(g.x, (unsigned)sp), b, 0, s, *st, *p, *a, fluid_bound, a->mode,
                           (const float4*)stab, (int32_t)sp, part);
        if (sp > 1)
          hipLaunchKernelGGL(k_adj_far_targets_finish, g, b, 0, s, *st, *p, *a, fluid_bound, (int32_t)sp,
                             (const float4*)part);
      } else {
        hipLaunchKernelGGL(k_adj_pack_targets, g, b, 0, s, *st, *a, fluid_bound, ws.samples);
        // target chunks over blockIdx.y: ~4096 waves, at least 64 targets each
        const int64_t sw = (M + 63) / 64;
        const int64_t ych = std::max<int64_t>(
            1, std::min<int64_t>({1024, (4096 + sw - 1) / sw, fluid_bound / 64, room / M}));
        hipLaunchKernelGGL(k_adj_far_sources, dim3(nblocks(a->mode), (unsigned)ych), b, 0, s, *st, *p, *a,
                           fluid_bound, (const float4*)stab, (const float4*)ws.samples, a->mode, part);
        hipLaunchKernelGGL(k_adj_far_sources_finish, dim3((unsigned)((M + 63) / 64)), dim3(1024), 0, s, *st, *a,
                           a->mode, (int32_t)ych, (const float4*)stab, (const float4*)part);
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
