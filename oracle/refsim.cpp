// Oracle: CPU restatement of the DIMCV-SPH particle step.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (dimcv-sph_amd/) links,
// loads or calls this file; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py do, as the checker.
//
// Each function restates one Taichi kernel of the reference with serial
// semantics (ascending slot order, stable counting sort, ascending atomic
// compaction) and the exact f32 operation order of the reference
// expressions.  Built with -ffp-contract=off (no fused multiply-add), IEEE
// division and sqrt.  Particle loops that only write their own slot run under
// OpenMP; this does not change any result.
//
// Build-defined deviations from the reference (each one is required by the
// host-fed-uniform, FP64-reduction design of this path or removes an
// unpinnable backend detail):
//  * exp/log/pow: refmath.h (ti.exp / ti.pow last-ulp behaviour is backend
//    defined);  ti.pow(t, 3.0) is evaluated as (t*t)*t.
//  * residual sums (dimcv_sph.py:407-419) and kden_sum (:738-751) are FP64
//    sums in a fixed tree order (canon_reduce) instead of f32 atomics;
//    vort2vel uses S_i = f32(kden_sum - f64(kden_nbr)).
//  * surface tension (sigma = 0, dimcv_sph.py:12,153-167) and boundary
//    viscosity (= 0, :182-190) contribute exact zeros and are not evaluated.
//  * the viscosity reaction on dynamic rigid neighbours (:191-193) is an exact
//    zero (boundary viscosity 0) and is not evaluated; the Jacobi updates'
//    reactions (:465-468, :546-549) are added serially in ascending p_i.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "refmath.h"

using refmath::expf_c;
using refmath::powf_c;

namespace {

struct V3 {
  float x, y, z;
};
inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 smul(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 vmuls(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 vdivs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// Taichi dot / norm_sqr: elementwise products summed left to right.
inline float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float vnorm(V3 a) { return std::sqrt(vdot(a, a)); }
inline V3 vcross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// Canonical FP64 reduction: groups of 256 consecutive values, each summed by
// a halving tree (a[t] += a[t+s], s = 128..1, zero padded), repeated on the
// group sums until one value remains.
double canon_reduce_vec(std::vector<double> cur) {
  if (cur.empty()) return 0.0;
  for (;;) {
    size_t G = (cur.size() + 255) / 256;
    std::vector<double> nxt(G);
    for (size_t g = 0; g < G; ++g) {
      double b[256];
      for (int t = 0; t < 256; ++t) {
        size_t i = g * 256 + t;
        b[t] = i < cur.size() ? cur[i] : 0.0;
      }
      for (int s = 128; s >= 1; s >>= 1)
        for (int t = 0; t < s; ++t) b[t] = b[t] + b[t + s];
      nxt[g] = b[0];
    }
    if (G == 1) return nxt[0];
    cur.swap(nxt);
  }
}

// Canonical FP64 inclusive scan.  Blocks of 4096 values; inside a block,
// 256 "threads" own 16 consecutive values each (serial prefix l0), threads
// form 16 groups of 16 (serial prefix l1 over thread totals), and the 16 group
// totals get a serial prefix L2.  L1[t] = L2[g-1] + l1[t] (g > 0),
// L[i] = L1[t-1] + l0[i] (t > 0).  Block totals are scanned recursively with
// the same definition and added last: out = P[b-1] + L[i] (b > 0).
void block_local_scan(const double* y, double* L, int64_t m) {
  double l0[4096], tt[256], l1[256], gt[16] = {}, L2[16], L1[256];
  int nthreads = (int)((m + 15) / 16);
  for (int t = 0; t < nthreads; ++t) {
    int64_t s = (int64_t)t * 16, e = std::min<int64_t>(s + 16, m);
    double acc = y[s];
    l0[s] = acc;
    for (int64_t i = s + 1; i < e; ++i) { acc = acc + y[i]; l0[i] = acc; }
    tt[t] = acc;
  }
  int ngroups = (nthreads + 15) / 16;
  for (int g = 0; g < ngroups; ++g) {
    int s = g * 16, e = std::min(s + 16, nthreads);
    double acc = tt[s];
    l1[s] = acc;
    for (int t = s + 1; t < e; ++t) { acc = acc + tt[t]; l1[t] = acc; }
    gt[g] = acc;
  }
  double acc = gt[0];
  L2[0] = acc;
  for (int g = 1; g < ngroups; ++g) { acc = acc + gt[g]; L2[g] = acc; }
  for (int t = 0; t < nthreads; ++t) {
    int g = t / 16;
    L1[t] = g == 0 ? l1[t] : L2[g - 1] + l1[t];
  }
  for (int64_t i = 0; i < m; ++i) {
    int t = (int)(i / 16);
    L[i] = t == 0 ? l0[i] : L1[t - 1] + l0[i];
  }
}

void canon_scan_rec(const double* x, double* out, int64_t n) {
  if (n <= 0) return;
  int64_t nb = (n + 4095) / 4096;
  std::vector<double> L(n);
  for (int64_t b = 0; b < nb; ++b) {
    int64_t s = b * 4096, m = std::min<int64_t>(4096, n - s);
    block_local_scan(x + s, L.data() + s, m);
  }
  if (nb == 1) {
    std::copy(L.begin(), L.end(), out);
    return;
  }
  std::vector<double> T(nb), P(nb);
  for (int64_t b = 0; b < nb; ++b) T[b] = L[std::min<int64_t>(b * 4096 + 4095, n - 1)];
  canon_scan_rec(T.data(), P.data(), nb);
  for (int64_t i = 0; i < n; ++i) {
    int64_t b = i / 4096;
    out[i] = b == 0 ? L[i] : P[b - 1] + L[i];
  }
}

}  // namespace

// Scene configuration as read by the reference constructors
// (particle_system.py:16-46,121-123; sph_base.py:9-20; dimcv_sph.py:13-64).
struct OrConfig {
  double domain_start[3];
  double domain_end[3];
  double particle_radius;
  double density0;
  double gravitation[3];
  double time_step;
  double volume_scale;
  double gen_a;
  double del_b;
  double vortex_start[3];
  double vortex_end[3];
  double transit[2];
  double viscosity;
  int64_t particle_max_num;
  int32_t has_vortex_domain;
  int32_t has_transit;
  int32_t num_emitters;
  int32_t periodic;  // build-defined periodic box (minimum image, wrap)
};

struct Sim {
  OrConfig cfg;
  int64_t N;
  int grid_num[3];
  int64_t ncells;
  // host f64 values (Python floats in the reference)
  double radius, diameter, support, inv_support, m_V0, vp_volume;
  // f32 constants as they enter the kernels
  float f_support, f_inv_support, f_grid_size, kW, kdW, f_density0, f_m_eps;
  float f_dstart[3], f_dend[3], f_vstart[3], f_vend[3], f_g[3];
  float f_visc_coef, f_visc_h2, f_transit_a, f_transit_ba, f_vp_factor;
  float f_eps, f_inv_eps, f_neg_a, f_neg_b;
  float f_box[3];
  float dt, f_inv_dt;
  // scalar fields
  int32_t particle_num, fluid_particle_num, solid_particle_num, num_samples;
  double kden_sum;
  int32_t err_key_range;
  std::vector<int32_t> num_particles_each_emitter;
  // ParticleSystem fields (particle_system.py:143-174)
  std::vector<int32_t> object_id, material, color, grid_ids;
  std::vector<float> x, x_0, v, acceleration, m_V, m, density, pressure;
  std::vector<uint8_t> is_dynamic, is_active, is_sample;
  std::vector<float> vorticity, vorticity_vis, dfsph_factor, density_adv, life_time;
  // sort buffers (:176-220)
  std::vector<int32_t> object_id_b, material_b, color_b, grid_ids_b, grid_ids_new;
  std::vector<float> x_b, x_0_b, v_b, acceleration_b, m_V_b, m_b, density_b, pressure_b;
  std::vector<uint8_t> is_dynamic_b, is_active_b, is_sample_b;
  std::vector<float> vorticity_b, dfsph_factor_b, density_adv_b, life_time_b;
  std::vector<int32_t> grid_particles_num, grid_particles_num_temp;
  // DIMCVSPHSolver fields (dimcv_sph.py:30-53)
  std::vector<float> d_vort, d_vort_smoothed, v_star, grad_v, x_temp, kvn, kernel_density;
  std::vector<int32_t> sample_idx;

  V3 X(int64_t i) const { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
  V3 Vv(int64_t i) const { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }
  static V3 get(const std::vector<float>& a, int64_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
  static void put(std::vector<float>& a, int64_t i, V3 q) {
    a[3 * i] = q.x; a[3 * i + 1] = q.y; a[3 * i + 2] = q.z;
  }
  static V3 get_ptr(const float* a, int64_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
  static void put_ptr(float* a, int64_t i, V3 q) {
    a[3 * i] = q.x; a[3 * i + 1] = q.y; a[3 * i + 2] = q.z;
  }
  bool is_fluid(int64_t i) const { return material[i] == 1; }
  bool is_solid(int64_t i) const { return material[i] == 0; }
  bool is_static_rigid(int64_t i) const { return material[i] == 0 && !is_dynamic[i]; }
  bool is_dynamic_rigid(int64_t i) const { return material[i] == 0 && is_dynamic[i]; }

  // sph_base.py:23-44
  float W(float r_norm) const {
    float res = 0.0f;
    float q = r_norm * f_inv_support;
    if (q <= 1.0f) {
      if (q <= 0.5f) {
        float q2 = q * q;
        float q3 = q2 * q;
        res = kW * (6.0f * q3 - 6.0f * q2 + 1.0f);
      } else {
        float t = 1.0f - q;
        res = kW * 2.0f * ((t * t) * t);
      }
    }
    return res;
  }
  // sph_base.py:46-68
  V3 gradW(V3 r) const {
    float rn = vnorm(r);
    float q = rn * f_inv_support;
    V3 res{0.0f, 0.0f, 0.0f};
    if (rn > 1e-5f && q <= 1.0f) {
      V3 gq = vmuls(vdivs(r, rn), f_inv_support);
      if (q <= 0.5f) {
        res = smul(kdW * q * (3.0f * q - 2.0f), gq);
      } else {
        float f = 1.0f - q;
        res = smul(kdW * (-f * f), gq);
      }
    }
    return res;
  }
  // x_i - x_j; in the periodic variant the minimum image
  V3 rel(V3 a, V3 b) const {
    V3 r = vsub(a, b);
    if (cfg.periodic) {
      float L[3] = {f_box[0], f_box[1], f_box[2]};
      float* q[3] = {&r.x, &r.y, &r.z};
      for (int d = 0; d < 3; ++d) {
        float hl = 0.5f * L[d];
        if (*q[d] > hl) *q[d] -= L[d];
        else if (*q[d] < -hl) *q[d] += L[d];
      }
    }
    return r;
  }
  // particle_system.py:382-402 (periodic variant: clamped to the last cell)
  void cell_of(V3 p, int c[3]) const {
    c[0] = (int)(p.x / f_grid_size);
    c[1] = (int)(p.y / f_grid_size);
    c[2] = (int)(p.z / f_grid_size);
    if (cfg.periodic)
      for (int d = 0; d < 3; ++d) c[d] = std::min(std::max(c[d], 0), grid_num[d] - 1);
  }
  int64_t flat(const int c[3]) const {
    return (int64_t)c[0] * grid_num[1] * grid_num[2] + (int64_t)c[1] * grid_num[2] + c[2];
  }
  // particle_system.py:562-586: 27 cells in ndrange order, range
  // [gpn[max(0,c-1)], gpn[c]) of the inclusive scan (cell 0 is never visited),
  // strict sqrt test, j != i.
  template <class F>
  void for_all_neighbors(int64_t p_i, F&& task) const {
    V3 xi = X(p_i);
    int cc[3];
    cell_of(xi, cc);
    for (int o0 = -1; o0 <= 1; ++o0)
      for (int o1 = -1; o1 <= 1; ++o1)
        for (int o2 = -1; o2 <= 1; ++o2) {
          int g[3] = {cc[0] + o0, cc[1] + o1, cc[2] + o2};
          bool ok = true;
          for (int d = 0; d < 3; ++d) {
            if (cfg.periodic) g[d] = (g[d] + grid_num[d]) % grid_num[d];
            else if (ok && (g[d] < 0 || g[d] >= grid_num[d])) ok = false;
          }
          if (!ok) continue;
          int64_t fl = flat(g);
          int32_t beg = grid_particles_num[std::max<int64_t>(0, fl - 1)];
          int32_t end = grid_particles_num[fl];
          for (int64_t p_j = beg; p_j < end; ++p_j)
            if (p_i != p_j && vnorm(rel(xi, X(p_j))) < f_support) task(p_j);
        }
  }
  // dimcv_sph.py:76-84
  bool in_vortex_domain(V3 p) const {
    if (!cfg.has_vortex_domain) return true;
    float q[3] = {p.x, p.y, p.z};
    bool flag = true;
    for (int d = 0; d < 3; ++d)
      if (flag && (q[d] < f_vstart[d] || q[d] > f_vend[d])) flag = false;
    return flag;
  }
  // dimcv_sph.py:401-405
  float transit_weight(float xpos) const {
    float w = (xpos - f_transit_a) / f_transit_ba;
    return std::fmin(std::fmax(w, 0.0f), 1.0f);
  }
};

extern "C" {

static void resize_all(Sim* s) {
  int64_t N = s->N;
  auto r1 = [N](auto& a) { a.assign(N, 0); };
  auto r3 = [N](auto& a) { a.assign(3 * N, 0); };
  r1(s->object_id); r1(s->material); r3(s->color); r1(s->grid_ids);
  r3(s->x); r3(s->x_0); r3(s->v); r3(s->acceleration);
  r1(s->m_V); r1(s->m); r1(s->density); r1(s->pressure);
  r1(s->is_dynamic); r1(s->is_active); r1(s->is_sample);
  r3(s->vorticity); r3(s->vorticity_vis); r1(s->dfsph_factor); r1(s->density_adv); r1(s->life_time);
  r1(s->object_id_b); r1(s->material_b); r3(s->color_b); r1(s->grid_ids_b); r1(s->grid_ids_new);
  r3(s->x_b); r3(s->x_0_b); r3(s->v_b); r3(s->acceleration_b);
  r1(s->m_V_b); r1(s->m_b); r1(s->density_b); r1(s->pressure_b);
  r1(s->is_dynamic_b); r1(s->is_active_b); r1(s->is_sample_b);
  r3(s->vorticity_b); r1(s->dfsph_factor_b); r1(s->density_adv_b); r1(s->life_time_b);
  r3(s->d_vort); r3(s->d_vort_smoothed); r3(s->v_star); s->grad_v.assign(9 * N, 0.0f);
  r3(s->x_temp); r1(s->kvn); r1(s->kernel_density); r1(s->sample_idx);
  s->grid_particles_num.assign(s->ncells, 0);
  s->grid_particles_num_temp.assign(s->ncells, 0);
}

Sim* or_new(const OrConfig* c) {
  Sim* s = new Sim();
  s->cfg = *c;
  s->N = c->particle_max_num;
  // particle_system.py:41-58
  s->radius = c->particle_radius;
  s->diameter = 2 * s->radius;
  s->support = s->radius * 4.0;
  s->inv_support = 1.0 / s->support;
  s->m_V0 = 0.8 * std::pow(s->diameter, 3.0);
  s->ncells = 1;
  for (int d = 0; d < 3; ++d) {
    double size = c->domain_end[d] - c->domain_start[d];
    s->grid_num[d] = (int)std::ceil(size / s->support);
    s->ncells *= s->grid_num[d];
  }
  // dimcv_sph.py:17-19
  s->vp_volume = c->volume_scale * s->m_V0;
  double inv3 = std::pow(s->inv_support, 3.0);
  s->f_support = (float)s->support;
  s->f_inv_support = (float)s->inv_support;
  s->f_grid_size = (float)s->support;
  s->kW = 2.5465f * (float)inv3;
  s->kdW = (6.0f * 2.5465f) * (float)inv3;
  s->f_density0 = (float)c->density0;
  s->f_m_eps = 1e-5f;
  for (int d = 0; d < 3; ++d) {
    s->f_dstart[d] = (float)c->domain_start[d];
    s->f_dend[d] = (float)c->domain_end[d];
    s->f_vstart[d] = (float)c->vortex_start[d];
    s->f_vend[d] = (float)c->vortex_end[d];
    s->f_g[d] = (float)c->gravitation[d];
    s->f_box[d] = (float)(c->domain_end[d] - c->domain_start[d]);
  }
  s->f_visc_coef = (float)(2 * (3 + 2) * c->viscosity);
  s->f_visc_h2 = (float)(0.01 * std::pow(s->support, 2.0));
  s->f_transit_a = (float)c->transit[0];
  s->f_transit_ba = (float)(c->transit[1] - c->transit[0]);
  s->f_vp_factor = (float)(0.1592 * s->vp_volume);
  s->f_eps = 1e-6f;
  s->f_inv_eps = (float)(1.0 / 1e-6);
  s->f_neg_a = (float)(-c->gen_a);
  s->f_neg_b = (float)(-c->del_b);
  s->dt = (float)c->time_step;
  s->f_inv_dt = (float)(1.0 / (double)s->dt);
  s->particle_num = 0;
  s->fluid_particle_num = 0;
  s->solid_particle_num = 0;
  s->num_samples = 0;
  s->kden_sum = 0.0;
  s->err_key_range = 0;
  s->num_particles_each_emitter.assign(std::max(1, c->num_emitters), 0);
  resize_all(s);
  return s;
}

void or_free(Sim* s) { delete s; }

// Field access for the Python layer.
void* or_ptr(Sim* s, const char* name) {
#define F(nm) if (!std::strcmp(name, #nm)) return (void*)s->nm.data();
  F(object_id) F(material) F(color) F(grid_ids) F(x) F(x_0) F(v) F(acceleration)
  F(m_V) F(m) F(density) F(pressure) F(is_dynamic) F(is_active) F(is_sample)
  F(vorticity) F(vorticity_vis) F(dfsph_factor) F(density_adv) F(life_time)
  F(grid_particles_num) F(grid_particles_num_temp) F(d_vort) F(d_vort_smoothed)
  F(v_star) F(grad_v) F(x_temp) F(kvn) F(kernel_density) F(sample_idx)
  F(num_particles_each_emitter)
#undef F
  if (!std::strcmp(name, "particle_num")) return &s->particle_num;
  if (!std::strcmp(name, "fluid_particle_num")) return &s->fluid_particle_num;
  if (!std::strcmp(name, "solid_particle_num")) return &s->solid_particle_num;
  if (!std::strcmp(name, "num_samples")) return &s->num_samples;
  if (!std::strcmp(name, "kden_sum")) return &s->kden_sum;
  if (!std::strcmp(name, "dt")) return &s->dt;
  if (!std::strcmp(name, "err_key_range")) return &s->err_key_range;
  return nullptr;
}

void or_info(Sim* s, int64_t* out) {
  out[0] = s->N; out[1] = s->grid_num[0]; out[2] = s->grid_num[1]; out[3] = s->grid_num[2];
  out[4] = s->ncells;
}

// particle_system.py:316-332,349-380 (silent drop beyond capacity, :361)
int32_t or_add_particles(Sim* s, int32_t object_id, int32_t n, const float* pos, const float* vel,
                         const float* density, const float* pressure, const int32_t* material,
                         const int32_t* is_dynamic, const int32_t* color) {
  int32_t added = 0;
  float mv0 = (float)s->m_V0;
  for (int64_t p = s->particle_num; p < (int64_t)s->particle_num + n; ++p) {
    if (p < s->N) {
      int64_t k = p - s->particle_num;
      s->object_id[p] = object_id;
      for (int d = 0; d < 3; ++d) {
        s->x[3 * p + d] = pos[3 * k + d];
        s->x_0[3 * p + d] = pos[3 * k + d];
        s->v[3 * p + d] = vel[3 * k + d];
        s->color[3 * p + d] = color[3 * k + d];
      }
      s->density[p] = density[k];
      s->m_V[p] = mv0;
      s->m[p] = mv0 * density[k];
      s->pressure[p] = pressure[k];
      s->material[p] = material[k];
      s->is_dynamic[p] = is_dynamic[k] != 0;
      s->is_active[p] = 1;
      if (material[k] == 1) s->fluid_particle_num += 1;
      ++added;
    }
  }
  s->particle_num += added;
  return added;
}

// karman_vortex.py:27-36
void or_init_cylinder(Sim* s, int32_t obj, float cx, float cy, float radius) {
  for (int64_t p = 0; p < s->particle_num; ++p) {
    if (s->object_id[p] == obj) {
      float dx = s->x[3 * p] - cx;
      float dy = s->x[3 * p + 1] - cy;
      if (std::sqrt(dx * dx + dy * dy) > radius) {
        s->x[3 * p] = 0.0f; s->x[3 * p + 1] = 0.0f; s->x[3 * p + 2] = 0.0f;
        s->is_active[p] = 0;
      }
    }
  }
}

// particle_system.py:413-427
void or_update_activity(Sim* s) {
  for (int64_t i = 0; i < s->particle_num; ++i) {
    if (s->material[i] == 0) {
      s->is_active[i] = 1;
    } else if (s->material[i] == 1 && s->cfg.periodic) {
      for (int j = 0; j < 3; ++j) {
        float& q = s->x[3 * i + j];
        if (q < s->f_dstart[j]) q += s->f_box[j];
        else if (q >= s->f_dend[j]) q -= s->f_box[j];
      }
    } else if (s->material[i] == 1) {
      // outside the domain box; or, build-defined, an active particle whose
      // flat cell index lies past the cell table (exactly on the upper x
      // face, or the last row's upper y/z face), where the reference indexes
      // out of bounds: it leaves as one past the face does
      bool leave = false;
      for (int j = 0; j < 3; ++j) {
        float xj = s->x[3 * i + j];
        if (xj < s->f_dstart[j] || xj > s->f_dend[j]) leave = true;
      }
      if (!leave && s->is_active[i]) {
        int c[3];
        s->cell_of(s->X(i), c);
        const int64_t g = s->flat(c);
        leave = g < 0 || g >= s->ncells;
      }
      if (leave) {
        s->is_active[i] = 0;
        s->fluid_particle_num -= 1;
        if (s->cfg.num_emitters > 0) {
          int32_t o = s->object_id[i];
          if (o >= 0 && o < (int32_t)s->num_particles_each_emitter.size())
            s->num_particles_each_emitter[o] -= 1;
        }
      }
    }
  }
}

// The 19 fields moved by particle_partition / counting_sort.
static void move_slot(Sim* s, int64_t dst, int64_t src, bool to_buffer) {
  if (to_buffer) {
    s->grid_ids_b[dst] = s->grid_ids[src];
    s->object_id_b[dst] = s->object_id[src];
    for (int d = 0; d < 3; ++d) {
      s->x_0_b[3 * dst + d] = s->x_0[3 * src + d];
      s->x_b[3 * dst + d] = s->x[3 * src + d];
      s->v_b[3 * dst + d] = s->v[3 * src + d];
      s->vorticity_b[3 * dst + d] = s->vorticity[3 * src + d];
      s->acceleration_b[3 * dst + d] = s->acceleration[3 * src + d];
      s->color_b[3 * dst + d] = s->color[3 * src + d];
    }
    s->m_V_b[dst] = s->m_V[src]; s->m_b[dst] = s->m[src];
    s->density_b[dst] = s->density[src]; s->pressure_b[dst] = s->pressure[src];
    s->material_b[dst] = s->material[src]; s->is_dynamic_b[dst] = s->is_dynamic[src];
    s->dfsph_factor_b[dst] = s->dfsph_factor[src]; s->density_adv_b[dst] = s->density_adv[src];
    s->is_active_b[dst] = s->is_active[src]; s->is_sample_b[dst] = s->is_sample[src];
    s->life_time_b[dst] = s->life_time[src];
  } else {
    s->grid_ids[dst] = s->grid_ids_b[src];
    s->object_id[dst] = s->object_id_b[src];
    for (int d = 0; d < 3; ++d) {
      s->x_0[3 * dst + d] = s->x_0_b[3 * src + d];
      s->x[3 * dst + d] = s->x_b[3 * src + d];
      s->v[3 * dst + d] = s->v_b[3 * src + d];
      s->vorticity[3 * dst + d] = s->vorticity_b[3 * src + d];
      s->acceleration[3 * dst + d] = s->acceleration_b[3 * src + d];
      s->color[3 * dst + d] = s->color_b[3 * src + d];
    }
    s->m_V[dst] = s->m_V_b[src]; s->m[dst] = s->m_b[src];
    s->density[dst] = s->density_b[src]; s->pressure[dst] = s->pressure_b[src];
    s->material[dst] = s->material_b[src]; s->is_dynamic[dst] = s->is_dynamic_b[src];
    s->dfsph_factor[dst] = s->dfsph_factor_b[src]; s->density_adv[dst] = s->density_adv_b[src];
    s->is_active[dst] = s->is_active_b[src]; s->is_sample[dst] = s->is_sample_b[src];
    s->life_time[dst] = s->life_time_b[src];
  }
}

// particle_system.py:429-488 (ascending compaction = serial atomic order)
void or_particle_partition(Sim* s) {
  int64_t new_idx = 0;
  for (int64_t i = 0; i < s->particle_num; ++i)
    if (s->is_active[i] == 1) move_slot(s, new_idx++, i, true);
  std::fill(s->grid_ids.begin(), s->grid_ids.end(), -1);
  std::fill(s->object_id.begin(), s->object_id.end(), -1);
  std::fill(s->x_0.begin(), s->x_0.end(), -1.0f);
  std::fill(s->x.begin(), s->x.end(), -1.0f);
  std::fill(s->v.begin(), s->v.end(), 0.0f);
  std::fill(s->acceleration.begin(), s->acceleration.end(), 0.0f);
  std::fill(s->color.begin(), s->color.end(), 255);
  std::fill(s->is_dynamic.begin(), s->is_dynamic.end(), 0);
  std::fill(s->is_active.begin(), s->is_active.end(), 0);
  std::fill(s->is_sample.begin(), s->is_sample.end(), 0);
  std::fill(s->life_time.begin(), s->life_time.end(), 0.0f);
  s->particle_num = s->fluid_particle_num + s->solid_particle_num;
  for (int64_t i = 0; i < s->particle_num; ++i) move_slot(s, i, i, false);
}

// particle_system.py:490-498.  An out-of-grid key is undefined behaviour in
// the reference; the oracle counts it and skips the histogram write.
void or_update_grid_id(Sim* s) {
  std::fill(s->grid_particles_num.begin(), s->grid_particles_num.end(), 0);
  for (int64_t I = 0; I < s->particle_num; ++I) {
    int c[3];
    s->cell_of(s->X(I), c);
    int64_t g = s->flat(c);
    s->grid_ids[I] = (int32_t)g;
    if (g < 0 || g >= s->ncells) { s->err_key_range += 1; continue; }
    s->grid_particles_num[g] += 1;
  }
  s->grid_particles_num_temp = s->grid_particles_num;
}

// particle_system.py:559 (PrefixSumExecutor: inclusive, in place)
void or_prefix_sum(Sim* s) {
  int64_t acc = 0;
  for (int64_t c = 0; c < s->ncells; ++c) {
    acc += s->grid_particles_num[c];
    s->grid_particles_num[c] = (int32_t)acc;
  }
}

// particle_system.py:500-553 (serial reverse loop: stable)
void or_counting_sort(Sim* s) {
  std::fill(s->grid_ids_new.begin(), s->grid_ids_new.end(), -1);
  for (int64_t i = 0; i < s->particle_num; ++i) {
    int64_t I = s->particle_num - 1 - i;
    int32_t g = s->grid_ids[I];
    int32_t base = 0;
    if (g - 1 >= 0) base = s->grid_particles_num[g - 1];
    s->grid_particles_num_temp[g] -= 1;
    s->grid_ids_new[I] = s->grid_particles_num_temp[g] + base;
  }
  for (int64_t I = 0; I < s->particle_num; ++I) move_slot(s, s->grid_ids_new[I], I, true);
  for (int64_t I = 0; I < s->particle_num; ++I) move_slot(s, I, I, false);
}

// particle_system.py:555-560
void or_initialize_particle_system(Sim* s) {
  or_update_activity(s);
  or_particle_partition(s);
  or_update_grid_id(s);
  or_prefix_sum(s);
  or_counting_sort(s);
}

// sph_base.py:91-103
void or_compute_static_boundary_volume(Sim* s) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t p_i = 0; p_i < s->N; ++p_i) {
    if (!s->is_static_rigid(p_i)) continue;
    float delta = s->W(0.0f);
    V3 xi = s->X(p_i);
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_solid(p_j)) delta += s->W(vnorm(s->rel(xi, s->X(p_j))));
    });
    s->m_V[p_i] = 1.0f / delta * 3.0f;
  }
}

// sph_base.py:105-112
void or_compute_moving_boundary_volume(Sim* s) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t p_i = 0; p_i < s->N; ++p_i) {
    if (!s->is_dynamic_rigid(p_i)) continue;
    float delta = s->W(0.0f);
    V3 xi = s->X(p_i);
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_solid(p_j)) delta += s->W(vnorm(s->rel(xi, s->X(p_j))));
    });
    s->m_V[p_i] = 1.0f / delta * 3.0f;
  }
}

// dimcv_sph.py:126-148
void or_compute_densities(Sim* s) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i);
    float dens = s->m_V[p_i] * s->W(0.0f);
    float den = 0.0f;
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j)) den += s->m_V[p_j] * s->W(vnorm(s->rel(xi, s->X(p_j))));
      else if (s->is_solid(p_j)) den += s->m_V[p_j] * s->W(vnorm(s->rel(xi, s->X(p_j))));
    });
    dens += den;
    dens *= s->f_density0;
    s->density[p_i] = dens;
  }
}

// dimcv_sph.py:86-124
void or_compute_DFSPH_factor(Sim* s) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i);
    float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f, r3 = 0.0f;
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j)) {
        V3 g = smul(s->m[p_j], s->gradW(s->rel(xi, s->X(p_j))));
        r3 += vdot(g, g);
        r0 += g.x; r1 += g.y; r2 += g.z;
      } else if (s->is_solid(p_j)) {
        V3 g = smul(s->m_V[p_j] * s->f_density0, s->gradW(s->rel(xi, s->X(p_j))));
        r0 += g.x; r1 += g.y; r2 += g.z;
      }
    });
    float sum = r3;
    V3 gp{r0, r1, r2};
    sum += vdot(gp, gp);
    float factor = 0.0f;
    if (sum > 1e-6f) factor = 1.0f / sum;
    else factor = 0.0f;
    s->dfsph_factor[p_i] = factor;
  }
}

// dimcv_sph.py:337-374
void or_compute_density_change(Sim* s) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i), vi = s->Vv(p_i);
    float dadv = 0.0f;
    int nn = 0;
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      V3 vj = s->Vv(p_j);
      if (s->is_fluid(p_j))
        dadv += s->m_V[p_j] * vdot(vsub(vi, vj), s->gradW(s->rel(xi, s->X(p_j))));
      else if (s->is_solid(p_j))
        dadv += s->m_V[p_j] * vdot(vsub(vi, vj), s->gradW(s->rel(xi, s->X(p_j))));
      nn += 1;
    });
    float d = std::fmax(dadv, 0.0f);
    if (nn < 20) d = 0.0f;
    s->density_adv[p_i] = d;
  }
}

// dimcv_sph.py:395-399.  which: 0 = dfsph_factor
void or_multiply_time_step(Sim* s, int32_t which, float ts) {
  std::vector<float>& f = s->dfsph_factor;
  (void)which;
  for (int64_t I = 0; I < s->particle_num; ++I)
    if (s->is_fluid(I)) f[I] *= ts;
}

// dimcv_sph.py:407-419 -> canonical FP64 sum over fluid slots (ascending).
double or_compute_density_error(Sim* s, float offset) {
  std::vector<double> terms;
  terms.reserve(s->particle_num);
  for (int64_t I = 0; I < s->particle_num; ++I) {
    if (!s->is_fluid(I)) continue;
    float t;
    if (!s->cfg.has_transit) t = s->density_adv[I] - offset;
    else t = s->transit_weight(s->x[3 * I]) * (s->density_adv[I] - offset);
    terms.push_back((double)t);
  }
  return canon_reduce_vec(std::move(terms));
}

// dimcv_sph.py:465-468, 546-549: a Jacobi update's reaction on the dynamic
// rigid neighbours of each fluid particle, acceleration[j] += -vel_change
// inv_dt rho_i / rho_j, added serially in ascending p_i (kappa_i as the update
// saw it: kappa(i) of the iteration's fields).
}  // extern "C"
template <class K>
static void rigid_reactions(Sim* s, K kappa) {
  const float dt = s->dt;
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    const float k_i = kappa(p_i);
    if (!(std::fabs(k_i) > s->f_m_eps)) continue;
    V3 xi = s->X(p_i);
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (!s->is_dynamic_rigid(p_j)) return;
      V3 g = smul(s->m_V[p_j] * s->f_density0, s->gradW(s->rel(xi, s->X(p_j))));
      V3 vc = smul(-dt * 1.0f * k_i, g);
      const float ri = s->density[p_i], rj = s->density[p_j];
      V3 add{((-vc.x) * s->f_inv_dt) * ri / rj, ((-vc.y) * s->f_inv_dt) * ri / rj, ((-vc.z) * s->f_inv_dt) * ri / rj};
      Sim::put(s->acceleration, p_j, vadd(Sim::get(s->acceleration, p_j), add));
    });
  }
}
extern "C" {

// dimcv_sph.py:421-468
void or_divergence_solver_iteration_kernel(Sim* s) {
  const float dt = s->dt;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i);
    float b_i = s->density_adv[p_i];
    float k_i = b_i * s->dfsph_factor[p_i] * s->density[p_i];
    V3 dv{0.0f, 0.0f, 0.0f};
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j)) {
        float b_j = s->density_adv[p_j];
        float k_j = b_j * s->dfsph_factor[p_j] * s->density[p_j];
        float k_sum = k_i + k_j;
        if (std::fabs(k_sum) > s->f_m_eps) {
          V3 g = smul(s->m[p_j], s->gradW(s->rel(xi, s->X(p_j))));
          dv = vsub(dv, smul(dt * k_sum, g));
        }
      } else if (s->is_solid(p_j)) {
        if (std::fabs(k_i) > s->f_m_eps) {
          V3 g = smul(s->m_V[p_j] * s->f_density0, s->gradW(s->rel(xi, s->X(p_j))));
          V3 vc = smul(-dt * 1.0f * k_i, g);
          dv = vadd(dv, vc);
        }
      }
    });
    V3 vi = s->Vv(p_i);
    if (!s->cfg.has_transit) vi = vadd(vi, dv);
    else vi = vadd(vi, smul(s->transit_weight(xi.x), dv));
    Sim::put(s->v, p_i, vi);
  }
  rigid_reactions(s, [&](int64_t i) { return s->density_adv[i] * s->dfsph_factor[i] * s->density[i]; });
}

// dimcv_sph.py:150-208 (surface tension and boundary viscosity are exact zeros)
void or_compute_non_pressure_forces(Sim* s) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (s->is_static_rigid(p_i)) {
      Sim::put(s->acceleration, p_i, {0.0f, 0.0f, 0.0f});
      continue;
    }
    V3 dv{s->f_g[0], s->f_g[1], s->f_g[2]};
    Sim::put(s->acceleration, p_i, dv);
    if (s->is_fluid(p_i)) {
      V3 xi = s->X(p_i), vi = s->Vv(p_i);
      s->for_all_neighbors(p_i, [&](int64_t p_j) {
        V3 r = s->rel(xi, s->X(p_j));
        float v_xy = vdot(vsub(vi, s->Vv(p_j)), r);
        if (s->is_fluid(p_j)) {
          float rn = vnorm(r);
          V3 fv = smul(s->f_visc_coef * s->m_V[p_j] * v_xy / (rn * rn + s->f_visc_h2), s->gradW(r));
          dv = vadd(dv, fv);
        }
      });
      Sim::put(s->acceleration, p_i, dv);
    }
  }
}

// dimcv_sph.py:210-218
void or_advect(Sim* s) {
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (s->is_dynamic[p_i]) {
      if (s->is_dynamic_rigid(p_i))
        Sim::put(s->v, p_i, vadd(s->Vv(p_i), smul(s->dt, Sim::get(s->acceleration, p_i))));
      Sim::put(s->x, p_i, vadd(s->X(p_i), smul(s->dt, s->Vv(p_i))));
    }
  }
}

// dimcv_sph.py:220-226
void or_predict_velocity(Sim* s) {
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i)
    if (s->is_dynamic[p_i] && s->is_fluid(p_i))
      Sim::put(s->v, p_i, vadd(s->Vv(p_i), smul(s->dt, Sim::get(s->acceleration, p_i))));
}

// dimcv_sph.py:264-271
void or_compute_v_star(Sim* s) {
  std::fill(s->v_star.begin(), s->v_star.end(), 0.0f);
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    Sim::put(s->v_star, p_i, vadd(s->Vv(p_i), vmuls(Sim::get(s->acceleration, p_i), s->dt)));
  }
}

// dimcv_sph.py:228-248,284-293
void or_compute_vorticity(Sim* s) {
  std::fill(s->vorticity.begin(), s->vorticity.end(), 0.0f);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i), vi = s->Vv(p_i);
    V3 vort{0.0f, 0.0f, 0.0f};
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j))
        vort = vadd(vort, smul(s->m_V[p_j], vcross(vsub(s->Vv(p_j), vi), s->gradW(s->rel(xi, s->X(p_j))))));
    });
    Sim::put(s->vorticity, p_i, vort);
  }
}

// dimcv_sph.py:250-262,273-282.  grad_v[i] row-major 3x3.
void or_compute_grad_v(Sim* s) {
  std::fill(s->grad_v.begin(), s->grad_v.end(), 0.0f);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i), vi = s->Vv(p_i);
    float G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j)) {
        V3 vji = vsub(s->Vv(p_j), vi);
        V3 kd = s->gradW(s->rel(xi, s->X(p_j)));
        V3 a = smul(s->m_V[p_j], vji);
        float av[3] = {a.x, a.y, a.z}, kv[3] = {kd.x, kd.y, kd.z};
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) G[3 * r + c] = G[3 * r + c] + av[r] * kv[c];
      }
    });
    for (int k = 0; k < 9; ++k) s->grad_v[9 * p_i + k] = G[k];
  }
}

// dimcv_sph.py:551-583
void or_compute_d_vorticity(Sim* s) {
  std::fill(s->d_vort.begin(), s->d_vort.end(), 0.0f);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i), vsi = Sim::get(s->v_star, p_i);
    V3 curl{0.0f, 0.0f, 0.0f};
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j))
        curl = vadd(curl, smul(s->m_V[p_j], vcross(vsub(Sim::get(s->v_star, p_j), vsi),
                                                    s->gradW(s->rel(xi, s->X(p_j))))));
    });
    Sim::put(s->d_vort, p_i, vsub(Sim::get(s->vorticity, p_i), curl));
  }
}

// dimcv_sph.py:660-677
void or_compute_all_d_vort_smoothed(Sim* s) {
  float w0 = s->W(0.0f);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i) || !s->in_vortex_domain(s->X(p_i))) continue;
    V3 xi = s->X(p_i);
    V3 acc = vmuls(smul(s->m_V[p_i], Sim::get(s->d_vort, p_i)), w0);
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j))
        acc = vadd(acc, vmuls(smul(s->m_V[p_j], Sim::get(s->d_vort, p_j)), s->W(vnorm(s->rel(xi, s->X(p_j))))));
    });
    Sim::put(s->d_vort_smoothed, p_i, acc);
  }
}

// dimcv_sph.py:295-324.  u1/u2: host-fed uniforms per slot (replace
// ti.random); with do_bernoulli = 0 only the KVN is written.
void or_compute_kinematic_vorticity_num_sample(Sim* s, const float* u1, const float* u2,
                                               int32_t do_bernoulli) {
#pragma omp parallel for schedule(static)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i) || !s->in_vortex_domain(s->X(p_i))) continue;
    const float* G = &s->grad_v[9 * p_i];
    float Wn = 0.0f, Sn = 0.0f;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        float w = G[3 * r + c] - G[3 * c + r];
        float st = G[3 * r + c] + G[3 * c + r];
        if (r == 0 && c == 0) { Wn = w * w; Sn = st * st; }
        else { Wn = Wn + w * w; Sn = Sn + st * st; }
      }
    float kvn = std::sqrt(Wn) / (std::sqrt(Sn) + 1e-8f);
    float t0 = std::fmin(std::fmax(s->life_time[p_i], 0.0f), 1.0f);
    float rr = 0.018f * expf_c(s->f_neg_a * t0) + 0.004f;
    float produce = 1.0f - 1.0f / powf_c(kvn + 1.0f, rr);
    float del = expf_c(s->f_neg_b * kvn);
    s->kvn[p_i] = kvn;
    if (!do_bernoulli) continue;
    if (s->is_sample[p_i] == 0 && u1[p_i] < produce) s->is_sample[p_i] = 1;
    else if (s->is_sample[p_i] == 1 && kvn < 2.0f && u2[p_i] < del) s->is_sample[p_i] = 0;
  }
}

// dimcv_sph.py:326-335 (ascending compaction)
void or_record_vort_part_sample_idx(Sim* s) {
  std::fill(s->sample_idx.begin(), s->sample_idx.end(), 0);
  int32_t idx = 0;
  for (int64_t p = 0; p < s->particle_num; ++p)
    if (s->is_sample[p] == 1) s->sample_idx[idx++] = (int32_t)p;
  s->num_samples = idx;
}

static float epan(const Sim* s, V3 r) {
  float ret = 1.0f;
  float q[3] = {std::fabs(r.x) * s->f_inv_support, std::fabs(r.y) * s->f_inv_support,
                std::fabs(r.z) * s->f_inv_support};
  for (int d = 0; d < 3; ++d) {
    if (q[d] < 1.0f) ret *= 0.75f * (1.0f - q[d] * q[d]);
    else ret *= 0.0f;
  }
  return ret;
}

// dimcv_sph.py:721-751; kden_sum is the canonical FP64 sum over fluid slots
// (out-of-domain fluid contributes 0).
void or_compute_kernel_density(Sim* s) {
  V3 zero{0.0f, 0.0f, 0.0f};
  float e0 = epan(s, zero);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i) || !s->in_vortex_domain(s->X(p_i))) continue;
    V3 xi = s->X(p_i);
    float kden = s->kvn[p_i] * e0;
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j)) kden += s->kvn[p_j] * epan(s, s->rel(xi, s->X(p_j)));
    });
    s->kernel_density[p_i] = kden;
  }
  std::vector<double> terms;
  for (int64_t p = 0; p < s->particle_num; ++p) {
    if (!s->is_fluid(p)) continue;
    terms.push_back(s->in_vortex_domain(s->X(p)) ? (double)s->kernel_density[p] : 0.0);
  }
  s->kden_sum = canon_reduce_vec(std::move(terms));
}

// Build-defined inverse-CDF importance sampler.  Weights over fluid slots in
// ascending order: kden (in-domain) or 0; canonical FP64 scan; for each u_k,
// t = u_k * total, index = first cdf > t (clamped to the first cdf == total).
// Writes sample_idx[0..M) (rest zero) and num_samples = M (0 if total <= 0).
int32_t or_cdf_sample(Sim* s, const double* u, int32_t M) {
  if (M < 0 || M > s->N) return -1;  // sample_idx holds particle_max_num entries
  std::vector<int32_t> slots;
  std::vector<double> w;
  for (int64_t p = 0; p < s->particle_num; ++p) {
    if (!s->is_fluid(p)) continue;
    slots.push_back((int32_t)p);
    w.push_back(s->in_vortex_domain(s->X(p)) ? (double)s->kernel_density[p] : 0.0);
  }
  std::fill(s->sample_idx.begin(), s->sample_idx.end(), 0);
  int64_t F = (int64_t)w.size();
  if (F == 0) { s->num_samples = 0; return 0; }
  std::vector<double> cdf(F);
  canon_scan_rec(w.data(), cdf.data(), F);
  double total = cdf[F - 1];
  if (!(total > 0.0)) { s->num_samples = 0; return 0; }
  for (int32_t k = 0; k < M; ++k) {
    double t = u[k] * total;
    int64_t idx = std::upper_bound(cdf.begin(), cdf.end(), t) - cdf.begin();
    if (idx >= F) idx = std::lower_bound(cdf.begin(), cdf.end(), total) - cdf.begin();
    s->sample_idx[k] = slots[idx];
  }
  s->num_samples = M;
  return 0;
}

// One target of vort2vel (dimcv_sph.py:649-658,679-711), in reference order:
// the near part (scaled), the far part (scaled by fac/added, zero if no
// sample was accepted) and the accepted-sample count.
static void vort2vel_one(const Sim* s, int64_t p_i, V3& near, V3& far, int& added) {
  const float fac = s->f_vp_factor;
  const float eps = s->f_eps, inv_eps = s->f_inv_eps;
  V3 xi = s->X(p_i);
  V3 vv{0.0f, 0.0f, 0.0f};
  float kns = s->kernel_density[p_i];
  s->for_all_neighbors(p_i, [&](int64_t p_j) {
    if (s->is_fluid(p_j)) {
      V3 r2 = s->rel(xi, s->X(p_j));
      V3 r{r2.x, r2.y, 0.0f};
      float rn = vnorm(r);
      float A = (1.0f - expf_c(-rn * inv_eps)) / (rn * (rn + eps * expf_c(-rn * inv_eps)));
      vv = vadd(vv, smul(A, vcross(r, Sim::get(s->d_vort_smoothed, p_j))));
      kns += s->kernel_density[p_j];
    }
  });
  near = vmuls(vv, fac);
  V3 vs{0.0f, 0.0f, 0.0f};
  added = 0;
  float S = (float)(s->kden_sum - (double)kns);
  for (int32_t j = 0; j < s->num_samples; ++j) {
    int32_t p_j = s->sample_idx[j];
    V3 r2 = s->rel(xi, s->X(p_j));
    float rn = vnorm(r2);
    if (rn > s->f_support) {
      V3 r{r2.x, r2.y, 0.0f};
      float A = (1.0f - expf_c(-rn * inv_eps)) / (rn * (rn + eps * expf_c(-rn * inv_eps)));
      V3 t = smul(A, vcross(r, Sim::get(s->d_vort_smoothed, p_j)));
      t = vmuls(t, S);
      t = vdivs(t, s->kernel_density[p_j] + 1e-8f);
      vs = vadd(vs, t);
      added += 1;
    }
  }
  if (added == 0) vs = {0.0f, 0.0f, 0.0f};
  else vs = vmuls(vs, fac / (float)added);
  far = vs;
}

// dimcv_sph.py:679-719.  v_vort_out (optional, 3N) receives near+far for
// every processed particle (analysis only).
void or_vort2vel(Sim* s, float* v_vort_out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i) || !s->in_vortex_domain(s->X(p_i))) continue;
    V3 near, far;
    int added;
    vort2vel_one(s, p_i, near, far, added);
    V3 vvort = vadd(far, near);
    if (v_vort_out) {
      v_vort_out[3 * p_i] = vvort.x;
      v_vort_out[3 * p_i + 1] = vvort.y;
      v_vort_out[3 * p_i + 2] = vvort.z;
    }
    V3 vi = s->Vv(p_i);
    if (vnorm(vvort) < vnorm(vi)) Sim::put(s->v, p_i, vadd(vi, vvort));
  }
}

// vort2vel's recovered velocity at selected slots without updating v
// (analysis): v_vort (near + far), far part and accepted count per target;
// slots that vort2vel skips (solid, outside the vortex domain) get zeros.
void or_vort2vel_targets(Sim* s, const int64_t* slots, int64_t n, float* v_vort_out, float* far_out,
                         int32_t* added_out) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < n; ++k) {
    const int64_t p_i = slots[k];
    V3 near{0.0f, 0.0f, 0.0f}, far{0.0f, 0.0f, 0.0f};
    int added = 0;
    if (p_i >= 0 && p_i < s->particle_num && s->is_fluid(p_i) && s->in_vortex_domain(s->X(p_i)))
      vort2vel_one(s, p_i, near, far, added);
    V3 vv = vadd(far, near);
    v_vort_out[3 * k] = vv.x; v_vort_out[3 * k + 1] = vv.y; v_vort_out[3 * k + 2] = vv.z;
    far_out[3 * k] = far.x; far_out[3 * k + 1] = far.y; far_out[3 * k + 2] = far.z;
    added_out[k] = added;
  }
}

// dimcv_sph.py:753-760
void or_update_particle_life_time(Sim* s) {
  for (int64_t p = 0; p < s->particle_num; ++p) {
    if (!s->is_fluid(p) || !s->in_vortex_domain(s->X(p))) continue;
    s->life_time[p] += s->dt;
  }
}

// dimcv_sph.py:470-493
void or_compute_density_adv(Sim* s) {
  const float dt = s->dt;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i), vi = s->Vv(p_i);
    float delta = 0.0f;
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      V3 vj = s->Vv(p_j);
      if (s->is_fluid(p_j))
        delta += s->m[p_j] * vdot(vsub(vi, vj), s->gradW(s->rel(xi, s->X(p_j))));
      else if (s->is_solid(p_j))
        delta += s->m_V[p_j] * s->f_density0 * vdot(vsub(vi, vj), s->gradW(s->rel(xi, s->X(p_j))));
    });
    float da = s->density[p_i] + dt * delta;
    s->density_adv[p_i] = std::fmax(da, s->f_density0);
  }
}

// dimcv_sph.py:495-549 (v_i updated in place per neighbour)
void or_pressure_solve_iteration_kernel(Sim* s) {
  const float dt = s->dt;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i);
    float b_i = s->density_adv[p_i] - s->f_density0;
    float k_i = b_i * s->dfsph_factor[p_i];
    V3 vi = s->Vv(p_i);
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (s->is_fluid(p_j)) {
        float b_j = s->density_adv[p_j] - s->f_density0;
        float k_j = b_j * s->dfsph_factor[p_j];
        float k_sum = k_i + k_j;
        if (std::fabs(k_sum) > s->f_m_eps) {
          V3 g = smul(s->m[p_j], s->gradW(s->rel(xi, s->X(p_j))));
          if (!s->cfg.has_transit) vi = vsub(vi, smul(dt * k_sum, g));
          else vi = vsub(vi, smul(s->transit_weight(xi.x) * dt * k_sum, g));
        }
      } else if (s->is_solid(p_j)) {
        if (std::fabs(k_i) > s->f_m_eps) {
          V3 g = smul(s->m_V[p_j] * s->f_density0, s->gradW(s->rel(xi, s->X(p_j))));
          V3 vc = smul(-dt * 1.0f * k_i, g);
          if (!s->cfg.has_transit) vi = vadd(vi, vc);
          else vi = vadd(vi, smul(s->transit_weight(xi.x), vc));
        }
      }
    });
    Sim::put(s->v, p_i, vi);
  }
  rigid_reactions(s, [&](int64_t i) { return (s->density_adv[i] - s->f_density0) * s->dfsph_factor[i]; });
}

// dimcv_sph.py:797-801
void or_copy_x_temp(Sim* s) {
  std::fill(s->x_temp.begin(), s->x_temp.end(), 0.0f);
  for (int64_t p = 0; p < s->particle_num; ++p) Sim::put(s->x_temp, p, s->X(p));
}

// dimcv_sph.py:778-795
void or_compute_vorticity_vis(Sim* s) {
  std::fill(s->vorticity_vis.begin(), s->vorticity_vis.end(), 0.0f);
  float w0 = s->W(0.0f);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i)) continue;
    V3 xi = s->X(p_i);
    float w = s->m_V[p_i] * w0;
    V3 vort = smul(w, Sim::get(s->vorticity, p_i));
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      float wj = s->m_V[p_j] * s->W(vnorm(s->rel(xi, s->X(p_j))));
      vort = vadd(vort, smul(wj, Sim::get(s->vorticity, p_j)));
      w += wj;
    });
    Sim::put(s->vorticity_vis, p_i, vdivs(vort, w + 1e-6f));
  }
}

// sph_base.py:118-123,149-179 (enforce_boundary_3D with particle_type =
// fluid, called by karman_vortex.py:226-230 when bounceBackBoundary is set).
// The bounds are f32(domain_size - padding) and f32(padding), padding =
// grid_size (particle_system.py:58); every test reads the position captured
// before the clamps.  Slots below particle_num (the rest hold nothing the
// step reads).
void or_enforce_boundary_3d(Sim* s) {
  const float pad = (float)s->support;
  float hi[3];
  for (int d = 0; d < 3; ++d)
    hi[d] = (float)((s->cfg.domain_end[d] - s->cfg.domain_start[d]) - s->support);
  for (int64_t p_i = 0; p_i < s->particle_num; ++p_i) {
    if (!s->is_fluid(p_i) || !s->is_dynamic[p_i]) continue;
    V3 pos = s->X(p_i);
    float q[3] = {pos.x, pos.y, pos.z};
    float n[3] = {0.0f, 0.0f, 0.0f};
    for (int d = 0; d < 3; ++d) {
      if (q[d] > hi[d]) {
        n[d] += 1.0f;
        s->x[3 * p_i + d] = hi[d];
      }
      if (q[d] <= pad) {
        n[d] += -1.0f;
        s->x[3 * p_i + d] = pad;
      }
    }
    V3 cn{n[0], n[1], n[2]};
    float len = vnorm(cn);
    if (len > 1e-6f) {
      V3 vec = vdivs(cn, len);
      V3 v = s->Vv(p_i);
      const float c_f = 0.5f;
      Sim::put(s->v, p_i, vsub(v, smul((1.0f + c_f) * vdot(v, vec), vec)));
    }
  }
}

// ---------------------------------------------------------------------------
// Segment solver (segment_solver.py): the O(N^2) regularised Biot-Savart sums
// and the RK4 endpoint update, serial j order, f32.
struct OrSegConst {
  float reg_radius, dt, dt6, inv4pi, gamma_decay;
  float u_inf[3];
};

// form 0: _compute_endpoint_velocity_bs (:378-418); form 1:
// _compute_velocity_at_factor (:468-521) with query offset (factor dt) k.
void or_segment_velocity(int32_t n, const float* xm, const float* xp, const float* gamma, const int32_t* active,
                         const OrSegConst* c, int32_t form, float factor, const float* km, const float* kp,
                         float* out_m, float* out_p) {
  const float R2 = c->reg_radius * c->reg_radius;
  const float qs = factor * c->dt;
  for (int32_t i = 0; i < n; ++i) {
    if (active[i] != 1) continue;
    V3 xmi = Sim::get_ptr(xm, i), xpi = Sim::get_ptr(xp, i);
    if (form == 1) {
      xmi = vadd(xmi, smul(qs, Sim::get_ptr(km, i)));
      xpi = vadd(xpi, smul(qs, Sim::get_ptr(kp, i)));
    }
    V3 um{c->u_inf[0], c->u_inf[1], c->u_inf[2]}, up = um;
    for (int32_t j = 0; j < n; ++j) {
      if (j == i || active[j] != 1) continue;
      V3 xmj = Sim::get_ptr(xm, j), xpj = Sim::get_ptr(xp, j);
      V3 cj = smul(0.5f, vadd(xmj, xpj));
      V3 dj = vsub(xpj, xmj);
      V3 w;
      if (form == 0) {
        float Lj = vnorm(dj) + 1e-8f;
        V3 tj = vdivs(dj, Lj);
        w = smul(gamma[j] * Lj, tj);
      } else {
        w = smul(gamma[j], dj);
      }
      V3 r = vsub(xmi, cj);
      float r2 = vdot(r, r) + R2;
      float den = r2 * std::sqrt(r2) + 1e-12f;
      um = vadd(um, vdivs(smul(c->inv4pi, vcross(w, r)), den));
      V3 rp = vsub(xpi, cj);
      float s2 = vdot(rp, rp) + R2;
      float denp = s2 * std::sqrt(s2) + 1e-12f;
      up = vadd(up, vdivs(smul(c->inv4pi, vcross(w, rp)), denp));
    }
    Sim::put_ptr(out_m, i, um);
    Sim::put_ptr(out_p, i, up);
  }
}

// advect_segments_rk4 (:420-447) with _update_endpoints_rk4 (:523-542);
// k: 24 n floats of stage scratch (k1..k4 minus, then plus).
void or_segment_rk4(int32_t n, float* xm, float* xp, float* gamma, const int32_t* active, float* age,
                    const OrSegConst* c, const float* vm, const float* vp, float* k) {
  const int64_t n3 = 3 * (int64_t)n;
  float* km[4] = {k, k + n3, k + 2 * n3, k + 3 * n3};
  float* kp[4] = {k + 4 * n3, k + 5 * n3, k + 6 * n3, k + 7 * n3};
  for (int32_t i = 0; i < n; ++i)
    if (active[i] == 1)
      for (int d = 0; d < 3; ++d) {
        km[0][3 * i + d] = vm[3 * i + d];
        kp[0][3 * i + d] = vp[3 * i + d];
      }
  const float fac[3] = {0.5f, 0.5f, 1.0f};
  for (int st = 0; st < 3; ++st)
    or_segment_velocity(n, xm, xp, gamma, active, c, 1, fac[st], km[st], kp[st], km[st + 1], kp[st + 1]);
  for (int32_t i = 0; i < n; ++i) {
    if (active[i] != 1) continue;
    V3 sm = vadd(vadd(vadd(Sim::get_ptr(km[0], i), smul(2.0f, Sim::get_ptr(km[1], i))),
                      smul(2.0f, Sim::get_ptr(km[2], i))), Sim::get_ptr(km[3], i));
    V3 sp = vadd(vadd(vadd(Sim::get_ptr(kp[0], i), smul(2.0f, Sim::get_ptr(kp[1], i))),
                      smul(2.0f, Sim::get_ptr(kp[2], i))), Sim::get_ptr(kp[3], i));
    Sim::put_ptr(xm, i, vadd(Sim::get_ptr(xm, i), smul(c->dt6, sm)));
    Sim::put_ptr(xp, i, vadd(Sim::get_ptr(xp, i), smul(c->dt6, sp)));
    gamma[i] *= c->gamma_decay;
    age[i] += c->dt;
  }
}

// ---------------------------------------------------------------------------
// Build-defined comparators (no reference function; see the module header).

// minimum image of an FP64 separation (periodic variant)
static void wrap_d(const Sim* s, double& rx, double& ry, double& rz) {
  double* q[3] = {&rx, &ry, &rz};
  for (int d = 0; d < 3; ++d) {
    double L = (double)s->f_box[d];
    if (*q[d] > 0.5 * L) *q[d] -= L;
    else if (*q[d] < -0.5 * L) *q[d] += L;
  }
}

// FP64 direct Biot-Savart: exactly the quantity vort2vel estimates, with the
// sample set = every in-domain fluid particle and weight 1.  Near part over
// the reference neighbour set with the 2-D norm; far part over |r| > h
// (f32 test, 3-D norm) with K = (1 - e^{-r/eps}) / (r (r + eps e^{-r/eps})).
// out: 3N doubles (v_vort per processed particle, zero elsewhere).
void or_direct_bs_f64(Sim* s, double* out) {
  const double fac = 0.1592 * s->vp_volume;
  const double eps = 1e-6;
  std::vector<int32_t> src;
  for (int64_t p = 0; p < s->particle_num; ++p)
    if (s->is_fluid(p) && s->in_vortex_domain(s->X(p))) src.push_back((int32_t)p);
  std::fill(out, out + 3 * s->N, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < (int64_t)src.size(); ++k) {
    int64_t p_i = src[k];
    V3 xi = s->X(p_i);
    double nx = 0, ny = 0, nz = 0;
    s->for_all_neighbors(p_i, [&](int64_t p_j) {
      if (!s->is_fluid(p_j)) return;
      double rx = (double)xi.x - (double)s->x[3 * p_j], ry = (double)xi.y - (double)s->x[3 * p_j + 1];
      if (s->cfg.periodic) { double rz = 0; wrap_d(s, rx, ry, rz); }
      double rn = std::sqrt(rx * rx + ry * ry);
      double e = std::exp(-rn / eps);
      double A = (1.0 - e) / (rn * (rn + eps * e));
      const float* w = &s->d_vort_smoothed[3 * p_j];
      nx += A * (ry * w[2]);
      ny += A * (-rx * w[2]);
      nz += A * (rx * w[1] - ry * w[0]);
    });
    double fx = 0, fy = 0, fz = 0;
    for (int32_t p_j : src) {
      V3 r2 = s->rel(xi, s->X(p_j));
      if (!(vnorm(r2) > s->f_support)) continue;
      double rx = (double)xi.x - (double)s->x[3 * p_j], ry = (double)xi.y - (double)s->x[3 * p_j + 1];
      double rz = (double)xi.z - (double)s->x[3 * p_j + 2];
      if (s->cfg.periodic) wrap_d(s, rx, ry, rz);
      double r2d = rx * rx + ry * ry + rz * rz;
      double A = 1.0 / r2d;
      const float* w = &s->d_vort_smoothed[3 * p_j];
      fx += A * (ry * w[2]);
      fy += A * (-rx * w[2]);
      fz += A * (rx * w[1] - ry * w[0]);
    }
    out[3 * p_i] = fac * (nx + fx);
    out[3 * p_i + 1] = fac * (ny + fy);
    out[3 * p_i + 2] = fac * (nz + fz);
  }
}

// Vortex-ring initial velocity: each ring (centre c, unit axis n, radius R,
// circulation G) is discretised into K points with line elements
// dl = (2 pi R / K) t_k; u(x) = sum G/(4 pi) dl x (x - p) / |x - p|^3 *
// (1 - exp(-|x - p|^2 / sigma^2)).  FP64, written as f32.
void or_ring_velocity(int32_t nrings, const double* centre, const double* axis, const double* R,
                      const double* gamma, double sigma, int32_t K, int64_t n, const float* x,
                      float* vel) {
  std::vector<double> px, py, pz, lx, ly, lz, gg;
  for (int r = 0; r < nrings; ++r) {
    double ax = axis[3 * r], ay = axis[3 * r + 1], az = axis[3 * r + 2];
    double an = std::sqrt(ax * ax + ay * ay + az * az);
    ax /= an; ay /= an; az /= an;
    // e1: unit vector orthogonal to the axis
    double tx, ty, tz;
    if (std::fabs(ax) < 0.9) { tx = 1; ty = 0; tz = 0; } else { tx = 0; ty = 1; tz = 0; }
    double d = tx * ax + ty * ay + tz * az;
    double e1x = tx - d * ax, e1y = ty - d * ay, e1z = tz - d * az;
    double e1n = std::sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
    e1x /= e1n; e1y /= e1n; e1z /= e1n;
    double e2x = ay * e1z - az * e1y, e2y = az * e1x - ax * e1z, e2z = ax * e1y - ay * e1x;
    double dl = 2.0 * M_PI * R[r] / K;
    for (int k = 0; k < K; ++k) {
      double th = 2.0 * M_PI * (k + 0.5) / K;
      double c = std::cos(th), sn = std::sin(th);
      px.push_back(centre[3 * r] + R[r] * (c * e1x + sn * e2x));
      py.push_back(centre[3 * r + 1] + R[r] * (c * e1y + sn * e2y));
      pz.push_back(centre[3 * r + 2] + R[r] * (c * e1z + sn * e2z));
      lx.push_back(dl * (-sn * e1x + c * e2x));
      ly.push_back(dl * (-sn * e1y + c * e2y));
      lz.push_back(dl * (-sn * e1z + c * e2z));
      gg.push_back(gamma[r] / (4.0 * M_PI));
    }
  }
  const double is2 = 1.0 / (sigma * sigma);
  int64_t P = (int64_t)px.size();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double X0 = x[3 * i], X1 = x[3 * i + 1], X2 = x[3 * i + 2];
    double u0 = 0, u1 = 0, u2 = 0;
    for (int64_t k = 0; k < P; ++k) {
      double rx = X0 - px[k], ry = X1 - py[k], rz = X2 - pz[k];
      double r2 = rx * rx + ry * ry + rz * rz;
      if (r2 <= 0.0) continue;
      double core = 1.0 - std::exp(-r2 * is2);
      double f = gg[k] * core / (r2 * std::sqrt(r2));
      u0 += f * (ly[k] * rz - lz[k] * ry);
      u1 += f * (lz[k] * rx - lx[k] * rz);
      u2 += f * (lx[k] * ry - ly[k] * rx);
    }
    vel[3 * i] = (float)u0; vel[3 * i + 1] = (float)u1; vel[3 * i + 2] = (float)u2;
  }
}

// Neighbour set of slot p_i in for_all_neighbors order; returns the count and
// writes up to cap slots.
int32_t or_neighbors(Sim* s, int64_t p_i, int32_t* out, int32_t cap) {
  int32_t k = 0;
  s->for_all_neighbors(p_i, [&](int64_t p_j) {
    if (k < cap) out[k] = (int32_t)p_j;
    ++k;
  });
  return k;
}

double or_canon_reduce(const double* a, int64_t n) {
  return canon_reduce_vec(std::vector<double>(a, a + n));
}
void or_canon_scan(const double* a, double* out, int64_t n) { canon_scan_rec(a, out, n); }

float or_expf(float x) { return expf_c(x); }
float or_logf(float x) { return refmath::logf_c(x); }
float or_powf(float x, float y) { return powf_c(x, y); }
float or_W(Sim* s, float r) { return s->W(r); }
void or_gradW(Sim* s, const float* r, float* out) {
  V3 g = s->gradW({r[0], r[1], r[2]});
  out[0] = g.x; out[1] = g.y; out[2] = g.z;
}

}  // extern "C"
