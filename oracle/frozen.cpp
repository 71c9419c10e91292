// Oracle (test infrastructure only): FP64 restatement of the DIMCV-SPH step
// with its discrete structure frozen, for central finite differences of the
// rollout adjoint.
//
// Each step takes, from a recorded FP32 forward: the particle order after
// the grid build (ids), the fluid neighbour lists in visiting order, the
// importance samples, the Jacobi iteration counts, and the FP32 values that
// decide every gate (b, rho*, rho, alpha, the update gate, the positions for
// the far-field cut-off).  Values are then evaluated in FP64 from the FP64
// state, gates from the FP32 record:
//   rho    = rho0 (mV W(0) + sum mV W)                       dimcv_sph.py:126-148
//   alpha0 = 1 / (sum |m gW|^2 + |sum m gW|^2)  (0 where the FP32 alpha is 0)  :86-124
//   b      = [b32 > 0] sum mV (v_i - v_j).gW                 :337-374
//   v     -= sum_{|k32_i + k32_j| > eps} dt (k_i + k_j) m gW, k = b alpha rho   :421-468
//   a      = sum c mV ((v_i - v_j).r) / (|r|^2 + 0.01 h^2) gW   :150-208
//   d_vort = curl(v) - curl(v + dt a); w = mV d W(0) + sum mV d W   :228-293, 551-583, 660-677
//   kvn    = |W| / (|S| + 1e-8) of grad v; kden = kvn E(0) + sum kvn E   :295-312, 721-751
//   v     += gate (near + far)                               :649-719
//   v     += dt a; da = [da32 > rho0] (rho + dt sum m (v_i - v_j).gW) else rho0;
//   v     -= pressure Jacobi with k = (da - rho0) alpha_p    :470-549
//   x     += dt v                                            :210-218
// This is the function the adjoint differentiates; at the recorded point it
// reproduces the FP32 forward to rounding.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct FrozenParams {
  double dt, rho0, mV, m, h, kernel_k, kernel_dk, visc_coef, visc_h2, eps, vp_factor;
  double inv_dt, inv_dt2;      // the FP32 multipliers of alpha, as doubles
  float far_t, m_eps, rho0_f;  // FP32 gate constants
};

struct FrozenStep {
  int32_t n, kv, kp, M;
  const int32_t* ids;        // [n] slot -> particle id
  const int32_t* nbr_off;    // [n + 1]
  const int32_t* nbr;        // fluid neighbours (slots), visiting order
  const float* x32;          // [n][3]
  const float* rho32;
  const float* alpha_div32;  // alpha0 * inv_dt (FP32)
  const float* alpha_p32;    // the pressure solve's alpha (FP32)
  const float* b32;          // [kv + 1][n]
  const float* da32;         // [kp + 1][n]
  const uint8_t* gate;       // [n]
  const int32_t* samples;    // [M] slots
};

struct D3 {
  double x, y, z;
};
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator*(double s, D3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(D3 a) { return std::sqrt(dot(a, a)); }

inline double W(const FrozenParams& P, double s) {
  const double q = s / P.h;
  if (q > 1.0) return 0.0;
  if (q <= 0.5) return P.kernel_k * (6.0 * q * q * q - 6.0 * q * q + 1.0);
  const double t = 1.0 - q;
  return P.kernel_k * 2.0 * t * t * t;
}

inline D3 gW(const FrozenParams& P, D3 r) {
  const double s = norm(r), q = s / P.h;
  if (!(s > 1e-5 && q <= 1.0)) return {0, 0, 0};
  const double g = q <= 0.5 ? P.kernel_dk * q * (3.0 * q - 2.0) : -P.kernel_dk * (1.0 - q) * (1.0 - q);
  return (g / (s * P.h)) * r;
}

inline double epan(const FrozenParams& P, D3 r) {
  const double q[3] = {std::fabs(r.x) / P.h, std::fabs(r.y) / P.h, std::fabs(r.z) / P.h};
  double e = 1.0;
  for (int d = 0; d < 3; ++d) e *= q[d] < 1.0 ? 0.75 * (1.0 - q[d] * q[d]) : 0.0;
  return e;
}

inline double A(const FrozenParams& P, double s) {
  const double E = std::exp(-s / P.eps);
  return (1.0 - E) / (s * (s + P.eps * E));
}

struct Ctx {
  const FrozenParams& P;
  const FrozenStep& S;
  std::vector<D3> X;
  Ctx(const FrozenParams& p, const FrozenStep& s) : P(p), S(s) {}
  template <class F>
  void nbrs(int i, F&& f) const {
    for (int k = S.nbr_off[i]; k < S.nbr_off[i + 1]; ++k) f(S.nbr[k]);
  }
  D3 r(int i, int j) const { return X[i] - X[j]; }
};

// sum_j w_j (v_i - v_j).gW_ij with the gate
void divergence(const Ctx& c, const std::vector<D3>& V, double w, const float* gate32, float thr, double base_add,
                const std::vector<double>* rho, double scale, std::vector<double>& out) {
  const int n = c.S.n;
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    double acc = 0.0;
    c.nbrs(i, [&](int j) { acc += w * dot(V[i] - V[j], gW(c.P, c.r(i, j))); });
    if (rho) out[i] = gate32[i] > thr ? (*rho)[i] + scale * acc : base_add;
    else out[i] = gate32[i] > thr ? acc : 0.0;
  }
}

void jacobi(const Ctx& c, std::vector<D3>& V, const std::vector<double>& kappa, const std::vector<float>& k32) {
  const int n = c.S.n;
  std::vector<D3> dv(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    D3 acc{0, 0, 0};
    c.nbrs(i, [&](int j) {
      const float ks = k32[i] + k32[j];
      if (!(std::fabs(ks) > c.P.m_eps)) return;
      acc = acc + (c.P.dt * (kappa[i] + kappa[j]) * c.P.m) * gW(c.P, c.r(i, j));
    });
    dv[i] = acc;
  }
  for (int i = 0; i < n; ++i) V[i] = V[i] - dv[i];
}

std::vector<D3> curl(const Ctx& c, const std::vector<D3>& U) {
  const int n = c.S.n;
  std::vector<D3> out(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    D3 acc{0, 0, 0};
    c.nbrs(i, [&](int j) { acc = acc + c.P.mV * cross(U[j] - U[i], gW(c.P, c.r(i, j))); });
    out[i] = acc;
  }
  return out;
}

// Optional per-pass values of the last step, in its slot order ([n] or [n][3]).
struct FrozenDebug {
  double *rho, *alpha0, *b0, *acc, *w, *kden, *vvort, *da0;
};

void put3(double* out, const std::vector<D3>& a) {
  if (!out) return;
  for (size_t i = 0; i < a.size(); ++i) {
    out[3 * i] = a[i].x;
    out[3 * i + 1] = a[i].y;
    out[3 * i + 2] = a[i].z;
  }
}
void put1(double* out, const std::vector<double>& a) {
  if (!out) return;
  for (size_t i = 0; i < a.size(); ++i) out[i] = a[i];
}

void step(const FrozenParams& P, const FrozenStep& S, std::vector<D3>& Xid, std::vector<D3>& Vid,
          const FrozenDebug* dbg) {
  const int n = S.n;
  Ctx c(P, S);
  c.X.resize(n);
  std::vector<D3> V(n);
  for (int s = 0; s < n; ++s) {
    c.X[s] = Xid[S.ids[s]];
    V[s] = Vid[S.ids[s]];
  }
  // density and DFSPH factor
  std::vector<double> rho(n), alpha0(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    double den = P.mV * W(P, 0.0);
    D3 G{0, 0, 0};
    double r3 = 0.0;
    c.nbrs(i, [&](int j) {
      const D3 r = c.r(i, j);
      den += P.mV * W(P, norm(r));
      const D3 g = P.m * gW(P, r);
      G = G + g;
      r3 += dot(g, g);
    });
    rho[i] = P.rho0 * den;
    alpha0[i] = S.alpha_div32[i] != 0.0f ? 1.0 / (r3 + dot(G, G)) : 0.0;
  }
  std::vector<double> a_div(n), a_p(n);
  for (int i = 0; i < n; ++i) {
    a_div[i] = alpha0[i] * P.inv_dt;
    a_p[i] = alpha0[i] * P.inv_dt * P.dt * P.inv_dt2;
  }
  // divergence solve
  std::vector<double> b(n), kap(n);
  std::vector<float> k32(n);
  divergence(c, V, P.mV, S.b32, 0.0f, 0.0, nullptr, 0.0, b);
  if (dbg) {
    put1(dbg->rho, rho);
    put1(dbg->alpha0, alpha0);
    put1(dbg->b0, b);
  }
  for (int t = 0; t < S.kv; ++t) {
    const float* bt = S.b32 + (int64_t)t * n;
    for (int i = 0; i < n; ++i) {
      kap[i] = b[i] * a_div[i] * rho[i];
      k32[i] = bt[i] * S.alpha_div32[i] * S.rho32[i];
    }
    jacobi(c, V, kap, k32);
    if (t + 1 < S.kv) divergence(c, V, P.mV, S.b32 + (int64_t)(t + 1) * n, 0.0f, 0.0, nullptr, 0.0, b);
  }
  // non-pressure acceleration
  std::vector<D3> acc(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    D3 s{0, 0, 0};
    c.nbrs(i, [&](int j) {
      const D3 r = c.r(i, j);
      const double beta = P.visc_coef * P.mV * dot(V[i] - V[j], r) / (dot(r, r) + P.visc_h2);
      s = s + beta * gW(P, r);
    });
    acc[i] = s;
  }
  // vorticity passes
  std::vector<D3> vstar(n);
  for (int i = 0; i < n; ++i) vstar[i] = V[i] + P.dt * acc[i];
  std::vector<D3> om = curl(c, V), cs = curl(c, vstar), dv(n), w(n);
  for (int i = 0; i < n; ++i) dv[i] = om[i] - cs[i];
  std::vector<double> kvn(n), kden(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    D3 s = (P.mV * W(P, 0.0)) * dv[i];
    double G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    c.nbrs(i, [&](int j) {
      const D3 r = c.r(i, j);
      s = s + (P.mV * W(P, norm(r))) * dv[j];
      const D3 d = P.mV * (V[j] - V[i]);
      const D3 g = gW(P, r);
      const double av[3] = {d.x, d.y, d.z}, gv[3] = {g.x, g.y, g.z};
      for (int rr = 0; rr < 3; ++rr)
        for (int cc = 0; cc < 3; ++cc) G[3 * rr + cc] += av[rr] * gv[cc];
    });
    w[i] = s;
    double Wn = 0.0, Sn = 0.0;
    for (int rr = 0; rr < 3; ++rr)
      for (int cc = 0; cc < 3; ++cc) {
        const double a1 = G[3 * rr + cc] - G[3 * cc + rr], s1 = G[3 * rr + cc] + G[3 * cc + rr];
        Wn += a1 * a1;
        Sn += s1 * s1;
      }
    kvn[i] = std::sqrt(Wn) / (std::sqrt(Sn) + 1e-8);
  }
  const double e0 = epan(P, {0, 0, 0});
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    double k = kvn[i] * e0;
    c.nbrs(i, [&](int j) { k += kvn[j] * epan(P, c.r(i, j)); });
    kden[i] = k;
  }
  double K = 0.0;
  for (int i = 0; i < n; ++i) K += kden[i];
  // velocity recovery
  std::vector<D3> vv(n);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    D3 near{0, 0, 0};
    double kns = kden[i];
    c.nbrs(i, [&](int j) {
      const D3 r3 = c.r(i, j), r{r3.x, r3.y, 0.0};
      near = near + A(P, norm(r)) * cross(r, w[j]);
      kns += kden[j];
    });
    const double Si = K - kns;
    D3 far{0, 0, 0};
    int added = 0;
    const float* xi32 = S.x32 + 3 * (int64_t)i;
    for (int q = 0; q < S.M; ++q) {
      const int sk = S.samples[q];
      const float* xs32 = S.x32 + 3 * (int64_t)sk;
      const float dx = xi32[0] - xs32[0], dy = xi32[1] - xs32[1], dz = xi32[2] - xs32[2];
      const float r2 = (dx * dx + dy * dy) + dz * dz;
      if (!(r2 > P.far_t)) continue;
      const D3 r3 = c.r(i, sk), r{r3.x, r3.y, 0.0};
      far = far + (A(P, norm(r3)) / (kden[sk] + 1e-8)) * cross(r, w[sk]);
      added += 1;
    }
    D3 vs{0, 0, 0};
    if (added > 0) vs = (P.vp_factor / added * Si) * far;
    vv[i] = P.vp_factor * near + vs;
  }
  if (dbg) {
    put3(dbg->acc, acc);
    put3(dbg->w, w);
    put1(dbg->kden, kden);
    put3(dbg->vvort, vv);
  }
  for (int i = 0; i < n; ++i) {
    if (S.gate[i]) V[i] = V[i] + vv[i];
    V[i] = V[i] + P.dt * acc[i];
  }
  // pressure solve
  std::vector<double> da(n);
  divergence(c, V, P.m, S.da32, P.rho0_f, P.rho0, &rho, P.dt, da);
  if (dbg) put1(dbg->da0, da);
  for (int t = 0; t < S.kp; ++t) {
    const float* dt32 = S.da32 + (int64_t)t * n;
    for (int i = 0; i < n; ++i) {
      kap[i] = (da[i] - P.rho0) * a_p[i];
      k32[i] = (dt32[i] - P.rho0_f) * S.alpha_p32[i];
    }
    jacobi(c, V, kap, k32);
    if (t + 1 < S.kp) divergence(c, V, P.m, S.da32 + (int64_t)(t + 1) * n, P.rho0_f, P.rho0, &rho, P.dt, da);
  }
  for (int s = 0; s < n; ++s) {
    Vid[S.ids[s]] = V[s];
    Xid[S.ids[s]] = c.X[s] + P.dt * V[s];
  }
}

}  // namespace

extern "C" {

// Runs nsteps frozen steps from (x0, v0) in particle-id order (n x 3 doubles
// each) and writes the final velocities (and positions if x_out); dbg
// (optional) receives the last step's per-pass values.
void or_frozen_rollout(const FrozenParams* P, int32_t nsteps, const FrozenStep* steps, int32_t n,
                       const double* x0, const double* v0, double* v_out, double* x_out, const FrozenDebug* dbg) {
  std::vector<D3> X(n), V(n);
  for (int p = 0; p < n; ++p) {
    X[p] = {x0[3 * p], x0[3 * p + 1], x0[3 * p + 2]};
    V[p] = {v0[3 * p], v0[3 * p + 1], v0[3 * p + 2]};
  }
  for (int k = 0; k < nsteps; ++k) step(*P, steps[k], X, V, k + 1 == nsteps ? dbg : nullptr);
  for (int p = 0; p < n; ++p) {
    v_out[3 * p] = V[p].x;
    v_out[3 * p + 1] = V[p].y;
    v_out[3 * p + 2] = V[p].z;
    if (x_out) {
      x_out[3 * p] = X[p].x;
      x_out[3 * p + 1] = X[p].y;
      x_out[3 * p + 2] = X[p].z;
    }
  }
}

}  // extern "C"
