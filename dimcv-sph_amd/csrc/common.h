// Shared device helpers and host plumbing for libdimcv (gfx950).
//
// Every f32 helper below performs the reference's operations in the
// reference's order (Taichi evaluates left to right, vector ops elementwise,
// dot/norm as left-to-right sums).  The library is compiled with
// -ffp-contract=off and IEEE-correct f32 division and sqrt, so the SPH passes
// reproduce the serial reference arithmetic bit for bit; kernels that trade
// that for speed use explicit fmaf / reciprocal builtins.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/dimcv.h"

#define DIMCV_BLOCK 256
#define DIMCV_SORT_TILE 4096  // elements per radix-sort workgroup (256 x 16)
#define DIMCV_SCAN_TILE 4096  // elements per canonical-scan workgroup (256 x 16)
#define DIMCV_MAX_SCAN_BLOCKS 4096

namespace dimcv {

// ---------------------------------------------------------------------------
// error plumbing (host)
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define DIMCV_HIP(call)                                   \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return ::dimcv::hip_fail(e_, #call); \
  } while (0)

#define DIMCV_LAUNCHED(what)                                  \
  do {                                                        \
    hipError_t e_ = hipGetLastError();                        \
    if (e_ != hipSuccess) return ::dimcv::hip_fail(e_, what); \
  } while (0)

inline unsigned nblocks(int64_t n, int per = DIMCV_BLOCK) {
  if (n <= 0) return 0;
  return (unsigned)((n + per - 1) / per);
}

// ---------------------------------------------------------------------------
// workspace layout (all offsets 256-byte aligned)
struct Workspace {
  uint32_t* keys[2];
  uint32_t* vals[2];
  uint32_t* radix_hist;   // 512 * sort blocks (digits of up to 9 bits)
  uint32_t* scan_tmp;     // block sums for the u32 scan
  uint32_t* compact_cnt;  // per-block counts for compaction
  double* partials;       // per-block FP64 partial sums
  double* cdf;            // canonical scan output
  double* cdf_weights;    // scan input
  double* block_tot;      // canonical scan block totals
  double* block_pre;      // scan of block totals
  float* near3;           // vort2vel near field per fluid entry (3 floats)
  float* scale_s;         // vort2vel S_i per fluid entry
  void* far_partials;     // split far-field partial sums (32 B each)
  int64_t far_partial_cap;
  float4* xm;             // per slot (x, y, z, m_V) of the current step (neighbour gathers)
  float4* xw;             // per slot (x, y, z, w), w = m (fluid) or m_V rho0 (the Jacobi passes)
  float* kap;             // per slot kappa of the current Jacobi iteration
  int32_t* tiny;          // step flag: some coordinate is nonzero and below 2^-76 in magnitude
  int32_t* bar_tree;      // group words of the persistent solves' grid barrier (grid_sync_tree)
  float4* samples;        // packed far-field sources: (x, y, z, -) (cx, cy, cz, -)
  int64_t max_samples;
  int64_t bytes;
};
int64_t workspace_layout(int64_t n_max, int64_t n_cells, int64_t max_samples, Workspace* ws,
                         void* base);
int bind_workspace(const dimcv_state* st, const dimcv_params* p, Workspace* ws, int64_t max_samples);

// ---------------------------------------------------------------------------
// device math
struct v3 {
  float x, y, z;
};
__device__ __forceinline__ v3 mk(float a, float b, float c) { return v3{a, b, c}; }
__device__ __forceinline__ v3 vsub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 vadd(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 smul(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ v3 vmuls(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 vdivs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Correctly rounded square root of x >= 0 (a sum of squares): the rounding
// steps of the IEEE square root hipcc emits for gfx950 (inputs below 2^-96
// scaled by 2^32, the hardware estimate, its neighbours one ulp down and up
// kept or dropped by the signs of their residuals fma(-s', s, x), the result
// scaled back by 2^-16), without that sequence's final class test for zero
// and infinity, whose results the steps already give for x = +0 and +inf.
__device__ __forceinline__ float sqrt_nonneg(float x) {
  const bool small = x < 0x1p-96f;
  const float xs = small ? x * 0x1p32f : x;
  const float s = __builtin_amdgcn_sqrtf(xs);
  const float dn = __uint_as_float(__float_as_uint(s) - 1u);
  const float up = __uint_as_float(__float_as_uint(s) + 1u);
  float r = fmaf(-dn, s, xs) <= 0.0f ? dn : s;
  r = fmaf(-up, s, xs) > 0.0f ? up : r;
  return small ? r * 0x1p-16f : r;
}
__device__ __forceinline__ float vnorm(v3 a) { return sqrt_nonneg(vdot(a, a)); }
__device__ __forceinline__ v3 vcross(v3 a, v3 b) {
  return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ v3 ld3(const float* __restrict__ a, int64_t i) {
  return mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]);
}
// Neighbour-field reads by slot through a 32-bit byte offset from the
// (wave-uniform) field base, so a gather is one scalar-base load with a
// vector offset instead of 64-bit address arithmetic per lane.  Slots stay
// below DIMCV_MAX_SLOTS, so 12 * slot fits in 32 bits.
// Byte offset j * 12 as (j + 2j) * 4: two full-rate shift-adds in place of
// the quarter-rate 32-bit integer multiply the compiler picks for j * 12 (the
// empty asm keeps the two from being folded back into that multiply).
__device__ __forceinline__ uint32_t off12(uint32_t j) {
  uint32_t t = j + (j << 1);
  asm("" : "+v"(t));
  return t << 2;
}
__device__ __forceinline__ v3 ld3s(const float* __restrict__ a, uint32_t j) {
  const float* q = (const float*)((const char*)a + off12(j));
  return mk(q[0], q[1], q[2]);
}
__device__ __forceinline__ float ld1s(const float* __restrict__ a, uint32_t j) {
  return *(const float*)((const char*)a + j * 4u);
}
__device__ __forceinline__ void st3(float* __restrict__ a, int64_t i, v3 q) {
  a[3 * i] = q.x;
  a[3 * i + 1] = q.y;
  a[3 * i + 2] = q.z;
}

// Cubic spline kernel, sph_base.py:23-44 ((1-q)^3 evaluated as (t*t)*t).
__device__ __forceinline__ float cubic_W(const dimcv_params& p, float r_norm) {
  const float q = r_norm * p.inv_support_radius;
  // both branches' values, then selects (see cubic_dW)
  const float q2 = q * q;
  const float q3 = q2 * q;
  const float wa = p.kernel_k * (6.0f * q3 - 6.0f * q2 + 1.0f);
  const float t = 1.0f - q;
  const float wb = p.kernel_k * 2.0f * ((t * t) * t);
  return q <= 1.0f ? (q <= 0.5f ? wa : wb) : 0.0f;
}

// a / b for the three components of a, rounded exactly as the IEEE division
// hipcc emits for gfx950 (v_div_scale, v_rcp, one refinement of the
// reciprocal, the quotient and two residual corrections, v_div_fmas,
// v_div_fixup), with the reciprocal refinement done once for the shared b.
// Bitwise equal to a.x / b etc. whenever the scale and fix-up steps change
// nothing: b a normal float in [2^-100, 2^100] and each component 0 or of
// magnitude in [2^-100, b]; the fix-up then only sets the sign, sign(a).
__device__ __forceinline__ float div_corrected(float x, float b, float r) {
  float q = x * r;
  const float e2 = fmaf(-b, q, x);
  q = fmaf(e2, r, q);
  const float e3 = fmaf(-b, q, x);
  q = fmaf(e3, r, q);
  return __int_as_float((__float_as_int(q) & 0x7fffffff) | (__float_as_int(x) & (int)0x80000000));
}
__device__ __forceinline__ v3 vdivs_shared(v3 a, float b) {
  float r = __builtin_amdgcn_rcpf(b);
  const float e = fmaf(-b, r, 1.0f);
  r = fmaf(e, r, r);
  return mk(div_corrected(a.x, b, r), div_corrected(a.y, b, r), div_corrected(a.z, b, r));
}

// Kernel gradient, sph_base.py:46-68.  With shared_div the three divisions
// r / |r| share their reciprocal (vdivs_shared); the caller passes it only
// when every coordinate of the step is 0 or at least 2^-76 in magnitude, so
// every nonzero component of r (a difference of two such floats, or one
// minus a box length) is at least 2^-99 and the quotients are bitwise the
// IEEE ones (|r| > 1e-5 and |r| <= h here).
__device__ __forceinline__ v3 cubic_dW(const dimcv_params& p, v3 r, bool shared_div = false) {
  float rn = vnorm(r);
  float q = rn * p.inv_support_radius;
  v3 res = mk(0.0f, 0.0f, 0.0f);
  if (rn > 1e-5f && q <= 1.0f) {
    v3 gq = vmuls(shared_div ? vdivs_shared(r, rn) : vdivs(r, rn), p.inv_support_radius);
    // both branches' factors, then a select: the lanes of a wave sit on both
    // sides of q = 1/2, where a branch would run both sides anyway
    const float fa = p.kernel_dk * q * (3.0f * q - 2.0f);
    const float f = 1.0f - q;
    const float fb = p.kernel_dk * (-f * f);
    res = smul(q <= 0.5f ? fa : fb, gq);
  }
  return res;
}

// f32 exp: Cody-Waite reduction x = k ln2 + r, degree-8 Taylor polynomial in
// r (Horner, no fused operations), scaling by 2^k through exponent bits.
__device__ __forceinline__ float pow2i(int k) { return __uint_as_float((uint32_t)(k + 127) << 23); }
__device__ inline float exp_f32(float x) {
  if (x != x) return x;
  if (x > 88.72283935546875f) return __int_as_float(0x7f800000);
  if (x < -103.97208404541015625f) return 0.0f;
  float kf = floorf(x * 1.44269502162933349609375f + 0.5f);
  float r = (x - kf * 0.693145751953125f) - kf * 1.428606765330187045e-06f;
  float q = 2.4801587301587302e-05f;
  q = q * r + 1.9841269841269841e-04f;
  q = q * r + 1.3888888888888889e-03f;
  q = q * r + 8.3333333333333332e-03f;
  q = q * r + 4.1666666666666664e-02f;
  q = q * r + 1.6666666666666666e-01f;
  q = q * r + 0.5f;
  q = q * r + 1.0f;
  q = q * r + 1.0f;
  int k = (int)kf;
  if (k > 127) return (q * pow2i(127)) * pow2i(k - 127);
  if (k < -126) return (q * pow2i(k + 64)) * pow2i(-64);
  return q * pow2i(k);
}

// f32 log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1)).
__device__ inline float log_f32(float x) {
  if (x != x) return x;
  if (x < 0.0f) return __int_as_float(0x7fc00000);
  if (x == 0.0f) return __int_as_float((int)0xff800000);
  if (x == __int_as_float(0x7f800000)) return x;
  int e = 0;
  if (x < 1.17549435e-38f) {
    x = x * 8388608.0f;
    e = -23;
  }
  uint32_t u = __float_as_uint(x);
  e += (int)((u >> 23) & 0xffu) - 127;
  float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
  if (m > 1.41421353816986083984375f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float t = 0.1111111111111111f;
  t = t * s2 + 0.14285714285714285f;
  t = t * s2 + 0.2f;
  t = t * s2 + 0.33333333333333331f;
  t = t * s2 + 1.0f;
  float lm = (2.0f * s) * t;
  float ef = (float)e;
  return ef * 0.693145751953125f + (ef * 1.428606765330187045e-06f + lm);
}

__device__ __forceinline__ float pow_f32(float x, float y) { return exp_f32(y * log_f32(x)); }

// dimcv_sph.py:76-84
__device__ __forceinline__ bool in_vortex_domain(const dimcv_params& p, v3 x) {
  if (!p.has_vortex_domain) return true;
  if (x.x < p.vortex_start[0] || x.x > p.vortex_end[0]) return false;
  if (x.y < p.vortex_start[1] || x.y > p.vortex_end[1]) return false;
  if (x.z < p.vortex_start[2] || x.z > p.vortex_end[2]) return false;
  return true;
}

// dimcv_sph.py:401-405
__device__ __forceinline__ float transit_weight(const dimcv_params& p, float xpos) {
  float w = (xpos - p.transit_a) / p.transit_ba;
  return fminf(fmaxf(w, 0.0f), 1.0f);
}

// Neighbour-list entry: slot in the low 31 bits, fluid flag in bit 31.
__device__ __forceinline__ int nb_slot(uint32_t e) { return (int)(e & 0x7fffffffu); }
__device__ __forceinline__ bool nb_fluid(uint32_t e) { return (e >> 31) != 0u; }

// The neighbour list of fluid-list entry f: entry k at nbr[k * n_max + f].
// Row k's base is wave-uniform (k is a loop counter), so a load is one
// scalar-base load with the lane's 32-bit byte offset 4 f.
struct NbrList {
  const uint32_t* __restrict__ base;
  uint32_t lane;
  __device__ __forceinline__ NbrList(const uint32_t* __restrict__ nbr, int64_t f)
      : base(nbr), lane((uint32_t)f * 4u) {}
  __device__ __forceinline__ uint32_t at(int64_t k, int64_t ns) const {
    return *(const uint32_t*)((const char*)(base + k * ns) + lane);
  }
};

// Register queue of the next NBR_Q list entries, loaded NBR_Q entries ahead
// of use: a group of four neighbours then waits only for its neighbours'
// data, not first for its own list entries (entries past nn read as 0 and
// are never used).
#define NBR_Q 16
struct NbrQueue {
  uint32_t q[NBR_Q];
  __device__ __forceinline__ NbrQueue(const NbrList& nl, int64_t ns, int nn) {
#pragma unroll
    for (int u = 0; u < NBR_Q; ++u) q[u] = u < nn ? nl.at(u, ns) : 0u;
  }
  __device__ __forceinline__ void advance4(const NbrList& nl, int64_t ns, int nn, int k) {
    advance<4>(nl, ns, nn, k);
  }
  // entries [k, k + S) have been used: shift and load entries k + NBR_Q ..
  template <int S>
  __device__ __forceinline__ void advance(const NbrList& nl, int64_t ns, int nn, int k) {
#pragma unroll
    for (int u = 0; u < NBR_Q - S; ++u) q[u] = q[u + S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const int kk = k + NBR_Q + u;
      q[NBR_Q - S + u] = kk < nn ? nl.at(kk, ns) : 0u;
    }
  }
};

// With nl = NbrList(st.nbr, f), ns = n_max and nn = the list length in scope:
// for k in [0, nn): e = list entry k, j = its slot, in list order (groups
// of four over the queue; `continue` moves to the next neighbour).
#define FOR_NBR(k, e, j)                                                        \
  if (NbrQueue Q_##k(nl, ns, nn); true)                                         \
    for (int k##0 = 0; k##0 < nn; Q_##k.advance4(nl, ns, nn, k##0), k##0 += 4)  \
      _Pragma("unroll") for (int k = k##0; k < k##0 + 4; ++k)                   \
        if (k < nn)                                                              \
          if (const uint32_t e = Q_##k.q[k - k##0]; true)                        \
            if (const int64_t j = nb_slot(e); true)

// Neighbour loop in two phases per group of GS list entries: gather(j)
// returns the neighbour's fields (a struct), called for all GS entries
// before body(k, e, fields) runs for the live ones, so the group's loads
// are in flight together (one memory latency per GS neighbours).  Entries
// past nn read slot 0 (a valid slot) and are never passed to body.  GS = 8
// suits passes with few gathered fields; passes with many keep 4 (register
// pressure).
template <int GS = 4, class Gather, class Body>
__device__ __forceinline__ void nbr_loop(const NbrList& nl, int64_t ns, int nn, Gather&& gather, Body&& body) {
  static_assert(GS <= NBR_Q, "group larger than the list queue");
  NbrQueue Q(nl, ns, nn);
  for (int k0 = 0; k0 < nn; Q.template advance<GS>(nl, ns, nn, k0), k0 += GS) {
    decltype(gather((uint32_t)0)) g[GS];
#pragma unroll
    for (int u = 0; u < GS; ++u) g[u] = gather((uint32_t)nb_slot(Q.q[u]));
#pragma unroll
    for (int u = 0; u < GS; ++u)
      if (k0 + u < nn) body(k0 + u, Q.q[u], g[u]);
  }
}

// Zeroes one counter (a kernel rather than a memset, so that the launch may
// sit in a captured graph).
static __global__ void k_zero_i32(int32_t* a) { *a = 0; }
// Clears n words; a kernel rather than a memset so that the clear can sit in
// a captured graph.
static __global__ void k_zero_i32n(int32_t* a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// The step's packed per-slot records (built by the substep; absent in the
// per-pass entry points): (x, m_V), (x, w) with w = m or m_V rho0, kappa,
// and the tiny-coordinate flag that gates the shared-reciprocal divisions.
struct Packed {
  const float4* xm;
  const float4* xw;
  float* kap;
  const int32_t* tiny;
};

// Two-phase neighbour loop for passes whose pairs mostly stop at a test on a
// few gathered values: gather1(j) for the group's GS entries; keep(e, a)
// selects the entries that need more; gather2(j) is issued for all of those
// together; body(k, e, a, b) runs for the kept entries in list order.  A
// group waits for two memory latencies however many entries pass the test.
template <int GS, class Gather1, class Keep, class Gather2, class Body>
__device__ __forceinline__ void nbr_loop2(const NbrList& nl, int64_t ns, int nn, Gather1&& gather1, Keep&& keep,
                                          Gather2&& gather2, Body&& body) {
  static_assert(GS <= NBR_Q, "group larger than the list queue");
  NbrQueue Q(nl, ns, nn);
  for (int k0 = 0; k0 < nn; Q.template advance<GS>(nl, ns, nn, k0), k0 += GS) {
    decltype(gather1((uint32_t)0)) a[GS];
#pragma unroll
    for (int u = 0; u < GS; ++u) a[u] = gather1((uint32_t)nb_slot(Q.q[u]));
    bool ok[GS];
#pragma unroll
    for (int u = 0; u < GS; ++u) ok[u] = k0 + u < nn && keep(Q.q[u], a[u]);
    decltype(gather2((uint32_t)0, 0u)) b[GS];
#pragma unroll
    for (int u = 0; u < GS; ++u)
      if (ok[u]) b[u] = gather2((uint32_t)nb_slot(Q.q[u]), Q.q[u]);
#pragma unroll
    for (int u = 0; u < GS; ++u)
      if (ok[u]) body(k0 + u, Q.q[u], a[u], b[u]);
  }
}

// Relative position x_i - x_j; with a periodic box the minimum image.
__device__ __forceinline__ v3 rel(const dimcv_params& p, v3 xi, v3 xj) {
  v3 r = vsub(xi, xj);
  if (p.periodic) {
    float hx = 0.5f * p.box_size[0], hy = 0.5f * p.box_size[1], hz = 0.5f * p.box_size[2];
    if (r.x > hx) r.x -= p.box_size[0];
    else if (r.x < -hx) r.x += p.box_size[0];
    if (r.y > hy) r.y -= p.box_size[1];
    else if (r.y < -hy) r.y += p.box_size[1];
    if (r.z > hz) r.z -= p.box_size[2];
    else if (r.z < -hz) r.z += p.box_size[2];
  }
  return r;
}

// ---------------------------------------------------------------------------
// XCD-aware block order.  gfx950 dispatches workgroups round-robin over its 8
// XCDs (workgroup b runs on XCD b % 8), each with its own L2.  Passes over the
// fluid list use the logical block xcd_block() instead of blockIdx.x, which
// hands every XCD one contiguous range of blocks: consecutive fluid-list
// entries are spatial neighbours (cell-sorted), so a particle's neighbour data
// is mostly already in that XCD's L2.  The map is a bijection, so every
// entry is processed exactly once whatever the dispatch order.
__device__ __forceinline__ int64_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, i = b >> 3;
  return (int64_t)(x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// ---------------------------------------------------------------------------
// Batched scenes.  A batch's fluid list holds one segment per scene starting
// at a multiple of 256; unused entries hold -1 and every pass skips them.
// block_scene[b] names the scene of 256-entry block b (-1: none).
__device__ __forceinline__ int block_scene_of(const dimcv_state& st) {
  return st.block_scene ? st.block_scene[xcd_block()] : 0;
}

// A native step whose build failed (DIMCV_C_STEP_ERROR): the kernels that
// write x, v, is_sample and life_time return without writing.
__device__ __forceinline__ bool step_failed(const dimcv_state& st) {
  return st.counters[DIMCV_C_STEP_ERROR] != 0;
}

// Jacobi gate: the solve of this block's scene (batched: done points at
// scene_counters + DIMCV_SC_DONE) or the single solve has converged.
// Launches using it are 256-wide over fluid-list entries, in xcd_block order.
// jacobi_gated_lb takes the logical block explicitly (persistent loops).
__device__ __forceinline__ bool jacobi_gated_lb(const dimcv_state& st, const int32_t* done, int64_t lb) {
  if (!done) return false;
  if (st.block_scene) {
    const int s = st.block_scene[lb];
    return s < 0 || done[(int64_t)s * DIMCV_SC_COUNT] != 0;
  }
  return *done != 0;
}
__device__ __forceinline__ bool jacobi_gated(const dimcv_state& st, const int32_t* done) {
  return jacobi_gated_lb(st, done, xcd_block());
}

// kden_sum of the scene of this block.
__device__ __forceinline__ double kden_sum_of(const dimcv_state& st) {
  if (st.block_scene)
    return st.scene_dscal[(int64_t)st.block_scene[xcd_block()] * DIMCV_SD_COUNT + DIMCV_SD_KDEN_SUM];
  return st.dscal[DIMCV_D_KDEN_SUM];
}

// Far-field source splits for a target set: about 16 waves per SIMD (256 CUs
// x 4 SIMDs) when the targets alone would leave the GPU idle, at most 16.
// A pure function of the target count, so that a scene's summation structure
// is the same alone and in a batch; splits * targets < 2^21 whenever
// splits > 1.
__host__ __device__ __forceinline__ int far_split_count(int64_t targets) {
  const int64_t want_waves = 256LL * 4 * 16;
  const int64_t waves = (targets + 63) / 64;
  int64_t s = waves > 0 ? (want_waves + waves - 1) / waves : 1;
  if (s > 16) s = 16;
  return s < 1 ? 1 : (int)s;
}

// Splits of the MC far field's sample range for a scene of `targets` fluid
// particles, the same in the single-scene and batched steps (so a batched
// scene sums in the order of a solo one): from 2^19 targets, twice the waves
// of far_split_count (two halves at 1M targets run 1.16 ms against 1.22 ms
// for one pass, the tail of the last round of waves being the difference),
// clamped so that splits * targets <= 2^21, the partial capacity of the solo
// workspace (min(16 n_max, 2^21)); the batched step sums a scene's splits in
// registers and keeps no partials.
#define DIMCV_FAR_PARTIAL_MAX (1LL << 21)
__host__ __device__ __forceinline__ int far_split_count_mc(int64_t targets) {
  int64_t s = far_split_count(targets >= (1 << 19) ? (targets + 1) / 2 : targets);
  const int64_t fit = targets > 0 ? DIMCV_FAR_PARTIAL_MAX / targets : s;
  if (s > fit) s = fit;
  return s < 1 ? 1 : (int)s;
}

// ---------------------------------------------------------------------------
// canonical FP64 block tree (256 values, a[t] += a[t + s], s = 128..1)
__device__ inline double block_tree_sum256(double v, double* lds) {
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
#pragma unroll
  for (int s = 128; s >= 1; s >>= 1) {
    if (t < s) lds[t] = lds[t] + lds[t + s];
    __syncthreads();
  }
  double r = lds[0];
  __syncthreads();
  return r;
}

// Canonical FP64 sum of scene s's block partials (fluid-list blocks of 256,
// reduced in place level by level with block_tree_sum256) into
// scene_dscal[s][which].  Every thread of the calling block takes part.
__device__ inline void batch_reduce_scene(const dimcv_state& st, double* __restrict__ partials, int which, int s,
                                          double* lds) {
  const int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
  double* a = partials + c[DIMCV_SC_FLUID_START] / 256;
  const int64_t n = ((int64_t)c[DIMCV_SC_FLUID] + 255) / 256;
  int64_t m = n;
  while (m > 1) {
    const int64_t groups = (m + 255) / 256;
    for (int64_t g = 0; g < groups; ++g) {
      const int64_t i = g * 256 + threadIdx.x;
      const double v = i < m ? a[i] : 0.0;
      const double r = block_tree_sum256(v, lds);
      if (threadIdx.x == 0) a[g] = r;
      __syncthreads();
    }
    m = groups;
  }
  if (threadIdx.x == 0) st.scene_dscal[(int64_t)s * DIMCV_SD_COUNT + which] = n > 0 ? a[0] : 0.0;
}

// Convergence test of scene s after an iteration (dimcv_sph.py:384-391 /
// 606-613: average = FP64 residual / the scene's fluid count; a scene without
// fluid counts as converged); returns whether the scene iterates on.
__device__ inline bool batch_scene_check(const dimcv_state& st, double eta, int32_t max_iter, int which, int s) {
  int32_t* c = st.scene_counters + (int64_t)s * DIMCV_SC_COUNT;
  if (c[DIMCV_SC_DONE]) return false;
  const int32_t it = ++c[DIMCV_SC_ITERS];
  const int32_t F = c[DIMCV_SC_FLUID];
  const double avg = F > 0 ? st.scene_dscal[(int64_t)s * DIMCV_SD_COUNT + DIMCV_SD_RESIDUAL] / (double)F : 0.0;
  if (F <= 0 || avg <= eta || it >= max_iter) c[DIMCV_SC_DONE] = 1;
  c[which ? DIMCV_SC_ITERS_P : DIMCV_SC_ITERS_V] = it;
  return c[DIMCV_SC_DONE] == 0;
}

// Grid-wide barrier of a persistent launch whose blocks are all resident
// (the single-scene and batched Jacobi solves), as a two-level tree: blocks
// arrive on their group's word (DIMCV_BAR_GROUP consecutive blocks per group,
// one group per 64 bytes of tree), the last block of a group arrives on the
// top word, the last group's arrival opens the next generation, and each
// group's last block then opens its group.  Through one word for every
// arrival and poll, a barrier of ~1,000 blocks cost tens of microseconds.
// Every wave waits for its stores to reach the XCD's L2, then thread 0
// releases them at agent scope (the L2s of the XCDs are not coherent with
// each other: the release writes the L2's dirty lines back) and arrives;
// after leaving, each block acquires (invalidating its CU's L1 and its
// XCD's L2).  A wait is bounded (about a second): past it the block raises
// abort, and every waiting block returns false at once, so the launch always
// ends; the caller reports the abort as an error.
// bar: top arrivals, generation, abort; tree[16 g]: arrivals of group g,
// tree[16 g + 1]: its generation (only incremented); gen / ggen: this
// block's copies of the two generations (grid_sync_tree_init at launch,
// held by thread 0).  Top and group arrival words must be 0 when a launch
// starts (cleared with the solve's reset, and left 0 by every completed
// barrier).
#define GRID_SPIN_LIMIT (1u << 20)
#define DIMCV_BAR_GROUP 16
#define DIMCV_BAR_GROUPS 256

__device__ inline void grid_sync_tree_init(const int32_t* bar, const int32_t* tree, int32_t& gen, int32_t& ggen) {
  if (threadIdx.x == 0) {
    gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ggen = __hip_atomic_load(&tree[16 * (blockIdx.x / DIMCV_BAR_GROUP) + 1], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Waits until *word differs from old; false when the barrier was aborted or
// the wait ran past GRID_SPIN_LIMIT polls (then it raises the abort flag).
__device__ inline bool bar_wait(const int32_t* word, int32_t old, int32_t* bar) {
  for (uint32_t spin = 0;; ++spin) {
    if (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != old) return true;
    if (__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
    if (spin >= GRID_SPIN_LIMIT) {
      atomicExch(&bar[2], 1);
      return false;
    }
    __builtin_amdgcn_s_sleep(8);
  }
}

__device__ inline bool grid_sync_tree(int32_t* bar, int32_t* tree, int32_t& gen, int32_t& ggen) {
  __shared__ int ok;
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this wave's stores have reached the L2
  __syncthreads();
  if (threadIdx.x == 0) {
    int r = 1;
    const int g = blockIdx.x / DIMCV_BAR_GROUP;
    const int ng = (gridDim.x + DIMCV_BAR_GROUP - 1) / DIMCV_BAR_GROUP;
    const int gsize = min((int)gridDim.x - g * DIMCV_BAR_GROUP, DIMCV_BAR_GROUP);
    int32_t* gw = tree + 16 * g;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the block's stores, from the L2 to memory
    if (atomicAdd(&gw[0], 1) == gsize - 1) {
      atomicExch(&gw[0], 0);
      if (atomicAdd(&bar[0], 1) == ng - 1) {
        atomicExch(&bar[0], 0);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        atomicAdd(&bar[1], 1);
      } else {
        r = bar_wait(&bar[1], gen, bar) ? 1 : 0;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // the group word's reset before the group opens
      if (r) atomicAdd(&gw[1], 1);
    } else {
      r = bar_wait(&gw[1], ggen, bar) ? 1 : 0;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    gen += 1;
    ggen += 1;
    ok = r;
  }
  __syncthreads();
  return ok != 0;
}

// Far-field partial of a split source range (A, Kahan compensation,
// rejected samples).
struct FarPartial {
  float a[3];
  float c[3];
  int rej;
  int pad;
};

// Batch workspace (dimcv_batch_workspace_bytes): unpadded fluid list, far
// field work order, misc counters, CDF tile totals.
enum BatchMisc {
  BM_UCOUNT = 0,       // unpadded fluid-list length
  BM_ITEMS = 1,        // far-field work items
  BM_QUEUE = 2,        // far-field work queue head
  BM_TILES = 3,        // CDF scan tiles
  BM_ERROR = 4,        // capacity error flags (1: fluid list)
  BM_BAR = 5,          // persistent Jacobi solve: barrier arrivals, generation, abort, two go flags
  BM_COUNT = 16
};
struct BatchWs {
  int32_t* ulist;       // [n_max]
  int32_t* order;       // [n_scenes]: far-field rank -> scene
  int32_t* item_start;  // [n_scenes + 1]: first work item of each rank
  int32_t* misc;        // [BM_COUNT]
  double* btot;         // [tile_bound]: CDF tile totals
  double* bpre;         // [tile_bound]: their per-scene scans
  int64_t tile_bound;   // n_max / 4096 + n_scenes + 1
  int64_t bytes;
};

// Last index r in [0, n) with a[r * stride] <= v for non-decreasing a (-1: none).
__device__ __forceinline__ int last_le(const int32_t* __restrict__ a, int n, int stride, int64_t v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[(int64_t)mid * stride] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}
int64_t batch_layout(int64_t n_scenes, int64_t n_max, BatchWs* bw, void* base);
int bind_batch(const dimcv_state* st, const dimcv_params* p, BatchWs* bw);

// the pressure update's reaction on dynamic rigid particles (grid.hip); no
// launch unless p->dyn_rigid
int launch_rigid_reaction(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound, const int32_t* done,
                          hipStream_t s);
// host helpers implemented in util.hip
int reduce_partials_f64(const double* partials, int64_t n, double* out, void* stream,
                        const int32_t* done = nullptr);
// reduce_partials_f64 in one launch (the last group finishes the tree),
// optionally followed by the Jacobi convergence test on the sum.
int reduce_tail_f64(double* partials, int64_t n, double* out, int32_t* counters, const int32_t* done, int check,
                    int32_t fluid_count, double eta, int32_t max_iter, hipStream_t s);
int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp, uint32_t* total,
                       void* stream);

}  // namespace dimcv
