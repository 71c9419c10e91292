// Library plumbing: error state, workspace layout, canonical FP64 reduction of
// block partials, device-wide exclusive u32 scan.
#include "common.h"

namespace dimcv {

static thread_local std::string g_error;

void set_error(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
}

int hip_fail(hipError_t e, const char* what) {
  set_error("%s failed: %s", what, hipGetErrorString(e));
  return DIMCV_ERR_HIP;
}

static int64_t align256(int64_t b) { return (b + 255) & ~int64_t(255); }

int64_t workspace_layout(int64_t n_max, int64_t n_cells, int64_t max_samples, Workspace* ws,
                         void* base) {
  (void)n_cells;
  char* b = (char*)base;
  int64_t off = 0;
  auto take = [&](int64_t bytes) -> void* {
    void* q = b ? (void*)(b + off) : nullptr;
    off += align256(bytes > 0 ? bytes : 1);
    return q;
  };
  int64_t n = n_max > 0 ? n_max : 1;
  int64_t sort_blocks = (n + DIMCV_SORT_TILE - 1) / DIMCV_SORT_TILE;
  int64_t nb256 = (n + DIMCV_BLOCK - 1) / DIMCV_BLOCK;
  int64_t scan_blocks = (n + DIMCV_SCAN_TILE - 1) / DIMCV_SCAN_TILE;
  Workspace w{};
  w.keys[0] = (uint32_t*)take(4 * n);
  w.keys[1] = (uint32_t*)take(4 * n);
  w.vals[0] = (uint32_t*)take(4 * n);
  w.vals[1] = (uint32_t*)take(4 * n);
  w.radix_hist = (uint32_t*)take(4 * 512 * sort_blocks);  // up to 9-bit digits
  w.scan_tmp = (uint32_t*)take(4 * (512 * sort_blocks / 1024 + 1024));
  w.compact_cnt = (uint32_t*)take(4 * (nb256 + 1024));
  w.partials = (double*)take(8 * (nb256 + nb256 / 256 + 512));
  w.cdf = (double*)take(8 * n);
  w.cdf_weights = (double*)take(8 * n);
  w.block_tot = (double*)take(8 * (scan_blocks + 16));
  w.block_pre = (double*)take(8 * (scan_blocks + 16));
  w.near3 = (float*)take(4 * 3 * n);
  w.scale_s = (float*)take(4 * n);
  w.xm = (float4*)take(16 * n);
  w.xw = (float4*)take(16 * n);
  w.kap = (float*)take(4 * n);
  w.tiny = (int32_t*)take(256);
  w.bar_tree = (int32_t*)take(4 * 16 * DIMCV_BAR_GROUPS);
  // split far-field partials: splits * targets of far_split_count for any
  // target count up to n (16 n when n < 65536, below 2^21 otherwise)
  w.far_partial_cap = n * 16 < DIMCV_FAR_PARTIAL_MAX ? n * 16 : DIMCV_FAR_PARTIAL_MAX;
  w.far_partials = take(32 * w.far_partial_cap);
  int64_t ms = max_samples > 0 ? max_samples : 1;
  w.samples = (float4*)take(32 * ms);
  w.max_samples = ms;
  w.bytes = off;
  if (ws) *ws = w;
  return off;
}

int bind_workspace(const dimcv_state* st, const dimcv_params* p, Workspace* ws, int64_t max_samples) {
  if (!st || !p) {
    set_error("null state or params");
    return DIMCV_ERR_ARG;
  }
  if (p->n_max > DIMCV_MAX_SLOTS) {
    set_error("n_max %lld exceeds DIMCV_MAX_SLOTS (%d)", (long long)p->n_max, DIMCV_MAX_SLOTS);
    return DIMCV_ERR_ARG;
  }
  // The sample table is sized by whatever the caller's workspace has room for.
  int64_t need0 = workspace_layout(p->n_max, p->n_cells, 1, nullptr, nullptr);
  if (st->workspace == nullptr || st->workspace_bytes < need0) {
    set_error("workspace too small: %lld < %lld bytes", (long long)st->workspace_bytes,
              (long long)need0);
    return DIMCV_ERR_WORKSPACE;
  }
  // the sample table is the last region; it starts at need0 - align256(32)
  int64_t room = (st->workspace_bytes - (need0 - 256)) / 32;
  workspace_layout(p->n_max, p->n_cells, room, ws, st->workspace);
  if (max_samples > ws->max_samples) {
    set_error("workspace holds %lld samples, %lld requested", (long long)ws->max_samples,
              (long long)max_samples);
    return DIMCV_ERR_WORKSPACE;
  }
  return DIMCV_OK;
}

// ---------------------------------------------------------------------------
// Canonical FP64 reduction of an array of partials: repeated 256-wide halving
// trees, one workgroup, level by level (same definition as the per-block
// trees that produced the partials).
__global__ void k_reduce_levels(double* a, int64_t n, double* out, const int32_t* done) {
  __shared__ double lds[256];
  if (done && *done) return;
  int64_t m = n;
  while (m > 1) {
    int64_t groups = (m + 255) / 256;
    for (int64_t g = 0; g < groups; ++g) {
      int64_t i = g * 256 + threadIdx.x;
      double v = i < m ? a[i] : 0.0;
      double s = block_tree_sum256(v, lds);
      if (threadIdx.x == 0) a[g] = s;
      __syncthreads();
    }
    m = groups;
  }
  if (threadIdx.x == 0) *out = (n > 0) ? a[0] : 0.0;
}

// One level of the canonical tree for all groups at once: group g of 256
// consecutive values -> b[g].
__global__ void k_reduce_groups(const double* __restrict__ a, int64_t n, double* __restrict__ b,
                                const int32_t* done) {
  __shared__ double lds[256];
  if (done && *done) return;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double s = block_tree_sum256(i < n ? a[i] : 0.0, lds);
  if (threadIdx.x == 0) b[blockIdx.x] = s;
}

// reduce_partials_f64 in one launch, optionally followed by the Jacobi
// convergence check.  Block g sums group g (level 1 of the tree) into
// b[g] = a[n + g]; the last block to finish (a ticket in counters, device-
// scope fences around it) runs the remaining levels exactly as
// k_reduce_levels does, writes *out, clears the ticket and, with check,
// applies the Jacobi loop's test (dimcv_sph.py:376-393, 597-614): one more
// iteration; done when the average residual is <= eta or max_iter is hit.
__global__ void k_reduce_tail(double* a, int64_t n, double* out, int32_t* counters, const int32_t* done,
                              int check, int32_t fluid_count, double eta, int32_t max_iter) {
  __shared__ double lds[256];
  __shared__ int last;
  if (done && *done) return;
  const int t = threadIdx.x;
  const int64_t G = gridDim.x;
  const int64_t i = (int64_t)blockIdx.x * 256 + t;
  double r = block_tree_sum256(i < n ? a[i] : 0.0, lds);
  if (G > 1) {
    double* b = a + n;
    if (t == 0) {
      b[blockIdx.x] = r;
      __threadfence();
      last = atomicAdd(&counters[DIMCV_C_REDUCE_TICKET], 1) == (int)G - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    int64_t m = G;
    while (m > 1) {
      const int64_t groups = (m + 255) / 256;
      for (int64_t g = 0; g < groups; ++g) {
        const int64_t k = g * 256 + t;
        const double v = k < m ? __hip_atomic_load(&b[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        const double s = block_tree_sum256(v, lds);
        if (t == 0) b[g] = s;
        __syncthreads();
      }
      m = groups;
    }
    r = __hip_atomic_load(&b[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t == 0) counters[DIMCV_C_REDUCE_TICKET] = 0;
  }
  if (t != 0) return;
  *out = n > 0 ? r : 0.0;
  if (check) {
    const int32_t it = ++counters[DIMCV_C_JACOBI_ITERS];
    // fluid_count 0: the fluid count of the last grid build, on the device
    const int32_t fc = fluid_count > 0 ? fluid_count : counters[DIMCV_C_FLUID_NUM];
    const double avg = r / (double)fc;
    if (avg <= eta || it >= max_iter) counters[DIMCV_C_JACOBI_DONE] = 1;
  }
}

int reduce_tail_f64(double* partials, int64_t n, double* out, int32_t* counters, const int32_t* done, int check,
                    int32_t fluid_count, double eta, int32_t max_iter, hipStream_t s) {
  const unsigned g = n > 0 ? (unsigned)((n + 255) / 256) : 1;
  hipLaunchKernelGGL(k_reduce_tail, dim3(g), dim3(256), 0, s, partials, n, out, counters, done, check, fluid_count,
                     eta, max_iter);
  DIMCV_LAUNCHED("k_reduce_tail");
  return DIMCV_OK;
}

int reduce_partials_f64(const double* partials, int64_t n, double* out, void* stream, const int32_t* done) {
  hipStream_t s = (hipStream_t)stream;
  double* a = const_cast<double*>(partials);
  if (n > 256) {
    // level 2 for all groups in parallel, written after the partials
    int64_t g = (n + 255) / 256;
    hipLaunchKernelGGL(k_reduce_groups, dim3((unsigned)g), dim3(256), 0, s, a, n, a + n, done);
    a = a + n;
    n = g;
  }
  hipLaunchKernelGGL(k_reduce_levels, dim3(1), dim3(256), 0, s, a, n, out, done);
  DIMCV_LAUNCHED("k_reduce_levels");
  return DIMCV_OK;
}

// ---------------------------------------------------------------------------
// Exclusive u32 scan: 1024 elements per workgroup (4 per thread), block sums
// scanned by one workgroup, then added back.
__device__ inline uint32_t block_excl_scan_u32(uint32_t v, uint32_t* lds_w, uint32_t* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if (lane == 63) lds_w[w] = incl;
  __syncthreads();
  uint32_t wpre = 0, tot = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
    if (i < w) wpre += lds_w[i];
    tot += lds_w[i];
  }
  __syncthreads();
  if (total) *total = tot;
  return wpre + incl - v;
}

// in may alias out (in-place scan): every thread reads its four values before
// writing them, and no two threads touch the same element.
__global__ void k_scan_u32_local(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* __restrict__ bsum) {
  __shared__ uint32_t lw[4];
  int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x * 4;
  uint32_t a[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = (base + k < n) ? in[base + k] : 0u;
    s += a[k];
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan_u32(s, lw, &tot);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = pre;
    pre += a[k];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Block b adds the sum of the block totals before it (each block forms that
// carry itself: integer sums, so every order gives the same value); the last
// block also writes the total.
__global__ void k_scan_u32_add(uint32_t* __restrict__ out, int64_t n, const uint32_t* __restrict__ bsum,
                               int64_t nb, uint32_t* total) {
  __shared__ uint32_t lw[4];
  uint32_t v = 0;
  for (int64_t i = threadIdx.x; i < (int64_t)blockIdx.x; i += 256) v += bsum[i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) lw[threadIdx.x >> 6] = v;
  __syncthreads();
  const uint32_t add = lw[0] + lw[1] + lw[2] + lw[3];
  if (total && threadIdx.x == 0 && (int64_t)blockIdx.x == nb - 1) *total = add + bsum[blockIdx.x];
  int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x * 4;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k < n) out[base + k] += add;
}

// n <= 4096: the whole scan in one workgroup of 1024 threads, four values
// each (one launch instead of three; integer sums, so the same result).
__global__ void __launch_bounds__(1024) k_scan_u32_one(const uint32_t* in, uint32_t* out, int64_t n,
                                                       uint32_t* total) {
  __shared__ uint32_t lw[16];
  const int64_t base = (int64_t)threadIdx.x * 4;
  uint32_t a[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = (base + k < n) ? in[base + k] : 0u;
    s += a[k];
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan_u32(s, lw, &tot);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = pre;
    pre += a[k];
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

int exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* tmp, uint32_t* total,
                       void* stream) {
  if (n <= 0) return DIMCV_OK;
  hipStream_t s = (hipStream_t)stream;
  if (n <= 4096) {
    hipLaunchKernelGGL(k_scan_u32_one, dim3(1), dim3(1024), 0, s, in, out, n, total);
    DIMCV_LAUNCHED("exclusive_scan_u32");
    return DIMCV_OK;
  }
  unsigned nb = nblocks(n, 1024);
  hipLaunchKernelGGL(k_scan_u32_local, dim3(nb), dim3(256), 0, s, in, out, n, tmp);
  hipLaunchKernelGGL(k_scan_u32_add, dim3(nb), dim3(256), 0, s, out, n, (const uint32_t*)tmp, (int64_t)nb, total);
  DIMCV_LAUNCHED("exclusive_scan_u32");
  return DIMCV_OK;
}

}  // namespace dimcv

extern "C" {

int dimcv_abi_version(void) { return DIMCV_ABI_VERSION; }

const char* dimcv_last_error(void) { return dimcv::g_error.c_str(); }

int64_t dimcv_workspace_bytes(int64_t n_max, int64_t n_cells, int64_t max_samples) {
  return dimcv::workspace_layout(n_max, n_cells, max_samples, nullptr, nullptr);
}

}  // extern "C"
