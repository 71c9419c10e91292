// The ASCII PLY rows of KarmanVortexSolver.export_ply (karman_vortex.py:
// 122-214), formatted on the device from the particle buffers.
//
// One row per active slot below particle_num, in slot order:
//   "{x:.7f} {y:.7f} {z:.7f} {wx:.7f} {wy:.7f} {wz:.7f} {r} {g} {b} {oid} {type}\n"
// with x from x_temp, w from vorticity, (r, g, b) the colour, oid the object
// id and type the particle-type code of the object.  A pass computes each
// row's length, an exclusive scan places the rows, a second pass writes them.
//
// Each float is written as Python's format(float(v), ".7f") writes the
// float32 value v: its exact binary value rounded to 7 decimals, ties to
// even, with the sign of v kept when the result is zero ("-0.0000000"),
// "nan" for any NaN, "inf" / "-inf".  The rounding is done in integer
// arithmetic on v = m 2^e (m < 2^24): the fraction m mod 2^-e scaled by 10^7
// fits 48 bits, and integers up to 2^128 are converted through 32-bit limbs.
#include "common.h"

namespace dimcv {

struct Sink {
  char* p;        // null: count only
  int64_t n;
  __host__ __device__ __forceinline__ void put(char c) {
    if (p) p[n] = c;
    ++n;
  }
};

__host__ __device__ void put_uint(Sink& o, uint64_t v) {
  char buf[20];
  int k = 0;
  do {
    buf[k++] = (char)('0' + v % 10);
    v /= 10;
  } while (v);
  while (k) o.put(buf[--k]);
}

__host__ __device__ void put_int(Sink& o, int32_t v) {
  if (v < 0) {
    o.put('-');
    put_uint(o, (uint64_t)(-(int64_t)v));
  } else {
    put_uint(o, (uint64_t)v);
  }
}

// m 2^e for e >= 0 (m < 2^24, e <= 104): decimal digits through base-10^9
// chunks of a 160-bit integer.
__host__ __device__ void put_big(Sink& o, uint32_t m, int e) {
  uint32_t L[5] = {0u, 0u, 0u, 0u, 0u};
  const int w = e >> 5, b = e & 31;
  const uint64_t sh = (uint64_t)m << b;
  L[w] = (uint32_t)sh;
  if (w + 1 < 5) L[w + 1] = (uint32_t)(sh >> 32);
  uint32_t chunk[6];
  int nc = 0;
  for (;;) {
    bool zero = true;
    for (int i = 0; i < 5; ++i) zero = zero && L[i] == 0u;
    if (zero) break;
    uint64_t rem = 0;
    for (int i = 4; i >= 0; --i) {
      const uint64_t cur = (rem << 32) | L[i];
      L[i] = (uint32_t)(cur / 1000000000ull);
      rem = cur % 1000000000ull;
    }
    chunk[nc++] = (uint32_t)rem;
  }
  if (nc == 0) {
    o.put('0');
    return;
  }
  put_uint(o, chunk[nc - 1]);
  for (int c = nc - 2; c >= 0; --c) {
    uint32_t v = chunk[c];
    char d[9];
    for (int i = 8; i >= 0; --i) {
      d[i] = (char)('0' + v % 10);
      v /= 10;
    }
    for (int i = 0; i < 9; ++i) o.put(d[i]);
  }
}

__host__ __device__ void put_f7(Sink& o, float v) {
  const uint32_t bits = __builtin_bit_cast(uint32_t, v);
  const uint32_t ex = (bits >> 23) & 0xffu, man = bits & 0x7fffffu;
  if (ex == 0xffu) {
    if (man) {
      o.put('n'); o.put('a'); o.put('n');
    } else {
      if (bits >> 31) o.put('-');
      o.put('i'); o.put('n'); o.put('f');
    }
    return;
  }
  if (bits >> 31) o.put('-');
  const uint64_t m = ex ? (man | 0x800000u) : man;
  const int e = ex ? (int)ex - 150 : -149;
  uint64_t q = 0;  // the seven decimals
  if (e >= 0) {
    put_big(o, (uint32_t)m, e);
  } else {
    const int k = -e;
    uint64_t I = k < 64 ? (m >> k) : 0;
    const uint64_t r = k < 64 ? m - (I << k) : m;
    const uint64_t N = r * 10000000ull;  // < 2^48
    uint64_t rem;
    if (k < 64) {
      q = N >> k;
      rem = N - (q << k);
    } else {
      q = 0;
      rem = N;
    }
    bool up = false;
    if (k <= 64) {
      const uint64_t half = 1ull << (k - 1);
      up = rem > half || (rem == half && (q & 1ull));
    }
    if (up && ++q == 10000000ull) {
      q = 0;
      ++I;
    }
    put_uint(o, I);
  }
  o.put('.');
  char d[7];
  for (int i = 6; i >= 0; --i) {
    d[i] = (char)('0' + q % 10);
    q /= 10;
  }
  for (int i = 0; i < 7; ++i) o.put(d[i]);
}

__device__ __forceinline__ bool ply_row(const dimcv_state& st, int64_t s, int32_t bound) {
  return s < bound && s < st.counters[DIMCV_C_PARTICLE_NUM] && st.is_active[s];
}

// karman_vortex.py:163-176: the type code of the particle's object
__device__ void put_row(Sink& o, const dimcv_state& st, int64_t s, const int32_t* obj_type, int32_t n_type) {
  for (int d = 0; d < 3; ++d) {
    put_f7(o, st.x_temp[3 * s + d]);
    o.put(' ');
  }
  for (int d = 0; d < 3; ++d) {
    put_f7(o, st.vorticity[3 * s + d]);
    o.put(' ');
  }
  for (int d = 0; d < 3; ++d) {
    put_int(o, st.color[3 * s + d]);
    o.put(' ');
  }
  const int32_t oid = st.object_id[s];
  put_int(o, oid);
  o.put(' ');
  put_int(o, oid >= 0 && oid < n_type ? obj_type[oid] : 4);
  o.put('\n');
}

__global__ void k_ply_len(dimcv_state st, int32_t bound, const int32_t* __restrict__ obj_type, int32_t n_type,
                          uint32_t* __restrict__ len, int32_t* __restrict__ rows) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= bound) return;
  const bool on = ply_row(st, s, bound);
  uint32_t n = 0;
  if (on) {
    Sink o{nullptr, 0};
    put_row(o, st, s, obj_type, n_type);
    n = (uint32_t)o.n;
  }
  len[s] = n;
  const uint64_t b = __ballot(on);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(rows, (int32_t)__popcll(b));
}

__global__ void k_ply_write(dimcv_state st, int32_t bound, const int32_t* __restrict__ obj_type, int32_t n_type,
                            const uint32_t* __restrict__ off, char* __restrict__ out, int64_t out_bytes) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= bound || !ply_row(st, s, bound)) return;
  Sink count{nullptr, 0};
  put_row(count, st, s, obj_type, n_type);
  if ((int64_t)off[s] + count.n > out_bytes) return;  // guarded by the caller's size check
  Sink o{out + off[s], 0};
  put_row(o, st, s, obj_type, n_type);
}

__global__ void k_ply_zero(int32_t* rows) { *rows = 0; }

// Row lengths, their exclusive scan (byte offsets, in the workspace) and the
// totals (rows, bytes) in the workspace's compaction counters.
static int ply_layout(const dimcv_state* st, const dimcv_params* p, int32_t bound, const int32_t* obj_type,
                      int32_t n_type, Workspace& ws, hipStream_t s) {
  int32_t* tot = (int32_t*)ws.compact_cnt;  // [0] rows, [1] bytes
  hipLaunchKernelGGL(k_ply_zero, dim3(1), dim3(1), 0, s, tot);
  hipLaunchKernelGGL(k_ply_zero, dim3(1), dim3(1), 0, s, tot + 1);
  if (bound <= 0) return DIMCV_OK;
  hipLaunchKernelGGL(k_ply_len, dim3(nblocks(bound)), dim3(256), 0, s, *st, bound, obj_type, n_type, ws.keys[0],
                     tot);
  DIMCV_LAUNCHED("k_ply_len");
  return exclusive_scan_u32(ws.keys[0], ws.keys[0], bound, ws.radix_hist, (uint32_t*)(tot + 1), s);
}

static int ply_args(const dimcv_state* st, const dimcv_params* p, int32_t bound, const int32_t* obj_type,
                    int32_t n_type) {
  if (!st || !p || bound < 0 || bound > p->n_max || n_type < 0 || (n_type > 0 && !obj_type)) {
    set_error("dimcv_export_ply: bad arguments");
    return DIMCV_ERR_ARG;
  }
  // byte offsets are 32-bit: a row is at most 6 * 48 + 5 * 12 bytes
  if ((int64_t)bound * 384 >= (1LL << 32)) {
    set_error("dimcv_export_ply: %d slots exceed one call's 4 GiB of rows", bound);
    return DIMCV_ERR_RANGE;
  }
  return DIMCV_OK;
}

}  // namespace dimcv

using namespace dimcv;

extern "C" {

int32_t dimcv_format_f7(float v, char* out) {
  if (!out) return -1;
  Sink o{out, 0};
  put_f7(o, v);
  out[o.n] = '\0';
  return (int32_t)o.n;
}

int dimcv_export_ply_size(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound,
                          const int32_t* obj_type, int32_t n_obj_type, int64_t* rows, int64_t* bytes,
                          void* stream) {
  int rc = ply_args(st, p, particle_bound, obj_type, n_obj_type);
  if (rc) return rc;
  if (!rows || !bytes) {
    set_error("dimcv_export_ply_size: null output");
    return DIMCV_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  rc = ply_layout(st, p, particle_bound, obj_type, n_obj_type, ws, s);
  if (rc) return rc;
  int32_t h[2];
  DIMCV_HIP(hipMemcpyAsync(h, ws.compact_cnt, sizeof(h), hipMemcpyDeviceToHost, s));
  DIMCV_HIP(hipStreamSynchronize(s));
  *rows = h[0];
  *bytes = (int64_t)(uint32_t)h[1];
  return DIMCV_OK;
}

int dimcv_export_ply_rows(const dimcv_state* st, const dimcv_params* p, int32_t particle_bound,
                          const int32_t* obj_type, int32_t n_obj_type, char* out, int64_t out_bytes,
                          void* stream) {
  int rc = ply_args(st, p, particle_bound, obj_type, n_obj_type);
  if (rc) return rc;
  if (!out && out_bytes > 0) {
    set_error("dimcv_export_ply_rows: null output");
    return DIMCV_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  Workspace ws;
  rc = bind_workspace(st, p, &ws, 0);
  if (rc) return rc;
  rc = ply_layout(st, p, particle_bound, obj_type, n_obj_type, ws, s);
  if (rc) return rc;
  if (particle_bound > 0)
    hipLaunchKernelGGL(k_ply_write, dim3(nblocks(particle_bound)), dim3(256), 0, s, *st, particle_bound, obj_type,
                       n_obj_type, ws.keys[0], out, out_bytes);
  DIMCV_LAUNCHED("k_ply_write");
  return DIMCV_OK;
}

}  // extern "C"
