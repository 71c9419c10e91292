// Vector-ALU issue-rate calibration (no counterpart in the reference): how
// many FP32 vector instructions per SIMD and shader-clock quad-cycle this
// device issues, and at which clock, under a dense load of independent
// operations.  The far field's roofline divides its instruction count by
// this ceiling rather than by the published peak, which assumes a 2.4 GHz
// clock and every cycle pairing two instructions.
//
// Two streams, each 4 resident waves per SIMD (one block of 4 waves per
// SIMD, at most 128 vector registers per wave) of long unrolled runs of 32
// independent fused multiply-adds: operands all in vector registers (two waves'
// instructions may issue in the same cycle), and one operand a scalar
// register (such instructions issue alone).  Block 0 reads the shader clock
// (s_memtime) and the 100 MHz constant clock (s_memrealtime) around its loop.
#include "common.h"

namespace dimcv {

#define PROBE_N 32
#define PROBE_UNROLL 4

template <int KIND>
__global__ void __launch_bounds__(256) k_valu_probe(float* __restrict__ sink, int iters, float a,
                                                    unsigned long long* __restrict__ clk) {
  float x[PROBE_N], y[PROBE_N], z[PROBE_N];
  const float t = (float)threadIdx.x * 1e-3f;
#pragma unroll
  for (int u = 0; u < PROBE_N; ++u) {
    x[u] = t + (float)u;
    y[u] = 0.999f + t * 1e-4f + (float)u * 1e-5f;  // per-lane values: no operand folds to a constant
    z[u] = t - (float)u;
  }
  const unsigned long long c0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < PROBE_UNROLL; ++r)
#pragma unroll
      for (int u = 0; u < PROBE_N; ++u) {
        if (KIND == 0) x[u] = fmaf(y[u], z[u], x[u]);
        else x[u] = fmaf(x[u], a, z[u]);
      }
  }
  const unsigned long long c1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
  float s = 0.0f;
#pragma unroll
  for (int u = 0; u < PROBE_N; ++u) s += x[u];
  sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    clk[0] = c1 - c0;
    clk[1] = r1 - r0;
  }
}

static int probe_run(int kind, unsigned blocks, int iters, float* sink, unsigned long long* clk, int simds,
                     hipStream_t s, double* rate, double* ghz) {
  hipEvent_t e[2] = {nullptr, nullptr};
  DIMCV_HIP(hipEventCreate(&e[0]));
  DIMCV_HIP(hipEventCreate(&e[1]));
  auto kern = kind == 0 ? k_valu_probe<0> : k_valu_probe<1>;
  for (int w = 0; w < 8; ++w)  // warm-up: the clock settles under the load (about 30 ms)
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, sink, iters, 0.999f, clk);
  DIMCV_HIP(hipEventRecord(e[0], s));
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, sink, iters, 0.999f, clk);
  DIMCV_HIP(hipEventRecord(e[1], s));
  DIMCV_LAUNCHED("k_valu_probe");
  DIMCV_HIP(hipEventSynchronize(e[1]));
  float ms = 0.0f;
  DIMCV_HIP(hipEventElapsedTime(&ms, e[0], e[1]));
  unsigned long long h[2] = {0, 0};
  DIMCV_HIP(hipMemcpy(h, clk, sizeof(h), hipMemcpyDeviceToHost));
  (void)hipEventDestroy(e[0]);
  (void)hipEventDestroy(e[1]);
  if (h[1] == 0 || ms <= 0.0f) {
    set_error("dimcv_valu_issue_probe: no clock reading");
    return DIMCV_ERR_HIP;
  }
  *ghz = (double)h[0] / ((double)h[1] / 100e6) / 1e9;
  const double wave_instr = (double)blocks * 4.0 * iters * PROBE_UNROLL * PROBE_N;
  *rate = wave_instr / simds / (ms * 1e-3 * *ghz * 1e9 / 4.0);
  return DIMCV_OK;
}

}  // namespace dimcv

using namespace dimcv;

extern "C" {

int64_t dimcv_valu_issue_probe_bytes(void) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    return -1;
  return (int64_t)cus * 4 * 256 * 4 + 256;  // 4 blocks of 256 threads per CU, then the clock words
}

int dimcv_valu_issue_probe(void* scratch, int64_t scratch_bytes, double* out, void* stream) {
  if (!scratch || !out) {
    set_error("dimcv_valu_issue_probe: null argument");
    return DIMCV_ERR_ARG;
  }
  const int64_t need = dimcv_valu_issue_probe_bytes();
  if (need < 0) {
    set_error("dimcv_valu_issue_probe: no device");
    return DIMCV_ERR_HIP;
  }
  if (scratch_bytes < need) {
    set_error("dimcv_valu_issue_probe: scratch of %lld bytes, %lld needed", (long long)scratch_bytes,
              (long long)need);
    return DIMCV_ERR_WORKSPACE;
  }
  const int64_t cus = (need - 256) / (4 * 256 * 4);
  const unsigned blocks = (unsigned)(cus * 4);  // 4 waves per SIMD, all resident at once
  float* sink = (float*)scratch;
  unsigned long long* clk = (unsigned long long*)((char*)scratch + (need - 256));
  hipStream_t s = (hipStream_t)stream;
  const int simds = (int)cus * 4;
  double rate0 = 0, ghz0 = 0, rate1 = 0, ghz1 = 0;
  int rc = probe_run(0, blocks, 4096, sink, clk, simds, s, &rate0, &ghz0);
  if (!rc) rc = probe_run(1, blocks, 4096, sink, clk, simds, s, &rate1, &ghz1);
  if (rc) return rc;
  out[0] = ghz0;
  out[1] = rate0;
  out[2] = ghz1;
  out[3] = rate1;
  return DIMCV_OK;
}

}  // extern "C"
