// valu_probe.hip -- FP32 VALU issue-rate probes for gfx950 (measurement only; built
// into libcfm_probe.so, not into the mobility library).
//
// Every device build disables packed FP32 instructions, so the ceiling for the FP32
// pair kernel is the rate at which the GPU issues ordinary FP32 VALU instructions.
// These kernels measure it for the instruction forms the pair kernel is made of:
// independent chains per lane (no latency bound), eight or more waves per SIMD, one
// store per thread.  The result is lane-instructions per second over the whole GPU.
#include <hip/hip_runtime.h>

#include "../../include/cfm_probe.h"

namespace cfm_probe {
namespace {

constexpr int kProbeBlocks = 256 * 32, kProbeThreads = 256, kProbeIters = 4096;

// v_fmac_f32 (two-operand form, VGPR sources): 16 accumulators, 4 multipliers.
__global__ void probe_fmac_kernel(float* out, int iters, float a) {
  float acc[16], x[4];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = threadIdx.x * 1e-3f + k;
#pragma unroll
  for (int k = 0; k < 4; ++k) x[k] = threadIdx.x * 2e-3f + k + a;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = __builtin_fmaf(x[k & 3], x[(k + 1) & 3], acc[k]);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 16; ++k) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s + x[0] + x[1] + x[2] + x[3];
}

// v_mul_f32 with a scalar operand.
__global__ void probe_mul_kernel(float* out, int iters, float a) {
  float acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = threadIdx.x * 1e-3f + k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = acc[k] * a;
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 16; ++k) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// v_add_f32 with a scalar operand.
__global__ void probe_add_kernel(float* out, int iters, float a) {
  float acc[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = threadIdx.x * 1e-3f + k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = acc[k] + a;
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 16; ++k) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// v_mul_f32 with VGPR operands only.
__global__ void probe_mulvv_kernel(float* out, int iters, float a) {
  float acc[16], x[4];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = threadIdx.x * 1e-3f + k;
#pragma unroll
  for (int k = 0; k < 4; ++k) x[k] = threadIdx.x * 2e-3f + k + a;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = acc[k] * x[k & 3];
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 16; ++k) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s + x[0] + x[1] + x[2] + x[3];
}

// v_fma_f32 (three-operand form) with VGPR operands only.
__global__ void probe_fma3_kernel(float* out, int iters, float a) {
  float acc[16], x[4];
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = threadIdx.x * 1e-3f + k;
#pragma unroll
  for (int k = 0; k < 4; ++k) x[k] = threadIdx.x * 2e-3f + k + a;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int k = 0; k < 16; ++k) acc[k] = __builtin_fmaf(acc[k], x[k & 3], x[(k + 1) & 3]);
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 16; ++k) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s + x[0] + x[1] + x[2] + x[3];
}

}  // namespace
}  // namespace cfm_probe

extern "C" int cfm_probe_valu_rate(int32_t device, int32_t form, double* lane_instr_per_s) {
  using namespace cfm_probe;
  if (!lane_instr_per_s || form < 0 || form > 4) return CFM_PROBE_EINVAL;
  *lane_instr_per_s = 0.0;
  if (hipSetDevice(device) != hipSuccess) return CFM_PROBE_EHIP;
  float* out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CFM_PROBE_EHIP;
  const size_t n = (size_t)kProbeBlocks * kProbeThreads;
  const double per_iter = 16.0;  // VALU instructions per loop iteration per lane
  if (hipMalloc(&out, sizeof(float) * n) == hipSuccess &&
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
      hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
    float best = 0.0f;
    bool ok = true;
    for (int rep = 0; rep < 3 && ok; ++rep) {  // first pass warms up; keep the fastest
      ok = hipEventRecord(e0, st) == hipSuccess;
      switch (form) {
        case 0: hipLaunchKernelGGL(probe_fmac_kernel, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, out, kProbeIters, 0.999f); break;
        case 1: hipLaunchKernelGGL(probe_mul_kernel, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, out, kProbeIters, 0.999f); break;
        case 3: hipLaunchKernelGGL(probe_mulvv_kernel, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, out, kProbeIters, 0.999f); break;
        case 4: hipLaunchKernelGGL(probe_fma3_kernel, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, out, kProbeIters, 0.999f); break;
        default: hipLaunchKernelGGL(probe_add_kernel, dim3(kProbeBlocks), dim3(kProbeThreads), 0, st, out, kProbeIters, 1e-3f); break;
      }
      ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, st) == hipSuccess &&
           hipEventSynchronize(e1) == hipSuccess;
      float ms = 0.0f;
      ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
      if (ok && rep > 0 && (best == 0.0f || ms < best)) best = ms;
    }
    if (ok && best > 0.0f) {
      *lane_instr_per_s = (double)n * kProbeIters * per_iter / (best * 1e-3);
      rc = CFM_PROBE_OK;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (out) (void)hipFree(out);
  return rc;
}
