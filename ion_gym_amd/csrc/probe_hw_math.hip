// The hardware FP32 transcendentals the FP32-basis flight kernels use
// (v_sin_f32 and v_cos_f32 on angles in revolutions, v_log_f32), evaluated
// on evenly spaced arguments over the ranges the kernels give them, for a
// test to hold their accuracy against the host's FP64 functions
// (measurement only: not part of the engine's C ABI).
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

// lane i: u = -1/2 + 3/2 i / n in [-1/2, 1) revolutions (the reduced drive
// phases and the random azimuths), v = (i + 1) / n in (0, 1] (the uniforms
// whose logarithms the kernels take)
// (the arguments are written out too, so the host compares at exactly
// the values the device used)
__global__ __launch_bounds__(256) void hw_math_kernel(int32_t n, float *__restrict__ out) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float u = -0.5f + 1.5f * ((float)i / (float)n);
  const float v = (float)(i + 1) / (float)n;
  out[i] = u;
  out[(size_t)n + i] = __builtin_amdgcn_sinf(u);
  out[2 * (size_t)n + i] = __builtin_amdgcn_cosf(u);
  out[3 * (size_t)n + i] = v;
  out[4 * (size_t)n + i] = __builtin_amdgcn_logf(v);
}

} // namespace

extern "C" {

// u, sin(2 pi u), cos(2 pi u), v, log2(v) at the arguments above into the
// caller's host array of 5 n floats (five rows of n); 0, or -1 on a HIP
// error
int igp_hw_math(int32_t n, float *out) {
  if (n < 1 || !out) return -1;
  float *d = nullptr;
  const size_t b = 5 * (size_t)n * sizeof(float);
  if (hipMalloc(&d, b) != hipSuccess) return -1;
  hipLaunchKernelGGL(hw_math_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, d);
  int rc = hipDeviceSynchronize() == hipSuccess ? 0 : -1;
  if (rc == 0) rc = hipMemcpy(out, d, b, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
  (void)hipFree(d);
  return rc;
}

} // extern "C"
