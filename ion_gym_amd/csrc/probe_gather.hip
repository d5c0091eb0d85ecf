// Gather-rate microbenchmark of an r-z FP32 field pack, for the benchmark's
// roofline (measurement only: not part of the engine's C ABI).
//
// Each lane walks one ion's sequence of sample points -- its position in
// node units and a constant per-step displacement, the ion's velocity times
// the flight step -- and at every point reads the four node records of its
// bilinear cell exactly as the flight kernel's FP32 gather does (records of
// S floats, node-major, rows i and i + 1, two adjacent records per row,
// 16-byte loads) and blends the 2 x NS values with the bilinear weights.
// Nothing else is computed, so the launch time is what the memory
// hierarchy needs to deliver these gathers in this order: with the
// engine's 16 MB funnel pack, the L1/L2/MALL gather rate of the flight's
// own access pattern.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

template <int S>
__global__ __launch_bounds__(256) void gather_probe_kernel(
    const float *__restrict__ g, int32_t nx, int32_t nu, const float *__restrict__ gx0,
    const float *__restrict__ gu0, const float *__restrict__ dgx, const float *__restrict__ dgu,
    int64_t n, int32_t n_iter, float *__restrict__ sink) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  float x = gx0[t], u = gu0[t];
  const float dx = dgx[t], du = dgu[t];
  float acc = 0.0f;
  for (int32_t k = 0; k < n_iter; ++k) {
    int i = (int)fminf(fmaxf(x, 0.0f), (float)(nx - 2));
    int j = (int)fminf(fmaxf(u, 0.0f), (float)(nu - 2));
    const float fx = fminf(fmaxf(x - (float)i, 0.0f), 1.0f);
    const float fu = fminf(fmaxf(u - (float)j, 0.0f), 1.0f);
    const float w00 = (1.0f - fx) * (1.0f - fu), w10 = fx * (1.0f - fu);
    const float w01 = (1.0f - fx) * fu, w11 = fx * fu;
    const float4 *p0 = reinterpret_cast<const float4 *>(g + ((int64_t)i * nu + j) * S);
    const float4 *p1 = p0 + (int64_t)nu * (S / 4);
#pragma unroll
    for (int q = 0; q < S / 4; ++q) {
      const float4 a = p0[q], b = p0[S / 4 + q], c = p1[q], d = p1[S / 4 + q];
      acc += w00 * (a.x + a.y + a.z + a.w) + w01 * (b.x + b.y + b.z + b.w) +
             w10 * (c.x + c.y + c.z + c.w) + w11 * (d.x + d.y + d.z + d.w);
    }
    x += dx;
    u += du;
  }
  sink[t] = acc;
}

template <int S>
void launch(const float *g, int32_t nx, int32_t nu, const float *bufs, int64_t n,
            int32_t n_iter, float *sink, hipStream_t st) {
  const unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(gather_probe_kernel<S>, dim3(grid), dim3(256), 0, st, g, nx, nu, bufs,
                     bufs + n, bufs + 2 * n, bufs + 3 * n, n, n_iter, sink);
}

// a plain device copy of known size, the calibration of the HBM traffic
// counters (FETCH_SIZE, WRITE_SIZE) a profiler reads for the flight kernel
__global__ __launch_bounds__(256) void copy_probe_kernel(const float4 *__restrict__ a,
                                                         float4 *__restrict__ b, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
    b[i] = a[i];
}

} // namespace

extern "C" {

// Copy n_floats floats (a multiple of 4) between two device buffers reps
// times; writes the mean time per copy (ms).  0, or a negative code.
int igp_copy(int64_t n_floats, int32_t reps, double *ms_out) {
  if (n_floats < 4 || n_floats % 4 || reps < 1 || !ms_out) return -1;
  float *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  const size_t bytes = (size_t)n_floats * sizeof(float);
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) rc = -2;
  if (!rc && (hipMemset(a, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
              hipEventCreate(&e1) != hipSuccess))
    rc = -3;
  if (!rc) {
    const int64_t n4 = n_floats / 4;
    const unsigned grid = 8192;
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, 0,
                         reinterpret_cast<const float4 *>(a), reinterpret_cast<float4 *>(b), n4);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
      rc = -4;
    } else {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      *ms_out = (double)ms / reps;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}


// Time `reps` launches (after one untimed) of the gather walk of n lanes x
// n_iter points over records (nx * nu * stride floats, stride 4, 8, 12 or
// 16).  pts holds four float arrays of n: start node coordinates (x, u)
// and per-point steps (dx, du).  Writes the mean launch time (ms) and
// returns 0, or a negative code (argument, allocation or launch error).
int igp_gather_rz(const float *records, int32_t nx, int32_t nu, int32_t stride,
                  const float *pts, int64_t n, int32_t n_iter, int32_t reps, double *ms_out) {
  if (!records || !pts || !ms_out || nx < 2 || nu < 2 || n < 1 || n_iter < 1 || reps < 1)
    return -1;
  if (stride != 4 && stride != 8 && stride != 12 && stride != 16) return -1;
  const size_t rec_bytes = (size_t)nx * nu * stride * sizeof(float);
  float *g = nullptr, *b = nullptr, *sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipMalloc(&g, rec_bytes) != hipSuccess || hipMalloc(&b, (size_t)n * 4 * sizeof(float)) !=
      hipSuccess || hipMalloc(&sink, (size_t)n * sizeof(float)) != hipSuccess)
    rc = -2;
  if (!rc && (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
              hipEventCreate(&e1) != hipSuccess))
    rc = -3;
  if (!rc && (hipMemcpy(g, records, rec_bytes, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(b, pts, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice) !=
                  hipSuccess))
    rc = -3;
  if (!rc) {
    auto go = [&]() {
      switch (stride) {
      case 4: launch<4>(g, nx, nu, b, n, n_iter, sink, st); break;
      case 8: launch<8>(g, nx, nu, b, n, n_iter, sink, st); break;
      case 12: launch<12>(g, nx, nu, b, n, n_iter, sink, st); break;
      default: launch<16>(g, nx, nu, b, n, n_iter, sink, st); break;
      }
    };
    go();
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) go();
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipGetLastError() != hipSuccess) {
      rc = -4;
    } else {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      *ms_out = (double)ms / reps;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (g) (void)hipFree(g);
  if (b) (void)hipFree(b);
  if (sink) (void)hipFree(sink);
  return rc;
}

} // extern "C"
