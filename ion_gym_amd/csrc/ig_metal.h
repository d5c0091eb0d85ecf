// The metal tests of the flight kernels, shared with the birth refusals.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ig {

// raster2d._metal_nn (raster2d.py:597-612)
__device__ __forceinline__ bool metal_nn(const uint8_t *__restrict__ m, double gx, double gy,
                                         int64_t nx, int64_t ny) {
  double fi = floor(gx + 0.5), fj = floor(gy + 0.5);
  if (!(fi >= 0.0) || !(fj >= 0.0) || fi > (double)(nx - 1) || fj > (double)(ny - 1)) return false;
  // inside the grid the node indices are whole numbers below 2^31: the
  // 32-bit conversion is exact (and one instruction)
  return m[(int64_t)(int32_t)fi * ny + (int32_t)fj] != 0;
}

// metal_nn split in two so the load can be issued early: probe() loads
// the nearest node's byte with no branch (node 0 when the point is
// outside the grid, so the load stays in bounds); hit() gives metal_nn's
// answer when the value is used
struct NodeProbe {
  uint32_t v;
  bool inside;
  __device__ __forceinline__ bool hit() const { return inside && v != 0; }
};
// (last_i, last_j: the last node indices nx - 1, ny - 1 as doubles)
__device__ __forceinline__ NodeProbe metal_nn_probe(const uint8_t *__restrict__ m, double gx,
                                                    double gy, double last_i, double last_j,
                                                    int64_t ny) {
  const double fi = floor(gx + 0.5), fj = floor(gy + 0.5);
  NodeProbe r;
  r.inside = fi >= 0.0 && fj >= 0.0 && fi <= last_i && fj <= last_j;
  const int32_t ci = (int32_t)(r.inside ? fi : 0.0), cj = (int32_t)(r.inside ? fj : 0.0);
  r.v = m[(int64_t)ci * ny + cj];
  return r;
}

// the hardware's FP64 -> int32 conversion, v_cvt_i32_f64: truncation
// toward zero, saturating at the int32 range (and 0 for NaN), one
// instruction (C's conversion is undefined out of range, so the compiler
// would need a clamp first)
__device__ __forceinline__ int32_t cvt_i32_sat(double v) {
  int32_t r;
  asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

// metal_nn_probe in 32-bit integer arithmetic (the FP32-basis kernels):
// the nearest node's indices by the saturating conversion, inside the grid
// iff both are at most the last index as unsigned numbers (a negative or
// saturated index is not), and a 32-bit node offset (2-D packs hold fewer
// than 2^31 nodes, checked at upload).  The same answer as metal_nn_probe
// for every finite point.
__device__ __forceinline__ NodeProbe metal_nn_probe_u32(const uint8_t *__restrict__ m, double gx,
                                                        double gy, uint32_t last_i,
                                                        uint32_t last_j, uint32_t ny) {
  const uint32_t ci = (uint32_t)cvt_i32_sat(floor(gx + 0.5));
  const uint32_t cj = (uint32_t)cvt_i32_sat(floor(gy + 0.5));
  NodeProbe r;
  r.inside = ci <= last_i && cj <= last_j;
  r.v = m[r.inside ? ci * ny + cj : 0u];
  return r;
}

// tracer3d._in_metal_3d (tracer3d.py:174-189)
__device__ __forceinline__ bool in_metal_3d(const uint8_t *__restrict__ m, double x, double y,
                                            double z, int64_t nx, int64_t ny, int64_t nz) {
  if (!(x >= 0.0) || !(y >= 0.0) || !(z >= 0.0) || x > nx - 1 || y > ny - 1 || z > nz - 1)
    return false;
  // inside the box each coordinate is below 2^31: the 32-bit truncation
  // is exact (one instruction; an int64 conversion is a long sequence)
  int64_t i = cvt_i32_sat(x), j = cvt_i32_sat(y), k = cvt_i32_sat(z);
  if (i > nx - 2) i = nx - 2;
  if (j > ny - 2) j = ny - 2;
  if (k > nz - 2) k = nz - 2;
  const uint8_t *a = m + (i * ny + j) * nz + k;
  const int64_t di = ny * nz, dj = nz;
  return a[0] && a[di] && a[dj] && a[di + dj] && a[1] && a[di + 1] && a[dj + 1] && a[di + dj + 1];
}

} // namespace ig
