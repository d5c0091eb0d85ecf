// Field composition on the device: phi = sum_t w_t * (phi_{b_t} / V_base)
// and E = -grad(phi) written into one set of a resident field pack.
//
// The gradient is numpy's np.gradient (edge_order 1) as the reference's
// field builders call it -- central differences in the interior (uniform
// spacing, or numpy's three-point coefficients on non-uniform coordinate
// arrays), one-sided at the array ends -- followed by the electrode-aware
// rule: at a metal node with vacuum on exactly one side along an axis, that
// axis derivative becomes the one-sided difference into the vacuum
// (ion_gym/physics/ionbench.py:44-96 for r-z on the mirror-unfolded grid,
// V/mm -> V/m; ion_gym/physics/tracer3d.py:54-117 for 3-D, V/mm).  Every
// quantity is the same expression in the same order as there, with
// contraction off, so FP64 results equal the host build bit for bit.
#pragma once
#pragma clang fp contract(off)
#include "ig_common.h"

namespace ig {

// one axis of np.gradient on n nodes
struct GradAxis {
  int64_t n;
  int32_t uniform;     // numpy found every coordinate step equal
  double dx;           // that step
  const double *abc;   // otherwise (n - 2, 3) interior coefficients (device)
  double dx_lo, dx_hi; // end steps
  double h_one;        // step of the one-sided metal-node difference
};

__device__ __forceinline__ double grad_axis(const GradAxis &g, int64_t i, double fm, double f0,
                                            double fp) {
  if (i == 0) return (fp - f0) / g.dx_lo;
  if (i == g.n - 1) return (f0 - fm) / g.dx_hi;
  if (g.uniform) return (fp - fm) / (2.0 * g.dx);
  const double *c = g.abc + 3 * (i - 1);
  return c[0] * fm + c[1] * f0 + c[2] * fp;
}

// electrode-aware override of one axis derivative at a metal node
__device__ __forceinline__ double one_sided(double d, bool metal, bool has_m, bool vac_m,
                                            bool has_p, bool vac_p, double fm, double f0,
                                            double fp, double h) {
  if (!metal) return d;
  const bool vp = has_p && vac_p, vm = has_m && vac_m;
  if (vp && !vm) return (fp - f0) / h;
  if (vm && !vp) return (f0 - fm) / h;
  return d;
}

__global__ void compose_kernel(const double *__restrict__ phi, int64_t nodes, int32_t n_terms,
                               const int32_t *__restrict__ bi, const double *__restrict__ w,
                               double base_volts, double *__restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nodes) return;
  double a = 0.0;
  for (int32_t t = 0; t < n_terms; ++t) {
    const double fa = phi[(int64_t)bi[t] * nodes + n] / base_volts;
    a = a + w[t] * fa;
  }
  out[n] = a;
}

template <typename T>
__device__ __forceinline__ void store_comp(void *pack, int64_t idx, double v, double inv_scale) {
  if constexpr (sizeof(T) == 8)
    reinterpret_cast<double *>(pack)[idx] = v;
  else
    reinterpret_cast<float *>(pack)[idx] = (float)(v * inv_scale);
}

// r-z: P and M on the half grid (nz, nr); the pack on the unfolded grid
// (nz, nu = 2 nr - 1), column j holding half-grid column |j - (nr - 1)|
template <typename T>
__global__ void grad_rz_kernel(const double *__restrict__ P, const uint8_t *__restrict__ M,
                               int64_t nz, int64_t nr, GradAxis gz, GradAxis gu, void *pack,
                               int64_t stride, int32_t set, double inv_scale) {
  const int64_t nu = 2 * nr - 1;
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nz * nu) return;
  const int64_t i = node / nu, j = node - i * nu;
  auto col = [&](int64_t jj) { return jj >= nr - 1 ? jj - (nr - 1) : (nr - 1) - jj; };
  const int64_t c0 = col(j);
  const double f0 = P[i * nr + c0];
  const bool metal = M[i * nr + c0] != 0;
  // z axis
  const bool hzm = i > 0, hzp = i + 1 < nz;
  const double zm = hzm ? P[(i - 1) * nr + c0] : 0.0, zp = hzp ? P[(i + 1) * nr + c0] : 0.0;
  double dz = grad_axis(gz, i, zm, f0, zp);
  dz = one_sided(dz, metal, hzm, hzm && M[(i - 1) * nr + c0] == 0, hzp,
                 hzp && M[(i + 1) * nr + c0] == 0, zm, f0, zp, gz.h_one);
  // u axis
  const bool hum = j > 0, hup = j + 1 < nu;
  const int64_t cm = hum ? col(j - 1) : 0, cp = hup ? col(j + 1) : 0;
  const double um = hum ? P[i * nr + cm] : 0.0, up = hup ? P[i * nr + cp] : 0.0;
  double du = grad_axis(gu, j, um, f0, up);
  du = one_sided(du, metal, hum, hum && M[i * nr + cm] == 0, hup, hup && M[i * nr + cp] == 0,
                 um, f0, up, gu.h_one);
  const double ez = -dz * 1e3, eu = -du * 1e3;
  store_comp<T>(pack, node * stride + 2 * set, ez, inv_scale);
  store_comp<T>(pack, node * stride + 2 * set + 1, eu, inv_scale);
}

template <typename T>
__global__ void grad_3d_kernel(const double *__restrict__ P, const uint8_t *__restrict__ M,
                               int64_t nx, int64_t ny, int64_t nz, GradAxis g, void *pack,
                               int64_t stride, int32_t set, double inv_scale) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= nx * ny * nz) return;
  const int64_t i = node / (ny * nz), r = node - i * ny * nz, j = r / nz, k = r - j * nz;
  const double f0 = P[node];
  const bool metal = M[node] != 0;
  const int64_t st[3] = {ny * nz, nz, 1};
  const int64_t ix[3] = {i, j, k}, nn[3] = {nx, ny, nz};
  double e[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    GradAxis ga = g;
    ga.n = nn[a];
    const bool hm = ix[a] > 0, hp = ix[a] + 1 < nn[a];
    const double fm = hm ? P[node - st[a]] : 0.0, fp = hp ? P[node + st[a]] : 0.0;
    double d = grad_axis(ga, ix[a], fm, f0, fp);
    d = one_sided(d, metal, hm, hm && M[node - st[a]] == 0, hp, hp && M[node + st[a]] == 0, fm,
                  f0, fp, ga.h_one);
    e[a] = -d;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) store_comp<T>(pack, node * stride + 3 * set + a, e[a], inv_scale);
}

// one component of one set of a pack, as doubles x the set's storage scale
template <typename T>
__global__ void read_comp_kernel(const void *pack, int64_t nodes, int64_t stride, int32_t q,
                                 double scale, double *__restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nodes) return;
  const T v = reinterpret_cast<const T *>(pack)[n * stride + q];
  out[n] = sizeof(T) == 8 ? (double)v : (double)v * scale;
}

} // namespace ig
