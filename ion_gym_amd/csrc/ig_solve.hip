// 3-D Laplace basis solver on the device: the red-black SOR of
// ion_gym/physics/solver3d.py:131-293 and the V-cycle multigrid of
// ion_gym/physics/multigrid3d.py:305-731, restated as stencil kernels over
// FP64 potentials in HBM.
//
// Every node update is the reference's expression in IEEE double
// (contraction off; float32 where it multiplies float32 leg fractions), a
// colour's nodes read only the other colour, and every reduction is a
// maximum, so the result does not depend on the launch shape and is the
// CPU restatement's bit for bit.  Reductions of |update| take the bit
// pattern of a non-negative double as a 64-bit integer (the orders agree):
// each workgroup writes its maximum to its own word of a partials array
// and a one-workgroup kernel reduces that array (no two lanes of the grid
// contend for one address).
//
// Iteration control stays on the device where it would otherwise cost a
// host round trip per sweep pair: SOR pairs are launched in chunks and a
// one-lane kernel after each pair records convergence in a flag that the
// next pair's kernels read (they return at once once it is set); a small
// coarsest error grid is relaxed to its tolerance inside one workgroup.
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ion_gym_amd.h"
#include "ig_solve.h"

namespace ig {
namespace {

constexpr int kBlock = 256;
constexpr int kSmallNodes = 1 << 16;  // coarsest grids relaxed in one workgroup
constexpr int kSorChunk = 64;         // SOR pairs per host check
constexpr double kErrRelTol = 0.05;   // multigrid3d.py:679 (_solve_err rel_tol)
constexpr int kErrCycles = 30;        // multigrid3d.py:553 (max_cycles)
constexpr int64_t kMinCoarse = 9;     // multigrid3d.py:554 (min_coarse)
constexpr int kBottomPairs = 20000;   // multigrid3d.py:572

struct Grid {
  int64_t nx, ny, nz;
  __host__ __device__ int64_t n() const { return nx * ny * nz; }
};

__device__ __forceinline__ int64_t ix(const Grid &g, int64_t i, int64_t j, int64_t k) {
  return (i * g.ny + j) * g.nz + k;
}

// neighbours with even reflection at every outer face and the node itself
// along a size-1 axis (solver3d.py:139-171)
__device__ __forceinline__ void nbrs(const double *__restrict__ p, const Grid &g, int64_t i,
                                     int64_t j, int64_t k, double (&v)[6]) {
  const double c = p[ix(g, i, j, k)];
  v[0] = i > 0 ? p[ix(g, i - 1, j, k)] : (g.nx > 1 ? p[ix(g, 1, j, k)] : c);
  v[1] = i < g.nx - 1 ? p[ix(g, i + 1, j, k)] : (g.nx > 1 ? p[ix(g, g.nx - 2, j, k)] : c);
  v[2] = j > 0 ? p[ix(g, i, j - 1, k)] : (g.ny > 1 ? p[ix(g, i, 1, k)] : c);
  v[3] = j < g.ny - 1 ? p[ix(g, i, j + 1, k)] : (g.ny > 1 ? p[ix(g, i, g.ny - 2, k)] : c);
  v[4] = k > 0 ? p[ix(g, i, j, k - 1)] : (g.nz > 1 ? p[ix(g, i, j, 1)] : c);
  v[5] = k < g.nz - 1 ? p[ix(g, i, j, k + 1)] : (g.nz > 1 ? p[ix(g, i, j, g.nz - 2)] : c);
}

// the stencil's numerator and denominator (solver3d.py:172-190)
__device__ __forceinline__ void stencil(const double (&v)[6], const float *__restrict__ th,
                                        int64_t n, int64_t q, int gl, double &num,
                                        double &den) {
  float t[6];
#pragma unroll
  for (int d = 0; d < 6; ++d) t[d] = th ? th[d * n + q] : 1.0f;
  if (gl) {
    const double wxm = 1.0 / (double)t[0], wxp = 1.0 / (double)t[1];
    const double wym = 1.0 / (double)t[2], wyp = 1.0 / (double)t[3];
    const double wzm = 1.0 / (double)t[4], wzp = 1.0 / (double)t[5];
    num = v[0] * wxm + v[1] * wxp + v[2] * wym + v[3] * wyp + v[4] * wzm + v[5] * wzp;
    den = wxm + wxp + wym + wyp + wzm + wzp;
  } else {
    const float sx = t[0] + t[1], sy = t[2] + t[3], sz = t[4] + t[5];
    const float ax = t[0] * sx, bx = t[1] * sx, ay = t[2] * sy, by = t[3] * sy;
    const float az = t[4] * sz, bz = t[5] * sz;
    num = v[0] / (double)ax + v[1] / (double)bx + v[2] / (double)ay + v[3] / (double)by +
          v[4] / (double)az + v[5] / (double)bz;
    const float px = t[0] * t[1], py = t[2] * t[3], pz = t[4] * t[5];
    den = 1.0 / (double)px + 1.0 / (double)py + 1.0 / (double)pz;
  }
}

// Launch shape of the grid kernels: one launch row (blockIdx.y) per
// x-plane i, the plane's (j, k) nodes (or, for a colour sweep, its (j, m)
// half-lattice slots) flattened along x, so the only index division is one
// 32-bit division per lane.

// the node of colour c of half-lattice slot t of a row list: rows (i, j),
// k of the row's parity stepping by 2; false past the last row or the
// row's end (a small grid: slots below 2^32)
__device__ __forceinline__ bool colour_node_flat(const Grid &g, int c, uint32_t t, int64_t &i,
                                                 int64_t &j, int64_t &k) {
  const uint32_t half = (uint32_t)((g.nz + 1) / 2), ny = (uint32_t)g.ny;
  const uint32_t row = t / half, m = t - row * half;
  const uint32_t ii = row / ny;
  i = ii;
  j = row - ii * ny;
  k = ((i + j + c) & 1) + 2 * (int64_t)m;
  return i < g.nx && k < g.nz;
}

// XCD-aware tile of a per-plane launch: the dispatcher deals workgroups
// to the 8 XCDs in turn, so workgroup b is renumbered to run r of XCD
// b % 8, each XCD taking a contiguous run of planes whose neighbours are
// then in its own L2
__device__ __forceinline__ void xcd_tile(uint32_t &bx, uint32_t &by) {
  const uint32_t nb = gridDim.x * gridDim.y;
  const uint32_t b = blockIdx.y * gridDim.x + blockIdx.x;
  const uint32_t q = nb / 8, r = nb % 8, x = b % 8, y = b / 8;
  const uint32_t L = x * q + (x < r ? x : r) + y;
  by = L / gridDim.x;
  bx = L - by * gridDim.x;
}

// the same in a per-plane launch: plane i = the tile's row
__device__ __forceinline__ bool colour_node(const Grid &g, int c, int64_t &i, int64_t &j,
                                            int64_t &k) {
  uint32_t bx, by;
  xcd_tile(bx, by);
  const uint32_t half = (uint32_t)((g.nz + 1) / 2);
  const uint32_t t = bx * blockDim.x + threadIdx.x;
  const uint32_t jj = t / half, m = t - jj * half;
  i = by;
  j = jj;
  k = ((i + j + c) & 1) + 2 * (int64_t)m;
  return j < g.ny && k < g.nz;
}

// node (i, j, k) of a per-plane full-grid launch; false past the plane
__device__ __forceinline__ bool plane_node(const Grid &g, int64_t &i, int64_t &j, int64_t &k) {
  const uint32_t nz = (uint32_t)g.nz;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t jj = t / nz;
  i = blockIdx.y;
  j = jj;
  k = t - jj * nz;
  return j < g.ny;
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long b) {
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long u = __shfl_xor(b, o);
    b = u > b ? u : b;
  }
  return b;
}

// the workgroup's maximum of d (every lane calls) to part[its linear
// workgroup index]
__device__ __forceinline__ void block_max(double d, unsigned long long *part) {
  __shared__ unsigned long long red[kBlock / 64];
  const unsigned long long b = wave_max((unsigned long long)__double_as_longlong(d));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
    part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = m;
  }
}

// the maximum of part[0, count) by one 1024-lane workgroup (every lane
// gets it)
__device__ __forceinline__ unsigned long long reduce_parts(const unsigned long long *part,
                                                           int64_t count) {
  __shared__ unsigned long long red[16];
  unsigned long long m = 0;
  for (int64_t q = threadIdx.x; q < count; q += blockDim.x) m = part[q] > m ? part[q] : m;
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  unsigned long long t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t = red[w] > t ? red[w] : t;
  return t;
}

// the first stage of a large pair maximum: workgroup b reduces its
// contiguous share of part[0, count) to out[b] (nothing once done is set)
__global__ __launch_bounds__(1024) void chunk_max_kernel(const unsigned long long *part,
                                                         int64_t count, const int32_t *done,
                                                         unsigned long long *out) {
  if (*done) return;
  const int64_t chunk = (count + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t len = lo < count ? (count - lo < chunk ? count - lo : chunk) : 0;
  const unsigned long long m = reduce_parts(part + lo, len);
  if (threadIdx.x == 0) out[blockIdx.x] = m;
}

__global__ __launch_bounds__(1024) void reduce_kernel(const unsigned long long *part,
                                                      int64_t count, unsigned long long *out) {
  const unsigned long long m = reduce_parts(part, count);
  if (threadIdx.x == 0) *out = m;
}

// an error level's tolerances from the maxima of |f| (multigrid3d.py:567-
// 569, 577): tolv[0] = rel max(max|f|, 1e-300) / 6, tolv[1] = 0.02 tolv[0]
__global__ __launch_bounds__(1024) void err_tol_kernel(const unsigned long long *part,
                                                       int64_t count, double rel,
                                                       double *tolv) {
  double scale = __longlong_as_double((long long)reduce_parts(part, count));
  if (threadIdx.x != 0) return;
  if (scale < 1e-300) scale = 1e-300;
  const double tol = rel * scale / 6.0;
  tolv[0] = tol;
  tolv[1] = 0.02 * tol;
}

// a cycle's decision data: out[0] = the maximum of part, out[1] = tol
__global__ __launch_bounds__(1024) void check_kernel(const unsigned long long *part,
                                                     int64_t count, const double *tolv,
                                                     double *out) {
  const unsigned long long m = reduce_parts(part, count);
  if (threadIdx.x != 0) return;
  out[0] = __longlong_as_double((long long)m);
  out[1] = tolv ? tolv[0] : 0.0;
}

// one colour sweep of the validated stencil (solver3d._sweep); done, when
// given, ends the launch at once (a converged SOR); snap, when given,
// receives each updated node's value before the update (a cycle's first
// pair fills the V-cycle's snapshot of the free nodes this way)
__global__ __launch_bounds__(kBlock) void sweep_kernel(double *__restrict__ phi,
                                                       const uint8_t *__restrict__ fixed,
                                                       const float *__restrict__ th,
                                                       const Grid g, double omega, int colour,
                                                       int gl, const int32_t *done,
                                                       unsigned long long *part,
                                                       double *__restrict__ snap) {
  if (done && *done) return;
  int64_t i, j, k;
  double d = 0.0;
  if (colour_node(g, colour, i, j, k)) {
    const int64_t q = ix(g, i, j, k);
    if (!fixed[q]) {
      double v[6], num, den;
      nbrs(phi, g, i, j, k, v);
      stencil(v, th, g.n(), q, gl, num, den);
      const double old = phi[q];
      const double nw = old + omega * (num / den - old);
      d = fabs(nw - old);
      phi[q] = nw;
      if (snap) snap[q] = old;
    }
  }
  if (part) block_max(d, part);
}

// the 2-D five-point stencil of solver2d.assemble_laplace
// (solver2d.py:31-124): u the radius when cyl (centre -6 and the first
// off-axis node 4 on the axis, 1 +- h/2r off it); at an open edge the
// missing neighbour's coefficient folds onto the opposite neighbour, or
// onto the centre for the edges in ghost (bits: 1 z0, 2 z1, 4 u0, 8 u1)
struct Lap2 {
  double h;
  int cyl;
  int ghost;
};

// one neighbour of the 2-D stencil (solver2d.py:97-117): its coefficient
// on the neighbour when it exists, else folded onto the centre (gl) or
// the opposite neighbour mq
__device__ __forceinline__ void link2(const double *__restrict__ phi, double coef, bool exists,
                                      int64_t tq, int64_t mq, int gl, double &acc,
                                      double &diag) {
  if (coef == 0.0) return;
  if (exists)
    acc += coef * phi[tq];
  else if (gl)
    diag += coef;
  else
    acc += coef * phi[mq];
}

// one colour sweep of SOR on that system, the grid held as (nz, 1, nu);
// the caller has refused free nodes on closed edges
__global__ __launch_bounds__(kBlock) void sweep2d_kernel(double *__restrict__ phi,
                                                         const uint8_t *__restrict__ fixed,
                                                         const Grid g, const Lap2 s, double omega,
                                                         int colour, const int32_t *done,
                                                         unsigned long long *part) {
  if (done && *done) return;
  int64_t i, j0, j;
  double d = 0.0;
  if (colour_node(g, colour, i, j0, j)) {
    const int64_t nz = g.nx, nu = g.nz;
    const int64_t q = i * nu + j;
    if (!fixed[q]) {
      double cup = 1.0, cdn = 1.0, centre = -4.0;
      if (s.cyl) {
        if (j == 0) {
          cup = 4.0;
          cdn = 0.0;
          centre = -6.0;
        } else {
          const double r = (double)j * s.h;
          cup = 1.0 + s.h / (2.0 * r);
          cdn = 1.0 - s.h / (2.0 * r);
        }
      }
      // (the kernel argument s is read into registers: a reference to it
      // would put a copy in scratch memory)
      const int ghost = s.ghost;
      double acc = 0.0, diag = centre;
      link2(phi, 1.0, i > 0, q - nu, q + nu, ghost & 1, acc, diag);
      link2(phi, 1.0, i < nz - 1, q + nu, q - nu, ghost & 2, acc, diag);
      link2(phi, cdn, j > 0, q - 1, q + 1, ghost & 4, acc, diag);
      link2(phi, cup, j < nu - 1, q + 1, q - 1, ghost & 8, acc, diag);
      const double old = phi[q];
      const double nw = old + omega * (-acc / diag - old);
      d = fabs(nw - old);
      phi[q] = nw;
    }
  }
  if (part) block_max(d, part);
}

// after pair s (the partials of both colour sweeps): convergence when its
// max |update| < tol, or < *tol_p when given (solver3d.py:281-287); last
// and pairs, when given, record the pair
__global__ __launch_bounds__(1024) void pair_finish_kernel(const unsigned long long *part,
                                                           int64_t count, int32_t *done,
                                                           int64_t s, double tol,
                                                           const double *tol_p, double *last,
                                                           int64_t *pairs) {
  if (*done) return;
  const double d = __longlong_as_double((long long)reduce_parts(part, count));
  if (threadIdx.x != 0) return;
  if (tol_p) tol = *tol_p;
  if (last) *last = d;
  if (pairs) *pairs = s + 1;
  if (d < tol) *done = 1;
}

__global__ void residual_kernel(const double *__restrict__ phi, const uint8_t *__restrict__ fixed,
                                const float *__restrict__ th, const Grid g, int gl,
                                double *__restrict__ out) {
  int64_t i, j, k;
  if (!plane_node(g, i, j, k)) return;
  const int64_t q = ix(g, i, j, k);
  if (fixed[q]) {
    out[q] = 0.0;
    return;
  }
  double v[6], num, den;
  nbrs(phi, g, i, j, k, v);
  stencil(v, th, g.n(), q, gl, num, den);
  out[q] = num - den * phi[q];
}

// the plain 7-point error equation sum(nbr) - 6 e = f (multigrid3d.py:357-391)
__device__ __forceinline__ double rhs_update(const double *__restrict__ e,
                                             const double *__restrict__ f, const Grid &g,
                                             int64_t i, int64_t j, int64_t k, int64_t q,
                                             double omega) {
  double v[6];
  nbrs(e, g, i, j, k, v);
  const double nv = (v[0] + v[1] + v[2] + v[3] + v[4] + v[5] - f[q]) / 6.0;
  const double old = e[q];
  return old + omega * (nv - old);
}

__global__ __launch_bounds__(kBlock) void sweep_rhs_kernel(double *__restrict__ e,
                                                           const uint8_t *__restrict__ fixed,
                                                           const double *__restrict__ f,
                                                           const Grid g, double omega,
                                                           int colour, const int32_t *done,
                                                           unsigned long long *part) {
  if (done && *done) return;
  int64_t i, j, k;
  double d = 0.0;
  if (colour_node(g, colour, i, j, k)) {
    const int64_t q = ix(g, i, j, k);
    if (!fixed[q]) {
      const double old = e[q];
      const double nw = rhs_update(e, f, g, i, j, k, q, omega);
      d = fabs(nw - old);
      e[q] = nw;
    }
  }
  if (part) block_max(d, part);
}

// a whole bottom relaxation (multigrid3d.py:573-578) in one workgroup: up
// to max_pairs pairs of colour sweeps, stopping after the first pair whose
// max |update| < thr; the workgroup barrier orders the colours.  With LDS
// the level's e and f are staged in dynamic LDS (16 bytes a node) for the
// whole relaxation and e written back at the end.
template <bool LDS>
__global__ __launch_bounds__(1024) void bottom_kernel(double *__restrict__ e_g,
                                                      const uint8_t *__restrict__ fixed,
                                                      const double *__restrict__ f_g,
                                                      const Grid g, double omega,
                                                      const double *thr_p, int max_pairs) {
  extern __shared__ double bsh[];
  const double thr = *thr_p;
  __shared__ unsigned long long red[16];
  __shared__ int stop;
  const int64_t n = g.n();
  double *e = e_g;
  const double *f = f_g;
  if constexpr (LDS) {
    for (int64_t q = threadIdx.x; q < n; q += blockDim.x) {
      bsh[q] = e_g[q];
      bsh[n + q] = f_g[q];
    }
    __syncthreads();
    e = bsh;
    f = bsh + n;
  }
  const int64_t half_total = g.nx * g.ny * ((g.nz + 1) / 2);
  for (int it = 0; it < max_pairs; ++it) {
    double dm = 0.0;
    for (int c = 0; c < 2; ++c) {
      for (uint32_t t = threadIdx.x; t < (uint32_t)half_total; t += blockDim.x) {
        int64_t i, j, k;
        if (!colour_node_flat(g, c, t, i, j, k)) continue;
        const int64_t q = ix(g, i, j, k);
        if (fixed[q]) continue;
        const double old = e[q];
        const double nw = rhs_update(e, f, g, i, j, k, q, omega);
        const double d = fabs(nw - old);
        dm = d > dm ? d : dm;
        e[q] = nw;
      }
      __syncthreads();
    }
    unsigned long long b = (unsigned long long)__double_as_longlong(dm);
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long u = __shfl_xor(b, o);
      b = u > b ? u : b;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long m = 0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = red[w] > m ? red[w] : m;
      stop = __longlong_as_double((long long)m) < thr;
    }
    __syncthreads();
    if (stop) break;
  }
  if constexpr (LDS) {
    for (int64_t q = threadIdx.x; q < n; q += blockDim.x) e_g[q] = bsh[q];
  }
}

__global__ void residual_rhs_kernel(const double *__restrict__ e,
                                    const uint8_t *__restrict__ fixed,
                                    const double *__restrict__ f, const Grid g,
                                    double *__restrict__ out) {
  int64_t i, j, k;
  if (!plane_node(g, i, j, k)) return;
  const int64_t q = ix(g, i, j, k);
  if (fixed[q]) {
    out[q] = 0.0;
    return;
  }
  double v[6];
  nbrs(e, g, i, j, k, v);
  out[q] = (v[0] + v[1] + v[2] + v[3] + v[4] + v[5] - 6.0 * e[q]) - f[q];
}

// full weighting (multigrid3d.py:420-455) followed by the caller's
// rc *= 4 and negation (multigrid3d.py:586-589, 676-679): -(4 R r) at
// coarse node (ic, jc, kc)
__device__ __forceinline__ double restrict_at(const double *__restrict__ fine, const Grid &g,
                                              int64_t ic, int64_t jc, int64_t kc) {
  const int64_t i = 2 * ic, j = 2 * jc, k = 2 * kc;
  double acc = 0.0;
  for (int di = -1; di <= 1; ++di) {
    int64_t ii = i + di;
    if (ii < 0) ii = 0;
    if (ii > g.nx - 1) ii = g.nx - 1;
    const double wi = di == 0 ? 1.0 : 0.5;
    for (int dj = -1; dj <= 1; ++dj) {
      int64_t jj = j + dj;
      if (jj < 0) jj = 0;
      if (jj > g.ny - 1) jj = g.ny - 1;
      const double wj = dj == 0 ? 1.0 : 0.5;
      for (int dk = -1; dk <= 1; ++dk) {
        int64_t kk = k + dk;
        if (kk < 0) kk = 0;
        if (kk > g.nz - 1) kk = g.nz - 1;
        const double wk = dk == 0 ? 1.0 : 0.5;
        acc += wi * wj * wk * fine[ix(g, ii, jj, kk)];
      }
    }
  }
  return -((acc / 8.0) * 4.0);
}

__global__ void restrict_kernel(const double *__restrict__ fine, const Grid g,
                                double *__restrict__ out, const Grid c) {
  int64_t ic, jc, kc;
  if (!plane_node(c, ic, jc, kc)) return;
  out[ix(c, ic, jc, kc)] = restrict_at(fine, g, ic, jc, kc);
}

__device__ __forceinline__ void pro_axis(int64_t i, int64_t c, int64_t &a0, int64_t &a1,
                                         double &t) {
  double xf = i * 0.5;
  if (xf > c - 1) xf = c - 1.0;
  int64_t x0 = (int64_t)xf;
  if (x0 > c - 2) x0 = c > 1 ? c - 2 : 0;
  a0 = x0;
  a1 = c > 1 ? x0 + 1 : x0;
  t = xf - x0;
}

// one axis of the cascade's prolongation (multigrid3d.py:144-156): fine
// index i sits at coarse coordinate min(i / 2, c - 1), its upper
// neighbour clamped to the last coarse node
__device__ __forceinline__ void cas_axis(int64_t i, int64_t c, int64_t &a0, int64_t &a1,
                                         double &t) {
  double xf = i / 2.0;
  if (xf > (double)(c - 1)) xf = (double)(c - 1);
  const int64_t x0 = (int64_t)floor(xf);
  a0 = x0;
  a1 = x0 + 1 < c - 1 ? x0 + 1 : c - 1;
  t = xf - (double)x0;
}

// the trilinear prolongation of co at fine node (i, j, k): the V-cycle's
// (multigrid3d.py:458-499) or, with CAS, the cascade's
// (multigrid3d.py:138-172); the same interpolant, x first, then y, then z
template <bool CAS = false>
__device__ __forceinline__ double prolong_at(const double *__restrict__ co, const Grid &c,
                                             int64_t i, int64_t j, int64_t k) {
  int64_t x0, x1, y0, y1, z0, z1;
  double tx, ty, tz;
  if (CAS) {
    cas_axis(i, c.nx, x0, x1, tx);
    cas_axis(j, c.ny, y0, y1, ty);
    cas_axis(k, c.nz, z0, z1, tz);
  } else {
    pro_axis(i, c.nx, x0, x1, tx);
    pro_axis(j, c.ny, y0, y1, ty);
    pro_axis(k, c.nz, z0, z1, tz);
  }
  auto C = [&](int64_t a, int64_t b, int64_t d) { return co[(a * c.ny + b) * c.nz + d]; };
  const double c00 = C(x0, y0, z0) * (1 - tx) + C(x1, y0, z0) * tx;
  const double c10 = C(x0, y1, z0) * (1 - tx) + C(x1, y1, z0) * tx;
  const double c01 = C(x0, y0, z1) * (1 - tx) + C(x1, y0, z1) * tx;
  const double c11 = C(x0, y1, z1) * (1 - tx) + C(x1, y1, z1) * tx;
  const double c0 = c00 * (1 - ty) + c10 * ty;
  const double c1 = c01 * (1 - ty) + c11 * ty;
  return c0 * (1 - tz) + c1 * tz;
}

// the prolongation added into the free nodes
__global__ void prolong_add_kernel(const double *__restrict__ co, const Grid c,
                                   double *__restrict__ out, const uint8_t *__restrict__ fixed,
                                   const Grid g) {
  int64_t i, j, k;
  if (!plane_node(g, i, j, k)) return;
  const int64_t q = ix(g, i, j, k);
  if (fixed[q]) return;
  out[q] += prolong_at(co, c, i, j, k);
}

// the cascade's starting guess: the prolongated coarser solution at every
// free node, the fixed nodes keeping their values
__global__ void prolongate_kernel(const double *__restrict__ co, const Grid c,
                                  double *__restrict__ out, const uint8_t *__restrict__ fixed,
                                  const Grid g) {
  int64_t i, j, k;
  if (!plane_node(g, i, j, k)) return;
  const int64_t q = ix(g, i, j, k);
  if (fixed[q]) return;
  out[q] = prolong_at<true>(co, c, i, j, k);
}

// the cascade's coarsening (multigrid3d.py:84-135): coarse node c takes
// fine node 2c by injection; when that node is free, c becomes fixed if any
// node of the fine 2 x 2 x 2 block at 2c is, with the value of the block's
// fixed node of largest |value| (the first in the order di, dj, dk, dk
// fastest).  pf holds the values at the fixed fine nodes; pc receives the
// coarse values at fixed nodes and zero at free ones (the coarsest
// solve's starting guess)
__global__ void coarsen_mask_kernel(const uint8_t *__restrict__ ff,
                                    const double *__restrict__ pf, const Grid g,
                                    uint8_t *__restrict__ fc, double *__restrict__ pc,
                                    const Grid c) {
  int64_t i, j, k;
  if (!plane_node(c, i, j, k)) return;
  const int64_t q = ix(c, i, j, k);
  const int64_t f0 = ix(g, 2 * i, 2 * j, 2 * k);
  uint8_t fx = ff[f0];
  double v = pf[f0];
  if (!fx) {
    double best = -1.0;
    for (int di = 0; di < 2; ++di)
      for (int dj = 0; dj < 2; ++dj)
        for (int dk = 0; dk < 2; ++dk) {
          const int64_t a = 2 * i + di, b = 2 * j + dj, e = 2 * k + dk;
          if (a >= g.nx || b >= g.ny || e >= g.nz) continue;
          const int64_t f = ix(g, a, b, e);
          if (!ff[f]) continue;
          const double m = fabs(pf[f]);
          if (m > best) {
            best = m;
            v = pf[f];
          }
        }
    fx = best >= 0.0;
  }
  fc[q] = fx;
  pc[q] = fx ? v : 0.0;
}

// coarse node fixed iff any fine node of its 2 x 2 x 2 block is
// (multigrid3d.py:507-526)
__global__ void coarsen_any_kernel(const uint8_t *__restrict__ fixed, const Grid g,
                                   uint8_t *__restrict__ out, const Grid c) {
  int64_t ic, jc, kc;
  if (!plane_node(c, ic, jc, kc)) return;
  const int64_t q = ix(c, ic, jc, kc);
  uint8_t a = 0;
  for (int di = 0; di < 2; ++di)
    for (int dj = 0; dj < 2; ++dj)
      for (int dk = 0; dk < 2; ++dk) {
        const int64_t i = 2 * ic + di, j = 2 * jc + dj, k = 2 * kc + dk;
        if (i < g.nx && j < g.ny && k < g.nz && fixed[ix(g, i, j, k)]) a = 1;
      }
  out[q] = a;
}

// phi <- (phi + flip(phi, axis)) * 0.5, each mirror pair from its lower
// index (multigrid3d.py:689-691)
__global__ void symmetrize_kernel(double *__restrict__ phi, const Grid g, int axis) {
  int64_t m[3];
  if (!plane_node(g, m[0], m[1], m[2])) return;
  const int64_t q = ix(g, m[0], m[1], m[2]);
  const int64_t dims[3] = {g.nx, g.ny, g.nz};
  m[axis] = dims[axis] - 1 - m[axis];
  const int64_t p = ix(g, m[0], m[1], m[2]);
  if (p < q) return;
  const double a = phi[q], b = phi[p];
  phi[q] = (a + b) * 0.5;
  phi[p] = (b + a) * 0.5;
}

__global__ __launch_bounds__(kBlock) void maxdiff_kernel(const double *__restrict__ a,
                                                         const double *__restrict__ b, int64_t n,
                                                         unsigned long long *part) {
  double m = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const double v = b ? fabs(a[q] - b[q]) : fabs(a[q]);
    m = v > m ? v : m;
  }
  block_max(m, part);
}

__global__ void fill_kernel(double *__restrict__ p, int64_t n, double v) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = v;
}

__global__ void zero_u64_kernel(unsigned long long *p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------- host

struct Buf {
  void *p = nullptr;
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  Buf(Buf &&o) noexcept : p(o.p) { o.p = nullptr; }
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

// pinned host words for readbacks
struct PinBuf {
  void *p = nullptr;
  PinBuf() = default;
  PinBuf(const PinBuf &) = delete;
  PinBuf &operator=(const PinBuf &) = delete;
  ~PinBuf() {
    if (p) (void)hipHostFree(p);
  }
};

struct Ctx {
  hipStream_t st;
  std::string *err;
  int rc = 0;
  bool ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    if (!rc) {
      rc = e == hipErrorOutOfMemory ? IG_E_NOMEM : IG_E_HIP;
      *err = std::string(what) + ": " + hipGetErrorString(e);
    }
    return false;
  }
  template <typename T>
  T *alloc(Buf &b, size_t count) {
    if (rc) return nullptr;
    if (!ok(hipMalloc(&b.p, count * sizeof(T) + 16), "hipMalloc")) return nullptr;
    return (T *)b.p;
  }
  bool launched() { return ok(hipGetLastError(), "kernel launch"); }
};

unsigned blocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
// per-plane launches: x covers one plane's nodes (or colour slots), y the
// planes
dim3 plane_grid(const Grid &g) { return dim3(blocks(g.ny * g.nz), (unsigned)g.nx); }
dim3 half_grid(const Grid &g) { return dim3(blocks(g.ny * ((g.nz + 1) / 2)), (unsigned)g.nx); }
int64_t n_blocks(const dim3 &d) { return (int64_t)d.x * d.y; }
constexpr unsigned kMaxDiffBlocks = 2048;

double host_omega(const Grid &g) {
  // solver3d.optimal_omega for unit spacing, as the reference's bottom
  // solve calls it (multigrid3d.py:571)
  const int64_t s[3] = {g.nx, g.ny, g.nz};
  double num = 0.0;
  int cnt = 0;
  for (int a = 0; a < 3; ++a)
    if (s[a] >= 2) {
      num = num + 1.0 * std::cos(M_PI / (double)(s[a] + 1));
      cnt += 1;
    }
  if (cnt == 0) return 1.0;
  double rho = num / (double)cnt;
  if (rho > 1.0 - 1e-15) rho = 1.0 - 1e-15;
  return 2.0 / (1.0 + std::sqrt(1.0 - rho * rho));
}

// read back the maximum of part[0, count) and, when given, tolv[0]: one
// 16-byte copy into pinned host memory
bool read_check(Ctx &x, const unsigned long long *part, int64_t count, const double *tolv,
                double *dev2, double *host2, double &m, double &tol) {
  hipLaunchKernelGGL(check_kernel, dim3(1), dim3(1024), 0, x.st, part, count, tolv, dev2);
  if (!x.launched()) return false;
  if (!x.ok(hipMemcpyAsync(host2, dev2, 16, hipMemcpyDeviceToHost, x.st), "read max") ||
      !x.ok(hipStreamSynchronize(x.st), "sync"))
    return false;
  m = host2[0];
  tol = host2[1];
  return true;
}

// the per-workgroup maxima of |a - b| (b may be null for |a|); returns
// how many
int64_t max_parts(Ctx &x, const double *a, const double *b, int64_t n,
                  unsigned long long *part) {
  const unsigned g = (unsigned)std::min<int64_t>(blocks(n), kMaxDiffBlocks);
  hipLaunchKernelGGL(maxdiff_kernel, dim3(g), dim3(kBlock), 0, x.st, a, b, n, part);
  return x.launched() ? (int64_t)g : -1;
}

struct SorOut {
  int64_t sweeps = 0;
  double delta = INFINITY;
  int status = 1;
};

// red-black SOR until a pair's max |update| < tol, at most max_sweeps
// pairs (solver3d.py:276-293): sweep(c, done, part) launches colour c;
// the pair maximum and the done flag are kept on the device, the host
// reads the flag every kSorChunk pairs.  ctl[4, 7) holds the done flag,
// the last pair's delta and the pair count; part takes both colours'
// per-workgroup maxima, nb each.  With every > 1 only every every-th pair
// (and the last allowed) is measured, so the pair count is a multiple of
// every unless the limit ends the run.
// A pair of more than kTwoStage partials is reduced in two stages when
// stage2 (kStage2 words) is given: kStage2 workgroups share the partials,
// one reads their maxima (one workgroup reading 51k partials of the 13M-node
// board takes 15 us).
constexpr int64_t kTwoStage = 4096;
constexpr unsigned kStage2 = 64;
template <class Sweep>
bool sor_drive(Ctx &x, Sweep sweep, int64_t nb, double tol, int64_t max_sweeps,
               unsigned long long *part, unsigned long long *ctl, double *host2, SorOut &o,
               int64_t every = 1, unsigned long long *stage2 = nullptr) {
  const bool two = stage2 && 2 * nb > kTwoStage;
  int32_t *done = (int32_t *)(ctl + 4);
  double *last = (double *)(ctl + 5);
  int64_t *pairs = (int64_t *)(ctl + 6);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, x.st, ctl + 4, 3);
  int32_t *hp = reinterpret_cast<int32_t *>(host2);
  for (int64_t base = 0; base < max_sweeps; base += kSorChunk) {
    const int64_t end = std::min<int64_t>(base + kSorChunk, max_sweeps);
    for (int64_t s = base; s < end; ++s) {
      const bool chk = (s + 1) % every == 0 || s + 1 == max_sweeps;
      for (int c = 0; c < 2; ++c) sweep(c, done, chk ? part + c * nb : nullptr);
      if (chk && two) {
        hipLaunchKernelGGL(chunk_max_kernel, dim3(kStage2), dim3(1024), 0, x.st, part, 2 * nb,
                           done, stage2);
        hipLaunchKernelGGL(pair_finish_kernel, dim3(1), dim3(1024), 0, x.st, stage2,
                           (int64_t)kStage2, done, s, tol, nullptr, last, pairs);
      } else if (chk) {
        hipLaunchKernelGGL(pair_finish_kernel, dim3(1), dim3(1024), 0, x.st, part, 2 * nb, done,
                           s, tol, nullptr, last, pairs);
      }
    }
    if (!x.launched()) return false;
    if (!x.ok(hipMemcpyAsync(hp, done, 4, hipMemcpyDeviceToHost, x.st), "read flag") ||
        !x.ok(hipStreamSynchronize(x.st), "sync"))
      return false;
    if (*hp) break;
  }
  // done, delta and pairs in one read
  unsigned long long *hw = reinterpret_cast<unsigned long long *>(host2);
  if (!x.ok(hipMemcpyAsync(hw, ctl + 4, 24, hipMemcpyDeviceToHost, x.st), "read result") ||
      !x.ok(hipStreamSynchronize(x.st), "sync"))
    return false;
  const int32_t fl = *reinterpret_cast<int32_t *>(hw);
  std::memcpy(&o.delta, hw + 1, 8);
  std::memcpy(&o.sweeps, hw + 2, 8);
  o.status = fl ? 0 : 1;
  if (o.sweeps == 0) o.delta = INFINITY;
  return true;
}

// solve3d's SOR of phi on the 3-D stencil
bool sor_run(Ctx &x, double *phi, const uint8_t *fixed, const float *th, const Grid &g, int gl,
             double omega, double tol, int64_t max_sweeps, unsigned long long *part,
             unsigned long long *ctl, double *host2, SorOut &o, unsigned long long *stage2) {
  const dim3 hg = half_grid(g);
  auto sweep = [&](int c, const int32_t *done, unsigned long long *pt) {
    hipLaunchKernelGGL(sweep_kernel, hg, dim3(kBlock), 0, x.st, phi, fixed, th, g, omega, c, gl,
                       done, pt, nullptr);
  };
  return sor_drive(x, sweep, n_blocks(hg), tol, max_sweeps, part, ctl, host2, o, 1, stage2);
}

// one level of the error hierarchy: its mask, right-hand side, solution
// and residual scratch
struct Level {
  Grid g;
  Buf fixed_b, f_b, e_b, r_b, tol_b;
  uint8_t *fixed = nullptr;
  double *f = nullptr, *e = nullptr, *r = nullptr;
  double *tolv = nullptr; // the level's tol and bottom threshold
};

struct Solver {
  Ctx &x;
  std::vector<Level> lv; // lv[0] is the fine grid (its fixed mask only)
  unsigned long long *part; // per-workgroup maxima
  double *dev2, *host2;     // a cycle's (max, tol) on the device / pinned host
  int32_t *done;
  int nu1, nu2;
  size_t lds_max; // dynamic LDS a bottom workgroup may take

  // _solve_err at level l (multigrid3d.py:553-599): lv[l].f given, lv[l].e
  // solved
  bool solve_err(size_t l) {
    Level &L = lv[l];
    const int64_t n = L.g.n();
    hipLaunchKernelGGL(fill_kernel, dim3(blocks(n)), dim3(kBlock), 0, x.st, L.e, n, 0.0);
    const int64_t np = max_parts(x, L.f, nullptr, n, part);
    if (np < 0) return false;
    hipLaunchKernelGGL(err_tol_kernel, dim3(1), dim3(1024), 0, x.st, part, np, kErrRelTol,
                       L.tolv);
    if (!x.launched()) return false;
    const int64_t mn = std::min(L.g.nx, std::min(L.g.ny, L.g.nz));
    if (mn <= kMinCoarse) return bottom(L);
    for (int cyc = 0; cyc < kErrCycles; ++cyc) {
      for (int s = 0; s < nu1; ++s)
        for (int c = 0; c < 2; ++c)
          hipLaunchKernelGGL(sweep_rhs_kernel, half_grid(L.g), dim3(kBlock), 0, x.st,
                             L.e, L.fixed, L.f, L.g, 1.0, c, nullptr, nullptr);
      hipLaunchKernelGGL(residual_rhs_kernel, plane_grid(L.g), dim3(kBlock), 0, x.st, L.e,
                         L.fixed, L.f, L.g, L.r);
      Level &C = lv[l + 1];
      hipLaunchKernelGGL(restrict_kernel, plane_grid(C.g), dim3(kBlock), 0, x.st, L.r,
                         L.g, C.f, C.g);
      if (!x.launched() || !solve_err(l + 1)) return false;
      hipLaunchKernelGGL(prolong_add_kernel, plane_grid(L.g), dim3(kBlock), 0, x.st, C.e, C.g,
                         L.e, L.fixed, L.g);
      // the last post-smoothing pair's max |update| decides
      const dim3 hg = half_grid(L.g);
      const int64_t nb = n_blocks(hg);
      for (int s = 0; s < nu2; ++s)
        for (int c = 0; c < 2; ++c)
          hipLaunchKernelGGL(sweep_rhs_kernel, hg, dim3(kBlock), 0, x.st, L.e, L.fixed, L.f, L.g,
                             1.0, c, nullptr, s == nu2 - 1 ? part + c * nb : nullptr);
      if (!x.launched()) return false;
      double d = 0.0, tol = 0.0;
      if (!read_check(x, part, nu2 > 0 ? 2 * nb : 0, L.tolv, dev2, host2, d, tol)) return false;
      if (d < tol) break;
    }
    return true;
  }

  // the coarsest level: SOR with the level's optimal omega until a pair's
  // max |update| < 0.02 tol, at most kBottomPairs pairs (multigrid3d.py:570-577)
  bool bottom(Level &L) {
    const double om = host_omega(L.g);
    const double *thr = L.tolv + 1;
    if (L.g.n() <= kSmallNodes) {
      const size_t bytes = (size_t)16 * L.g.n();
      if (bytes <= lds_max) {
        hipLaunchKernelGGL(bottom_kernel<true>, dim3(1), dim3(1024), bytes, x.st, L.e, L.fixed,
                           L.f, L.g, om, thr, kBottomPairs);
      } else {
        hipLaunchKernelGGL(bottom_kernel<false>, dim3(1), dim3(1024), 0, x.st, L.e, L.fixed,
                           L.f, L.g, om, thr, kBottomPairs);
      }
      return x.launched();
    }
    hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, x.st,
                       reinterpret_cast<unsigned long long *>(done), 1);
    const dim3 hg = half_grid(L.g);
    const int64_t nb = n_blocks(hg);
    for (int64_t base = 0; base < kBottomPairs; base += kSorChunk) {
      const int64_t end = std::min<int64_t>(base + kSorChunk, kBottomPairs);
      for (int64_t s = base; s < end; ++s) {
        for (int c = 0; c < 2; ++c)
          hipLaunchKernelGGL(sweep_rhs_kernel, hg, dim3(kBlock), 0, x.st, L.e, L.fixed, L.f,
                             L.g, om, c, done, part + c * nb);
        hipLaunchKernelGGL(pair_finish_kernel, dim3(1), dim3(1024), 0, x.st, part, 2 * nb, done,
                           s, 0.0, thr, nullptr, nullptr);
      }
      if (!x.launched()) return false;
      int32_t *h = reinterpret_cast<int32_t *>(host2);
      if (!x.ok(hipMemcpyAsync(h, done, 4, hipMemcpyDeviceToHost, x.st), "read flag") ||
          !x.ok(hipStreamSynchronize(x.st), "sync"))
        return false;
      if (*h) break;
    }
    return true;
  }
};

} // namespace

int solve3d_run(hipStream_t st, const ig_solve3d_desc *d, double *phi_out,
                ig_solve3d_result *res, std::string &err) {
  Ctx x{st, &err};
  const Grid g{d->n[0], d->n[1], d->n[2]};
  const int64_t n = g.n();
  const int gl = d->stencil == IG_STENCIL_GHOST_LINEAR;
  // the initial iterate: phi0 or zeros, fixed nodes at their values
  std::vector<double> init(n);
  for (int64_t q = 0; q < n; ++q) {
    const double v0 = d->phi0 ? d->phi0[q] : 0.0;
    init[q] = d->fixed[q] ? d->val[q] : v0;
  }
  Buf phi_b, fixed_b, th_b, ctl_b;
  double *phi = x.alloc<double>(phi_b, n);
  uint8_t *fixed = x.alloc<uint8_t>(fixed_b, n);
  float *th = d->theta ? x.alloc<float>(th_b, 6 * n) : nullptr;
  // control words: [2, 4) a cycle's (max, tol), the done flag, last
  // delta, pairs
  unsigned long long *ctl = x.alloc<unsigned long long>(ctl_b, 8);
  PinBuf pin;
  if (!x.rc && !x.ok(hipHostMalloc(&pin.p, 64, hipHostMallocDefault), "hipHostMalloc"))
    return x.rc;
  double *host2 = (double *)pin.p;
  // per-workgroup maxima: both colours of a fine sweep pair at most
  const int64_t nb_fine = n_blocks(half_grid(g));
  Buf part_b;
  const int64_t n_part = 2 * std::max<int64_t>(nb_fine, kMaxDiffBlocks);
  unsigned long long *part = x.alloc<unsigned long long>(part_b, n_part + kStage2);
  if (x.rc) return x.rc;
  double *dev2 = (double *)(ctl + 2);
  int32_t *done = (int32_t *)(ctl + 4);
  x.ok(hipMemcpyAsync(phi, init.data(), n * 8, hipMemcpyHostToDevice, st), "upload phi");
  x.ok(hipMemcpyAsync(fixed, d->fixed, n, hipMemcpyHostToDevice, st), "upload mask");
  if (th) x.ok(hipMemcpyAsync(th, d->theta, 6 * n * 4, hipMemcpyHostToDevice, st), "upload theta");
  x.ok(hipMemsetAsync(ctl, 0, 64, st), "clear control");
  if (x.rc) return x.rc;
  res->sweeps = 0;
  res->delta = INFINITY;
  res->status = 1;
  res->level = 0;

  if (d->method == IG_SOLVE_SOR) {
    SorOut o;
    if (!sor_run(x, phi, fixed, th, g, gl, d->omega, d->tol, d->max_sweeps, part, ctl, host2, o,
                 part + n_part))
      return x.rc;
    res->sweeps = o.sweeps;
    res->delta = o.delta;
    res->status = o.status;
  } else if (d->method == IG_SOLVE_CASCADE) {
    // the hierarchy, fine -> coarse (multigrid3d.py:194-201); each grid's
    // phi holds its fixed values and zeros until its solve
    struct Grd {
      Grid g;
      Buf f_b, p_b;
      uint8_t *fixed = nullptr;
      double *phi = nullptr;
    };
    std::vector<Grd> gr(1);
    gr[0].g = g;
    gr[0].fixed = fixed;
    gr[0].phi = phi;
    while ((int64_t)gr.size() < d->max_levels) {
      const Grid p = gr.back().g;
      const int64_t mn = std::min(p.nx, std::min(p.ny, p.nz));
      if (mn <= d->min_coarse || (mn + 1) / 2 < 3) break;
      Grd L;
      L.g = Grid{(p.nx + 1) / 2, (p.ny + 1) / 2, (p.nz + 1) / 2};
      L.fixed = x.alloc<uint8_t>(L.f_b, L.g.n());
      L.phi = x.alloc<double>(L.p_b, L.g.n());
      if (x.rc) return x.rc;
      hipLaunchKernelGGL(coarsen_mask_kernel, plane_grid(L.g), dim3(kBlock), 0, st,
                         gr.back().fixed, gr.back().phi, p, L.fixed, L.phi, L.g);
      if (!x.launched()) return x.rc;
      gr.push_back(std::move(L));
    }
    // coarsest to tol from zeros, the middle grids to tol * factor from
    // the prolongated coarser solution, the fine grid with its fractions
    // (multigrid3d.py:207-229)
    const int nl = (int)gr.size();
    SorOut o;
    for (int l = nl - 1; l >= 0; --l) {
      Grd &L = gr[l];
      if (l < nl - 1)
        hipLaunchKernelGGL(prolongate_kernel, plane_grid(L.g), dim3(kBlock), 0, st, gr[l + 1].phi,
                           gr[l + 1].g, L.phi, L.fixed, L.g);
      const double lt = (l == nl - 1 || l == 0) ? d->tol : d->tol * d->level_tol_factor;
      const double om = l == 0 ? d->omega : host_omega(L.g);
      if (!sor_run(x, L.phi, L.fixed, l == 0 ? th : nullptr, L.g, gl, om, lt, d->max_sweeps, part,
                   ctl, host2, o, part + n_part))
        return x.rc;
      res->level = l;
      if (o.status) break;
    }
    res->sweeps = o.sweeps;
    res->delta = o.delta;
    res->status = o.status;
  } else {
    // the error hierarchy, fine -> coarse, down to the first level with an
    // axis of at most kMinCoarse nodes (multigrid3d.py:569-582)
    // the LDS a workgroup may allocate, less the bottom kernel's own words
    int dev = 0, lds = 0;
    if (!x.ok(hipGetDevice(&dev), "hipGetDevice") ||
        !x.ok(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
              "LDS size"))
      return x.rc;
    size_t lds_max = lds > 1024 ? (size_t)lds - 1024 : 0;
    if (lds_max > 64512 &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(&bottom_kernel<true>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_max) != hipSuccess) {
      (void)hipGetLastError(); // keep to the 64 KiB every workgroup may take
      lds_max = 64512;
    }
    Solver S{x, {}, part, dev2, host2, done, d->nu1, d->nu2, lds_max};
    S.lv.emplace_back();
    S.lv[0].g = g;
    S.lv[0].fixed = fixed;
    for (;;) {
      const Grid &p = S.lv.back().g;
      const bool is_fine = S.lv.size() == 1;
      if (!is_fine && std::min(p.nx, std::min(p.ny, p.nz)) <= kMinCoarse) break;
      Level L;
      L.g = Grid{(p.nx + 1) / 2, (p.ny + 1) / 2, (p.nz + 1) / 2};
      const int64_t m = L.g.n();
      L.fixed = x.alloc<uint8_t>(L.fixed_b, m);
      L.f = x.alloc<double>(L.f_b, m);
      L.e = x.alloc<double>(L.e_b, m);
      L.r = x.alloc<double>(L.r_b, m);
      L.tolv = x.alloc<double>(L.tol_b, 2);
      if (x.rc) return x.rc;
      hipLaunchKernelGGL(coarsen_any_kernel, plane_grid(L.g), dim3(kBlock), 0, st,
                         S.lv.back().fixed, p, L.fixed, L.g);
      if (!x.launched()) return x.rc;
      S.lv.push_back(std::move(L));
    }
    Buf c0_b, r_b;
    double *c0 = x.alloc<double>(c0_b, n);
    double *r = x.alloc<double>(r_b, n);
    if (x.rc) return x.rc;
    double prev = INFINITY, cd = INFINITY;
    int64_t sw = 0;
    int status = 1;
    const Level &C = S.lv[1];
    // the cycle-start snapshot: fixed nodes never change, so one copy of
    // the whole field up front, then each cycle's first pre-smoothing pair
    // writes the free nodes' values before it updates them (a whole copy
    // per cycle when there is no pre-smoothing)
    if (!x.ok(hipMemcpyAsync(c0, phi, n * 8, hipMemcpyDeviceToDevice, st), "snapshot"))
      return x.rc;
    for (int cyc = 0; cyc < d->max_cycles; ++cyc) {
      if (cyc > 0 && d->nu1 == 0 &&
          !x.ok(hipMemcpyAsync(c0, phi, n * 8, hipMemcpyDeviceToDevice, st), "snapshot"))
        return x.rc;
      for (int s = 0; s < d->nu1; ++s, ++sw)
        for (int c = 0; c < 2; ++c)
          hipLaunchKernelGGL(sweep_kernel, half_grid(g), dim3(kBlock), 0, st, phi, fixed,
                             th, g, 1.0, c, gl, nullptr, nullptr, s == 0 ? c0 : nullptr);
      hipLaunchKernelGGL(residual_kernel, plane_grid(g), dim3(kBlock), 0, st, phi, fixed, th, g,
                         gl, r);
      hipLaunchKernelGGL(restrict_kernel, plane_grid(C.g), dim3(kBlock), 0, st, r, g, C.f,
                         C.g);
      if (!x.launched() || !S.solve_err(1)) return x.rc;
      hipLaunchKernelGGL(prolong_add_kernel, plane_grid(g), dim3(kBlock), 0, st, C.e, C.g, phi,
                         fixed, g);
      for (int s = 0; s < d->nu2; ++s, ++sw)
        for (int c = 0; c < 2; ++c)
          hipLaunchKernelGGL(sweep_kernel, half_grid(g), dim3(kBlock), 0, st, phi, fixed,
                             th, g, 1.0, c, gl, nullptr, nullptr, nullptr);
      for (int a = 0; a < 3; ++a)
        if ((d->sym_axes >> a) & 1)
          hipLaunchKernelGGL(symmetrize_kernel, plane_grid(g), dim3(kBlock), 0, st, phi, g, a);
      const int64_t np = max_parts(x, phi, c0, n, part);
      double unused;
      if (np < 0 || !read_check(x, part, np, nullptr, dev2, host2, cd, unused)) return x.rc;
      if (cd < d->tol) {
        status = 0;
        break;
      }
      if (cd >= prev) {
        status = 2;
        break;
      }
      prev = cd;
    }
    res->sweeps = sw;
    res->delta = cd;
    res->status = status;
  }
  if (!x.ok(hipMemcpyAsync(phi_out, phi, n * 8, hipMemcpyDeviceToHost, st), "download phi") ||
      !x.ok(hipStreamSynchronize(st), "sync"))
    return x.rc;
  return 0;
}

int solve2d_run(hipStream_t st, const ig_solve2d_desc *d, double *phi_out,
                ig_solve3d_result *res, std::string &err) {
  Ctx x{st, &err};
  const Grid g{d->nz, 1, d->nu};
  const int64_t n = g.n();
  std::vector<double> init(n);
  for (int64_t q = 0; q < n; ++q) {
    const double v0 = d->phi0 ? d->phi0[q] : 0.0;
    init[q] = d->fixed[q] ? d->val[q] : v0;
  }
  Buf phi_b, fixed_b, ctl_b, part_b;
  double *phi = x.alloc<double>(phi_b, n);
  uint8_t *fixed = x.alloc<uint8_t>(fixed_b, n);
  unsigned long long *ctl = x.alloc<unsigned long long>(ctl_b, 8);
  const dim3 hg = half_grid(g);
  const int64_t nb = n_blocks(hg);
  unsigned long long *part = x.alloc<unsigned long long>(part_b, 2 * nb);
  PinBuf pin;
  if (!x.rc && !x.ok(hipHostMalloc(&pin.p, 64, hipHostMallocDefault), "hipHostMalloc"))
    return x.rc;
  if (x.rc) return x.rc;
  x.ok(hipMemcpyAsync(phi, init.data(), n * 8, hipMemcpyHostToDevice, st), "upload phi");
  x.ok(hipMemcpyAsync(fixed, d->fixed, n, hipMemcpyHostToDevice, st), "upload mask");
  if (x.rc) return x.rc;
  const Lap2 lap{d->h, d->symmetry == IG_SYM_CYLINDRICAL, d->ghost_linear};
  auto sweep = [&](int c, const int32_t *done, unsigned long long *pt) {
    hipLaunchKernelGGL(sweep2d_kernel, hg, dim3(kBlock), 0, st, phi, fixed, g, lap, d->omega, c,
                       done, pt);
  };
  SorOut o;
  // the reference's LU has no pair count to match: measure every 8th pair
  if (!sor_drive(x, sweep, nb, d->tol, d->max_sweeps, part, ctl, (double *)pin.p, o, 8))
    return x.rc;
  res->sweeps = o.sweeps;
  res->delta = o.delta;
  res->status = o.status;
  res->level = 0;
  if (!x.ok(hipMemcpyAsync(phi_out, phi, n * 8, hipMemcpyDeviceToHost, st), "download phi") ||
      !x.ok(hipStreamSynchronize(st), "sync"))
    return x.rc;
  return 0;
}

} // namespace ig
