// Births on the device (see ig_births.h for the scheme): PCG64 words,
// the per-position draw count, binary lifting over it, and the births.
#pragma clang fp contract(off)
#include "ig_births.h"

#include <cmath>

namespace ig {
namespace {

// 128-bit unsigned arithmetic modulo 2^128 (high, low words)
struct U128 {
  uint64_t hi, lo;
};

__device__ __forceinline__ U128 mul128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo * b.lo;
  r.hi = __umul64hi(a.lo, b.lo) + a.hi * b.lo + a.lo * b.hi;
  return r;
}

__device__ __forceinline__ U128 add128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo + b.lo;
  r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
  return r;
}

// the PCG64 multiplier (O'Neill's 128-bit LCG constant, numpy's PCG64)
__device__ __forceinline__ U128 pcg_mult() { return U128{0x2360ED051FC65DA4ull, 0x4385DF649FCCF645ull}; }

// XSL-RR output of a 128-bit state
__device__ __forceinline__ uint64_t pcg_out(U128 s) {
  const unsigned rot = (unsigned)(s.hi >> 58);
  const uint64_t v = s.hi ^ s.lo;
  return (v >> rot) | (v << ((64u - rot) & 63u));
}

// the state `delta` steps after s (Brown's jump: the affine map s -> a s + c
// raised to the power delta by squaring)
__device__ U128 pcg_advance(U128 s, U128 inc, uint64_t delta) {
  U128 acc_m{0, 1}, acc_c{0, 0}, cur_m = pcg_mult(), cur_c = inc;
  while (delta) {
    if (delta & 1u) {
      acc_m = mul128(acc_m, cur_m);
      acc_c = add128(mul128(acc_c, cur_m), cur_c);
    }
    cur_c = mul128(add128(cur_m, U128{0, 1}), cur_c);
    cur_m = mul128(cur_m, cur_m);
    delta >>= 1;
  }
  return add128(mul128(acc_m, s), acc_c);
}

constexpr int kChunk = 256; // words per thread in the word pass

__global__ void words_kernel(U128 s0, U128 inc, uint64_t *__restrict__ W, int64_t nW) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t start = t * kChunk;
  if (start >= nW) return;
  // numpy steps the state, then outputs: word j comes from state j + 1
  U128 s = pcg_advance(s0, inc, (uint64_t)start);
  const int64_t end = start + kChunk < nW ? start + kChunk : nW;
  for (int64_t j = start; j < end; ++j) {
    s = add128(mul128(s, pcg_mult()), inc);
    W[j] = pcg_out(s);
  }
}

// one consumer of the word stream
struct Draw {
  const uint64_t *W;
  int64_t nW, p;
  bool over;
  __device__ uint64_t word() {
    if (p >= nW) {
      over = true;
      return 0;
    }
    return W[p++];
  }
  // numpy's next_double: the top 53 bits times 2^-53
  __device__ double dbl() { return (double)(word() >> 11) * (1.0 / 9007199254740992.0); }
};

constexpr double kZigR = 3.6541528853610088;    // the 256-layer ziggurat's r
constexpr double kZigInvR = 0.27366123732975828; // 1 / r

// numpy's standard normal (random_standard_normal) on the word stream
__device__ double std_normal(Draw &d, const BirthsK &b) {
  for (;;) {
    uint64_t r = d.word();
    if (d.over) return 0.0;
    const int idx = (int)(r & 0xffu);
    r >>= 8;
    const uint64_t sign = r & 1u;
    const uint64_t rabs = (r >> 1) & 0x000fffffffffffffull;
    double x = (double)rabs * b.zw[idx];
    if (sign) x = -x;
    if (rabs < b.zk[idx]) return x;
    if (idx == 0) {
      for (;;) {
        const double xx = -kZigInvR * log1p(-d.dbl());
        const double yy = -log1p(-d.dbl());
        if (d.over) return 0.0;
        if (yy + yy > xx * xx) return ((rabs >> 8) & 1u) ? -(kZigR + xx) : kZigR + xx;
      }
    } else {
      const double u = d.dbl();
      if (d.over) return 0.0;
      if ((b.zf[idx - 1] - b.zf[idx]) * u + b.zf[idx] < exp(-0.5 * x * x)) return x;
    }
  }
}

// Generator.normal() / normal(size=3) elements: loc + scale * z
__device__ __forceinline__ double normal01(Draw &d, const BirthsK &b) {
  return 0.0 + 1.0 * std_normal(d, b);
}

// Generator.uniform(low, high): low + (high - low) * next_double
__device__ __forceinline__ double uniform(Draw &d, double lo, double hi) {
  return lo + (hi - lo) * d.dbl();
}

constexpr double kKB = 1.380649e-23;       // ion_gym/physics/constants.py:19
constexpr double kEChg = 1.602176634e-19;  // ion_gym/physics/constants.py:18
constexpr double kKgAmu = 1.66053906660e-27; // ion_gym/physics/constants.py:23
constexpr double kTwoPi = 6.283185307179586; // 2 * math.pi

// One ion's birth (the loop body of sim_build.py:144-275) from word p on:
// returns the words it took, -1 when the stream ran out first, -2 when a
// truncated-Gaussian axis exhausted its tries.  VALUES = false computes
// only what the draw count depends on.
template <bool VALUES>
__device__ int64_t birth_eval(const BirthsK &b, const uint64_t *W, int64_t nW, int64_t p,
                              int64_t i, double *o) {
  Draw d{W, nW, p, false};
  double pos[3] = {b.p0[0], b.p0[1], b.p0[2]};
  const int ax = b.axis;
  const int ea = ax == 0 ? 1 : 0, eb = ax == 2 ? 1 : 2; // the disc plane's axes
  switch (b.dist) {
  case IG_BIRTH_DISC: {
    const double u = uniform(d, 0.0, 1.0);
    const double th = uniform(d, 0.0, kTwoPi);
    if (VALUES) {
      const double rr = b.r_mm * sqrt(u);
      pos[ea] += rr * cos(th);
      pos[eb] += rr * sin(th);
    }
    break;
  }
  case IG_BIRTH_GRID:
    if (VALUES) {
      const int64_t g = b.grid_g, ib = i % b.n_per;
      const int64_t gi = ib % g, gj = ib / g;
      const double dd = (double)(g - 1 > 1 ? g - 1 : 1);
      const double u = ((double)gi / dd - 0.5) * 2.0 * b.r_mm;
      const double v = ((double)gj / dd - 0.5) * 2.0 * b.r_mm;
      const double rr = hypot(u, v), th = atan2(v, u);
      pos[ea] += rr * cos(th);
      pos[eb] += rr * sin(th);
    }
    break;
  case IG_BIRTH_LINE:
    if (VALUES) {
      const int64_t ib = i % b.n_per;
      pos[ax] += b.len_mm * ((double)ib / (double)(b.n_per - 1 > 1 ? b.n_per - 1 : 1));
    }
    break;
  case IG_BIRTH_BOX:
    for (int k = 0; k < 3; ++k) pos[k] += b.box[k] * (d.dbl() - 0.5);
    break;
  case IG_BIRTH_GAUSSIAN:
    for (int k = 0; k < 3; ++k) {
      if (b.gsig[k] <= 0.0) continue;
      bool ok = false;
      double dpos = 0.0;
      for (int32_t t = 0; t < b.gtries; ++t) {
        dpos = normal01(d, b) * b.gsig[k];
        if (d.over) return -1;
        if (fabs(dpos) <= b.gtrunc[k]) {
          ok = true;
          break;
        }
      }
      if (!ok) return -2;
      pos[k] += dpos;
    }
    break;
  default: // IG_BIRTH_POINT: born at the centre
    break;
  }
  const double mz = b.mz[(i / b.n_per) % b.n_mz];
  const double m_kg = mz * kKgAmu;
  double v[3] = {0.0, 0.0, 0.0};
  if (b.has_ke) {
    const double ke = uniform(d, b.ke_lo, b.ke_hi);
    if (VALUES && b.has_dir) {
      const double sp = sqrt(2.0 * ke * kEChg / m_kg) / 1000.0;
      for (int k = 0; k < 3; ++k) v[k] = v[k] + sp * b.dir[k];
    }
  }
  if (b.temp_k > 0.0) {
    const double sig = VALUES ? sqrt(kKB * b.temp_k / m_kg) / 1000.0 : 0.0;
    for (int k = 0; k < 3; ++k) {
      const double z = normal01(d, b);
      v[k] = v[k] + z * sig;
    }
  }
  if (b.has_dv)
    for (int k = 0; k < 3; ++k) {
      const double z = normal01(d, b);
      v[k] = v[k] + z * b.dvs[k];
    }
  if (b.has_drift)
    for (int k = 0; k < 3; ++k) v[k] = v[k] + b.vd[k];
  const double tob = uniform(d, 0.0, b.tob_span);
  if (d.over) return -1;
  if (VALUES) {
    o[0] = pos[0];
    o[1] = pos[1];
    o[2] = pos[2];
    o[3] = v[0];
    o[4] = v[1];
    o[5] = v[2];
    o[6] = tob;
  }
  return d.p - p;
}

// level 0 of the jump table: where an ion starting at p ends; P and P + 1
// absorb (P: past the stream, P + 1: a truncation that failed)
__global__ void next_kernel(const BirthsK b, const uint64_t *__restrict__ W, int64_t nW,
                            int32_t *__restrict__ J0, int64_t P) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P + 1) return;
  if (p >= P) {
    J0[p] = (int32_t)p;
    return;
  }
  const int64_t c = birth_eval<false>(b, W, nW, p, 0, nullptr);
  int64_t q;
  if (c == -2)
    q = P + 1;
  else if (c < 0 || p + c >= P)
    q = P;
  else
    q = p + c;
  J0[p] = (int32_t)q;
}

__global__ void jump_kernel(const int32_t *__restrict__ Jk, int32_t *__restrict__ Jn, int64_t P) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p > P + 1) return;
  Jn[p] = Jk[Jk[p]];
}

__global__ void flag_init_kernel(int64_t *flag) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    flag[0] = 0;
    flag[1] = INT64_MAX;
  }
}

__global__ void ions_kernel(const BirthsK b, const uint64_t *__restrict__ W, int64_t nW,
                            const int32_t *__restrict__ J, int64_t P, int32_t levels,
                            const BirthsOut out, int64_t *flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.n) return;
  int64_t o = 0;
  for (int32_t k = 0; k < levels; ++k)
    if ((i >> k) & 1) o = J[(int64_t)k * (P + 2) + o];
  double r[7];
  int64_t c = o == P + 1 ? -3 : (o >= P ? -1 : birth_eval<true>(b, W, nW, o, i, r));
  if (c < 0) {
    // -3: an earlier ion's truncation failed; the failing ion reports it
    if (c == -1) atomicOr((unsigned long long *)&flag[0], 1ull);
    if (c == -2) {
      atomicOr((unsigned long long *)&flag[0], 2ull);
      atomicMin((long long *)&flag[1], (long long)i);
    }
    for (int q = 0; q < 7; ++q) r[q] = __builtin_nan("");
  }
  for (int q = 0; q < 7; ++q) out.col[q][i * out.stride] = r[q];
}

} // namespace

int births_launch(const BirthsK &b, uint64_t s_hi, uint64_t s_lo, uint64_t i_hi, uint64_t i_lo,
                  uint64_t *W, int64_t nW, int32_t *J, int64_t P, int32_t levels,
                  const BirthsOut &out, int64_t *flag, hipStream_t stream) {
  const U128 s0{s_hi, s_lo}, inc{i_hi, i_lo};
  const int bs = 256;
  hipLaunchKernelGGL(flag_init_kernel, dim3(1), dim3(64), 0, stream, flag);
  const int64_t nthr = (nW + kChunk - 1) / kChunk;
  hipLaunchKernelGGL(words_kernel, dim3((unsigned)((nthr + bs - 1) / bs)), dim3(bs), 0, stream,
                     s0, inc, W, nW);
  const unsigned gp = (unsigned)((P + 2 + bs - 1) / bs);
  hipLaunchKernelGGL(next_kernel, dim3(gp), dim3(bs), 0, stream, b, (const uint64_t *)W, nW, J, P);
  for (int32_t k = 1; k < levels; ++k)
    hipLaunchKernelGGL(jump_kernel, dim3(gp), dim3(bs), 0, stream,
                       (const int32_t *)(J + (int64_t)(k - 1) * (P + 2)),
                       J + (int64_t)k * (P + 2), P);
  if (b.n > 0)
    hipLaunchKernelGGL(ions_kernel, dim3((unsigned)((b.n + bs - 1) / bs)), dim3(bs), 0, stream, b,
                       (const uint64_t *)W, nW, (const int32_t *)J, P, levels, out, flag);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

} // namespace ig
