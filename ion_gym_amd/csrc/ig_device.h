// Device helpers shared by the r-z and 3-D flight kernels: the Philox
// stream, the reference's polar-method normals, the hard-sphere and SDS
// gas steps, and the drive waveforms.  All FP64 and written with
// contraction off (the including translation unit sets
// `#pragma clang fp contract(off)`), so the arithmetic follows the
// reference's evaluation order operation by operation.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "ig_common.h"

namespace ig {

// per-lane stage of the flight kernels' step machines
enum : int { kStep = 0, kCollide = 1, kPost = 2, kDone = 3 };

#define IG_KB 1.380649e-23         // ion_gym/physics/constants.py:19
#define IG_E_CHG 1.602176634e-19   // ion_gym/physics/constants.py:18
#define IG_KG_AMU 1.66053906660e-27 // ion_gym/physics/constants.py:23
#define IG_TWO_PI 6.283185307179586
#define IG_PI 3.141592653589793

// ---------------------------------------------------------------- Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC11).  One ion's stream is
// keyed by its seed XOR the run key; draw n of the stream takes block n/2
// and its word pair (0,1) or (2,3).  The stream is a function of (seed,
// run key, n) only, so an ion draws the same numbers whatever lane, batch,
// window or request mix it flies in.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // full 64-bit products: one multiply gives both halves
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    uint32_t n0 = hi1 ^ c[1] ^ k0;
    uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

struct Rng {
  uint32_t k0, k1;
  uint64_t n;
  int32_t has_gauss;
  double gauss;
  // words (2, 3) of block (n - 1) / 2 when n is odd: the next draw's pair,
  // kept from the block that produced the previous draw (cached = true)
  uint32_t w2, w3;
  bool cached;
};

// a double in [0, 1) from two 32-bit words, 27 + 26 bits (the construction
// numba's np.random.random() uses)
__device__ __forceinline__ double words_to_u01(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}

__device__ __forceinline__ double rng_uniform(Rng &r) {
  const uint64_t blk = r.n >> 1;
  const bool hi = (r.n & 1u) != 0;
  r.n += 1;
  if (hi && r.cached) {
    r.cached = false;
    return words_to_u01(r.w2, r.w3);
  }
  uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), 0u, 0u};
  philox4x32_10(c, r.k0, r.k1);
  if (hi) return words_to_u01(c[2], c[3]);
  r.w2 = c[2];
  r.w3 = c[3];
  r.cached = true;
  return words_to_u01(c[0], c[1]);
}

// polar-method normal with the second value of each pair cached, the
// algorithm numba's np.random.normal() uses (collision3d.py:295-297)
__device__ __forceinline__ double rng_normal(Rng &r) {
  if (r.has_gauss) {
    r.has_gauss = 0;
    return r.gauss;
  }
  double x1, x2, r2;
  // acceptance is pi/4 per pair: 200 rejections in a row has probability
  // 1e-134, so the bound never binds for a working stream and guarantees
  // the loop ends for any input
  int tries = 0;
  do {
    x1 = 2.0 * rng_uniform(r) - 1.0;
    x2 = 2.0 * rng_uniform(r) - 1.0;
    r2 = x1 * x1 + x2 * x2;
  } while ((r2 >= 1.0 || r2 == 0.0) && ++tries < 200);
  if (r2 >= 1.0 || r2 == 0.0) r2 = 0.5;
  double f = sqrt(-2.0 * log(r2) / r2);
  r.gauss = f * x1;
  r.has_gauss = 1;
  return f * x2;
}

// ------------------------------------------------------------- hard sphere
// collision3d.py:265-272
__device__ __forceinline__ double c_bar_rel(double speed, double c_star, double c_bar) {
  double s = speed / c_star;
  if (s < 1e-12) return c_bar;
  const double sqrt_pi = 1.7724538509055159; // math.sqrt(math.pi)
  return c_bar * ((s + 1.0 / (2.0 * s)) * 0.5 * sqrt_pi * erf(s) + 0.5 * exp(-s * s));
}

// collision3d.py:281-330; returns false when no partner is accepted in
// MAX_PARTNER_TRIES draws (the reference raises there)
__device__ __noinline__ bool hs_collide(Rng &rng, double &vx, double &vy, double &vz, double m_ion,
                                        double m_gas, double sig1d, double speed_rel) {
  double scale = speed_rel + sig1d * 5.196152422706632;
  double rx = 0, ry = 0, rz = 0, rlen = 0;
  bool accepted = false;
  for (int tr = 0; tr < 10000; ++tr) {
    double gx = rng_normal(rng) * sig1d;
    double gy = rng_normal(rng) * sig1d;
    double gz = rng_normal(rng) * sig1d;
    rx = vx - gx;
    ry = vy - gy;
    rz = vz - gz;
    rlen = sqrt(rx * rx + ry * ry + rz * rz);
    if (rng_uniform(rng) < rlen / scale) {
      accepted = true;
      break;
    }
  }
  if (!accepted) return false;
  if (rlen < 1e-12) return true;
  double alpha = asin(sqrt(0.999999999 * rng_uniform(rng)));
  double theta = 2.0 * IG_PI * rng_uniform(rng);
  double ux = rx / rlen, uy = ry / rlen, uz = rz / rlen;
  double px, py, pz;
  if (fabs(ux) < 0.9) {
    px = 0.0;
    py = -uz;
    pz = uy;
  } else {
    px = -uz;
    py = 0.0;
    pz = ux;
  }
  double pl = sqrt(px * px + py * py + pz * pz);
  px = px / pl;
  py = py / pl;
  pz = pz / pl;
  double qx = uy * pz - uz * py;
  double qy = uz * px - ux * pz;
  double qz = ux * py - uy * px;
  double ca = cos(alpha), sa = sin(alpha);
  double ct = cos(theta), st = sin(theta);
  double rhx = ca * ux + sa * (ct * px + st * qx);
  double rhy = ca * uy + sa * (ct * py + st * qy);
  double rhz = ca * uz + sa * (ct * pz + st * qz);
  double f = 2.0 * m_gas / (m_ion + m_gas) * (rx * rhx + ry * rhy + rz * rhz);
  vx = vx - f * rhx;
  vy = vy - f * rhy;
  vz = vz - f * rhz;
  return true;
}

// collision3d.py:333-374, in two halves so a kernel can defer the second:
// hs_test draws the step's collision test in the gas-flow frame,
// hs_scatter performs the collision (the velocity is unchanged in
// between, so both halves see the same w and sp the reference computes
// once).
__device__ __forceinline__ bool hs_test(Rng &rng, double vx, double vy, double vz,
                                        const FlightK &p, double dt, double dt_to_us,
                                        double sigma) {
  double wx = vx - p.vg[0], wy = vy - p.vg[1], wz = vz - p.vg[2];
  double sp = sqrt(wx * wx + wy * wy + wz * wz);
  if (sp < 1e-7) sp = 1e-7;
  double cbr = c_bar_rel(sp, p.c_star, p.c_bar);
  double lam = 1000.0 * IG_KB * p.T_k * (sp / cbr) / (p.P_pa * sigma);
  double u = rng_uniform(rng);
  return u < 1.0 - exp(-sp * dt * dt_to_us / lam);
}

// returns false when no partner is accepted (the reference raises)
__device__ __forceinline__ bool hs_scatter(Rng &rng, double &vx, double &vy, double &vz,
                                           const FlightK &p, double m_ion) {
  double wx = vx - p.vg[0], wy = vy - p.vg[1], wz = vz - p.vg[2];
  double sp = sqrt(wx * wx + wy * wy + wz * wz);
  if (sp < 1e-7) sp = 1e-7;
  if (!hs_collide(rng, wx, wy, wz, m_ion, p.m_gas, p.sig1d, sp)) return false;
  vx = wx + p.vg[0];
  vy = wy + p.vg[1];
  vz = wz + p.vg[2];
  return true;
}

// one collision opportunity over the step.  Returns 1 collided, 0 not, -1
// partner draw refused.
__device__ __forceinline__ int hs_step(Rng &rng, double &vx, double &vy, double &vz,
                                       const FlightK &p, double dt, double dt_to_us,
                                       double sigma, double m_ion) {
  if (!hs_test(rng, vx, vy, vz, p, dt, dt_to_us, sigma)) return 0;
  return hs_scatter(rng, vx, vy, vz, p, m_ion) ? 1 : -1;
}

// ------------------------------------------------------------------ SDS
// sds.py:134-151 (_icdf_steps_at) on the (5, 1002) jump table
__device__ __forceinline__ double icdf_steps_at(const double *__restrict__ st, double log_mr,
                                                double u) {
  double x = u * (1002 - 2);
  int64_t i = (int64_t)x;
  double w = x - (double)i;
  int64_t k = (int64_t)floor(log_mr);
  if (k < 0)
    k = 0;
  else if (k > 3)
    k = 3;
  double t = log_mr - (double)k;
  const double *a = st + k * 1002, *b = a + 1002;
  double lo = a[i] + w * (a[i + 1] - a[i]);
  double hi = b[i] + w * (b[i + 1] - b[i]);
  double llo = log10(lo);
  return pow(10.0, llo + t * (log10(hi) - llo));
}

// tracer_rz.py:264-267 with sds.py:154-169
__device__ __forceinline__ void sds_jump(Rng &rng, const double *__restrict__ st, double log_mr,
                                         double V, double mfp, double dt, double &jx, double &jy,
                                         double &jz) {
  double dsteps = icdf_steps_at(st, log_mr, rng_uniform(rng));
  double nc = V / mfp * dt;
  double rj = sqrt(nc / 100000.0) * dsteps * mfp;
  double c = 2.0 * rng_uniform(rng) - 1.0;
  double ph = 2.0 * IG_PI * rng_uniform(rng);
  double one_m = 1.0 - c * c;
  double s = rj * sqrt(one_m > 0.0 ? one_m : 0.0);
  jx = s * cos(ph);
  jy = s * sin(ph);
  jz = rj * c;
}

// -------------------------------------------------------------- waveforms
__device__ __forceinline__ double py_mod1(double a) {
  double m = fmod(a, 1.0);
  if (m != 0.0) {
    if (m < 0.0) m += 1.0;
  } else {
    m = 0.0;
  }
  return m;
}

// Drive waveform of channel k at lab time t (us).  affine = false is the
// r-z unit waveform (build_planar.py:1059-1096: the amplitude is baked into
// the channel grid); affine = true the 3-D amp*base(t)+off
// (tracer3d.py:234-276).
template <bool AFFINE>
__device__ __forceinline__ double wave_eval(const Waves *__restrict__ W, const FieldsK &f, int k,
                                            double t) {
  int kind = W->kind[k];
  double om = W->om[k], ph = W->ph[k];
  if (kind == IG_WAVE_SIN) {
    double s = sin(om * t + ph);
    return AFFINE ? W->amp[k] * s + W->off[k] : s;
  }
  if (kind == IG_WAVE_COS) {
    double s = cos(om * t + ph);
    return AFFINE ? W->amp[k] * s + W->off[k] : s;
  }
  if (kind == IG_WAVE_SQUARE) {
    double base;
    if (W->duty[k] == 0.5)
      base = sin(om * t + ph) >= 0.0 ? 1.0 : -1.0;
    else
      base = py_mod1((om * t + ph) / (2.0 * IG_PI)) < W->duty[k] ? 1.0 : -1.0;
    return AFFINE ? W->amp[k] * base + W->off[k] : base;
  }
  int64_t o0 = W->tab_off[k], o1 = W->tab_off[k + 1];
  const double *tt = f.tab_t, *tv = f.tab_v;
  if (o1 - o0 == 0) return AFFINE ? W->off[k] : 0.0;
  double v;
  if (t <= tt[o0]) {
    v = tv[o0];
  } else if (t >= tt[o1 - 1]) {
    v = tv[o1 - 1];
  } else {
    int64_t lo = o0, hi = o1 - 1;
    while (hi - lo > 1) {
      int64_t mid = (lo + hi) / 2;
      if (tt[mid] <= t)
        lo = mid;
      else
        hi = mid;
    }
    if (kind == IG_WAVE_TAB_HOLD) {
      v = tv[lo];
    } else {
      double fr = (t - tt[lo]) / (tt[hi] - tt[lo]);
      v = tv[lo] + fr * (tv[hi] - tv[lo]);
    }
  }
  return AFFINE ? W->amp[k] * v + W->off[k] : v;
}

// ------------------------------------------- FP32-basis mode (IG_PREC_F32)
// The FP32 mode keeps FP64 for positions, velocities, time and the drive
// phase argument; the quantities that only enter as a bounded weight or a
// probability are computed in FP32 from FP64 inputs.  Their error against
// the FP64 path is stated by the tests (a few 1e-7 relative per value).

// Sines, cosines and logarithms of the FP32 mode use the hardware's FP32
// instructions: v_sin_f32 and v_cos_f32 take the angle in revolutions
// (measured on the MI355X over [-1/2, 1) revolutions: absolute error below
// 1.3e-7), v_log_f32 is log2 (relative error below 1.2e-7; every argument
// here is a normal number in (0, 1]).

// the phase om t + ph in revolutions, reduced to [-1/2, 1/2] in FP64
__device__ __forceinline__ float phase_rev(double arg) {
  const double q = arg * (1.0 / IG_TWO_PI);
  return (float)(q - rint(q));
}

// the natural logarithm of a normal x > 0
__device__ __forceinline__ float ln_f(float x) {
  return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}

// the general waveform, out of line: tables and square waves of other
// duty are rare, and inlined per channel they would crowd the step loop
template <bool AFFINE>
__device__ __noinline__ double wave_eval_slow(const Waves *__restrict__ W, const FieldsK &f, int k,
                                              double t) {
  return wave_eval<AFFINE>(W, f, k, t);
}

// the unit waveform in FP32 from the FP64-reduced phase om t + ph: kind
// IG_WAVE_SIN, IG_WAVE_COS or IG_WAVE_SQUARE at duty 0.5 (sin >= 0, i.e.
// the reduced phase >= 0)
__device__ __forceinline__ float fast_unit(int kind, double om, double ph, double t) {
  const float r = phase_rev(om * t + ph);
  if (kind == IG_WAVE_SQUARE) return r >= 0.0f ? 1.0f : -1.0f;
  if (kind == IG_WAVE_COS) return __builtin_amdgcn_cosf(r);
  return __builtin_amdgcn_sinf(r);
}

// An affine drive weight (a w(om t + ph) + b) sc at the lab times
// t0 + m hh, m < M (a 3-D step's t, t + h/2 and t + h with hh = h/2), for w
// sin, cos or square at duty 0.5 (mode): the phase reduced in revolutions
// in FP64 at t0 and advanced by om hh / 2 pi in FP32 (the same phases to
// the FP32 rounding; the square's sign taken from the hardware sine, so a
// phase past half a revolution needs no second reduction), the affine map
// in FP32.
template <int M>
__device__ __forceinline__ void affine_weights_fast(int mode, double om, double ph, double a,
                                                    double b, double sc, double t0, double hh,
                                                    float (&o)[M]) {
  float r = phase_rev(om * t0 + ph);
  const float dr = (float)(om * hh * (1.0 / IG_TWO_PI));
  const float af = (float)(a * sc), bf = (float)(b * sc);
#pragma unroll
  for (int m = 0; m < M; ++m) {
    float u;
    if (mode == IG_WAVE_SQUARE)
      u = __builtin_amdgcn_sinf(r) >= 0.0f ? 1.0f : -1.0f;
    else if (mode == IG_WAVE_COS)
      u = __builtin_amdgcn_cosf(r);
    else
      u = __builtin_amdgcn_sinf(r);
    o[m] = __builtin_fmaf(af, u, bf);
    r = r + dr;
  }
}

// The tune row every active lane of the wave flies (its index read from
// the first active lane, so the row's address derives from the table's
// with a wave-uniform offset and loads through it are scalar), or -1 when
// the lanes fly different rows.
__device__ __forceinline__ int32_t wave_shared_row(int32_t row) {
  const int32_t row0 = __builtin_amdgcn_readfirstlane(row);
  return __all(row == row0) ? row0 : -1;
}

// FP32-basis waveform: sin, cos and square at duty 0.5 by fast_unit; the
// rest as in FP64
template <bool AFFINE>
__device__ __forceinline__ double wave_eval_fast(const Waves *__restrict__ W, const FieldsK &f,
                                                 int k, double t) {
  const int kind = W->kind[k];
  const bool sq = kind == IG_WAVE_SQUARE && W->duty[k] == 0.5;
  if (kind == IG_WAVE_SIN || kind == IG_WAVE_COS || sq) {
    const float v = fast_unit(kind, W->om[k], W->ph[k], t);
    return AFFINE ? W->amp[k] * (double)v + W->off[k] : (double)v;
  }
  return wave_eval_slow<AFFINE>(W, f, k, t);
}

// ---- FP32-basis random draws
// The FP32-basis mode draws whole Philox blocks: its counter n counts
// 32-bit words and every consumer takes the next block(s), n advancing by
// 4 per block (the FP64 modes' double draws take a word pair each).
// Uniforms take the top 24 bits of a word, the FP32 significand; the low 8
// bits of three words make one more 24-bit uniform (the bits of a Philox
// word are independent).
__device__ __forceinline__ void rng_block(Rng &r, uint32_t *w) {
  const uint64_t blk = (r.n + 3) >> 2;
  w[0] = (uint32_t)blk;
  w[1] = (uint32_t)(blk >> 32);
  w[2] = 0u;
  w[3] = 0u;
  philox4x32_10(w, r.k0, r.k1);
  r.n = (blk + 1) << 2;
}

// uniform in [0, 1) and in (0, 1] from a word's top 24 bits
__device__ __forceinline__ float u24(uint32_t w) { return (float)(w >> 8) * 5.9604644775390625e-8f; }
__device__ __forceinline__ float u24p(uint32_t w) {
  return (float)((w >> 8) + 1u) * 5.9604644775390625e-8f;
}
// uniform strictly inside (0, 1): the midpoints (k + 1/2) 2^-24
__device__ __forceinline__ float u24h(uint32_t w) {
  return ((float)(w >> 8) + 0.5f) * 5.9604644775390625e-8f;
}
// uniform in [0, 1) from the low bytes of three words
__device__ __forceinline__ float u24_low(uint32_t a, uint32_t b, uint32_t c) {
  return (float)((a & 0xFFu) | ((b & 0xFFu) << 8) | ((c & 0xFFu) << 16)) *
         5.9604644775390625e-8f;
}

// (sin, cos) of the uniformly random angle 2 pi U, U in [0, 1)
__device__ __forceinline__ void azimuth(float u, float &sn, float &cs) {
  sn = __builtin_amdgcn_sinf(u);
  cs = __builtin_amdgcn_cosf(u);
}

// ---- hard sphere, FP32-basis mode, flattened into the lane state machine
//
// The FP32 mode samples the reference's collision model (collision3d.py:
// 265-374) exactly in distribution with fewer draws and shorter branches:
//
// * Collision times.  The reference tests once per step: collide iff
//   U < 1 - exp(-h_n), h_n = sp dt / lam(sp) = dt c_rel(sp) P sigma /
//   (1000 kB T) (the ion speed cancels), at most one collision per step.
//   Here each ion carries a remaining hazard E ~ Exp(1); every step
//   subtracts h_n and the collision is due in the step that takes E to
//   zero or below, after which a fresh E is drawn.  Given survival to step
//   n, P(due at n) = 1 - exp(-h_n) by the memorylessness of Exp(1): the
//   same law of collision steps, with no exponential and no draw per step.
//   c_rel is the exact Maxwell mean relative speed, evaluated in FP32 by a
//   fitted polynomial (relative error below 5e-7, c_rel_ratio_f).
// * Partner velocity.  The reference draws g ~ N(0, sig1d^2 I) and accepts
//   with |w - g| / (sp + 3 sqrt(3) sig1d): the accepted g has density
//   proportional to |w - g| N(g).  Here the proposal is the mixture
//   q(g) ~ (sp + |g|) N(g) -- with weight sp / (sp + <|g|>) a Maxwell
//   draw, otherwise a speed-weighted Maxwell draw (|g| = sig1d
//   sqrt(chi^2_4) = sig1d sqrt(-2 ln(U1 U2)), isotropic direction) -- and
//   the acceptance |w - g| / (sp + |g|) <= 1, so the accepted density is
//   again proportional to |w - g| N(g).  The acceptance rate
//   c_rel(sp) / (sp + <|g|>) is at least 0.69 at every speed (the
//   reference's is about 0.3 for heavy ions), so a due collision takes
//   1.0-1.45 tries on average.
// * Scattering: the reference's impact angle asin(sqrt(U)), uniform
//   azimuth and line-of-centres attenuation, in FP32 unit vectors.
// Logarithms, reciprocals, square roots, sines and cosines use the
// hardware's FP32 instructions (v_log_f32, v_rcp_f32, v_rsq_f32,
// v_sqrt_f32, v_sin_f32, v_cos_f32, each within about one ulp or 1.3e-7
// absolute).  A lane performs one partner try per pass of its kernel loop
// while the other lanes keep stepping.
struct HsTry {
  float wx, wy, wz; // ion velocity in the gas frame
  float sp;         // max(|w|, 1e-7)
  int tries;
};

// per-ion constant of the collision exponent, dt P sigma c_bar / (1000 kB
// T): the step's hazard is this times c_rel(sp) / c_bar
__device__ __forceinline__ float hs_rate_const(const FlightK &p, double dt, double dt_to_us,
                                               double sigma) {
  return (float)(dt * dt_to_us * p.P_pa * sigma / (1000.0 * IG_KB * p.T_k) * p.c_bar);
}

// an Exp(1) variate from a word
__device__ __forceinline__ float exp1_of(uint32_t w) { return -ln_f(u24p(w)); }

// c_rel(sp) / c_bar as a function of s = sp / c* (collision3d.py:265-272),
// from fits by ion_gym_amd/csrc/fit_c_rel.py: for s < 1 (heavy ions in a
// light gas nearly always; a wave whose lanes all are takes only this
// branch) the degree-5 polynomial in y = 2 s^2 - 1
__device__ __forceinline__ float c_rel_small_f(float s) {
  const float y = __builtin_fmaf(2.0f * s, s, -1.0f);
  float q = 1.876325837e-06f;
  q = __builtin_fmaf(q, y, -3.035061573e-05f);
  q = __builtin_fmaf(q, y, 4.534179993e-04f);
  q = __builtin_fmaf(q, y, -6.771401594e-03f);
  q = __builtin_fmaf(q, y, 1.516326654e-01f);
  return __builtin_fmaf(q, y, 1.158889719e+00f);
}

// for 1 <= s < 4 the degree-12 polynomial in x = s^2 / 8 - 1
__device__ __forceinline__ float c_rel_mid_f(float s) {
  const float x = __builtin_fmaf(s * s, 0.125f, -1.0f);
  float q = -7.843976244e-04f;
  q = __builtin_fmaf(q, x, 1.642305600e-03f);
  q = __builtin_fmaf(q, x, -1.029022523e-03f);
  q = __builtin_fmaf(q, x, 1.949520542e-03f);
  q = __builtin_fmaf(q, x, -6.028446949e-03f);
  q = __builtin_fmaf(q, x, 1.072184697e-02f);
  q = __builtin_fmaf(q, x, -1.747645726e-02f);
  q = __builtin_fmaf(q, x, 3.065115064e-02f);
  q = __builtin_fmaf(q, x, -5.550525405e-02f);
  q = __builtin_fmaf(q, x, 1.078872557e-01f);
  q = __builtin_fmaf(q, x, -2.546248822e-01f);
  q = __builtin_fmaf(q, x, 1.174990659e+00f);
  return __builtin_fmaf(q, x, 2.663291561e+00f);
}

// above s = 4, (sqrt(pi)/2)(s + 1/(2 s)), relative error below 3e-11
__device__ __forceinline__ float c_rel_ratio_f(float s) {
  if (s < 1.0f) return c_rel_small_f(s);
  if (s >= 4.0f) return 0.886226925452758f * (s + 0.5f * __builtin_amdgcn_rcpf(s));
  return c_rel_mid_f(s);
}

// One step of the hazard clock.  hz is the remaining Exp(1) hazard (NaN:
// none drawn yet).  Returns true when the collision is due this step and
// fills the try state.
__device__ __forceinline__ bool hs_hazard_step(Rng &rng, float &hz, double vx, double vy,
                                               double vz, const FlightK &p, float kconst,
                                               float inv_c_star, HsTry &st) {
  if (!(hz >= 0.0f)) {
    uint32_t w[4];
    rng_block(rng, w);
    hz = exp1_of(w[0]);
  }
  // (in still gas v - 0 is v exactly)
  const float wx = (float)(vx - p.vg[0]);
  const float wy = (float)(vy - p.vg[1]);
  const float wz = (float)(vz - p.vg[2]);
  const float sp = fmaxf(__builtin_amdgcn_sqrtf(wx * wx + wy * wy + wz * wz), 1e-7f);
  hz -= kconst * c_rel_ratio_f(sp * inv_c_star);
  if (hz > 0.0f) return false;
  st.wx = wx;
  st.wy = wy;
  st.wz = wz;
  st.sp = sp;
  st.tries = 0;
  return true;
}

// One partner try on two Philox blocks (8 words): w0 acceptance, w1-w4
// the proposal (two Box-Muller pairs for the Maxwell branch; U1, U2,
// cos theta and azimuth for the speed-weighted one), w5 the next hazard,
// w6 and w7 the scattered direction (its cosine and azimuth), and the low
// bytes of w1-w3 the mixture choice.  Returns 1 when the collision is complete (velocity
// updated, a fresh hazard drawn), 0 to try again, -1 when
// MAX_PARTNER_TRIES (collision3d.py:72) ran out.
__device__ __forceinline__ int hs_try_flat(Rng &rng, HsTry &st, float &hz, double &vx,
                                           double &vy, double &vz, const FlightK &p,
                                           double m_ion) {
  uint32_t w[8];
  rng_block(rng, w);
  rng_block(rng, w + 4);
  const float sig = (float)p.sig1d;
  const float gmean = sig * 1.5957691216057308f; // <|g|> = sig1d sqrt(8 / pi)
  float gx, gy, gz, gl;
  if (u24_low(w[1], w[2], w[3]) * (st.sp + gmean) < st.sp) {
    // Maxwell: three normals from two Box-Muller pairs
    const float r1 = sig * __builtin_amdgcn_sqrtf(-2.0f * ln_f(u24p(w[1])));
    const float r2 = sig * __builtin_amdgcn_sqrtf(-2.0f * ln_f(u24p(w[3])));
    float s1, c1, s2, c2;
    azimuth(u24(w[2]), s1, c1);
    azimuth(u24(w[4]), s2, c2);
    gx = r1 * c1;
    gy = r1 * s1;
    gz = r2 * c2;
    gl = __builtin_amdgcn_sqrtf(gx * gx + gy * gy + gz * gz);
  } else {
    // speed-weighted Maxwell: |g| = sig1d sqrt(chi^2_4), isotropic
    gl = sig * __builtin_amdgcn_sqrtf(-2.0f * ln_f(u24p(w[1]) * u24p(w[2])));
    const float c = __builtin_fmaf(2.0f, u24(w[3]), -1.0f);
    float sn, cs;
    azimuth(u24(w[4]), sn, cs);
    const float one_m = 1.0f - c * c;
    const float sr = gl * __builtin_amdgcn_sqrtf(one_m > 0.0f ? one_m : 0.0f);
    gx = sr * cs;
    gy = sr * sn;
    gz = gl * c;
  }
  const float rx = st.wx - gx, ry = st.wy - gy, rz = st.wz - gz;
  const float rlen = __builtin_amdgcn_sqrtf(rx * rx + ry * ry + rz * rz);
  if (!(u24(w[0]) * (st.sp + gl) < rlen)) {
    return ++st.tries >= 10000 ? -1 : 0;
  }
  hz = exp1_of(w[5]);
  if (rlen < 1e-12f) return 1;
  // The reference reflects the relative velocity r about the line of
  // centres rh, drawn at angle asin(sqrt(U)) from r with a uniform azimuth
  // (collision3d.py:310-330): r' = r - 2 (r . rh) rh.  Its angle chi to r
  // then has cos chi = 2U - 1, uniform on [-1, 1] (the reference scales U
  // by 0.999999999), and a uniform azimuth:
  // r' is |r| times a uniform unit vector (hard spheres scatter
  // isotropically in the centre-of-mass frame), drawn here directly from
  // the same two uniforms.  The ion takes m_gas / (m_ion + m_gas) of the
  // change, the reference's v - f rh with f = 2 m_gas / (m_ion + m_gas)
  // (r . rh).
  const float c = __builtin_fmaf(2.0f, u24(w[6]), -1.0f);
  const float one_m = 1.0f - c * c;
  const float s = rlen * __builtin_amdgcn_sqrtf(one_m > 0.0f ? one_m : 0.0f);
  float sth, cth;
  azimuth(u24(w[7]), sth, cth);
  const double mf = p.m_gas / (m_ion + p.m_gas);
  vx = (vx - p.vg[0] + mf * (double)(s * cth - rx)) + p.vg[0];
  vy = (vy - p.vg[1] + mf * (double)(s * sth - ry)) + p.vg[1];
  vz = (vz - p.vg[2] + mf * (double)(rlen * c - rz)) + p.vg[2];
  return 1;
}

// the SDS jump with the table interpolation, the log-log mass-ratio blend
// and the direction in FP32, on one Philox block per step; rj0 is the
// ion's step-independent factor sqrt(V / mfp dt / 1e5) mfp of the jump
// length (sds.py:401, computed once per ion by sds_rj0).  The log-log
// blend 10^(log10 lo + t (log10 hi - log10 lo)) is the same function in
// base 2, on the hardware's v_log_f32 and v_exp_f32.
__device__ __forceinline__ double sds_rj0(double V, double mfp, double dt) {
  return sqrt(V / mfp * dt / 100000.0) * mfp;
}
__device__ __forceinline__ void sds_jump_fast(Rng &rng, const double *__restrict__ st,
                                              double log_mr, double rj0, double &jx, double &jy,
                                              double &jz) {
  uint32_t wd[4];
  rng_block(rng, wd);
  // a quantile strictly inside (0, 1): the table's end quantiles are 0
  // and its largest step count, and the logarithm of a 0 quantile is -inf
  float x = u24h(wd[0]) * (float)(1002 - 2);
  int i = (int)x;
  if (i > 999) i = 999;
  float w = x - (float)i;
  int k = (int)floor(log_mr);
  k = k < 0 ? 0 : (k > 3 ? 3 : k);
  float t = (float)(log_mr - (double)k);
  const double *a = st + k * 1002, *b = a + 1002;
  float lo = (float)a[i] + w * ((float)a[i + 1] - (float)a[i]);
  float hi = (float)b[i] + w * ((float)b[i + 1] - (float)b[i]);
  const float llo = __builtin_amdgcn_logf(lo);
  const float dsteps = __builtin_amdgcn_exp2f(__builtin_fmaf(t, __builtin_amdgcn_logf(hi) - llo, llo));
  double rj = rj0 * (double)dsteps;
  float c = __builtin_fmaf(2.0f, u24(wd[1]), -1.0f);
  float one_m = 1.0f - c * c;
  float sn, cs;
  azimuth(u24(wd[2]), sn, cs);
  double s = rj * (double)__builtin_sqrtf(one_m > 0.0f ? one_m : 0.0f);
  jx = s * (double)cs;
  jy = s * (double)sn;
  jz = rj * (double)c;
}

} // namespace ig
