// r-z flight kernel for gfx950: the batched, device-resident restatement of
// ion_gym/physics/tracer_rz.py:113-446 (_fly_rec_full).
//
// One ion per lane.  The ion's state lives in registers for a whole window
// of steps and in HBM (structure of arrays) between windows.  The kick-drift-
// kick leapfrog, the field composition E_A + sum_k w_k(t) E_k (+ gate), the
// record cadence, the 40-step nearest-node impact bisection, the box,
// station and bound tests and their order are the reference's, with FP64
// arithmetic evaluated in the reference's order (contraction off).
//
// Field evaluation reuse: in vacuum and hard-sphere flight the first field
// evaluation of step n+1 is at the same point as the second evaluation of
// step n (tracer_rz.py:216-217 after :234-235; a collision changes only the
// velocity).  The kernel keeps the interpolated per-set values of the last
// gather and re-gathers only when the sample point differs bitwise, and
// keeps the waveform weights and re-evaluates them only when the lab time
// differs bitwise.  Every value the reference computes is therefore
// computed from identical inputs: the reuse changes no result, and it halves
// the gathers per step.
#pragma once
#pragma clang fp contract(off)
#include "ig_device.h"
#include "ig_metal.h"
#include <type_traits>

namespace ig {

template <int PREC>
struct StoreT {
  using T = double;
};
template <>
struct StoreT<IG_PREC_F32_STORE> {
  using T = float;
};
template <>
struct StoreT<IG_PREC_F32> {
  using T = float;
};

// values per node record for NS sets of NC components, a whole number of
// 16-byte units (the host lays records out with the same rule)
template <typename T, int NS, int NC>
struct RecLayout {
  static constexpr int kVals = NS * NC;
  static constexpr int kPer16 = 16 / (int)sizeof(T);
  static constexpr int kStride = ((kVals + kPer16 - 1) / kPer16) * kPer16;
};

template <typename T, int N>
__device__ __forceinline__ void load_vals(const T *__restrict__ p, T (&v)[N]) {
  if constexpr (sizeof(T) == 4) {
    static_assert(N % 4 == 0, "float record");
#pragma unroll
    for (int q = 0; q < N / 4; ++q) {
      float4 w = reinterpret_cast<const float4 *>(p)[q];
      v[4 * q] = w.x;
      v[4 * q + 1] = w.y;
      v[4 * q + 2] = w.z;
      v[4 * q + 3] = w.w;
    }
  } else {
    static_assert(N % 2 == 0, "double record");
#pragma unroll
    for (int q = 0; q < N / 2; ++q) {
      double2 w = reinterpret_cast<const double2 *>(p)[q];
      v[2 * q] = w.x;
      v[2 * q + 1] = w.y;
    }
  }
}

// interp._bilin (interp.py:19-36) for every set of the records around
// (gx, gu): index clamped to the interior, fraction clamped to [0, 1].
template <int PREC, int NS, typename OT>
__device__ __forceinline__ void rz_gather(const FieldsK &f, double gx, double gu, OT (&cz)[NS],
                                          OT (&cr)[NS]) {
  using T = typename StoreT<PREC>::T;
  using L = RecLayout<T, NS, 2>;
  const int64_t nx = f.n0, nu = f.n1;
  if constexpr (PREC == IG_PREC_F32) {
    // int(g) by the saturating hardware conversion, clamped to the
    // interior in integers (the same indices as the FP64 path's pre-clamp,
    // conversion and clamp); the fractions converted to FP32, then clamped
    // to [0, 1] (the same values as clamping in FP64 first: the rounding is
    // monotone and 0 and 1 are exact); 32-bit byte offsets (an FP32 2-D
    // pack stays below 4 GiB, checked at upload)
    const int32_t i = min(max(cvt_i32_sat(gx), 0), (int32_t)(nx - 2));
    const int32_t j = min(max(cvt_i32_sat(gu), 0), (int32_t)(nu - 2));
    const float fxf = __builtin_amdgcn_fmed3f((float)(gx - (double)i), 0.0f, 1.0f);
    const float fuf = __builtin_amdgcn_fmed3f((float)(gu - (double)j), 0.0f, 1.0f);
    constexpr uint32_t kRec = L::kStride * sizeof(T);
    const uint32_t o0 = ((uint32_t)i * (uint32_t)nu + (uint32_t)j) * kRec;
    const uint32_t o1 = o0 + (uint32_t)nu * kRec;
    const char *g = reinterpret_cast<const char *>(f.grid);
    T a00[L::kStride], a01[L::kStride], a10[L::kStride], a11[L::kStride];
    load_vals(reinterpret_cast<const T *>(g + o0), a00);
    load_vals(reinterpret_cast<const T *>(g + o0 + kRec), a01);
    load_vals(reinterpret_cast<const T *>(g + o1), a10);
    load_vals(reinterpret_cast<const T *>(g + o1 + kRec), a11);
    const float w00 = (1.0f - fxf) * (1.0f - fuf), w10 = fxf * (1.0f - fuf);
    const float w01 = (1.0f - fxf) * fuf, w11 = fxf * fuf;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      float vz = __builtin_fmaf(a11[2 * s], w11,
                                __builtin_fmaf(a01[2 * s], w01,
                                               __builtin_fmaf(a10[2 * s], w10, a00[2 * s] * w00)));
      float vr = __builtin_fmaf(
          a11[2 * s + 1], w11,
          __builtin_fmaf(a01[2 * s + 1], w01,
                         __builtin_fmaf(a10[2 * s + 1], w10, a00[2 * s + 1] * w00)));
      cz[s] = (OT)(vz * (float)f.scale[s]);
      cr[s] = (OT)(vr * (float)f.scale[s]);
    }
    return;
  }
  // the pre-clamp keeps the conversion defined and gives int(g) clamped as
  // the reference clamps it; within [-1, n] a 32-bit index is exact
  int32_t i = (int32_t)fmin(fmax(gx, -1.0), f.dn[0]);
  int32_t j = (int32_t)fmin(fmax(gu, -1.0), f.dn[1]);
  // (node counts are below 2^31, checked at upload: 32-bit compares)
  const int32_t imax = (int32_t)(nx - 2), jmax = (int32_t)(nu - 2);
  if (i < 0) i = 0;
  if (j < 0) j = 0;
  if (i > imax) i = imax;
  if (j > jmax) j = jmax;
  double fx = gx - (double)i, fu = gu - (double)j;
  if (fx < 0.0) fx = 0.0;
  if (fx > 1.0) fx = 1.0;
  if (fu < 0.0) fu = 0.0;
  if (fu > 1.0) fu = 1.0;
  const T *g = reinterpret_cast<const T *>(f.grid);
  const T *p0 = g + ((int64_t)i * nu + j) * (int64_t)L::kStride; // (i, j), (i, j+1)
  const T *p1 = p0 + nu * (int64_t)L::kStride;         // (i+1, j), (i+1, j+1)
  T a00[L::kStride], a01[L::kStride], a10[L::kStride], a11[L::kStride];
  load_vals(p0, a00);
  load_vals(p0 + L::kStride, a01);
  load_vals(p1, a10);
  load_vals(p1 + L::kStride, a11);
  const double ax = 1 - fx, au = 1 - fu;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int q = 2 * s + c;
      double v = (double)a00[q] * ax * au + (double)a10[q] * fx * au +
                 (double)a01[q] * ax * fu + (double)a11[q] * fx * fu;
      if constexpr (PREC == IG_PREC_F32_STORE) v = v * f.scale[s];
      if (c == 0)
        cz[s] = v;
      else
        cr[s] = v;
    }
  }
}

// tracer_rz._write_row (tracer_rz.py:40-111)
__device__ __noinline__ void rz_write_row(double *row, uint32_t mask, double t, double x, double y,
                                          double z, double vx, double vy, double vz, double ez,
                                          double er, int64_t ncol, double m_ion, double path,
                                          double eat, double ket) {
  row[0] = t;
  row[1] = x;
  row[2] = y;
  row[3] = z;
  row[4] = vx;
  row[5] = vy;
  row[6] = vz;
  int c = 7;
  double sp = sqrt(vx * vx + vy * vy + vz * vz);
  double r = sqrt(y * y + z * z);
  if (mask & IG_REC_SPEED) row[c++] = sp;
  if (mask & IG_REC_KE) {
    double s = sp * 1e3;
    row[c++] = 0.5 * m_ion * IG_KG_AMU * (s * s) / IG_E_CHG;
  }
  double kes = 0.5 * m_ion * IG_KG_AMU * 1e6 / IG_E_CHG;
  if (mask & IG_REC_KE_X) row[c++] = kes * vx * vx;
  if (mask & IG_REC_KE_Y) row[c++] = kes * vy * vy;
  if (mask & IG_REC_KE_Z) row[c++] = kes * vz * vz;
  if (mask & IG_REC_E_FIELD) row[c++] = sqrt(ez * ez + er * er);
  if (mask & IG_REC_E_AXIAL) row[c++] = ez;
  if (mask & IG_REC_E_RADIAL) row[c++] = er;
  if (mask & IG_REC_RADIUS) row[c++] = r;
  if (mask & IG_REC_N_COL) row[c++] = (double)ncol;
  if (mask & IG_REC_PATH) row[c++] = path;
  if (mask & IG_REC_E_AXIAL_TINT) row[c++] = eat;
  if (mask & IG_REC_KE_TINT) row[c++] = ket;
}


// The flight kernel.  Each lane runs a small state machine per step:
//   kStep    the leapfrog step (kick, drift, kick), t += dt, the record
//            accumulators, then the SDS jump or the hard-sphere test;
//   kCollide a hard-sphere collision is due (the test fired);
//   kPost    record, metal + bisection, box, stations, bounds, x-high
//            (tracer_rz.py:279-440), then back to kStep or done.
// Hard-sphere collisions are rare per lane (dt ~ tau/20) but the
// scattering with its rejection-sampled partner is long; run inline, some
// lane of a 64-lane wave would pay it on nearly every step.  A lane whose
// test fires therefore parks in kCollide while the other lanes keep
// stepping, and the wave scatters all parked lanes together once
// p.coll_batch of them wait (or no lane can step).  Each ion still
// executes exactly the reference's sequence of operations and draws --
// step, test, scatter, checks, step -- only lanes drift apart in step
// count, which the per-lane window bound absorbs.
// OCC bounds the register budget so that at least OCC waves share a SIMD
// (1 = the compiler's choice); the collision pass then keeps what does not
// fit in scratch.  Results do not depend on it.
template <int PREC, int NS, int GAS, bool REC, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) void rz_fly_kernel(
    const FieldsK f, const FlightK p, const EnsK e) {
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t i;
  if (e.idx) {
    if (tid >= (int64_t)*e.n_live) return;
    i = e.idx[tid];
  } else {
    if (tid >= e.n) return;
    i = tid;
  }
  if (e.fate[i] != IG_FATE_FLYING) return;

  double x = e.x[i], y = e.y[i], z = e.z[i];
  double vx = e.vx[i], vy = e.vy[i], vz = e.vz[i];
  double t = e.t[i];
  const double tob = e.tob[i], acc = e.acc[i], m_ion = e.m_amu[i];
  const double mm = f.h, u0 = f.u0;
  // the ion's tune: the pack's waveforms and the flight's scalars, or its
  // row of the ensemble's tune table
  const Waves *W = f.waves;
  double dt = p.dt, t_max = p.t_max;
  uint64_t rkey = p.rng_key;
  const double *tscale = nullptr;
  if (e.tunes) {
    const Tune *T = e.tunes + e.tune[i];
    W = &T->w;
    dt = T->dt;
    t_max = T->t_max;
    rkey = T->rng_key;
    tscale = T->scale;
  }
  // drive parameters every lane of the wave shares are read through one
  // address (scalar loads): the pack's waveforms when the flight is
  // untuned, the tune row when the wave's lanes all fly the same row
  // (rows that differ take a vector load per parameter)
  const int32_t row0 = e.tunes ? wave_shared_row(e.tune[i]) : -1;
  const Tune *const Tu = row0 >= 0 ? e.tunes + row0 : nullptr;
  const Waves *const Wu = e.tunes ? (row0 >= 0 ? &e.tunes[row0].w : nullptr) : f.waves;
  const int64_t nx = f.n0, nu = f.n1;
  const bool gate = f.has_gate && f.tau_gate >= 0.0;
  int64_t ncol = e.ncol[i];
  const int64_t steps0 = e.steps[i];
  int32_t nrec = e.nrec[i];
  Rng rng;
  {
    uint64_t key = (uint64_t)e.seed[i] ^ rkey;
    rng.k0 = (uint32_t)key;
    rng.k1 = (uint32_t)(key >> 32);
    rng.n = e.rng_n[i];
    rng.has_gauss = e.has_gauss[i];
    rng.gauss = e.gauss[i];
    rng.cached = false;
  }
  double path = 0.0, eat = 0.0, ket = 0.0;
  if constexpr (REC) {
    path = e.path[i];
    eat = e.eat[i];
    ket = e.ket[i];
  }
  const double fscale = acc != 0 ? 1e-3 / (1e-9 * acc) : 0.0;
  double sds_damp = 0.0, sds_mfp = 0.0, sds_V = 0.0, sds_logmr = 0.0, sds_fac = 1.0;
  if constexpr (GAS == IG_GAS_SDS) {
    sds_damp = e.sds_damp[i];
    sds_mfp = e.sds_mfp[i];
    sds_V = e.sds_V[i];
    sds_logmr = e.sds_logmr[i];
    if (sds_damp > 0.0) {
      double tt = sds_damp * dt;
      sds_fac = (1.0 - exp(-tt)) / tt;
    }
  }
  // the FP32-basis jump's per-ion length factor
  const double sds_rj = GAS == IG_GAS_SDS && sds_mfp > 0.0 ? sds_rj0(sds_V, sds_mfp, dt) : 0.0;
  const double sigma = GAS == IG_GAS_HS ? e.sigma[i] : 0.0;
  // FP32-basis hard sphere: collision exponent constant and the
  // flattened partner-try state
  constexpr bool kFlat = PREC == IG_PREC_F32 && GAS == IG_GAS_HS;
  float hs_k = 0.0f, hs_inv_cstar = 0.0f, hz = 0.0f;
  HsTry hst{};
  if constexpr (kFlat) {
    hs_k = hs_rate_const(p, dt, 1.0, sigma);
    hs_inv_cstar = (float)(1.0 / p.c_star);
    hz = e.hz[i];
  }

  // cached gather and waveform weights (see the file comment); the
  // FP32-basis mode keeps and composes them in FP32
  using CT = typename std::conditional<PREC == IG_PREC_F32, float, double>::type;
  CT cz[NS], cr[NS];
  double cgx = __builtin_nan(""), cgu = __builtin_nan("");
  CT wk[NS > 1 ? NS - 1 : 1];
  const double acc9 = 1e-9 * acc;
  double wtt = __builtin_nan("");
  // Node coordinates, the radius and the radial direction: the FP64 modes
  // divide and take FP64 square roots in the reference's order; the
  // FP32-basis mode multiplies by the reciprocal pitch and takes the radius
  // and its reciprocal in FP32 with the hardware square root and reciprocal
  // (relative error ~1e-7, the size of the FP32 field values they weight)
  const double inv_mm = 1.0 / mm;
  auto node = [&](double v) { return PREC == IG_PREC_F32 ? v * inv_mm : v / mm; };
  auto radius = [&](double a, double b) {
    if constexpr (PREC == IG_PREC_F32)
      return (double)__builtin_amdgcn_sqrtf((float)(a * a + b * b));
    else
      return sqrt(a * a + b * b);
  };
  // the radial kick components er * y / rr and er * z / rr
  auto radial = [&](double er_, double rr_, double &ky, double &kz) {
    if constexpr (PREC == IG_PREC_F32) {
      const double q = er_ * (double)__builtin_amdgcn_rcpf((float)rr_);
      ky = q * y;
      kz = q * z;
    } else {
      ky = er_ * y / rr_;
      kz = er_ * z / rr_;
    }
  };
  // FP32-basis mode: the last evaluation's radial component in FP32 (the
  // radial kick is taken from it; er is converted only where it is kept)
  float erf_c = 0.0f;
  const float acc9f = (float)acc9;
  // fresh: the point and lab time are known to differ from the cached ones
  // (the step's second evaluation), so the comparisons are skipped
  auto efield = [&](double px, double pr, double tt, double &ez_o, double &er_o, bool fresh) {
    const double gx = node(px);
    const double gu = node(pr - u0);
    if (fresh || !(gx == cgx && gu == cgu)) {
      rz_gather<PREC, NS>(f, gx, gu, cz, cr);
      if (tscale) {
        if (Tu) {
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            cz[q] = cz[q] * (CT)Tu->scale[q];
            cr[q] = cr[q] * (CT)Tu->scale[q];
          }
        } else {
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            cz[q] = cz[q] * (CT)tscale[q];
            cr[q] = cr[q] * (CT)tscale[q];
          }
        }
      }
      cgx = gx;
      cgu = gu;
    }
    if (fresh || !(tt == wtt)) {
      auto wave_at = [&](const Waves *Wp, int k) {
        return PREC == IG_PREC_F32 ? (CT)wave_eval_fast<false>(Wp, f, k, tt)
                                   : (CT)wave_eval<false>(Wp, f, k, tt);
      };
      const int nch = f.has_gate ? NS - 2 : NS - 1; // K, per the pack's set count
      if (Wu) {
#pragma unroll
        for (int k = 0; k < NS - 1; ++k)
          if (k < nch) wk[k] = wave_at(Wu, k);
      } else {
#pragma unroll
        for (int k = 0; k < NS - 1; ++k)
          if (k < nch) wk[k] = wave_at(W, k);
      }
      wtt = tt;
    }
    CT zz = cz[0], rz = cr[0];
    auto add_channel = [&](int k) {
      if constexpr (PREC == IG_PREC_F32) {
        zz = __builtin_fmaf(wk[k], cz[1 + k], zz);
        rz = __builtin_fmaf(wk[k], cr[1 + k], rz);
      } else {
        zz = zz + wk[k] * cz[1 + k];
        rz = rz + wk[k] * cr[1 + k];
      }
    };
    // the pack's sets are 1 + K + has_gate: without a gate every set after
    // the static one is a drive channel (a uniform branch, so the channel
    // count is known in each arm)
    if (!f.has_gate) {
#pragma unroll
      for (int k = 0; k < NS - 1; ++k) add_channel(k);
    } else {
#pragma unroll
      for (int k = 0; k < NS - 2; ++k) add_channel(k);
      if (gate && tt >= f.tau_gate) {
        zz = zz + cz[NS - 1];
        rz = rz + cr[NS - 1];
      }
    }
    if constexpr (PREC == IG_PREC_F32) {
      ez_o = (double)(zz * acc9f);
      erf_c = rz * acc9f;
      (void)er_o;
    } else {
      ez_o = zz * 1e-9 * acc;
      er_o = rz * 1e-9 * acc;
    }
  };
  // er of the last evaluation, in every mode
  auto er_last = [&](double er_) { return PREC == IG_PREC_F32 ? (double)erf_c : er_; };
  // the radial kick components at radius r from the last evaluation: in
  // FP32 from the FP32 radial field and the hardware reciprocal
  auto kick_r = [&](double er_, double r_, double &ky, double &kz) {
    if constexpr (PREC == IG_PREC_F32) {
      const float q = erf_c * __builtin_amdgcn_rcpf(fmaxf((float)r_, 1e-9f));
      ky = (double)q * y;
      kz = (double)q * z;
    } else {
      const double rr = r_ > 1e-9 ? r_ : 1e-9;
      radial(er_, rr, ky, kz);
    }
  };

  double ez, er;
  if (steps0 == 0) {
    // field at the birth point for row 0 (tracer_rz.py:184-202)
    double r0 = radius(y, z);
    efield(x, r0, tob, ez, er, false);
    er = er_last(er);
    if constexpr (REC) {
      rz_write_row(e.rec + (size_t)i * p.rec_rows * p.rec_cols, p.rec_mask, tob, x, y, z, vx, vy,
                   vz, ez * fscale, er_last(er) * fscale, ncol, m_ion, path, eat, ket);
      nrec = 1;
    }
  } else {
    ez = e.e_last[3 * i];
    er = e.e_last[3 * i + 1];
    erf_c = (float)er; // (an FP32 value, kept as a double)
  }

  int fate = IG_FATE_FLYING, face = -1, label = 0;
  int64_t s = 0;
  const int64_t lim = p.max_steps;
  const int coll_batch = p.coll_batch;
  int stage = kStep;
  // sqrt(y^2 + z^2) of the current position while it stays valid (the
  // reference takes it three times per step at two positions; only the
  // drift, the SDS jump and the terminal interpolations move the ion)
  bool rvalid = false;
  double rcur = 0.0;
  // the radial kick components of the last field evaluation; kfresh while
  // the ion has not moved since (only the step's drift, the SDS jump and
  // the terminal interpolations move it), so that a step starting at the
  // same lab time reuses the field and the kick
  bool kfresh = false;
  double ky_last = 0.0, kz_last = 0.0;
  NodeProbe metal_next{0u, false};
  const double half_dt = 0.5 * dt; // the FP32-basis kicks: one FMA each
  // the step's start point, for the impact bisection and the crossings
  double xo = x, yo = y, zo = z, vxo = vx, vyo = vy, vzo = vz, to = t;
  for (;;) {
    if (stage == kStep) {
      if (!(t < t_max)) {
        fate = IG_FATE_TIMEOUT;
        stage = kDone;
      } else if (lim > 0 && s >= lim) {
        stage = kDone; // window ends, the ion keeps flying
      } else {
        xo = x;
        yo = y;
        zo = z;
        vxo = vx;
        vyo = vy;
        vzo = vz;
        to = t;
        double r = rvalid ? rcur : radius(y, z);
        double ky, kz;
        if (kfresh && tob + t == wtt) {
          // the point and lab time of the previous step's second field
          // evaluation: the same inputs, so the same field and kick
          ky = ky_last;
          kz = kz_last;
        } else {
          efield(x, r, tob + t, ez, er, false);
          kick_r(er, r, ky, kz);
        }
        if constexpr (GAS == IG_GAS_SDS) {
          vx += 0.5 * sds_fac * (ez - (vx - p.vg[0]) * sds_damp) * dt;
          vy += 0.5 * sds_fac * (ky - (vy - p.vg[1]) * sds_damp) * dt;
          vz += 0.5 * sds_fac * (kz - (vz - p.vg[2]) * sds_damp) * dt;
        } else if constexpr (PREC == IG_PREC_F32) {
          vx = __builtin_fma(ez, half_dt, vx);
          vy = __builtin_fma(ky, half_dt, vy);
          vz = __builtin_fma(kz, half_dt, vz);
        } else {
          vx += 0.5 * ez * dt;
          vy += 0.5 * ky * dt;
          vz += 0.5 * kz * dt;
        }
        if constexpr (PREC == IG_PREC_F32) {
          x = __builtin_fma(vx, dt, x);
          y = __builtin_fma(vy, dt, y);
          z = __builtin_fma(vz, dt, z);
        } else {
          x += vx * dt;
          y += vy * dt;
          z += vz * dt;
        }
        if constexpr (REC)
          path += sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo));
        r = radius(y, z);
        rcur = r;
        rvalid = true;
        // the post-step metal test reads the nearest node of this same
        // point; issued beside the field gathers, its load overlaps theirs
        // (without an SDS jump the ion does not move before that test)
        if constexpr (GAS != IG_GAS_SDS)
          metal_next = PREC == IG_PREC_F32
                           ? metal_nn_probe_u32(f.metal, node(x), node(r - u0),
                                                (uint32_t)(nx - 1), (uint32_t)(nu - 1),
                                                (uint32_t)nu)
                           : metal_nn_probe(f.metal, node(x), node(r - u0), f.dnm1[0],
                                            f.dnm1[1], nu);
        efield(x, r, tob + t + dt, ez, er, true);
        kick_r(er, r, ky, kz);
        ky_last = ky;
        kz_last = kz;
        kfresh = true;
        if constexpr (GAS == IG_GAS_SDS) {
          vx += 0.5 * sds_fac * (ez - (vx - p.vg[0]) * sds_damp) * dt;
          vy += 0.5 * sds_fac * (ky - (vy - p.vg[1]) * sds_damp) * dt;
          vz += 0.5 * sds_fac * (kz - (vz - p.vg[2]) * sds_damp) * dt;
        } else if constexpr (PREC == IG_PREC_F32) {
          vx = __builtin_fma(ez, half_dt, vx);
          vy = __builtin_fma(ky, half_dt, vy);
          vz = __builtin_fma(kz, half_dt, vz);
        } else {
          vx += 0.5 * ez * dt;
          vy += 0.5 * ky * dt;
          vz += 0.5 * kz * dt;
        }
        t += dt;
        s += 1;
        if constexpr (REC) {
          const double ke_scale = 0.5 * m_ion * IG_KG_AMU * 1e6 / IG_E_CHG;
          double sp = sqrt(vx * vx + vy * vy + vz * vz);
          eat += ez * fscale * dt;
          ket += ke_scale * sp * sp * dt;
        }
        stage = kPost;
        if constexpr (GAS == IG_GAS_SDS) {
          if (sds_mfp > 0.0) {
            double pre = 0.0;
            if constexpr (REC)
              pre = sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo));
            double jx, jy, jz;
            if constexpr (PREC == IG_PREC_F32)
              sds_jump_fast(rng, p.sds_stats, sds_logmr, sds_rj, jx, jy, jz);
            else
              sds_jump(rng, p.sds_stats, sds_logmr, sds_V, sds_mfp, dt, jx, jy, jz);
            x += jx;
            y += jy;
            z += jz;
            rvalid = false;
            kfresh = false;
            if constexpr (REC)
              path += sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo)) - pre;
          }
        } else if constexpr (kFlat) {
          if (hs_hazard_step(rng, hz, vx, vy, vz, p, hs_k, hs_inv_cstar, hst)) stage = kCollide;
        } else if constexpr (GAS == IG_GAS_HS) {
          if (hs_test(rng, vx, vy, vz, p, dt, 1.0, sigma)) stage = kCollide;
        }
      }
    }
    if constexpr (kFlat) {
      // one partner try per pass for every lane whose collision is due,
      // once coll_batch of them are parked, or once fewer than coll_batch
      // lanes can step (at a window's end: waiting for those few would
      // only delay the parked lanes' own remaining steps)
      const uint64_t parked = __ballot(stage == kCollide);
      if (parked) {
        const uint64_t moving = __ballot(stage == kStep || stage == kPost);
        if (__popcll(parked) >= coll_batch || __popcll(moving) < coll_batch) {
          if (stage == kCollide) {
            const int rc = hs_try_flat(rng, hst, hz, vx, vy, vz, p, m_ion);
            if (rc > 0) {
              ncol += 1;
              stage = kPost;
            } else if (rc < 0) {
              fate = -3; // no partner accepted: refused, as the reference raises
              stage = kDone;
            }
          }
        }
      }
    } else if constexpr (GAS == IG_GAS_HS) {
      const uint64_t parked = __ballot(stage == kCollide);
      if (parked) {
        const uint64_t moving = __ballot(stage == kStep || stage == kPost);
        if (__popcll(parked) >= coll_batch || __popcll(moving) < coll_batch) {
          if (stage == kCollide) {
            if (hs_scatter(rng, vx, vy, vz, p, m_ion)) {
              ncol += 1;
              stage = kPost;
            } else {
              fate = -3; // no partner accepted: refused, as the reference raises
              stage = kDone;
            }
          }
        }
      }
    }
    if (stage == kPost) {
      stage = kStep;
      if constexpr (REC) {
        if ((steps0 + s) % p.rec_every == 0 && nrec < p.rec_rows - 1) {
          rz_write_row(e.rec + ((size_t)i * p.rec_rows + nrec) * p.rec_cols, p.rec_mask, tob + t,
                       x, y, z, vx, vy, vz, ez * fscale, er_last(er) * fscale, ncol, m_ion, path, eat,
                       ket);
          nrec += 1;
        }
      }
      const double r = rvalid ? rcur : radius(y, z);
      const bool hit =
          GAS != IG_GAS_SDS ? metal_next.hit() : metal_nn(f.metal, node(x), node(r - u0), nx, nu);
      if (hit) {
        double f0 = 0.0, f1 = 1.0;
        for (int b = 0; b < 40; ++b) {
          double fm = 0.5 * (f0 + f1);
          double ym = yo + fm * (y - yo);
          double zm = zo + fm * (z - zo);
          double rm = radius(ym, zm);
          if (metal_nn(f.metal, node(xo + fm * (x - xo)), node(rm - u0), nx, nu))
            f1 = fm;
          else
            f0 = fm;
        }
        if constexpr (REC)
          path -=
              (1.0 - f1) * sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo));
        x = xo + f1 * (x - xo);
        y = yo + f1 * (y - yo);
        z = zo + f1 * (z - zo);
        vx = vxo + f1 * (vx - vxo);
        vy = vyo + f1 * (vy - vyo);
        vz = vzo + f1 * (vz - vzo);
        t = to + f1 * dt;
        fate = IG_FATE_IMPACT;
        if (f.label) {
          double re = radius(y, z);
          double li = floor(node(x) + 0.5), lj = floor(node(re - u0) + 0.5);
          if (li >= 0.0 && lj >= 0.0 && li <= f.dnm1[0] && lj <= f.dnm1[1])
            label = f.label[(int64_t)li * nu + (int64_t)lj];
        }
        stage = kDone;
      } else if (x < 0 || node(r) > f.dnm2[1]) {
        fate = IG_FATE_BOX;
        stage = kDone;
      } else {
        for (int ip = 0; ip < p.n_pl; ++ip) {
          const int pcol = p.pl_col[ip];
          const double cn = pcol == 0 ? x : (pcol == 1 ? y : z);
          const double co = pcol == 0 ? xo : (pcol == 1 ? yo : zo);
          const double sgn = p.pl_sgn[ip], val = p.pl_val[ip];
          bool crossed;
          if (sgn == 0.0)
            crossed = (cn - val) * (co - val) <= 0.0 && cn != co;
          else
            crossed = (cn - val) * sgn >= 0.0 && (co - val) * sgn < 0.0;
          if (!crossed) continue;
          const double den = cn - co;
          double fr = den == 0.0 ? 0.0 : (val - co) / den;
          if (fr < 0.0) fr = 0.0;
          if (fr > 1.0) fr = 1.0;
          const double xc = xo + fr * (x - xo), yc = yo + fr * (y - yo), zc = zo + fr * (z - zo);
          const double w1 = pcol == 0 ? yc : xc;
          const double w2 = pcol == 2 ? yc : zc;
          const bool ins = p.pl_w[ip][0] <= w1 && w1 <= p.pl_w[ip][1] && p.pl_w[ip][2] <= w2 &&
                           w2 <= p.pl_w[ip][3];
          if (p.pl_kind[ip] == IG_FATE_DETECT) {
            if (!ins) continue;
          } else if (ins) {
            continue;
          }
          if constexpr (REC)
            path -= (1.0 - fr) *
                    sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo));
          x = xc;
          y = yc;
          z = zc;
          vx = vxo + fr * (vx - vxo);
          vy = vyo + fr * (vy - vyo);
          vz = vzo + fr * (vz - vzo);
          t = to + fr * dt;
          fate = p.pl_kind[ip];
          face = ip;
          stage = kDone;
          break;
        }
        if (stage != kDone) {
#pragma unroll
          for (int fc = 0; fc < 6; ++fc) {
            if (stage == kDone || !p.bnd_on[fc]) continue;
            const double v = fc < 2 ? x : (fc < 4 ? y : z);
            if ((fc & 1) ? v > p.bnd_val[fc] : v < p.bnd_val[fc]) {
              fate = IG_FATE_BOUND;
              face = fc;
              stage = kDone;
            }
          }
        }
        if (stage != kDone && node(x) > f.dnm1[0]) {
          fate = IG_FATE_BOX;
          stage = kDone;
        }
      }
    }
    if (__ballot(stage != kDone) == 0) break;
  }
  if constexpr (REC) {
    if (fate != IG_FATE_FLYING) {
      rz_write_row(e.rec + ((size_t)i * p.rec_rows + nrec) * p.rec_cols, p.rec_mask, tob + t, x,
                   y, z, vx, vy, vz, ez * fscale, er_last(er) * fscale, ncol, m_ion, path, eat, ket);
      nrec += 1;
    }
  }
  e.x[i] = x;
  e.y[i] = y;
  e.z[i] = z;
  e.vx[i] = vx;
  e.vy[i] = vy;
  e.vz[i] = vz;
  e.t[i] = t;
  e.fate[i] = fate;
  e.face[i] = face;
  e.label[i] = label;
  e.ncol[i] = ncol;
  e.steps[i] = steps0 + s;
  e.nrec[i] = nrec;
  e.rng_n[i] = rng.n;
  e.has_gauss[i] = rng.has_gauss;
  e.gauss[i] = rng.gauss;
  e.e_last[3 * i] = ez;
  e.e_last[3 * i + 1] = er_last(er);
  if constexpr (REC) {
    e.path[i] = path;
    e.eat[i] = eat;
    e.ket[i] = ket;
  }
  if constexpr (kFlat) e.hz[i] = hz;
}

template <int PREC, int NS, int GAS>
static void launch_rz_gas(const FieldsK &f, const FlightK &p, const EnsK &e, dim3 grid, int block,
                          hipStream_t st) {
  if (p.rec_rows > 0 && e.rec) {
    hipLaunchKernelGGL((rz_fly_kernel<PREC, NS, GAS, true, 1>), grid, dim3(block), 0, st, f, p,
                       e);
  } else if (p.occupancy == 2) {
    hipLaunchKernelGGL((rz_fly_kernel<PREC, NS, GAS, false, 2>), grid, dim3(block), 0, st, f, p,
                       e);
  } else if (p.occupancy >= 3) {
    hipLaunchKernelGGL((rz_fly_kernel<PREC, NS, GAS, false, 3>), grid, dim3(block), 0, st, f, p,
                       e);
  } else {
    hipLaunchKernelGGL((rz_fly_kernel<PREC, NS, GAS, false, 1>), grid, dim3(block), 0, st, f, p,
                       e);
  }
}

template <int PREC, int NS>
static int launch_rz_ns(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch,
                        int block, hipStream_t st) {
  dim3 grid((unsigned)((n_launch + block - 1) / block));
  if (p.gas == IG_GAS_HS)
    launch_rz_gas<PREC, NS, IG_GAS_HS>(f, p, e, grid, block, st);
  else if (p.gas == IG_GAS_SDS)
    launch_rz_gas<PREC, NS, IG_GAS_SDS>(f, p, e, grid, block, st);
  else
    launch_rz_gas<PREC, NS, IG_GAS_VACUUM>(f, p, e, grid, block, st);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// r-z packs carry 1 + K (+ gate) sets; 1..4 covers K <= 3, or K <= 2 with a
// gate, which is every r-z drive the reference builds for the configs
template <int PREC>
int launch_rz_prec(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch, int block,
                   void *stream) {
  hipStream_t st = (hipStream_t)stream;
  switch (f.n_sets) {
  case 1: return launch_rz_ns<PREC, 1>(f, p, e, n_launch, block, st);
  case 2: return launch_rz_ns<PREC, 2>(f, p, e, n_launch, block, st);
  case 3: return launch_rz_ns<PREC, 3>(f, p, e, n_launch, block, st);
  case 4: return launch_rz_ns<PREC, 4>(f, p, e, n_launch, block, st);
  default: return -2;
  }
}

} // namespace ig
