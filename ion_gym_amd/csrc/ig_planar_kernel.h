// Planar flight kernel for gfx950: the batched, device-resident restatement
// of ion_gym/physics/build_planar.py:1111-1429 (_fly_planar).
//
// A 2-D transverse field E(x, y, t) = E_A + sum_k w_k(t) E_k on a node grid
// (V/m), no field along z; the ion state is 3-D.  Per step: kick-drift-
// kick leapfrog (SDS: the damped acceleration factor (E - (v - v_gas)
// gamma) in each half kick, z relaxing under the gas only), t += dt, the
// SDS jump or the hard-sphere step on the full velocity, then nearest-node
// metal with the 40-step bisection (x, y, vx, vy, t interpolated), the box,
// the station planes, the six bounds with their face, and the record when
// int(t / dt) % rec_every == 0.  FP64 arithmetic in the reference's order
// (contraction off); the FP32-basis mode as in the r-z kernel (FP32
// interpolation and waveforms, hazard-clock hard sphere, reciprocal
// pitch).  As in the r-z kernel, the first field evaluation of a step is
// the second of the previous one (same point and lab time), so it is
// reused bitwise.
#pragma once
#pragma clang fp contract(off)
#include "ig_rz_kernel.h"

namespace ig {

// the planar record row (build_planar.py:1162-1214)
__device__ __noinline__ void planar_write_row(double *row, uint32_t mask, double t, double x,
                                              double y, double z, double vx, double vy, double vz,
                                              double exv, double eyv, int64_t ncol, double m_ion,
                                              double axx, double axy) {
  row[0] = t;
  row[1] = x;
  row[2] = y;
  row[3] = z;
  row[4] = vx;
  row[5] = vy;
  row[6] = vz;
  int c = 7;
  const double sp = sqrt(vx * vx + vy * vy + vz * vz);
  if (mask & IG_REC_SPEED) row[c++] = sp;
  if (mask & IG_REC_KE) {
    const double s = sp * 1e3;
    row[c++] = 0.5 * m_ion * IG_KG_AMU * (s * s) / IG_E_CHG;
  }
  if (mask & IG_REC_KE_X) {
    const double s = vx * 1e3;
    row[c++] = 0.5 * m_ion * IG_KG_AMU * (s * s) / IG_E_CHG;
  }
  if (mask & IG_REC_KE_Y) {
    const double s = vy * 1e3;
    row[c++] = 0.5 * m_ion * IG_KG_AMU * (s * s) / IG_E_CHG;
  }
  if (mask & IG_REC_KE_Z) {
    const double s = vz * 1e3;
    row[c++] = 0.5 * m_ion * IG_KG_AMU * (s * s) / IG_E_CHG;
  }
  if (mask & IG_REC_E_FIELD) row[c++] = sqrt(exv * exv + eyv * eyv);
  if (mask & IG_REC_E_AXIAL) row[c++] = exv;
  if (mask & IG_REC_E_RADIAL) row[c++] = eyv;
  if (mask & IG_REC_E_X) row[c++] = exv;
  if (mask & IG_REC_E_Y) row[c++] = eyv;
  if (mask & IG_REC_E_Z) row[c++] = 0.0;
  if (mask & IG_REC_RADIUS) row[c++] = sqrt((x - axx) * (x - axx) + (y - axy) * (y - axy));
  if (mask & IG_REC_N_COL) row[c++] = (double)ncol;
}

template <int PREC, int NS, int GAS, bool REC, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) void planar_fly_kernel(
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
  const double mm = f.h;
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
  const int64_t nx = f.n0, ny = f.n1;
  const int K = f.K;
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
  double sds_damp = 0.0, sds_mfp = 0.0, sds_V = 0.0, sds_logmr = 0.0, sds_fac = 1.0;
  if constexpr (GAS == IG_GAS_SDS) {
    sds_damp = e.sds_damp[i];
    sds_mfp = e.sds_mfp[i];
    sds_V = e.sds_V[i];
    sds_logmr = e.sds_logmr[i];
    const double tt = sds_damp * dt;
    sds_fac = (1.0 - exp(-tt)) / tt;
  }
  // the FP32-basis jump's per-ion length factor
  const double sds_rj = GAS == IG_GAS_SDS && sds_mfp > 0.0 ? sds_rj0(sds_V, sds_mfp, dt) : 0.0;
  const double sigma = GAS == IG_GAS_HS ? e.sigma[i] : 0.0;
  constexpr bool kFlat = PREC == IG_PREC_F32 && GAS == IG_GAS_HS;
  float hs_k = 0.0f, hs_inv_cstar = 0.0f, hz = 0.0f;
  HsTry hst{};
  if constexpr (kFlat) {
    hs_k = hs_rate_const(p, dt, 1.0, sigma);
    hs_inv_cstar = (float)(1.0 / p.c_star);
    hz = e.hz[i];
  }
  // the radius channel's axis: the domain centre (build_planar.py:1140-1141)
  const double axx = 0.5 * (nx - 1) * mm, axy = 0.5 * (ny - 1) * mm;

  using CT = typename std::conditional<PREC == IG_PREC_F32, float, double>::type;
  CT cz[NS], cr[NS];
  double cgx = __builtin_nan(""), cgy = __builtin_nan("");
  CT wk[NS > 1 ? NS - 1 : 1];
  double wtt = __builtin_nan("");
  const double inv_mm = 1.0 / mm;
  auto node = [&](double v) { return PREC == IG_PREC_F32 ? v * inv_mm : v / mm; };
  // _efield (build_planar.py:1147-1157): V/m at (px, py), lab time tt
  auto efield = [&](double px, double py, double tt, double &ex_o, double &ey_o) {
    const double gx = node(px), gy = node(py);
    if (!(gx == cgx && gy == cgy)) {
      rz_gather<PREC, NS>(f, gx, gy, cz, cr);
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
      cgy = gy;
    }
    if (!(tt == wtt)) {
      auto wave_at = [&](const Waves *Wp, int k) {
        return PREC == IG_PREC_F32 ? (CT)wave_eval_fast<false>(Wp, f, k, tt)
                                   : (CT)wave_eval<false>(Wp, f, k, tt);
      };
      if (Wu) {
#pragma unroll
        for (int k = 0; k < NS - 1; ++k)
          if (k < K) wk[k] = wave_at(Wu, k);
      } else {
#pragma unroll
        for (int k = 0; k < NS - 1; ++k)
          if (k < K) wk[k] = wave_at(W, k);
      }
      wtt = tt;
    }
    CT ex = cz[0], ey = cr[0];
#pragma unroll
    for (int k = 0; k < NS - 1; ++k) {
      if (k < K) {
        if constexpr (PREC == IG_PREC_F32) {
          ex = __builtin_fmaf(wk[k], cz[1 + k], ex);
          ey = __builtin_fmaf(wk[k], cr[1 + k], ey);
        } else {
          ex = ex + wk[k] * cz[1 + k];
          ey = ey + wk[k] * cr[1 + k];
        }
      }
    }
    ex_o = (double)ex;
    ey_o = (double)ey;
  };
  const double sc = 1e-9 * acc; // ex *= 1e-9 * acc
  // the field at the current position for the record, V/m
  double fxr, fyr;
  if (steps0 == 0) {
    efield(x, y, tob, fxr, fyr);
    if constexpr (REC) {
      planar_write_row(e.rec + (size_t)i * p.rec_rows * p.rec_cols, p.rec_mask, tob, x, y, z, vx,
                       vy, vz, fxr * 1e-3, fyr * 1e-3, ncol, m_ion, axx, axy);
      nrec = 1;
    }
  } else {
    fxr = e.e_last[3 * i];
    fyr = e.e_last[3 * i + 1];
  }

  int fate = IG_FATE_FLYING, face = -1;
  int64_t s = 0;
  const int64_t lim = p.max_steps;
  int stage = kStep;
  double xo = x, yo = y, zo = z, vxo = vx, vyo = vy, to = t;
  NodeProbe metal_next{0u, false};
  for (;;) {
    if (stage == kStep) {
      if (!(t < t_max)) {
        fate = IG_FATE_TIMEOUT;
        stage = kDone;
      } else if (lim > 0 && s >= lim) {
        stage = kDone;
      } else {
        xo = x;
        yo = y;
        zo = z;
        vxo = vx;
        vyo = vy;
        to = t;
        double ex, ey;
        efield(x, y, tob + t, ex, ey);
        ex = ex * sc;
        ey = ey * sc;
        double az = 0.0;
        if constexpr (GAS == IG_GAS_SDS) {
          ex = sds_fac * (ex - (vx - p.vg[0]) * sds_damp);
          ey = sds_fac * (ey - (vy - p.vg[1]) * sds_damp);
          az = -sds_fac * (vz - p.vg[2]) * sds_damp;
        }
        vx += 0.5 * ex * dt;
        vy += 0.5 * ey * dt;
        vz += 0.5 * az * dt;
        x += vx * dt;
        y += vy * dt;
        z += vz * dt;
        // the post-step metal test's node load, issued beside the field
        // gathers of the same point (without an SDS jump the ion does not
        // move before that test)
        if constexpr (GAS != IG_GAS_SDS)
          metal_next = metal_nn_probe(f.metal, node(x), node(y), f.dnm1[0], f.dnm1[1], ny);
        efield(x, y, tob + t + dt, ex, ey);
        fxr = ex;
        fyr = ey;
        ex = ex * sc;
        ey = ey * sc;
        if constexpr (GAS == IG_GAS_SDS) {
          ex = sds_fac * (ex - (vx - p.vg[0]) * sds_damp);
          ey = sds_fac * (ey - (vy - p.vg[1]) * sds_damp);
          az = -sds_fac * (vz - p.vg[2]) * sds_damp;
        }
        vx += 0.5 * ex * dt;
        vy += 0.5 * ey * dt;
        vz += 0.5 * az * dt;
        t += dt;
        s += 1;
        stage = kPost;
        if constexpr (GAS == IG_GAS_SDS) {
          if (sds_mfp > 0.0) {
            double jx, jy, jz;
            if constexpr (PREC == IG_PREC_F32)
              sds_jump_fast(rng, p.sds_stats, sds_logmr, sds_rj, jx, jy, jz);
            else
              sds_jump(rng, p.sds_stats, sds_logmr, sds_V, sds_mfp, dt, jx, jy, jz);
            x += jx;
            y += jy;
            z += jz;
          }
        } else if constexpr (kFlat) {
          if (hs_hazard_step(rng, hz, vx, vy, vz, p, hs_k, hs_inv_cstar, hst)) stage = kCollide;
        } else if constexpr (GAS == IG_GAS_HS) {
          if (hs_test(rng, vx, vy, vz, p, dt, 1.0, sigma)) stage = kCollide;
        }
      }
    }
    if constexpr (kFlat) {
      const uint64_t parked = __ballot(stage == kCollide);
      if (parked) {
        const uint64_t moving = __ballot(stage == kStep || stage == kPost);
        if (__popcll(parked) >= p.coll_batch || __popcll(moving) < p.coll_batch) {
          if (stage == kCollide) {
            const int rc = hs_try_flat(rng, hst, hz, vx, vy, vz, p, m_ion);
            if (rc > 0) {
              ncol += 1;
              stage = kPost;
            } else if (rc < 0) {
              fate = -3;
              stage = kDone;
            }
          }
        }
      }
    } else if constexpr (GAS == IG_GAS_HS) {
      const uint64_t parked = __ballot(stage == kCollide);
      if (parked) {
        const uint64_t moving = __ballot(stage == kStep || stage == kPost);
        if (__popcll(parked) >= p.coll_batch || __popcll(moving) < p.coll_batch) {
          if (stage == kCollide) {
            if (hs_scatter(rng, vx, vy, vz, p, m_ion)) {
              ncol += 1;
              stage = kPost;
            } else {
              fate = -3;
              stage = kDone;
            }
          }
        }
      }
    }
    if (stage == kPost) {
      stage = kStep;
      const bool hit =
          GAS != IG_GAS_SDS ? metal_next.hit() : metal_nn(f.metal, node(x), node(y), nx, ny);
      if (hit) {
        double f0 = 0.0, f1 = 1.0;
        for (int b = 0; b < 40; ++b) {
          const double fm = 0.5 * (f0 + f1);
          if (metal_nn(f.metal, node(xo + fm * (x - xo)), node(yo + fm * (y - yo)), nx, ny))
            f1 = fm;
          else
            f0 = fm;
        }
        x = xo + f1 * (x - xo);
        y = yo + f1 * (y - yo);
        vx = vxo + f1 * (vx - vxo);
        vy = vyo + f1 * (vy - vyo);
        t = to + f1 * dt;
        fate = IG_FATE_IMPACT;
        stage = kDone;
      } else if (x < 0 || y < 0 || node(x) > f.dnm1[0] || node(y) > f.dnm1[1]) {
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
          x = xc;
          y = yc;
          z = zc;
          vx = vxo + fr * (vx - vxo);
          vy = vyo + fr * (vy - vyo);
          t = to + fr * dt;
          fate = p.pl_kind[ip];
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
        if constexpr (REC) {
          if (stage != kDone && ((int64_t)(t / dt)) % p.rec_every == 0 && nrec < p.rec_rows - 1) {
            planar_write_row(e.rec + ((size_t)i * p.rec_rows + nrec) * p.rec_cols, p.rec_mask,
                             tob + t, x, y, z, vx, vy, vz, fxr * 1e-3, fyr * 1e-3, ncol, m_ion,
                             axx, axy);
            nrec += 1;
          }
        }
      }
    }
    if (__ballot(stage != kDone) == 0) break;
  }
  if constexpr (REC) {
    if (fate != IG_FATE_FLYING) {
      planar_write_row(e.rec + ((size_t)i * p.rec_rows + nrec) * p.rec_cols, p.rec_mask, tob + t,
                       x, y, z, vx, vy, vz, fxr * 1e-3, fyr * 1e-3, ncol, m_ion, axx, axy);
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
  e.label[i] = 0;
  e.ncol[i] = ncol;
  e.steps[i] = steps0 + s;
  e.nrec[i] = nrec;
  e.rng_n[i] = rng.n;
  e.has_gauss[i] = rng.has_gauss;
  e.gauss[i] = rng.gauss;
  e.e_last[3 * i] = fxr;
  e.e_last[3 * i + 1] = fyr;
  if constexpr (kFlat) e.hz[i] = hz;
}

template <int PREC, int NS, int GAS>
static void launch_planar_gas(const FieldsK &f, const FlightK &p, const EnsK &e, dim3 grid,
                              int block, hipStream_t st) {
  if (p.rec_rows > 0 && e.rec)
    hipLaunchKernelGGL((planar_fly_kernel<PREC, NS, GAS, true, 1>), grid, dim3(block), 0, st, f,
                       p, e);
  else if (p.occupancy == 2)
    hipLaunchKernelGGL((planar_fly_kernel<PREC, NS, GAS, false, 2>), grid, dim3(block), 0, st,
                       f, p, e);
  else if (PREC == IG_PREC_F32 && GAS == IG_GAS_VACUUM && p.occupancy == 4)
    // a 128-VGPR budget, for the FP32-basis vacuum step only (measured:
    // 10 % faster on the planar guide; the gas kernels spill at it)
    hipLaunchKernelGGL((planar_fly_kernel<PREC, NS, GAS, false,
                                          PREC == IG_PREC_F32 && GAS == IG_GAS_VACUUM ? 4 : 3>),
                       grid, dim3(block), 0, st, f, p, e);
  else if (p.occupancy >= 3)
    hipLaunchKernelGGL((planar_fly_kernel<PREC, NS, GAS, false, 3>), grid, dim3(block), 0, st,
                       f, p, e);
  else
    hipLaunchKernelGGL((planar_fly_kernel<PREC, NS, GAS, false, 1>), grid, dim3(block), 0, st,
                       f, p, e);
}

template <int PREC, int NS>
static int launch_planar_ns(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch,
                            int block, hipStream_t st) {
  dim3 grid((unsigned)((n_launch + block - 1) / block));
  if (p.gas == IG_GAS_HS)
    launch_planar_gas<PREC, NS, IG_GAS_HS>(f, p, e, grid, block, st);
  else if (p.gas == IG_GAS_SDS)
    launch_planar_gas<PREC, NS, IG_GAS_SDS>(f, p, e, grid, block, st);
  else
    launch_planar_gas<PREC, NS, IG_GAS_VACUUM>(f, p, e, grid, block, st);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// planar packs carry 1 + K sets: 1..5 covers up to four drive channels
template <int PREC>
int launch_planar_prec(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch,
                       int block, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  switch (f.n_sets) {
  case 1: return launch_planar_ns<PREC, 1>(f, p, e, n_launch, block, st);
  case 2: return launch_planar_ns<PREC, 2>(f, p, e, n_launch, block, st);
  case 3: return launch_planar_ns<PREC, 3>(f, p, e, n_launch, block, st);
  case 4: return launch_planar_ns<PREC, 4>(f, p, e, n_launch, block, st);
  case 5: return launch_planar_ns<PREC, 5>(f, p, e, n_launch, block, st);
  default: return -2;
  }
}

} // namespace ig
