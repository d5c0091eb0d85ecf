// 3-D flight kernel for gfx950: the batched, device-resident restatement of
// ion_gym/physics/tracer3d.py:280-575 (_fly3d).
//
// One ion per lane; classic RK4 with positions in mm, velocities in m/s
// and the step in s, exactly as the reference carries them; field
// E_A + sum_k w_k(t) E_k with the affine waveforms; hard-sphere gas in the
// flow frame; all-eight-corners impact test with 40-step bisection; planes
// and box.  FP64 arithmetic in the reference's order (contraction off).
// The two mid-step RK4 evaluations share one lab time
// (tracer3d.py:403, 408), so their waveform weights are computed once.
// The periodic transporter (tracer3d.py:433-458) shifts an ion crossing
// its accept plane.  The 3-D SDS flight (_fly3d_sds) is xyz_sds_kernel
// below.
#pragma once
#pragma clang fp contract(off)
#include "ig_device.h"
#include "ig_metal.h"

namespace ig {

template <int PREC>
struct StoreT3 {
  using T = double;
};
template <>
struct StoreT3<IG_PREC_F32_STORE> {
  using T = float;
};
template <>
struct StoreT3<IG_PREC_F32> {
  using T = float;
};

template <typename T, int NS>
struct RecLayout3 {
  static constexpr int kPer16 = 16 / (int)sizeof(T);
  static constexpr int kStride = ((NS * 3 + kPer16 - 1) / kPer16) * kPer16;
};

// tracer3d._tri (tracer3d.py:121-151) for all sets, combined with the
// waveform weights as _acc/_efield combine them (tracer3d.py:351-386):
// e = tri(E_A) then e = e + w_k * tri(E_k) for k = 1..K.
template <int PREC, int NS>
__device__ __forceinline__ void xyz_field(const FieldsK &f, double gx, double gy, double gz,
                                          const double *w, int K, const double *tscale,
                                          double &ex, double &ey, double &ez) {
  using T = typename StoreT3<PREC>::T;
  using L = RecLayout3<T, NS>;
  const int64_t nx = f.n0, ny = f.n1, nz = f.n2;
  double x = gx, y = gy, z = gz;
  if (x < 0.0) x = 0.0;
  if (y < 0.0) y = 0.0;
  if (z < 0.0) z = 0.0;
  if (x > nx - 1) x = nx - 1.0;
  if (y > ny - 1) y = ny - 1.0;
  if (z > nz - 1) z = nz - 1.0;
  // the clamped coordinates lie in [0, n - 1] (n below 2^31) or are NaN,
  // which passes the clamps above: the hardware's 32-bit truncation gives
  // the same index, and 0 for NaN (one instruction each; a conversion to
  // int64 is a long sequence)
  int64_t i = cvt_i32_sat(x), j = cvt_i32_sat(y), k = cvt_i32_sat(z);
  if (i > nx - 2) i = nx - 2;
  if (j > ny - 2) j = ny - 2;
  if (k > nz - 2) k = nz - 2;
  const double tx = x - i, ty = y - j, tz = z - k;
  const T *g = reinterpret_cast<const T *>(f.grid);
  const int64_t S = L::kStride;
  const T *p00 = g + ((i * ny + j) * nz + k) * S; // (i, j, k), (i, j, k+1)
  const T *p10 = p00 + ny * nz * S;               // (i+1, j, .)
  const T *p01 = p00 + nz * S;                    // (i, j+1, .)
  const T *p11 = p10 + nz * S;                    // (i+1, j+1, .)
  double e[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    if (s > K) continue;
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int q = 3 * s + c;
      if constexpr (PREC == IG_PREC_F32) {
        const float txf = (float)tx, tyf = (float)ty, tzf = (float)tz;
        float c00 = __builtin_fmaf(p10[q] - p00[q], txf, p00[q]);
        float c10 = __builtin_fmaf(p11[q] - p01[q], txf, p01[q]);
        float c01 = __builtin_fmaf(p10[S + q] - p00[S + q], txf, p00[S + q]);
        float c11 = __builtin_fmaf(p11[S + q] - p01[S + q], txf, p01[S + q]);
        float c0 = __builtin_fmaf(c10 - c00, tyf, c00);
        float c1 = __builtin_fmaf(c11 - c01, tyf, c01);
        v[c] = (double)__builtin_fmaf(c1 - c0, tzf, c0) * f.scale[s];
      } else {
        double c00 = (double)p00[q] * (1 - tx) + (double)p10[q] * tx;
        double c10 = (double)p01[q] * (1 - tx) + (double)p11[q] * tx;
        double c01 = (double)p00[S + q] * (1 - tx) + (double)p10[S + q] * tx;
        double c11 = (double)p01[S + q] * (1 - tx) + (double)p11[S + q] * tx;
        double c0 = c00 * (1 - ty) + c10 * ty;
        double c1 = c01 * (1 - ty) + c11 * ty;
        double r = c0 * (1 - tz) + c1 * tz;
        if constexpr (PREC == IG_PREC_F32_STORE) r = r * f.scale[s];
        v[c] = r;
      }
    }
    if (tscale) {
      v[0] = v[0] * tscale[s];
      v[1] = v[1] * tscale[s];
      v[2] = v[2] * tscale[s];
    }
    if (s == 0) {
      e[0] = v[0];
      e[1] = v[1];
      e[2] = v[2];
    } else {
      e[0] = e[0] + w[s - 1] * v[0];
      e[1] = e[1] + w[s - 1] * v[1];
      e[2] = e[2] + w[s - 1] * v[2];
    }
  }
  ex = e[0];
  ey = e[1];
  ez = e[2];
}

// The cell and fractions of _tri (tracer3d.py:121-151) from FP64 grid
// units: coordinates clamped to the box, the index clamped to n - 2, the
// fractions rounded to FP32 for the FP32-basis blend.
struct Cell3 {
  int32_t i, j, k; // a 3-D pack holds fewer than 2^31 nodes
  float tx, ty, tz;
};

// In integer and FP32 arithmetic: the index by the saturating hardware
// conversion clamped to [0, n - 2], the fraction converted to FP32 and
// clamped to [0, 1].  For every finite point these are the index and
// fraction of clamping the coordinate to [0, n - 1] first (a coordinate
// below 0 gives index 0 and fraction 0, one above n - 1 index n - 2 and
// fraction 1, and the FP32 rounding is monotone with 0 and 1 exact),
// without the FP64 selects and the int64 conversions.
__device__ __forceinline__ int32_t cell_index(double g, int64_t n) {
  return min(max(cvt_i32_sat(g), 0), (int32_t)(n - 2));
}
__device__ __forceinline__ float cell_frac(double g, int32_t i) {
  return __builtin_amdgcn_fmed3f((float)(g - (double)i), 0.0f, 1.0f);
}
__device__ __forceinline__ Cell3 cell3(const FieldsK &f, double gx, double gy, double gz) {
  Cell3 c;
  c.i = cell_index(gx, f.n0);
  c.j = cell_index(gy, f.n1);
  c.k = cell_index(gz, f.n2);
  c.tx = cell_frac(gx, c.i);
  c.ty = cell_frac(gy, c.j);
  c.tz = cell_frac(gz, c.k);
  return c;
}

// trilinear weight of corner (i + di, j + dj, k + dk)
__device__ __forceinline__ float corner_w(const Cell3 &c, int di, int dj, int dk) {
  return (di ? c.tx : 1.0f - c.tx) * (dj ? c.ty : 1.0f - c.ty) * (dk ? c.tz : 1.0f - c.tz);
}

// One corner record (3 NS floats, as float4s) composed with per-set
// weights: acc[comp] = sum over sets of w[set] * E[set][comp], FMAs in
// record order.  compose_vals takes the record from registers,
// compose_rec loads it first.
template <int NS, int M>
__device__ __forceinline__ void compose_vals(const float4 (&rec)[RecLayout3<float, NS>::kStride / 4],
                                             const float (*const (&w)[M])[NS],
                                             float (&acc)[M][3]) {
  constexpr int Q = RecLayout3<float, NS>::kStride / 4;
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m][0] = acc[m][1] = acc[m][2] = 0.0f;
#pragma unroll
  for (int q4 = 0; q4 < Q; ++q4) {
    const float4 v = rec[q4];
    const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int q = 4 * q4 + l;
      if (q < 3 * NS) {
#pragma unroll
        for (int m = 0; m < M; ++m)
          acc[m][q % 3] = __builtin_fmaf((*w[m])[q / 3], vv[l], acc[m][q % 3]);
      }
    }
  }
}
template <int NS, int M>
__device__ __forceinline__ void compose_rec(const float4 *rec, const float (*const (&w)[M])[NS],
                                            float (&acc)[M][3]) {
  float4 r[RecLayout3<float, NS>::kStride / 4];
#pragma unroll
  for (int q4 = 0; q4 < RecLayout3<float, NS>::kStride / 4; ++q4) r[q4] = rec[q4];
  compose_vals<NS, M>(r, w, acc);
}

// The FP32-basis field: each corner record is read as float4s once and
// its sets are composed with the per-set weights wt (waveform x scale)
// before the trilinear blend, so an evaluation is 8 x 3 NS FMAs over
// 8 x S/4 vector loads whatever the number of sets.  The clamps and cell
// index are those of _tri in FP64; the sums are FP32 in a different order
// than the reference's, within the FP32 mode's stated error.
template <int NS>
__device__ __forceinline__ void xyz_field_f32(const FieldsK &f, double gx, double gy, double gz,
                                              const float (&wt)[NS], double &ex, double &ey,
                                              double &ez) {
  constexpr int Q = RecLayout3<float, NS>::kStride / 4;
  const int64_t ny = f.n1, nz = f.n2;
  const Cell3 cl = cell3(f, gx, gy, gz);
  const float4 *g = reinterpret_cast<const float4 *>(f.grid);
  const int64_t r00 = (((int64_t)cl.i * ny + cl.j) * nz + cl.k) * Q, dI = ny * nz * Q,
                dJ = nz * Q;
  const float (*const ws[1])[NS] = {&wt};
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  // the four (i, j) columns one after another, each as its two z-adjacent
  // records (contiguous, 2 S floats): the loads of one column are in flight
  // together without holding all eight corners in registers
#pragma unroll 1
  for (int col = 0; col < 4; ++col) {
#pragma unroll
    for (int dk = 0; dk < 2; ++dk) {
      const int di = (col >> 1) & 1, dj = col & 1;
      const float wc = corner_w(cl, di, dj, dk);
      float c[1][3];
      compose_rec<NS, 1>(g + r00 + di * dI + dj * dJ + dk * Q, ws, c);
      ax = __builtin_fmaf(wc, c[0][0], ax);
      ay = __builtin_fmaf(wc, c[0][1], ay);
      az = __builtin_fmaf(wc, c[0][2], az);
    }
  }
  ex = ax;
  ey = ay;
  ez = az;
}

// One RK4 step's gathers from one cell (FP32 basis): the eight corner
// records are read once and each is composed with the weights of the
// step's three lab times t, t + h/2 and t + h.  The first stage's field,
// at the cell's own point, is returned; the corner sums at t + h/2 and
// t + h go to the lane's slice of LDS (kc[(6 n + m) bs], m = 0..2 for
// t + h/2 and 3..5 for t + h, bs the block size, so a wave's accesses are
// consecutive words) for the later stages, which blend them when their
// point lies in the same cell (an ion moves a small fraction of a cell
// per step).  The FMA chains and the corner order are those of
// xyz_field_f32, so the field is bitwise the one it returns.
constexpr int kCellLdsFloats = 48; // per lane: 8 corners x 2 times x 3
template <int NS>
__device__ __forceinline__ void xyz_cell_f32(const FieldsK &f, const Cell3 &cl,
                                             const float (&w0)[NS], const float (&w1)[NS],
                                             const float (&w2)[NS], double &ex, double &ey,
                                             double &ez, float *kc, int bs) {
  constexpr int Q = RecLayout3<float, NS>::kStride / 4;
  const int64_t ny = f.n1, nz = f.n2;
  const float4 *g = reinterpret_cast<const float4 *>(f.grid);
  const int64_t r00 = (((int64_t)cl.i * ny + cl.j) * nz + cl.k) * Q, dI = ny * nz * Q,
                dJ = nz * Q;
  const float (*const ws[3])[NS] = {&w0, &w1, &w2};
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  // one corner record in flight at a time (the three weight sets and nine
  // sums leave no room for a second), its float4s all loaded before the
  // first is used
#pragma unroll 1
  for (int n = 0; n < 8; ++n) {
    const int col = n >> 1, dk = n & 1, di = (col >> 1) & 1, dj = col & 1;
    const float wc = corner_w(cl, di, dj, dk);
    float c[3][3];
    compose_rec<NS, 3>(g + r00 + di * dI + dj * dJ + dk * Q, ws, c);
    ax = __builtin_fmaf(wc, c[0][0], ax);
    ay = __builtin_fmaf(wc, c[0][1], ay);
    az = __builtin_fmaf(wc, c[0][2], az);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      kc[(6 * n + m) * bs] = c[1][m];
      kc[(6 * n + 3 + m) * bs] = c[2][m];
    }
  }
  ex = ax;
  ey = ay;
  ez = az;
}

// a later stage's field from the kept corner sums (t + h when last, else
// t + h/2), blended at the stage's own fractions in xyz_field_f32's
// corner order
__device__ __forceinline__ void blend_cell_f32(const Cell3 &cl, const float *kc, int bs,
                                               bool last, double &ex, double &ey, double &ez) {
  const float *b = kc + (last ? 3 : 0) * bs;
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
#pragma unroll
  for (int n = 0; n < 8; ++n) {
    const int col = n >> 1, dk = n & 1, di = (col >> 1) & 1, dj = col & 1;
    const float wc = corner_w(cl, di, dj, dk);
    ax = __builtin_fmaf(wc, b[(6 * n + 0) * bs], ax);
    ay = __builtin_fmaf(wc, b[(6 * n + 1) * bs], ay);
    az = __builtin_fmaf(wc, b[(6 * n + 2) * bs], az);
  }
  ex = ax;
  ey = ay;
  ez = az;
}

template <int PREC, int NS, int GAS, int OCC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(OCC))) void xyz_fly_kernel(
    const FieldsK f, const FlightK p, const EnsK e) {
  // FP32 mode: the lane's slice of the corner-sum cache, kCellLdsFloats
  // words per lane of dynamic LDS
  extern __shared__ float xyz_lds[];
  float *const kc = xyz_lds + threadIdx.x;
  const int bs = blockDim.x;
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
  const int64_t steps0 = e.steps[i];
  double x = e.x[i], y = e.y[i], z = e.z[i];
  double vx = e.vx[i], vy = e.vy[i], vz = e.vz[i];
  if (e.nrec[i] == 0) {
    // first entry: births arrive in mm/us; the kernel state is m/s
    // (tracer3d.py:774-775: v0_mm_us * 1e3)
    vx = vx * 1e3;
    vy = vy * 1e3;
    vz = vz * 1e3;
  }
  double t = e.t[i];
  const double tob = e.tob[i], qm = e.acc[i], m_amu = e.m_amu[i];
  const double h = f.h, inv_h = f.inv_h;
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
  const int64_t nx = f.n0, ny = f.n1, nz = f.n2;
  const int K = f.K;
  const bool rec = p.rec_rows > 0 && e.rec != nullptr;
  int64_t ncol = e.ncol[i];
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
  const double sigma = GAS == IG_GAS_HS ? e.sigma[i] : 0.0;
  constexpr int NW = NS > 1 ? NS - 1 : 1;
  double w[NW];
  double wtt = __builtin_nan("");
  // FP32 mode: the per-set composition weights at lab times t0 + m hh,
  // m < M (a step's t, t + h/2, t + h with hh = h/2; M = 1 for one time).
  // A tuned ion reads its tune's packed records (the tune's set scale
  // already in a and b) -- through scalar loads when every lane of the
  // wave flies the same row -- an untuned one the pack's waveforms through
  // the kernel argument (scalar loads); sin, cos and square at duty 0.5 by
  // affine_weights_fast, other drives (tables, other duties) in FP64
  const Tune *tun = e.tunes ? e.tunes + e.tune[i] : nullptr;
  const int32_t row0 = e.tunes ? wave_shared_row(e.tune[i]) : -1;
  const Tune *const tunU = row0 >= 0 ? e.tunes + row0 : nullptr;
  auto weights_f32 = [&](auto &o, double t0, double hh) {
    constexpr int M = sizeof(o) / sizeof(o[0]);
    const Tune *T = tunU ? tunU : tun;
    const Waves *Wq = tunU ? &tunU->w : W;
    const double *sq = tunU ? tunU->scale : tscale;
    uint32_t mws[4] = {0u, 0u, 0u, 0u};
    if (T) {
      const uint4 mw = *reinterpret_cast<const uint4 *>(T->wmode);
      mws[0] = mw.x;
      mws[1] = mw.y;
      mws[2] = mw.z;
      mws[3] = mw.w;
    }
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      float v[M];
      if (q == 0) {
        const float a0 = (float)((T ? T->wf[0].a : 1.0) * f.scale[0]);
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = a0;
      } else if (q > K) {
#pragma unroll
        for (int m = 0; m < M; ++m) v[m] = 0.0f;
      } else {
        int mode;
        double om, ph, a, b;
        if (T) {
          mode = wave_mode(mws, q);
          const WaveF w = T->wf[q];
          om = w.om;
          ph = w.ph;
          a = w.a;
          b = w.b;
        } else {
          const int kind = f.waves->kind[q - 1];
          mode = (kind == IG_WAVE_SIN || kind == IG_WAVE_COS ||
                  (kind == IG_WAVE_SQUARE && f.waves->duty[q - 1] == 0.5))
                     ? kind
                     : kWaveSlow;
          om = f.waves->om[q - 1];
          ph = f.waves->ph[q - 1];
          a = f.waves->amp[q - 1];
          b = f.waves->off[q - 1];
        }
        if (mode == kWaveSlow) {
#pragma unroll
          for (int m = 0; m < M; ++m)
            v[m] = (float)(wave_eval_slow<true>(T ? Wq : f.waves, f, q - 1, t0 + m * hh) *
                           (T ? sq[q] : 1.0) * f.scale[q]);
        } else {
          affine_weights_fast<M>(mode, om, ph, a, b, f.scale[q], t0, hh, v);
        }
      }
#pragma unroll
      for (int m = 0; m < M; ++m) o[m][q] = v[m];
    }
  };
  // FP64 modes: the waveform weights, kept while the lab time repeats
  auto weights = [&](double tt) {
    if (!(tt == wtt)) {
      // an untuned flight reads the pack's waveforms through the kernel
      // argument (scalar loads); a tune row is a per-lane address
      if (e.tunes) {
#pragma unroll
        for (int k = 0; k < NW; ++k)
          if (k < K) w[k] = wave_eval<true>(W, f, k, tt);
      } else {
#pragma unroll
        for (int k = 0; k < NW; ++k)
          if (k < K) w[k] = wave_eval<true>(f.waves, f, k, tt);
      }
      wtt = tt;
    }
  };
  auto efield = [&](double px, double py, double pz, double tt, double &a, double &b,
                    double &c) {
    if constexpr (PREC == IG_PREC_F32) {
      // weights made afresh: FP32 steps gather through the cell cache, so
      // this path (a stage that left its cell, a record row) is rare and
      // keeps no weights live across steps
      float wl[1][NS];
      weights_f32(wl, tt, 0.0);
      xyz_field_f32<NS>(f, px * inv_h, py * inv_h, pz * inv_h, wl[0], a, b, c);
    } else {
      weights(tt);
      xyz_field<PREC, NS>(f, px * inv_h, py * inv_h, pz * inv_h, w, K, tscale, a, b, c);
    }
  };
  int32_t tpp = p.tp_col >= 0 ? e.tp_pass[i] : 0;
  // a record row; with the transporter the 11th column is the pass count
  // (0 on a station-plane row, which the reference leaves unset)
  auto write = [&](int row, double tus, int32_t wrap) {
    double *r = e.rec + ((size_t)i * p.rec_rows + row) * p.rec_cols;
    r[0] = tus;
    r[1] = x;
    r[2] = y;
    r[3] = z;
    r[4] = vx * 1e-3;
    r[5] = vy * 1e-3;
    r[6] = vz * 1e-3;
    double a, b, c;
    efield(x, y, z, tus, a, b, c);
    r[7] = a;
    r[8] = b;
    r[9] = c;
    if (p.rec_cols > 10) r[10] = (double)wrap;
  };
  if (nrec == 0) {
    if (rec) write(0, tob, 0);
    nrec = 1;
  }
  const int R = rec ? p.rec_rows : 0;
  const double c = 1e3;
  int fate = IG_FATE_FLYING, hit_ip = -1;
  int64_t s = 0;
  const int64_t lim = p.max_steps;
  // the step machine of the r-z kernel (kStep / kCollide / kPost): hard-
  // sphere collisions are scattered together by the wave (FP64, the
  // reference's algorithm) or one partner try per pass (FP32-basis mode)
  constexpr bool kFlat = PREC == IG_PREC_F32 && GAS == IG_GAS_HS;
  float hs_k = 0.0f, hs_inv_cstar = 0.0f, hz = 0.0f;
  HsTry hst{};
  if constexpr (kFlat) {
    hs_k = hs_rate_const(p, dt, 1e6, sigma);
    hs_inv_cstar = (float)(1.0 / p.c_star);
    hz = e.hz[i];
  }
  const int coll_batch = p.coll_batch;
  int stage = kStep;
  double xo = x, yo = y, zo = z, vxo = vx, vyo = vy, vzo = vz, to = t;
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
        vzo = vz;
        to = t;
        const double t_us = tob + t * 1e6;
        const double h_us = dt * 1e6;
        // the four RK4 stages as one loop around one field evaluation
        // (tracer3d.py:395-417): stage s is evaluated at point p_s and lab
        // time t_s; k_1 = v, p_{s+1} = x + cf_s dt k_s c and
        // k_{s+1} = v + cf_s dt a_s with cf = 1/2, 1/2, 1; the weighted sums
        // k_1 + 2 k_2 + 2 k_3 + k_4 accumulate left to right, the
        // reference's order
        double px = x, py = y, pz = z;
        double kx = vx, ky = vy, kz = vz;
        double skx = 0.0, sky = 0.0, skz = 0.0, sax = 0.0, say = 0.0, saz = 0.0;
        // FP32 mode: the first stage reads the cell's corner records once
        // for all three lab times; later stages in the same cell blend
        // the kept sums, and only a stage that has left the cell gathers
        Cell3 cl0{};
#pragma unroll 1
        for (int st = 0; st < 4; ++st) {
          const double tt = st == 0 ? t_us : (st == 3 ? t_us + h_us : t_us + 0.5 * h_us);
          double ea, eb, ec;
          if constexpr (PREC == IG_PREC_F32) {
            const Cell3 cl = cell3(f, px * inv_h, py * inv_h, pz * inv_h);
            if (st == 0) {
              float ww[3][NS];
              weights_f32(ww, t_us, 0.5 * h_us);
              xyz_cell_f32<NS>(f, cl, ww[0], ww[1], ww[2], ea, eb, ec, kc, bs);
              cl0 = cl;
            } else if (cl.i == cl0.i && cl.j == cl0.j && cl.k == cl0.k) {
              blend_cell_f32(cl, kc, bs, st == 3, ea, eb, ec);
            } else {
              efield(px, py, pz, tt, ea, eb, ec);
            }
          } else {
            efield(px, py, pz, tt, ea, eb, ec);
          }
          const double ax = qm * 1e3 * ea, ay = qm * 1e3 * eb, az = qm * 1e3 * ec;
          if (st == 0) {
            skx = kx;
            sky = ky;
            skz = kz;
            sax = ax;
            say = ay;
            saz = az;
          } else if (st == 3) {
            skx = skx + kx;
            sky = sky + ky;
            skz = skz + kz;
            sax = sax + ax;
            say = say + ay;
            saz = saz + az;
          } else {
            skx = skx + 2 * kx;
            sky = sky + 2 * ky;
            skz = skz + 2 * kz;
            sax = sax + 2 * ax;
            say = say + 2 * ay;
            saz = saz + 2 * az;
          }
          if (st < 3) {
            const double cf = st < 2 ? 0.5 : 1.0;
            px = x + cf * dt * kx * c;
            py = y + cf * dt * ky * c;
            pz = z + cf * dt * kz * c;
            kx = vx + cf * dt * ax;
            ky = vy + cf * dt * ay;
            kz = vz + cf * dt * az;
          }
        }
        x = x + dt / 6.0 * skx * c;
        y = y + dt / 6.0 * sky * c;
        z = z + dt / 6.0 * skz * c;
        vx = vx + dt / 6.0 * sax;
        vy = vy + dt / 6.0 * say;
        vz = vz + dt / 6.0 * saz;
        t += dt;
        s += 1;
        stage = kPost;
        // HS in mm/us (tracer3d.py:424-431): the kernel carries m/s
        if constexpr (kFlat) {
          if (hs_hazard_step(rng, hz, vx * 1e-3, vy * 1e-3, vz * 1e-3, p, hs_k, hs_inv_cstar,
                             hst))
            stage = kCollide;
        } else if constexpr (GAS == IG_GAS_HS) {
          if (hs_test(rng, vx * 1e-3, vy * 1e-3, vz * 1e-3, p, dt, 1e6, sigma)) stage = kCollide;
        }
      }
    }
    if constexpr (kFlat) {
      const uint64_t parked = __ballot(stage == kCollide);
      if (parked) {
        const uint64_t moving = __ballot(stage == kStep || stage == kPost);
        if (__popcll(parked) >= coll_batch || __popcll(moving) < coll_batch) {
          if (stage == kCollide) {
            double sx = vx * 1e-3, sy = vy * 1e-3, sz = vz * 1e-3;
            const int rc = hs_try_flat(rng, hst, hz, sx, sy, sz, p, m_amu);
            if (rc > 0) {
              vx = sx * 1e3;
              vy = sy * 1e3;
              vz = sz * 1e3;
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
        if (__popcll(parked) >= coll_batch || __popcll(moving) < coll_batch) {
          if (stage == kCollide) {
            double sx = vx * 1e-3, sy = vy * 1e-3, sz = vz * 1e-3;
            if (hs_scatter(rng, sx, sy, sz, p, m_amu)) {
              vx = sx * 1e3;
              vy = sy * 1e3;
              vz = sz * 1e3;
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
      // the transporter plane, before any terminal test: a crossing of
      // the accept plane in its direction counts a pass and shifts the
      // ion and the step's start point by emit - accept; the last
      // allowed pass ends the flight unshifted
      if (p.tp_col >= 0) {
        const double cn = p.tp_col == 0 ? x : (p.tp_col == 1 ? y : z);
        const double co = p.tp_col == 0 ? xo : (p.tp_col == 1 ? yo : zo);
        if ((cn - p.tp_accept) * p.tp_sgn >= 0.0 && (co - p.tp_accept) * p.tp_sgn < 0.0) {
          tpp += 1;
          if (tpp >= p.tp_maxp) {
            fate = IG_FATE_TRANSPORTER;
            if (nrec < R) write(nrec, tob + t * 1e6, tpp);
            if (rec && nrec < R) nrec++;
            stage = kDone;
          } else if (p.tp_col == 0) {
            x += p.tp_d;
            xo += p.tp_d;
          } else if (p.tp_col == 1) {
            y += p.tp_d;
            yo += p.tp_d;
          } else {
            z += p.tp_d;
            zo += p.tp_d;
          }
        }
      }
      if (stage != kDone && in_metal_3d(f.metal, x * inv_h, y * inv_h, z * inv_h, nx, ny, nz)) {
        double f0 = 0.0, f1 = 1.0;
        for (int b = 0; b < 40; ++b) {
          double fm = 0.5 * (f0 + f1);
          if (in_metal_3d(f.metal, (xo + fm * (x - xo)) * inv_h, (yo + fm * (y - yo)) * inv_h,
                          (zo + fm * (z - zo)) * inv_h, nx, ny, nz))
            f1 = fm;
          else
            f0 = fm;
        }
        const double fr = f1;
        x = xo + fr * (x - xo);
        y = yo + fr * (y - yo);
        z = zo + fr * (z - zo);
        vx = vxo + fr * (vx - vxo);
        vy = vyo + fr * (vy - vyo);
        vz = vzo + fr * (vz - vzo);
        t = to + fr * dt;
        fate = IG_FATE_IMPACT;
        if (nrec < R) write(nrec, tob + t * 1e6, tpp);
        if (rec && nrec < R) nrec++;
        stage = kDone;
      }
      if (stage != kDone) {
        for (int ip = 0; ip < p.n_pl; ++ip) {
          const int col = p.pl_col[ip];
          const double cn = col == 0 ? x : (col == 1 ? y : z);
          const double co = col == 0 ? xo : (col == 1 ? yo : zo);
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
          const double t1 = col == 0 ? yc : xc;
          const double t2 = col == 2 ? yc : zc;
          const bool ins = p.pl_w[ip][0] <= t1 && t1 <= p.pl_w[ip][1] && p.pl_w[ip][2] <= t2 &&
                           t2 <= p.pl_w[ip][3];
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
          vz = vzo + fr * (vz - vzo);
          t = to + fr * dt;
          fate = p.pl_kind[ip];
          hit_ip = ip;
          if (nrec < R) write(nrec, tob + t * 1e6, 0);
          if (rec && nrec < R) nrec++;
          stage = kDone;
          break;
        }
      }
      if (stage != kDone && (x < 0.0 || y < 0.0 || z < 0.0 || x > (nx - 1) * h ||
                             y > (ny - 1) * h || z > (nz - 1) * h)) {
        fate = IG_FATE_BOX;
        if (nrec < R) write(nrec, tob + t * 1e6, tpp);
        if (rec && nrec < R) nrec++;
        stage = kDone;
      }
      if (stage != kDone && (steps0 + s) % p.rec_every == 0 && nrec < R) {
        write(nrec, tob + t * 1e6, tpp);
        nrec++;
      }
    }
    if (__ballot(stage != kDone) == 0) break;
  }
  e.x[i] = x;
  e.y[i] = y;
  e.z[i] = z;
  e.vx[i] = vx;
  e.vy[i] = vy;
  e.vz[i] = vz;
  e.t[i] = t;
  e.fate[i] = fate;
  e.face[i] = hit_ip;
  e.ncol[i] = ncol;
  e.steps[i] = steps0 + s;
  e.nrec[i] = nrec;
  e.rng_n[i] = rng.n;
  e.has_gauss[i] = rng.has_gauss;
  e.gauss[i] = rng.gauss;
  if constexpr (kFlat) e.hz[i] = hz;
  if (p.tp_col >= 0) e.tp_pass[i] = tpp;
}

// tracer3d._fly3d_sds (tracer3d.py:851-1025) with _sds_field
// (tracer3d.py:792-812): the 3-D SDS flight in mm/us.  Per step: E at
// (x, tob + t); a_f = qm 1e-6 E; the single damped kick v += factor (a_f -
// (v - v_gas) gamma) dt with factor = (1 - e^{-gamma dt}) / (gamma dt)
// (plain a_f without damping); x += v dt; the ICDF diffusion jump when the
// ion's mean free path is positive; t += dt; then metal (all eight corners,
// no bisection), the box, and the station/bound planes crossed over the
// whole step (advection plus jump), each ending the flight with a record;
// else the record every rec_every steps.  Record rows are (tob + t, x, y, z,
// vx, vy, vz, Ex, Ey, Ez) with the field at the recorded point and time.
// The flight clock is microseconds (the reference's dt_ns * 1e-3) and the
// velocities stay in mm/us.  The FP32-basis mode reads its field and draws
// its jump as the other kernels do in that mode.
template <int PREC, int NS>
__global__ __launch_bounds__(256) void xyz_sds_kernel(const FieldsK f, const FlightK p,
                                                      const EnsK e) {
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
  const int64_t steps0 = e.steps[i];
  double x = e.x[i], y = e.y[i], z = e.z[i];
  double vx = e.vx[i], vy = e.vy[i], vz = e.vz[i];
  double t = e.t[i];
  const double tob = e.tob[i], qm = e.acc[i];
  const double h = f.h, inv_h = f.inv_h;
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
  const int64_t nx = f.n0, ny = f.n1, nz = f.n2;
  const int K = f.K;
  const bool rec = p.rec_rows > 0 && e.rec != nullptr;
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
  const double damp = e.sds_damp[i], mfp = e.sds_mfp[i], V = e.sds_V[i], logmr = e.sds_logmr[i];
  double factor = 1.0;
  if (damp > 0.0) {
    const double tterm = damp * dt;
    factor = (1.0 - exp(-tterm)) / tterm;
  }
  // the FP32-basis jump's per-ion length factor
  const double sds_rj = mfp > 0.0 ? sds_rj0(V, mfp, dt) : 0.0;
  constexpr int NW = NS > 1 ? NS - 1 : 1;
  double w[NW];
  float wt[NS];
  double wtt = __builtin_nan("");
  auto efield = [&](double px, double py, double pz, double tt, double &a, double &b,
                    double &c) {
    if (!(tt == wtt)) {
      // an untuned flight reads the pack's waveforms through the kernel
      // argument (scalar loads); a tune row is a per-lane address
      const bool tuned = e.tunes != nullptr;
      if constexpr (PREC == IG_PREC_F32) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          double v = q == 0 ? 1.0 : 0.0;
          if (q > 0 && q <= K) {
            if (tuned)
              v = wave_eval_fast<true>(W, f, q - 1, tt);
            else
              v = wave_eval_fast<true>(f.waves, f, q - 1, tt);
          }
          v = v * f.scale[q];
          if (tscale) v = v * tscale[q];
          wt[q] = (float)v;
        }
      } else if (tuned) {
#pragma unroll
        for (int k = 0; k < NW; ++k)
          if (k < K) w[k] = wave_eval<true>(W, f, k, tt);
      } else {
#pragma unroll
        for (int k = 0; k < NW; ++k)
          if (k < K) w[k] = wave_eval<true>(f.waves, f, k, tt);
      }
      wtt = tt;
    }
    if constexpr (PREC == IG_PREC_F32)
      xyz_field_f32<NS>(f, px * inv_h, py * inv_h, pz * inv_h, wt, a, b, c);
    else
      xyz_field<PREC, NS>(f, px * inv_h, py * inv_h, pz * inv_h, w, K, tscale, a, b, c);
  };
  const int R = rec ? p.rec_rows : 0;
  auto write = [&](int row) {
    double *r = e.rec + ((size_t)i * p.rec_rows + row) * p.rec_cols;
    const double tus = tob + t;
    r[0] = tus;
    r[1] = x;
    r[2] = y;
    r[3] = z;
    r[4] = vx;
    r[5] = vy;
    r[6] = vz;
    double a, b, c;
    efield(x, y, z, tus, a, b, c);
    r[7] = a;
    r[8] = b;
    r[9] = c;
  };
  if (nrec == 0) {
    if (rec) write(0);
    nrec = 1;
  }
  int fate = IG_FATE_FLYING, hit_ip = -1;
  int64_t s = 0;
  const int64_t lim = p.max_steps;
  for (;;) {
    if (!(t < t_max)) {
      fate = IG_FATE_TIMEOUT;
      break;
    }
    if (lim > 0 && s >= lim) break;
    double ex, ey, ez;
    efield(x, y, z, tob + t, ex, ey, ez);
    const double afx = qm * 1e-6 * ex, afy = qm * 1e-6 * ey, afz = qm * 1e-6 * ez;
    double aex = afx, aey = afy, aez = afz;
    if (damp > 0.0) {
      aex = factor * (afx - (vx - p.vg[0]) * damp);
      aey = factor * (afy - (vy - p.vg[1]) * damp);
      aez = factor * (afz - (vz - p.vg[2]) * damp);
    }
    vx = vx + aex * dt;
    vy = vy + aey * dt;
    vz = vz + aez * dt;
    const double xo = x, yo = y, zo = z, vxo = vx, vyo = vy, vzo = vz, to = t;
    x = x + vx * dt;
    y = y + vy * dt;
    z = z + vz * dt;
    if (mfp > 0.0) {
      double jx, jy, jz;
      if constexpr (PREC == IG_PREC_F32)
        sds_jump_fast(rng, p.sds_stats, logmr, sds_rj, jx, jy, jz);
      else
        sds_jump(rng, p.sds_stats, logmr, V, mfp, dt, jx, jy, jz);
      x = x + jx;
      y = y + jy;
      z = z + jz;
    }
    t += dt;
    s += 1;
    if (in_metal_3d(f.metal, x * inv_h, y * inv_h, z * inv_h, nx, ny, nz)) {
      fate = IG_FATE_IMPACT;
      if (nrec < R) write(nrec);
      if (rec && nrec < R) nrec++;
      break;
    }
    if (x < 0.0 || y < 0.0 || z < 0.0 || x > (nx - 1) * h || y > (ny - 1) * h ||
        z > (nz - 1) * h) {
      fate = IG_FATE_BOX;
      if (nrec < R) write(nrec);
      if (rec && nrec < R) nrec++;
      break;
    }
    bool hit = false;
    for (int ip = 0; ip < p.n_pl; ++ip) {
      const int col = p.pl_col[ip];
      const double cn = col == 0 ? x : (col == 1 ? y : z);
      const double co = col == 0 ? xo : (col == 1 ? yo : zo);
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
      const double t1 = col == 0 ? yc : xc;
      const double t2 = col == 2 ? yc : zc;
      const bool ins = p.pl_w[ip][0] <= t1 && t1 <= p.pl_w[ip][1] && p.pl_w[ip][2] <= t2 &&
                       t2 <= p.pl_w[ip][3];
      if (p.pl_kind[ip] == IG_FATE_DETECT) {
        if (!ins) continue;
      } else if (p.pl_kind[ip] == IG_FATE_PLANE && ins) {
        continue;
      }
      x = xc;
      y = yc;
      z = zc;
      vx = vxo + fr * (vx - vxo);
      vy = vyo + fr * (vy - vyo);
      vz = vzo + fr * (vz - vzo);
      t = to + fr * (t - to);
      fate = p.pl_kind[ip];
      hit_ip = ip;
      hit = true;
      break;
    }
    if (hit) {
      if (nrec < R) write(nrec);
      if (rec && nrec < R) nrec++;
      break;
    }
    if ((steps0 + s) % p.rec_every == 0 && nrec < R) {
      write(nrec);
      nrec++;
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
  e.face[i] = hit_ip;
  e.steps[i] = steps0 + s;
  e.nrec[i] = nrec;
  e.rng_n[i] = rng.n;
  e.has_gauss[i] = rng.has_gauss;
  e.gauss[i] = rng.gauss;
}

template <int PREC, int NS>
static int launch_xyz_ns(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch,
                         int block, hipStream_t st) {
  dim3 grid((unsigned)((n_launch + block - 1) / block));
  const size_t lds = PREC == IG_PREC_F32 ? (size_t)kCellLdsFloats * sizeof(float) * block : 0;
  if (p.gas == IG_GAS_HS) {
    if (p.occupancy == 2)
      hipLaunchKernelGGL((xyz_fly_kernel<PREC, NS, IG_GAS_HS, 2>), grid, dim3(block), lds, st, f,
                         p, e);
    else if (p.occupancy >= 3)
      hipLaunchKernelGGL((xyz_fly_kernel<PREC, NS, IG_GAS_HS, 3>), grid, dim3(block), lds, st, f,
                         p, e);
    else
      hipLaunchKernelGGL((xyz_fly_kernel<PREC, NS, IG_GAS_HS, 1>), grid, dim3(block), lds, st, f,
                         p, e);
  } else if (p.gas == IG_GAS_VACUUM) {
    if (p.occupancy == 2)
      hipLaunchKernelGGL((xyz_fly_kernel<PREC, NS, IG_GAS_VACUUM, 2>), grid, dim3(block), lds,
                         st, f, p, e);
    else if (p.occupancy >= 3)
      hipLaunchKernelGGL((xyz_fly_kernel<PREC, NS, IG_GAS_VACUUM, 3>), grid, dim3(block), lds,
                         st, f, p, e);
    else
      hipLaunchKernelGGL((xyz_fly_kernel<PREC, NS, IG_GAS_VACUUM, 1>), grid, dim3(block), lds,
                         st, f, p, e);
  } else {
    hipLaunchKernelGGL((xyz_sds_kernel<PREC, NS>), grid, dim3(block), 0, st, f, p, e);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// 3-D packs: static only (1), two RF channels (3: quadrupole), ten drive
// channels (11: traveling-wave board)
template <int PREC>
int launch_xyz_prec(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch, int block,
                    void *stream) {
  hipStream_t st = (hipStream_t)stream;
  switch (f.n_sets) {
  case 1: return launch_xyz_ns<PREC, 1>(f, p, e, n_launch, block, st);
  case 2: return launch_xyz_ns<PREC, 2>(f, p, e, n_launch, block, st);
  case 3: return launch_xyz_ns<PREC, 3>(f, p, e, n_launch, block, st);
  case 7:
    // the 11-set SLIM board with its linked channels folded (FP32 basis)
    if constexpr (PREC == IG_PREC_F32) return launch_xyz_ns<PREC, 7>(f, p, e, n_launch, block, st);
    return -2;
  case 11: return launch_xyz_ns<PREC, 11>(f, p, e, n_launch, block, st);
  default: return -2;
  }
}

} // namespace ig
