// Structures shared by the engine's host code and its gfx950 kernels.
//
// Everything a flight kernel reads is passed by value in three kernel
// arguments (well under the 4 KiB kernarg limit): the field pack
// (FieldsK), the flight parameters (FlightK) and the device-resident ion
// state (EnsK).
#pragma once
#include <stdint.h>

#include "../../include/ion_gym_amd.h"

namespace ig {

// Field pack on the device.  Channel grids are interleaved node-major: the
// record of node n holds, for every set s (static, drive channels, gate),
// its n_comp components, padded to `stride` values so that a record is a
// whole number of 16-byte units.  One bilinear or trilinear gather then
// reads a few contiguous records instead of 2(1+K) or 3(1+K) separate
// arrays.
// Drive waveforms of the K channels (device memory).
struct Waves {
  int32_t kind[IG_MAX_CHANNELS];
  double om[IG_MAX_CHANNELS], ph[IG_MAX_CHANNELS], duty[IG_MAX_CHANNELS];
  double amp[IG_MAX_CHANNELS], off[IG_MAX_CHANNELS];
  int64_t tab_off[IG_MAX_CHANNELS + 1];
};

// A tune's drive of one field set packed for the FP32-basis 3-D weights:
// the phase terms, and the affine amplitude and offset already multiplied
// by the tune's set scale, in two 16-byte words; the sets' modes
// (IG_WAVE_SIN / _COS / _SQUARE at duty 0.5, kWaveSlow, kWaveConst) sit
// four bits each in one 16-byte word of the tune.
constexpr int kWaveSlow = 5;  // a square of duty != 0.5 (wave_eval_slow)
constexpr int kWaveConst = 6; // the static set: weight a
struct alignas(16) WaveF {
  double om, ph;
  double a, b;
};

__host__ __device__ inline int wave_mode(const uint32_t *m, int q) {
  return (int)((m[q >> 3] >> (4 * (q & 7))) & 15u);
}

// One tune: the per-request / per-ensemble parameters an ion of a batched
// launch flies with (ig_tune of the C ABI, on the device).
struct Tune {
  Waves w;
  double dt, t_max;
  double scale[IG_MAX_CHANNELS + 2]; // multiplies each set's field
  uint64_t rng_key;
  WaveF wf[IG_MAX_CHANNELS + 1];     // set 0 (static) and the drive channels
  alignas(16) uint32_t wmode[4];     // wf's modes, four bits a set
};

struct FieldsK {
  int32_t route;
  int32_t prec;
  int64_t n0, n1, n2;
  // the node counts and their last two indices as doubles, kernel
  // arguments so the kernels compare FP64 coordinates with them straight
  // from scalar registers
  double dn[3], dnm1[3], dnm2[3];
  double h, inv_h, u0;
  int32_t K;        // drive channels
  int32_t has_gate; // r-z gate set present (set K+1)
  double tau_gate;
  int32_t n_sets;   // 1 + K + has_gate
  int32_t n_comp;   // 2 (r-z) or 3 (3-D)
  int32_t stride;   // values per node record (>= n_sets*n_comp)
  const void *grid; // double or float records
  const uint8_t *metal;
  const uint8_t *label;
  double scale[IG_MAX_CHANNELS + 2]; // per-set multiplier (FP32 storage)
  const struct Waves *waves;         // the pack's drive waveforms (device)
  const double *tab_t, *tab_v;
};

struct FlightK {
  double dt, t_max;
  int64_t max_steps;
  int32_t rec_every, rec_rows, rec_cols;
  uint32_t rec_mask;
  int32_t gas;
  double T_k, P_pa, c_star, c_bar, sig1d, m_gas;
  double vg[3];
  const double *sds_stats;
  int32_t bnd_on[6];
  double bnd_val[6];
  int32_t n_pl;
  int32_t pl_col[IG_MAX_PLANES];
  double pl_val[IG_MAX_PLANES], pl_sgn[IG_MAX_PLANES];
  double pl_w[IG_MAX_PLANES][4];
  int32_t pl_kind[IG_MAX_PLANES];
  uint64_t rng_key;
  int32_t coll_batch; // parked hard-sphere lanes a wave scatters together
  int32_t occupancy;  // register-budget variant of the kernel (1, 2, 3)
  int32_t tp_col;     // transporter axis, -1 none
  double tp_accept, tp_d, tp_sgn; // accept plane, emit - accept, direction
  int64_t tp_maxp;
};

// Device-resident ion state, structure of arrays.  t is the elapsed flight
// time in the route's clock (r-z us, 3-D s); tob the lab birth time (us).
// Velocities are mm/us on both routes (the 3-D kernel converts to m/s in
// registers exactly as the reference wrapper does on entry).
struct EnsK {
  int64_t n;
  double *x, *y, *z, *vx, *vy, *vz, *t, *tob;
  double *m_amu, *acc, *sigma;
  double *sds_damp, *sds_mfp, *sds_V, *sds_logmr;
  int64_t *seed;
  uint64_t *rng_n;     // draws consumed
  double *gauss;       // polar-method cache value
  int32_t *has_gauss;
  float *hz;           // FP32-basis hard sphere: remaining Exp(1) hazard
                       // (NaN until the first draw)
  int32_t *fate, *face, *label, *nrec;
  int64_t *ncol, *steps;
  double *path, *eat, *ket; // record accumulators
  double *e_last;      // (n, 3) last evaluated field (acceleration units)
  double *rec;         // (n, rec_rows, rec_cols) or null
  int32_t *tp_pass;    // transporter crossings so far
  const int32_t *idx;        // optional compacted list of live ions
  const uint32_t *n_live;    // its length (device), with idx
  const int32_t *tune; // per-ion tune index (with tunes)
  const Tune *tunes;   // per-ensemble tune table, or null: the pack's
                       // waveforms and the flight's dt/t_max/rng_key
};

// values per node record: n_vals rounded up to a whole number of 16-byte
// units (the kernels' RecLayout uses the same rule at compile time)
inline int record_stride(int prec, int n_vals) {
  const int per16 = prec == IG_PREC_F64 ? 2 : 4;
  return ((n_vals + per16 - 1) / per16) * per16;
}

// launchers, instantiated once per precision in their own translation units
// (ig_rz_*.hip, ig_xyz_*.hip); return 0, -1 launch error, -2 no kernel for
// this pack
template <int PREC>
int launch_rz_prec(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch, int block,
                   void *stream);
template <int PREC>
int launch_planar_prec(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch,
                       int block, void *stream);
template <int PREC>
int launch_xyz_prec(const FieldsK &f, const FlightK &p, const EnsK &e, int64_t n_launch,
                    int block, void *stream);


} // namespace ig
