/*
 * ig_oracle.h -- CPU FP64 oracle of ion_gym's flight kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU engine is
 * compared against and the timed CPU baseline of the benchmark; nothing in
 * the product path links, loads or calls it.
 *
 * It restates, in plain C with strict IEEE double arithmetic
 * (-O3 -fno-fast-math -ffp-contract=off), the reference kernels:
 *   ion_gym/physics/tracer_rz.py:113-446   _fly_rec_full (r-z leapfrog)
 *   ion_gym/physics/tracer3d.py:280-575    _fly3d        (3-D RK4)
 *   ion_gym/physics/build_planar.py:1111-1429 _fly_planar (planar leapfrog)
 *   ion_gym/physics/interp.py:19-36        _bilin
 *   ion_gym/physics/tracer3d.py:121-151    _tri
 *   ion_gym/physics/raster2d.py:597-612    _metal_nn
 *   ion_gym/physics/tracer3d.py:174-189    _in_metal_3d
 *   ion_gym/physics/build_planar.py:1059-1096, tracer3d.py:234-276 waves
 *   ion_gym/physics/collision3d.py:265-374 HS collision step
 *   ion_gym/physics/sds.py:134-169         SDS jump
 * with two random streams:
 *   OR_RNG_MT     numba's per-thread np.random as seeded by the reference
 *                 (np.random.seed(seed) = MT19937 init_genrand, doubles
 *                 from two 32-bit words, polar-method normals with a cached
 *                 second value), reproducing the reference's draws;
 *   OR_RNG_PHILOX the engine's counter-based stream (Philox4x32-10 keyed by
 *                 the ion seed), so the GPU can be compared ion by ion.
 */
#ifndef IG_ORACLE_H
#define IG_ORACLE_H
#include <stdint.h>

#define OR_RNG_MT 0
#define OR_RNG_PHILOX 1

typedef struct {
  /* grids in the reference layout */
  int64_t nx, nu;          /* r-z: axial nodes, unfolded radial width */
  const double *EzA, *EuA; /* (nx, nu) V/m */
  const double *EzK, *EuK; /* (K, nx, nu) */
  const double *EzG, *EuG; /* (nx, nu) or NULL */
  double tau_gate;
  const double *ele;       /* (nx, nu) 0/1 mirrored metal */
  const uint8_t *label;    /* (nx, nu) labels or NULL */
  double u0, mm;
  int32_t K;
  const int64_t *ch_kind;
  const double *ch_om, *ch_ph, *ch_duty;
  const double *tab_t, *tab_v;
  const int64_t *tab_off;
} or_rz_fields;

typedef struct {
  /* 3-D grids (nx, ny, nz), V/mm */
  int64_t nx, ny, nz;
  const double *EAx, *EAy, *EAz;
  const double *ExK, *EyK, *EzK; /* (K, nx, ny, nz) */
  const uint8_t *ele;
  double h_mm;
  int32_t K;
  const int64_t *ch_kind;
  const double *ch_om, *ch_ph, *ch_amp, *ch_off, *ch_duty;
  const double *tab_t, *tab_v;
  const int64_t *tab_off;
} or_xyz_fields;

typedef struct {
  /* planar grids (nx, ny), V/m (build_planar.py:1111-1118) */
  int64_t nx, ny;
  const double *ExA, *EyA;
  const double *ExK, *EyK; /* (K, nx, ny) */
  const double *ele;       /* (nx, ny) 0/1 metal */
  double mm;
  int32_t K;
  const int64_t *ch_kind;
  const double *ch_om, *ch_ph, *ch_duty;
  const double *tab_t, *tab_v;
  const int64_t *tab_off;
} or_planar_fields;

typedef struct {
  double dt;        /* r-z: us; 3-D: s */
  double t_max;     /* r-z: us; 3-D: s */
  int64_t max_steps; /* <= 0: fly to fate */
  int64_t rec_every;
  int64_t rec_rows;  /* rows per ion (the reference's rec.shape[0]) */
  int64_t rec_cols;
  int32_t flags[13]; /* r-z optional channel flags (build_rz.py:569-571) */
  double T_k, P_pa, c_star, c_bar, sig1d, m_gas;
  int32_t gas;      /* 0 vacuum, 1 HS, 2 SDS */
  double vg[3];
  const double *sds_stats; /* (5, 1002) */
  int32_t bnd_on[6];
  double bnd_val[6];
  int32_t n_pl;
  const int64_t *pl_col;
  const double *pl_val, *pl_sgn, *pl_w; /* pl_w (n_pl, 4) */
  const int64_t *pl_kind;
  int32_t rng;
  uint64_t rng_key;
  /* 3-D transporter (tracer3d.py:433-458): tp_col -1 none, else the axis */
  int32_t tp_col;
  double tp_accept, tp_emit, tp_sgn;
  int64_t tp_maxp;
} or_params;

typedef struct {
  int64_t n;
  const double *x, *y, *z, *vx, *vy, *vz, *tob, *m_ion, *acc, *sigma;
  const int64_t *seed;
  const double *sds; /* (n, 4) or NULL */
} or_ions;

typedef struct {
  double *x, *y, *z, *vx, *vy, *vz, *t; /* final row (t = lab time) */
  int32_t *kind, *face, *label;
  int64_t *ncol, *steps, *nrec;
  double *rec; /* (n, rec_rows, rec_cols) or NULL */
} or_out;

#ifdef __cplusplus
extern "C" {
#endif
int or_rz_fly(const or_rz_fields *f, const or_params *p, const or_ions *ions,
              or_out *out, int n_threads);
int or_xyz_fly(const or_xyz_fields *f, const or_params *p, const or_ions *ions,
               or_out *out, int n_threads);
int or_planar_fly(const or_planar_fields *f, const or_params *p,
                  const or_ions *ions, or_out *out, int n_threads);
/* primitives exposed for the known-answer tests */
double or_bilin(const double *F, double gx, double gu, int64_t nx, int64_t nu);
double or_tri(const double *F, double x, double y, double z, int64_t nx,
              int64_t ny, int64_t nz);
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]);
/* n draws of the reference stream after np.random.seed(seed):
   kind 0 = random(), kind 1 = normal() */
void or_mt_draws(uint32_t seed, int kind, int64_t n, double *out);
/* the reference's _collide driven by the MT stream, as the pinned test
   tests/test_collide_bound.py:25-35 runs it */
void or_collide_run(uint32_t seed, int64_t n, double vx, double vy, double vz,
                    double m_ion, double m_gas, double sig1d, double *out);
double or_c_bar_rel(double s_mm_us, double c_star, double c_bar);
double or_icdf_steps_at(const double *stats, double log_mr, double u);
double or_wave_rz(int64_t kind, double om, double ph, double duty,
                  const double *tab_t, const double *tab_v, int64_t o0,
                  int64_t o1, double t);
#ifdef __cplusplus
}
#endif
#endif
