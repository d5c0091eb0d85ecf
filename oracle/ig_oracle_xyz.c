/*
 * 3-D flight of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * A restatement, in strict IEEE double arithmetic, of
 * ion_gym/physics/tracer3d.py:280-575 (_fly3d: classic RK4 with positions
 * in mm, velocities in m/s and the step in s) with _tri (:121-151),
 * _in_metal_3d (:174-189) and the affine waveform _wave_eval (:234-276).
 * The periodic transporter (:434-458) is not restated: the engine's 3-D
 * route refuses it.  Record rows are [t_us, x, y, z, vx, vy, vz (mm/us),
 * ex, ey, ez (V/mm)], the ten record arrays of fly3d (:709-717).
 */
#include <math.h>

#include "ig_oracle.h"
#include "ig_oracle_gas.h"
#include "ig_oracle_rng.h"

double or_tri(const double *F, double x, double y, double z, int64_t nx,
              int64_t ny, int64_t nz) {
  if (x < 0.0) x = 0.0;
  if (y < 0.0) y = 0.0;
  if (z < 0.0) z = 0.0;
  if (x > nx - 1) x = nx - 1.0;
  if (y > ny - 1) y = ny - 1.0;
  if (z > nz - 1) z = nz - 1.0;
  int64_t i = (int64_t)x, j = (int64_t)y, k = (int64_t)z;
  if (i > nx - 2) i = nx - 2;
  if (j > ny - 2) j = ny - 2;
  if (k > nz - 2) k = nz - 2;
  double tx = x - i, ty = y - j, tz = z - k;
#define F3(a, b, c) F[((a)*ny + (b)) * nz + (c)]
  double c00 = F3(i, j, k) * (1 - tx) + F3(i + 1, j, k) * tx;
  double c10 = F3(i, j + 1, k) * (1 - tx) + F3(i + 1, j + 1, k) * tx;
  double c01 = F3(i, j, k + 1) * (1 - tx) + F3(i + 1, j, k + 1) * tx;
  double c11 = F3(i, j + 1, k + 1) * (1 - tx) + F3(i + 1, j + 1, k + 1) * tx;
#undef F3
  double c0 = c00 * (1 - ty) + c10 * ty;
  double c1 = c01 * (1 - ty) + c11 * ty;
  return c0 * (1 - tz) + c1 * tz;
}

static int in_metal_3d(const uint8_t *e, double x, double y, double z, int64_t nx,
                       int64_t ny, int64_t nz) {
  if (x < 0.0 || y < 0.0 || z < 0.0 || x > nx - 1 || y > ny - 1 || z > nz - 1) return 0;
  int64_t i = (int64_t)x, j = (int64_t)y, k = (int64_t)z;
  if (i > nx - 2) i = nx - 2;
  if (j > ny - 2) j = ny - 2;
  if (k > nz - 2) k = nz - 2;
#define E3(a, b, c) e[((a)*ny + (b)) * nz + (c)]
  return E3(i, j, k) && E3(i + 1, j, k) && E3(i, j + 1, k) && E3(i + 1, j + 1, k) &&
         E3(i, j, k + 1) && E3(i + 1, j, k + 1) && E3(i, j + 1, k + 1) &&
         E3(i + 1, j + 1, k + 1);
#undef E3
}

static double py_mod1(double a) {
  double m = fmod(a, 1.0);
  if (m != 0.0) {
    if (m < 0.0) m += 1.0;
  } else {
    m = 0.0;
  }
  return m;
}

static double wave3d(int64_t kind, double om, double ph, double amp, double off,
                     double duty, const double *tab_t, const double *tab_v,
                     int64_t o0, int64_t o1, double t) {
  if (kind == 0) return amp * sin(om * t + ph) + off;
  if (kind == 1) return amp * cos(om * t + ph) + off;
  if (kind == 2) {
    double base;
    if (duty == 0.5)
      base = sin(om * t + ph) >= 0.0 ? 1.0 : -1.0;
    else
      base = py_mod1((om * t + ph) / (2.0 * M_PI)) < duty ? 1.0 : -1.0;
    return amp * base + off;
  }
  int64_t n = o1 - o0;
  if (n == 0) return off;
  if (t <= tab_t[o0]) return amp * tab_v[o0] + off;
  if (t >= tab_t[o1 - 1]) return amp * tab_v[o1 - 1] + off;
  int64_t lo = o0, hi = o1 - 1;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) / 2;
    if (tab_t[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  if (kind == 3) return amp * tab_v[lo] + off;
  double f = (t - tab_t[lo]) / (tab_t[hi] - tab_t[lo]);
  return amp * (tab_v[lo] + f * (tab_v[hi] - tab_v[lo])) + off;
}

static void efield3(const or_xyz_fields *f, double px, double py, double pz, double tt,
                    double *ex, double *ey, double *ez) {
  int64_t nx = f->nx, ny = f->ny, nz = f->nz, vol = nx * ny * nz;
  double inv_h = 1.0 / f->h_mm;
  double gx = px * inv_h, gy = py * inv_h, gz = pz * inv_h;
  double a = or_tri(f->EAx, gx, gy, gz, nx, ny, nz);
  double b = or_tri(f->EAy, gx, gy, gz, nx, ny, nz);
  double c = or_tri(f->EAz, gx, gy, gz, nx, ny, nz);
  for (int k = 0; k < f->K; ++k) {
    double w = wave3d(f->ch_kind[k], f->ch_om[k], f->ch_ph[k], f->ch_amp[k], f->ch_off[k],
                      f->ch_duty[k], f->tab_t, f->tab_v, f->tab_off[k], f->tab_off[k + 1], tt);
    a = a + w * or_tri(f->ExK + k * vol, gx, gy, gz, nx, ny, nz);
    b = b + w * or_tri(f->EyK + k * vol, gx, gy, gz, nx, ny, nz);
    c = c + w * or_tri(f->EzK + k * vol, gx, gy, gz, nx, ny, nz);
  }
  *ex = a;
  *ey = b;
  *ez = c;
}

/* a record row (tracer3d.py:441-451 and the other record sites); with
 * an 11th column, the transporter pass count (wrs), which the station-
 * plane row leaves at 0 */
static void rec_row(double *row, double t_us, double x, double y, double z, double vx,
                    double vy, double vz, const or_xyz_fields *f, int64_t C, int64_t wrap) {
  row[0] = t_us;
  row[1] = x;
  row[2] = y;
  row[3] = z;
  row[4] = vx * 1e-3;
  row[5] = vy * 1e-3;
  row[6] = vz * 1e-3;
  efield3(f, x, y, z, t_us, row + 7, row + 8, row + 9);
  if (C > 10) row[10] = (double)wrap;
}

static void fly_one3(const or_xyz_fields *f, const or_params *p, const or_ions *io,
                     or_out *out, int64_t ii) {
  double qm = io->acc[ii], m_amu = io->m_ion[ii];
  double x = io->x[ii], y = io->y[ii], z = io->z[ii];
  double vx = io->vx[ii] * 1e3, vy = io->vy[ii] * 1e3, vz = io->vz[ii] * 1e3;
  double tob = io->tob[ii], dt = p->dt, t_max = p->t_max;
  double sigma = io->sigma ? io->sigma[ii] : 0.0;
  int64_t nx = f->nx, ny = f->ny, nz = f->nz;
  double h = f->h_mm, inv_h = 1.0 / h;
  int collide_on = p->gas == 1;
  or_rng rng;
  if (p->rng == OR_RNG_MT)
    or_mt_seed(&rng, (uint32_t)io->seed[ii]);
  else
    or_philox_seed(&rng, (uint64_t)io->seed[ii], p->rng_key);
  double *rec = p->rec_rows > 0 && out->rec ? out->rec + ii * p->rec_rows * p->rec_cols : 0;
  int64_t R = p->rec_rows, C = p->rec_cols;
  const double c = 1e3;
  if (rec) rec_row(rec, tob, x, y, z, vx, vy, vz, f, C, 0);
  int64_t nrec = 1, step = 0, ncol = 0, tp_pass = 0;
  const double tp_d = p->tp_emit - p->tp_accept;
  int kind = 2, hit_ip = -1;
  double t = 0.0;
  int64_t limit = p->max_steps;
#define ACC(px, py, pz, tt, ax, ay, az)           \
  do {                                            \
    double e0, e1, e2;                            \
    efield3(f, px, py, pz, tt, &e0, &e1, &e2);    \
    ax = qm * 1e3 * e0;                           \
    ay = qm * 1e3 * e1;                           \
    az = qm * 1e3 * e2;                           \
  } while (0)
  while (t < t_max) {
    if (limit > 0 && step >= limit) {
      kind = -2;
      break;
    }
    double xo = x, yo = y, zo = z, vxo = vx, vyo = vy, vzo = vz, to = t;
    double t_us = tob + t * 1e6;
    double h_us = dt * 1e6;
    double ax1, ay1, az1, ax2, ay2, az2, ax3, ay3, az3, ax4, ay4, az4;
    ACC(x, y, z, t_us, ax1, ay1, az1);
    double k1x = vx, k1y = vy, k1z = vz;
    double px = x + 0.5 * dt * k1x * c, py = y + 0.5 * dt * k1y * c, pz = z + 0.5 * dt * k1z * c;
    ACC(px, py, pz, t_us + 0.5 * h_us, ax2, ay2, az2);
    double k2x = vx + 0.5 * dt * ax1, k2y = vy + 0.5 * dt * ay1, k2z = vz + 0.5 * dt * az1;
    px = x + 0.5 * dt * k2x * c;
    py = y + 0.5 * dt * k2y * c;
    pz = z + 0.5 * dt * k2z * c;
    ACC(px, py, pz, t_us + 0.5 * h_us, ax3, ay3, az3);
    double k3x = vx + 0.5 * dt * ax2, k3y = vy + 0.5 * dt * ay2, k3z = vz + 0.5 * dt * az2;
    px = x + dt * k3x * c;
    py = y + dt * k3y * c;
    pz = z + dt * k3z * c;
    ACC(px, py, pz, t_us + h_us, ax4, ay4, az4);
    double k4x = vx + dt * ax3, k4y = vy + dt * ay3, k4z = vz + dt * az3;
    x = x + dt / 6.0 * (k1x + 2 * k2x + 2 * k3x + k4x) * c;
    y = y + dt / 6.0 * (k1y + 2 * k2y + 2 * k3y + k4y) * c;
    z = z + dt / 6.0 * (k1z + 2 * k2z + 2 * k3z + k4z) * c;
    vx = vx + dt / 6.0 * (ax1 + 2 * ax2 + 2 * ax3 + ax4);
    vy = vy + dt / 6.0 * (ay1 + 2 * ay2 + 2 * ay3 + ay4);
    vz = vz + dt / 6.0 * (az1 + 2 * az2 + 2 * az3 + az4);
    t += dt;
    step += 1;
    if (collide_on) {
      double sx = vx * 1e-3, sy = vy * 1e-3, sz = vz * 1e-3;
      int hit = or_hs_step(&rng, &sx, &sy, &sz, p->vg[0], p->vg[1], p->vg[2], dt, 1e6, p->T_k,
                           p->P_pa, sigma, p->c_star, p->c_bar, m_amu, p->m_gas, p->sig1d);
      if (hit < 0) {
        kind = -3;
        break;
      }
      if (hit) {
        vx = sx * 1e3;
        vy = sy * 1e3;
        vz = sz * 1e3;
        ncol += 1;
      }
    }
    if (p->tp_col >= 0) {
      double cn, co;
      if (p->tp_col == 0) {
        cn = x;
        co = xo;
      } else if (p->tp_col == 1) {
        cn = y;
        co = yo;
      } else {
        cn = z;
        co = zo;
      }
      if ((cn - p->tp_accept) * p->tp_sgn >= 0.0 && (co - p->tp_accept) * p->tp_sgn < 0.0) {
        tp_pass += 1;
        if (tp_pass >= p->tp_maxp) {
          kind = 4;
          if (rec && nrec < R) rec_row(rec + nrec * C, tob + t * 1e6, x, y, z, vx, vy, vz, f, C,
                                       tp_pass);
          if (nrec < R) nrec++;
          break;
        }
        if (p->tp_col == 0) {
          x += tp_d;
          xo += tp_d;
        } else if (p->tp_col == 1) {
          y += tp_d;
          yo += tp_d;
        } else {
          z += tp_d;
          zo += tp_d;
        }
      }
    }
    if (in_metal_3d(f->ele, x * inv_h, y * inv_h, z * inv_h, nx, ny, nz)) {
      double f0 = 0.0, f1 = 1.0;
      for (int b = 0; b < 40; ++b) {
        double fm = 0.5 * (f0 + f1);
        if (in_metal_3d(f->ele, (xo + fm * (x - xo)) * inv_h, (yo + fm * (y - yo)) * inv_h,
                        (zo + fm * (z - zo)) * inv_h, nx, ny, nz))
          f1 = fm;
        else
          f0 = fm;
      }
      double fr = f1;
      x = xo + fr * (x - xo);
      y = yo + fr * (y - yo);
      z = zo + fr * (z - zo);
      vx = vxo + fr * (vx - vxo);
      vy = vyo + fr * (vy - vyo);
      vz = vzo + fr * (vz - vzo);
      t = to + fr * dt;
      kind = 0;
      if (rec && nrec < R)
        rec_row(rec + nrec * C, tob + t * 1e6, x, y, z, vx, vy, vz, f, C, tp_pass);
      if (nrec < R) nrec++;
      break;
    }
    if (p->n_pl > 0) {
      int hit_pl = 0;
      for (int ip = 0; ip < p->n_pl; ++ip) {
        int64_t col = p->pl_col[ip];
        double cn, co;
        if (col == 0) {
          cn = x;
          co = xo;
        } else if (col == 1) {
          cn = y;
          co = yo;
        } else {
          cn = z;
          co = zo;
        }
        double sgn = p->pl_sgn[ip], val = p->pl_val[ip];
        int crossed;
        if (sgn == 0.0)
          crossed = (cn - val) * (co - val) <= 0.0 && cn != co;
        else
          crossed = (cn - val) * sgn >= 0.0 && (co - val) * sgn < 0.0;
        if (!crossed) continue;
        double den = cn - co, fr;
        fr = den == 0.0 ? 0.0 : (val - co) / den;
        if (fr < 0.0) fr = 0.0;
        if (fr > 1.0) fr = 1.0;
        double xc = xo + fr * (x - xo), yc = yo + fr * (y - yo), zc = zo + fr * (z - zo);
        double t1, t2;
        if (col == 0) {
          t1 = yc;
          t2 = zc;
        } else if (col == 1) {
          t1 = xc;
          t2 = zc;
        } else {
          t1 = xc;
          t2 = yc;
        }
        const double *pw = p->pl_w + 4 * ip;
        int ins = pw[0] <= t1 && t1 <= pw[1] && pw[2] <= t2 && t2 <= pw[3];
        if (p->pl_kind[ip] == 6) {
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
        kind = (int)p->pl_kind[ip];
        hit_ip = ip;
        hit_pl = 1;
        break;
      }
      if (hit_pl) {
        if (rec && nrec < R) rec_row(rec + nrec * C, tob + t * 1e6, x, y, z, vx, vy, vz, f, C, 0);
        if (nrec < R) nrec++;
        break;
      }
    }
    if (x < 0.0 || y < 0.0 || z < 0.0 || x > (nx - 1) * h || y > (ny - 1) * h ||
        z > (nz - 1) * h) {
      kind = 1;
      if (rec && nrec < R)
        rec_row(rec + nrec * C, tob + t * 1e6, x, y, z, vx, vy, vz, f, C, tp_pass);
      if (nrec < R) nrec++;
      break;
    }
    if (step % p->rec_every == 0 && nrec < R) {
      if (rec) rec_row(rec + nrec * C, tob + t * 1e6, x, y, z, vx, vy, vz, f, C, tp_pass);
      nrec++;
    }
  }
#undef ACC
  out->x[ii] = x;
  out->y[ii] = y;
  out->z[ii] = z;
  out->vx[ii] = vx * 1e-3;
  out->vy[ii] = vy * 1e-3;
  out->vz[ii] = vz * 1e-3;
  out->t[ii] = tob + t * 1e6;
  out->kind[ii] = kind;
  out->face[ii] = hit_ip;
  out->label[ii] = 0;
  out->ncol[ii] = ncol;
  out->steps[ii] = step;
  out->nrec[ii] = nrec;
}

/* tracer3d._fly3d_sds (tracer3d.py:851-1025): mm, us, mm/us throughout;
 * dt and t_max in us; one field evaluation per step; the damped kick,
 * the drift, the diffusion jump when mfp > 0, then metal (no bisection),
 * box and planes crossed over the whole step; records (t_us, x, y, z,
 * vx, vy, vz, Ex, Ey, Ez) with the field at the record point and time
 * (_sds_field, tracer3d.py:792-812). */
static void rec_row_sds(double *row, double t_us, double x, double y, double z, double vx,
                        double vy, double vz, const or_xyz_fields *f) {
  row[0] = t_us;
  row[1] = x;
  row[2] = y;
  row[3] = z;
  row[4] = vx;
  row[5] = vy;
  row[6] = vz;
  efield3(f, x, y, z, t_us, row + 7, row + 8, row + 9);
}

static void fly_one3_sds(const or_xyz_fields *f, const or_params *p, const or_ions *io,
                         or_out *out, int64_t ii) {
  const double qm = io->acc[ii];
  double x = io->x[ii], y = io->y[ii], z = io->z[ii];
  double vx = io->vx[ii], vy = io->vy[ii], vz = io->vz[ii];
  const double tob = io->tob[ii], dt = p->dt, t_max = p->t_max;
  const double damping = io->sds ? io->sds[4 * ii] : 0.0;
  const double mfp = io->sds ? io->sds[4 * ii + 1] : 0.0;
  const double V = io->sds ? io->sds[4 * ii + 2] : 0.0;
  const double logmr = io->sds ? io->sds[4 * ii + 3] : 0.0;
  const int64_t nx = f->nx, ny = f->ny, nz = f->nz;
  const double h = f->h_mm, inv_h = 1.0 / h;
  or_rng rng;
  if (p->rng == OR_RNG_MT)
    or_mt_seed(&rng, (uint32_t)io->seed[ii]);
  else
    or_philox_seed(&rng, (uint64_t)io->seed[ii], p->rng_key);
  double *rec = p->rec_rows > 0 && out->rec ? out->rec + ii * p->rec_rows * p->rec_cols : 0;
  const int64_t R = p->rec_rows, C = p->rec_cols;
  double t = 0.0;
  if (rec) rec_row_sds(rec, tob, x, y, z, vx, vy, vz, f);
  int64_t nrec = 1, step = 0;
  int kind = 2, hit_ip = -1;
  const int64_t limit = p->max_steps;
  while (t < t_max) {
    if (limit > 0 && step >= limit) {
      kind = -2;
      break;
    }
    double ex, ey, ez;
    efield3(f, x, y, z, tob + t, &ex, &ey, &ez);
    const double afx = qm * 1e-6 * ex, afy = qm * 1e-6 * ey, afz = qm * 1e-6 * ez;
    double aex, aey, aez;
    if (damping > 0.0) {
      const double tterm = damping * dt;
      const double factor = (1.0 - exp(-tterm)) / tterm;
      aex = factor * (afx - (vx - p->vg[0]) * damping);
      aey = factor * (afy - (vy - p->vg[1]) * damping);
      aez = factor * (afz - (vz - p->vg[2]) * damping);
    } else {
      aex = afx;
      aey = afy;
      aez = afz;
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
      or_sds_jump(&rng, p->sds_stats, logmr, V, mfp, dt, &jx, &jy, &jz);
      x = x + jx;
      y = y + jy;
      z = z + jz;
    }
    t += dt;
    step += 1;
    if (in_metal_3d(f->ele, x * inv_h, y * inv_h, z * inv_h, nx, ny, nz)) {
      kind = 0;
      if (rec && nrec < R) rec_row_sds(rec + nrec * C, tob + t, x, y, z, vx, vy, vz, f);
      if (nrec < R) nrec++;
      break;
    }
    if (x < 0.0 || y < 0.0 || z < 0.0 || x > (nx - 1) * h || y > (ny - 1) * h ||
        z > (nz - 1) * h) {
      kind = 1;
      if (rec && nrec < R) rec_row_sds(rec + nrec * C, tob + t, x, y, z, vx, vy, vz, f);
      if (nrec < R) nrec++;
      break;
    }
    int hit_pl = 0;
    for (int ip = 0; ip < p->n_pl; ++ip) {
      const int64_t col = p->pl_col[ip];
      const double cn = col == 0 ? x : (col == 1 ? y : z);
      const double co = col == 0 ? xo : (col == 1 ? yo : zo);
      const double sgn = p->pl_sgn[ip], val = p->pl_val[ip];
      int crossed;
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
      const double w1 = col == 0 ? yc : xc, w2 = col == 2 ? yc : zc;
      const double *pw = p->pl_w + 4 * ip;
      const int ins = pw[0] <= w1 && w1 <= pw[1] && pw[2] <= w2 && w2 <= pw[3];
      if (p->pl_kind[ip] == 6) {
        if (!ins) continue;
      } else if (p->pl_kind[ip] == 5 && ins) {
        continue;
      }
      x = xc;
      y = yc;
      z = zc;
      vx = vxo + fr * (vx - vxo);
      vy = vyo + fr * (vy - vyo);
      vz = vzo + fr * (vz - vzo);
      t = to + fr * (t - to);
      kind = (int)p->pl_kind[ip];
      hit_ip = ip;
      hit_pl = 1;
      break;
    }
    if (hit_pl) {
      if (rec && nrec < R) rec_row_sds(rec + nrec * C, tob + t, x, y, z, vx, vy, vz, f);
      if (nrec < R) nrec++;
      break;
    }
    if (step % p->rec_every == 0 && nrec < R) {
      if (rec) rec_row_sds(rec + nrec * C, tob + t, x, y, z, vx, vy, vz, f);
      nrec++;
    }
  }
  out->x[ii] = x;
  out->y[ii] = y;
  out->z[ii] = z;
  out->vx[ii] = vx;
  out->vy[ii] = vy;
  out->vz[ii] = vz;
  out->t[ii] = tob + t;
  out->kind[ii] = kind;
  out->face[ii] = hit_ip;
  out->label[ii] = 0;
  out->ncol[ii] = 0;
  out->steps[ii] = step;
  out->nrec[ii] = nrec;
}

int or_xyz_fly(const or_xyz_fields *f, const or_params *p, const or_ions *ions,
               or_out *out, int n_threads) {
  if (!f || !p || !ions || !out || f->nx < 2 || f->ny < 2 || f->nz < 2 || p->rec_every < 1)
    return -1;
  if (p->gas == 2 && !p->sds_stats) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t i = 0; i < ions->n; ++i) {
    if (p->gas == 2)
      fly_one3_sds(f, p, ions, out, i);
    else
      fly_one3(f, p, ions, out, i);
  }
  return 0;
}
