/*
 * Planar flight of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * A line-by-line restatement, in strict IEEE double arithmetic, of
 * ion_gym/physics/build_planar.py:1111-1429 (_fly_planar, its record row
 * :1162-1214 and its _efield :1147-1157) with the helpers _bilin
 * (build_planar.py:1039-1056), raster2d._metal_nn (raster2d.py:597-612), the
 * 2-D unit waveform _wave_eval (build_planar.py:1059-1096), the shared
 * hard-sphere step and the SDS jump.  The record flags follow the planar
 * channel order: speed, ke_ev, ke_x, ke_y, ke_z, e_field, e_axial,
 * e_radial, e_x, e_y, e_z, radius, n_col.
 */
#include <math.h>
#include <string.h>

#include "ig_oracle.h"
#include "ig_oracle_gas.h"
#include "ig_oracle_rng.h"

#define E_CHG 1.602176634e-19   /* ion_gym/physics/constants.py:18 */
#define KG_AMU 1.66053906660e-27 /* ion_gym/physics/constants.py:23 */

static int metal_nn(const double *ele, double gx, double gy, int64_t nx, int64_t ny) {
  double fi = floor(gx + 0.5), fj = floor(gy + 0.5);
  if (!(fi >= 0.0) || !(fj >= 0.0) || fi > (double)(nx - 1) || fj > (double)(ny - 1)) return 0;
  return ele[(int64_t)fi * ny + (int64_t)fj] > 0.5;
}

static void efield(const or_planar_fields *f, double px, double py, double tt, double *ex,
                   double *ey) {
  int64_t nx = f->nx, ny = f->ny, plane = nx * ny;
  double gx = px / f->mm;
  double gy = py / f->mm;
  double a = or_bilin(f->ExA, gx, gy, nx, ny);
  double b = or_bilin(f->EyA, gx, gy, nx, ny);
  for (int k = 0; k < f->K; ++k) {
    double w = or_wave_rz(f->ch_kind[k], f->ch_om[k], f->ch_ph[k], f->ch_duty[k], f->tab_t,
                          f->tab_v, f->tab_off[k], f->tab_off[k + 1], tt);
    a = a + w * or_bilin(f->ExK + k * plane, gx, gy, nx, ny);
    b = b + w * or_bilin(f->EyK + k * plane, gx, gy, nx, ny);
  }
  *ex = a;
  *ey = b;
}

static void write_row(double *row, const int32_t *fl, double t, double x, double y, double z,
                      double vx, double vy, double vz, double exv, double eyv, int64_t nc,
                      double m_ion, double axx, double axy) {
  row[0] = t;
  row[1] = x;
  row[2] = y;
  row[3] = z;
  row[4] = vx;
  row[5] = vy;
  row[6] = vz;
  int c = 7;
  double sp = sqrt(vx * vx + vy * vy + vz * vz);
  if (fl[0]) row[c++] = sp;
  if (fl[1]) {
    double s = sp * 1e3;
    row[c++] = 0.5 * m_ion * KG_AMU * (s * s) / E_CHG;
  }
  if (fl[2]) {
    double s = vx * 1e3;
    row[c++] = 0.5 * m_ion * KG_AMU * (s * s) / E_CHG;
  }
  if (fl[3]) {
    double s = vy * 1e3;
    row[c++] = 0.5 * m_ion * KG_AMU * (s * s) / E_CHG;
  }
  if (fl[4]) {
    double s = vz * 1e3;
    row[c++] = 0.5 * m_ion * KG_AMU * (s * s) / E_CHG;
  }
  if (fl[5]) row[c++] = sqrt(exv * exv + eyv * eyv);
  if (fl[6]) row[c++] = exv;
  if (fl[7]) row[c++] = eyv;
  if (fl[8]) row[c++] = exv;
  if (fl[9]) row[c++] = eyv;
  if (fl[10]) row[c++] = 0.0;
  if (fl[11]) row[c++] = sqrt((x - axx) * (x - axx) + (y - axy) * (y - axy));
  if (fl[12]) row[c++] = (double)nc;
}

static void fly_one(const or_planar_fields *f, const or_params *p, const or_ions *io,
                    or_out *out, int64_t ii) {
  double x = io->x[ii], y = io->y[ii], z = io->z[ii];
  double vx = io->vx[ii], vy = io->vy[ii], vz = io->vz[ii];
  double tob = io->tob[ii], m_ion = io->m_ion[ii], acc = io->acc[ii];
  double sigma = io->sigma ? io->sigma[ii] : 0.0;
  double dt = p->dt, t_max = p->t_max;
  int sds_on = p->gas == 2;
  int collide_on = p->gas == 1 && p->P_pa > 0.0;
  double damp = 0, mfp = 0, V = 0, logmr = 0;
  if (sds_on && io->sds) {
    damp = io->sds[4 * ii];
    mfp = io->sds[4 * ii + 1];
    V = io->sds[4 * ii + 2];
    logmr = io->sds[4 * ii + 3];
  }
  double gfx = p->vg[0], gfy = p->vg[1], gfz = p->vg[2];
  int64_t nx = f->nx, ny = f->ny;
  double mm = f->mm;
  double axx = 0.5 * (nx - 1) * mm, axy = 0.5 * (ny - 1) * mm;
  or_rng rng;
  if (p->rng == OR_RNG_MT)
    or_mt_seed(&rng, (uint32_t)io->seed[ii]);
  else
    or_philox_seed(&rng, (uint64_t)io->seed[ii], p->rng_key);
  double *rec = p->rec_rows > 0 && out->rec ? out->rec + ii * p->rec_rows * p->rec_cols : 0;
  double t = 0.0;
  int64_t ncol = 0, step = 0;
  double ex, ey;
  efield(f, x, y, tob, &ex, &ey);
  double fxr = ex, fyr = ey;
  if (rec) write_row(rec, p->flags, tob, x, y, z, vx, vy, vz, fxr * 1e-3, fyr * 1e-3, ncol, m_ion,
                     axx, axy);
  int64_t nrec = 1;
  int kind = 2, face = -1;
  int64_t limit = p->max_steps;
  while (t < t_max) {
    if (limit > 0 && step >= limit) {
      kind = -2; /* window ended: still flying */
      break;
    }
    double xo = x, yo = y, zo = z, vxo = vx, vyo = vy, to = t;
    efield(f, x, y, tob + t, &ex, &ey);
    ex *= 1e-9 * acc;
    ey *= 1e-9 * acc;
    double az = 0.0;
    if (sds_on) {
      double tterm = damp * dt;
      double factor = (1.0 - exp(-tterm)) / tterm;
      ex = factor * (ex - (vx - gfx) * damp);
      ey = factor * (ey - (vy - gfy) * damp);
      az = -factor * (vz - gfz) * damp;
    }
    vx += 0.5 * ex * dt;
    vy += 0.5 * ey * dt;
    vz += 0.5 * az * dt;
    x += vx * dt;
    y += vy * dt;
    z += vz * dt;
    efield(f, x, y, tob + t + dt, &ex, &ey);
    fxr = ex;
    fyr = ey;
    ex *= 1e-9 * acc;
    ey *= 1e-9 * acc;
    if (sds_on) {
      double tterm = damp * dt;
      double factor = (1.0 - exp(-tterm)) / tterm;
      ex = factor * (ex - (vx - gfx) * damp);
      ey = factor * (ey - (vy - gfy) * damp);
      az = -factor * (vz - gfz) * damp;
    }
    vx += 0.5 * ex * dt;
    vy += 0.5 * ey * dt;
    vz += 0.5 * az * dt;
    t += dt;
    step += 1;
    if (sds_on) {
      double jx, jy, jz;
      or_sds_jump(&rng, p->sds_stats, logmr, V, mfp, dt, &jx, &jy, &jz);
      x += jx;
      y += jy;
      z += jz;
    } else if (collide_on) {
      int hit = or_hs_step(&rng, &vx, &vy, &vz, gfx, gfy, gfz, dt, 1.0, p->T_k, p->P_pa, sigma,
                           p->c_star, p->c_bar, m_ion, p->m_gas, p->sig1d);
      if (hit < 0) {
        kind = -3;
        break;
      }
      if (hit) ncol += 1;
    }
    if (metal_nn(f->ele, x / mm, y / mm, nx, ny)) {
      double f0 = 0.0, f1 = 1.0;
      for (int bs = 0; bs < 40; ++bs) {
        double fm = 0.5 * (f0 + f1);
        if (metal_nn(f->ele, (xo + fm * (x - xo)) / mm, (yo + fm * (y - yo)) / mm, nx, ny))
          f1 = fm;
        else
          f0 = fm;
      }
      x = xo + f1 * (x - xo);
      y = yo + f1 * (y - yo);
      vx = vxo + f1 * (vx - vxo);
      vy = vyo + f1 * (vy - vyo);
      t = to + f1 * dt;
      kind = 0;
      break;
    }
    if (x < 0 || y < 0 || x / mm > (double)(nx - 1) || y / mm > (double)(ny - 1)) {
      kind = 1;
      break;
    }
    if (p->n_pl > 0) {
      int hit_pl = 0;
      for (int ip = 0; ip < p->n_pl; ++ip) {
        int64_t pcol = p->pl_col[ip];
        double cn = pcol == 0 ? x : (pcol == 1 ? y : z);
        double co = pcol == 0 ? xo : (pcol == 1 ? yo : zo);
        double sgn = p->pl_sgn[ip], val = p->pl_val[ip];
        int crossed;
        if (sgn == 0.0)
          crossed = (cn - val) * (co - val) <= 0.0 && cn != co;
        else
          crossed = (cn - val) * sgn >= 0.0 && (co - val) * sgn < 0.0;
        if (!crossed) continue;
        double den = cn - co, fr = den == 0.0 ? 0.0 : (val - co) / den;
        if (fr < 0.0) fr = 0.0;
        if (fr > 1.0) fr = 1.0;
        double xc = xo + fr * (x - xo), yc = yo + fr * (y - yo), zc = zo + fr * (z - zo);
        double w1 = pcol == 0 ? yc : xc;
        double w2 = pcol == 2 ? yc : zc;
        const double *pw = p->pl_w + 4 * ip;
        int ins = pw[0] <= w1 && w1 <= pw[1] && pw[2] <= w2 && w2 <= pw[3];
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
        t = to + fr * dt;
        kind = (int)p->pl_kind[ip];
        hit_pl = 1;
        break;
      }
      if (hit_pl) break;
    }
    {
      int stop = 0;
      for (int fc = 0; fc < 6; ++fc) {
        if (!p->bnd_on[fc]) continue;
        double v = fc < 2 ? x : (fc < 4 ? y : z);
        if ((fc & 1) ? v > p->bnd_val[fc] : v < p->bnd_val[fc]) {
          kind = 3;
          face = fc;
          stop = 1;
          break;
        }
      }
      if (stop) break;
    }
    if (rec && ((int64_t)(t / dt)) % p->rec_every == 0 && nrec < p->rec_rows - 1) {
      write_row(rec + nrec * p->rec_cols, p->flags, tob + t, x, y, z, vx, vy, vz, fxr * 1e-3,
                fyr * 1e-3, ncol, m_ion, axx, axy);
      nrec += 1;
    }
  }
  if (rec && kind != -2)
    write_row(rec + nrec * p->rec_cols, p->flags, tob + t, x, y, z, vx, vy, vz, fxr * 1e-3,
              fyr * 1e-3, ncol, m_ion, axx, axy);
  if (kind != -2) nrec += 1;
  out->x[ii] = x;
  out->y[ii] = y;
  out->z[ii] = z;
  out->vx[ii] = vx;
  out->vy[ii] = vy;
  out->vz[ii] = vz;
  out->t[ii] = tob + t;
  out->kind[ii] = kind;
  out->face[ii] = face;
  out->label[ii] = 0;
  out->ncol[ii] = ncol;
  out->steps[ii] = step;
  out->nrec[ii] = nrec;
}

int or_planar_fly(const or_planar_fields *f, const or_params *p, const or_ions *ions,
                  or_out *out, int n_threads) {
  if (!f || !p || !ions || !out || f->nx < 2 || f->ny < 2 || p->rec_every < 1) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t i = 0; i < ions->n; ++i) fly_one(f, p, ions, out, i);
  return 0;
}
