/*
 * r-z flight of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * A line-by-line restatement, in strict IEEE double arithmetic, of
 * ion_gym/physics/tracer_rz.py:113-446 (_fly_rec_full and _write_row
 * :40-111) with its helpers interp._bilin (interp.py:19-36),
 * raster2d._metal_nn (raster2d.py:597-612) and the 2-D unit waveform
 * build_planar._wave_eval (build_planar.py:1059-1096).  Each ion is
 * independent; OpenMP spreads ions over threads with a dynamic schedule.
 */
#include <math.h>
#include <string.h>

#include "ig_oracle.h"
#include "ig_oracle_gas.h"
#include "ig_oracle_rng.h"

#define E_CHG 1.602176634e-19   /* ion_gym/physics/constants.py:18 */
#define KG_AMU 1.66053906660e-27 /* ion_gym/physics/constants.py:23 */

double or_bilin(const double *F, double gx, double gu, int64_t nx, int64_t nu) {
  int64_t i = (int64_t)gx, j = (int64_t)gu;
  if (i < 0) i = 0;
  if (j < 0) j = 0;
  if (i > nx - 2) i = nx - 2;
  if (j > nu - 2) j = nu - 2;
  double fx = gx - (double)i, fu = gu - (double)j;
  if (fx < 0.0) fx = 0.0;
  if (fx > 1.0) fx = 1.0;
  if (fu < 0.0) fu = 0.0;
  if (fu > 1.0) fu = 1.0;
  const double *r0 = F + i * nu, *r1 = F + (i + 1) * nu;
  return r0[j] * (1 - fx) * (1 - fu) + r1[j] * fx * (1 - fu) +
         r0[j + 1] * (1 - fx) * fu + r1[j + 1] * fx * fu;
}

static int metal_nn(const double *ele, double gx, double gy, int64_t nx,
                    int64_t ny) {
  int64_t i = (int64_t)floor(gx + 0.5), j = (int64_t)floor(gy + 0.5);
  if (i < 0 || j < 0 || i > nx - 1 || j > ny - 1) return 0;
  return ele[i * ny + j] > 0.5;
}

static double py_mod1(double a) {
  /* Python's float a % 1.0 (result carries the divisor's sign) */
  double m = fmod(a, 1.0);
  if (m != 0.0) {
    if (m < 0.0) m += 1.0;
  } else {
    m = copysign(0.0, 1.0);
  }
  return m;
}

double or_wave_rz(int64_t kind, double om, double ph, double duty,
                  const double *tab_t, const double *tab_v, int64_t o0,
                  int64_t o1, double t) {
  if (kind == 0) return sin(om * t + ph);
  if (kind == 1) return cos(om * t + ph);
  if (kind == 2) {
    if (duty == 0.5) return sin(om * t + ph) >= 0.0 ? 1.0 : -1.0;
    double frac = py_mod1((om * t + ph) / (2.0 * M_PI));
    return frac < duty ? 1.0 : -1.0;
  }
  int64_t n = o1 - o0;
  if (n == 0) return 0.0;
  if (t <= tab_t[o0]) return tab_v[o0];
  if (t >= tab_t[o1 - 1]) return tab_v[o1 - 1];
  int64_t lo = o0, hi = o1 - 1;
  while (hi - lo > 1) {
    int64_t mid = (lo + hi) / 2;
    if (tab_t[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  if (kind == 3) return tab_v[lo];
  double f = (t - tab_t[lo]) / (tab_t[hi] - tab_t[lo]);
  return tab_v[lo] + f * (tab_v[hi] - tab_v[lo]);
}

typedef struct {
  const or_rz_fields *f;
  double acc;
} efctx;

static void efield(const efctx *c, double px, double pr, double tt, double *ez,
                   double *er) {
  const or_rz_fields *f = c->f;
  int64_t nx = f->nx, nu = f->nu, plane = nx * nu;
  double gx = px / f->mm;
  double gu = (pr - f->u0) / f->mm;
  double z = or_bilin(f->EzA, gx, gu, nx, nu);
  double r = or_bilin(f->EuA, gx, gu, nx, nu);
  for (int k = 0; k < f->K; ++k) {
    double w = or_wave_rz(f->ch_kind[k], f->ch_om[k], f->ch_ph[k], f->ch_duty[k],
                          f->tab_t, f->tab_v, f->tab_off[k], f->tab_off[k + 1], tt);
    z = z + w * or_bilin(f->EzK + k * plane, gx, gu, nx, nu);
    r = r + w * or_bilin(f->EuK + k * plane, gx, gu, nx, nu);
  }
  if (f->tau_gate >= 0.0 && tt >= f->tau_gate && f->EzG) {
    z = z + or_bilin(f->EzG, gx, gu, nx, nu);
    r = r + or_bilin(f->EuG, gx, gu, nx, nu);
  }
  *ez = z * 1e-9 * c->acc;
  *er = r * 1e-9 * c->acc;
}

static void write_row(double *row, const int32_t *fl, double t, double x,
                      double y, double z, double vx, double vy, double vz,
                      double ez, double er, int64_t ncol, double m_ion,
                      double path, double eat, double ket) {
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
  if (fl[0]) row[c++] = sp;
  if (fl[1]) {
    double s = sp * 1e3;
    row[c++] = 0.5 * m_ion * KG_AMU * (s * s) / E_CHG;
  }
  double kes = 0.5 * m_ion * KG_AMU * 1e6 / E_CHG;
  if (fl[2]) row[c++] = kes * vx * vx;
  if (fl[3]) row[c++] = kes * vy * vy;
  if (fl[4]) row[c++] = kes * vz * vz;
  if (fl[5]) row[c++] = sqrt(ez * ez + er * er);
  if (fl[6]) row[c++] = ez;
  if (fl[7]) row[c++] = er;
  if (fl[8]) row[c++] = r;
  if (fl[9]) row[c++] = (double)ncol;
  if (fl[10]) row[c++] = path;
  if (fl[11]) row[c++] = eat;
  if (fl[12]) row[c++] = ket;
}

static void fly_one(const or_rz_fields *f, const or_params *p,
                    const or_ions *io, or_out *out, int64_t ii) {
  double x = io->x[ii], y = io->y[ii], z = io->z[ii];
  double vx = io->vx[ii], vy = io->vy[ii], vz = io->vz[ii];
  double tob = io->tob[ii], m_ion = io->m_ion[ii], acc = io->acc[ii];
  double sigma = io->sigma ? io->sigma[ii] : 0.0;
  double dt = p->dt, t_max = p->t_max;
  int sds_on = p->gas == 2;
  double sds_damping = 0, sds_mfp = 0, sds_V = 0, sds_logmr = 0;
  if (sds_on && io->sds) {
    sds_damping = io->sds[4 * ii];
    sds_mfp = io->sds[4 * ii + 1];
    sds_V = io->sds[4 * ii + 2];
    sds_logmr = io->sds[4 * ii + 3];
  }
  double vgx = p->vg[0], vgy = p->vg[1], vgz = p->vg[2];
  double P_pa = p->gas == 1 ? p->P_pa : 0.0;
  int64_t nx = f->nx, nu = f->nu;
  double mm = f->mm, u0 = f->u0;
  or_rng rng;
  if (p->rng == OR_RNG_MT)
    or_mt_seed(&rng, (uint32_t)io->seed[ii]);
  else
    or_philox_seed(&rng, (uint64_t)io->seed[ii], p->rng_key);
  efctx ec = {f, acc};
  double *rec = p->rec_rows > 0 && out->rec ? out->rec + ii * p->rec_rows * p->rec_cols : 0;
  double t = 0.0;
  int64_t ncol = 0, step = 0;
  double path = 0.0, eat = 0.0, ket = 0.0;
  double ke_scale = 0.5 * m_ion * KG_AMU * 1e6 / E_CHG;
  double sds_fac = 1.0;
  if (sds_on && sds_damping > 0.0) {
    double tt = sds_damping * dt;
    sds_fac = (1.0 - exp(-tt)) / tt;
  }
  double r0 = sqrt(y * y + z * z);
  double ez0, er0;
  efield(&ec, x, r0, tob, &ez0, &er0);
  double fscale = acc != 0 ? 1e-3 / (1e-9 * acc) : 0.0;
  if (rec) write_row(rec, p->flags, tob, x, y, z, vx, vy, vz, ez0 * fscale, er0 * fscale,
                     ncol, m_ion, path, eat, ket);
  int64_t nrec = 1;
  int kind = 2, face = -1, label = 0;
  double ez = ez0, er = er0;
  int64_t limit = p->max_steps;
  while (t < t_max) {
    if (limit > 0 && step >= limit) {
      kind = -2; /* window ended: still flying */
      break;
    }
    double xo = x, yo = y, zo = z, vxo = vx, vyo = vy, vzo = vz, to = t;
    double r = sqrt(y * y + z * z);
    efield(&ec, x, r, tob + t, &ez, &er);
    double rr = r > 1e-9 ? r : 1e-9;
    if (sds_on) {
      vx += 0.5 * sds_fac * (ez - (vx - vgx) * sds_damping) * dt;
      vy += 0.5 * sds_fac * (er * y / rr - (vy - vgy) * sds_damping) * dt;
      vz += 0.5 * sds_fac * (er * z / rr - (vz - vgz) * sds_damping) * dt;
    } else {
      vx += 0.5 * ez * dt;
      vy += 0.5 * er * y / rr * dt;
      vz += 0.5 * er * z / rr * dt;
    }
    x += vx * dt;
    y += vy * dt;
    z += vz * dt;
    path += sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo));
    r = sqrt(y * y + z * z);
    efield(&ec, x, r, tob + t + dt, &ez, &er);
    rr = r > 1e-9 ? r : 1e-9;
    if (sds_on) {
      vx += 0.5 * sds_fac * (ez - (vx - vgx) * sds_damping) * dt;
      vy += 0.5 * sds_fac * (er * y / rr - (vy - vgy) * sds_damping) * dt;
      vz += 0.5 * sds_fac * (er * z / rr - (vz - vgz) * sds_damping) * dt;
    } else {
      vx += 0.5 * ez * dt;
      vy += 0.5 * er * y / rr * dt;
      vz += 0.5 * er * z / rr * dt;
    }
    t += dt;
    step += 1;
    double sp = sqrt(vx * vx + vy * vy + vz * vz);
    eat += ez * fscale * dt;
    ket += ke_scale * sp * sp * dt;
    if (sds_on) {
      if (sds_mfp > 0.0) {
        double pre = sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo));
        double jx, jy, jz;
        or_sds_jump(&rng, p->sds_stats, sds_logmr, sds_V, sds_mfp, dt, &jx, &jy, &jz);
        x += jx;
        y += jy;
        z += jz;
        path += sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo)) - pre;
      }
    } else if (P_pa > 0.0) {
      int hit = or_hs_step(&rng, &vx, &vy, &vz, vgx, vgy, vgz, dt, 1.0, p->T_k, P_pa, sigma,
                           p->c_star, p->c_bar, m_ion, p->m_gas, p->sig1d);
      if (hit < 0) {
        kind = -3; /* partner draw refused */
        break;
      }
      if (hit) ncol += 1;
    }
    if (rec && step % p->rec_every == 0 && nrec < p->rec_rows - 1) {
      write_row(rec + nrec * p->rec_cols, p->flags, tob + t, x, y, z, vx, vy, vz, ez * fscale,
                er * fscale, ncol, m_ion, path, eat, ket);
      nrec += 1;
    }
    r = sqrt(y * y + z * z);
    if (metal_nn(f->ele, x / mm, (r - u0) / mm, nx, nu)) {
      double f0 = 0.0, f1 = 1.0;
      for (int bs = 0; bs < 40; ++bs) {
        double fm = 0.5 * (f0 + f1);
        double ym = yo + fm * (y - yo);
        double zm = zo + fm * (z - zo);
        double rm = sqrt(ym * ym + zm * zm);
        if (metal_nn(f->ele, (xo + fm * (x - xo)) / mm, (rm - u0) / mm, nx, nu))
          f1 = fm;
        else
          f0 = fm;
      }
      path -= (1.0 - f1) * sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo));
      x = xo + f1 * (x - xo);
      y = yo + f1 * (y - yo);
      z = zo + f1 * (z - zo);
      vx = vxo + f1 * (vx - vxo);
      vy = vyo + f1 * (vy - vyo);
      vz = vzo + f1 * (vz - vzo);
      t = to + f1 * dt;
      kind = 0;
      if (f->label) {
        double re = sqrt(y * y + z * z);
        int64_t li = (int64_t)floor(x / mm + 0.5), lj = (int64_t)floor((re - u0) / mm + 0.5);
        if (li >= 0 && lj >= 0 && li < nx && lj < nu) label = f->label[li * nu + lj];
      }
      break;
    }
    if (x < 0 || r / mm > (double)(nu - 2)) {
      kind = 1;
      break;
    }
    if (p->n_pl > 0) {
      int hit_pl = 0;
      for (int ip = 0; ip < p->n_pl; ++ip) {
        int64_t pcol = p->pl_col[ip];
        double cn, co;
        if (pcol == 0) {
          cn = x;
          co = xo;
        } else if (pcol == 1) {
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
        if (den == 0.0)
          fr = 0.0;
        else
          fr = (val - co) / den;
        if (fr < 0.0) fr = 0.0;
        if (fr > 1.0) fr = 1.0;
        double xc = xo + fr * (x - xo), yc = yo + fr * (y - yo), zc = zo + fr * (z - zo);
        double w1, w2;
        if (pcol == 0) {
          w1 = yc;
          w2 = zc;
        } else if (pcol == 1) {
          w1 = xc;
          w2 = zc;
        } else {
          w1 = xc;
          w2 = yc;
        }
        const double *pw = p->pl_w + 4 * ip;
        int ins = pw[0] <= w1 && w1 <= pw[1] && pw[2] <= w2 && w2 <= pw[3];
        if (p->pl_kind[ip] == 6) {
          if (!ins) continue;
        } else if (ins) {
          continue;
        }
        path -= (1.0 - fr) * sqrt((x - xo) * (x - xo) + (y - yo) * (y - yo) + (z - zo) * (z - zo));
        x = xc;
        y = yc;
        z = zc;
        vx = vxo + fr * (vx - vxo);
        vy = vyo + fr * (vy - vyo);
        vz = vzo + fr * (vz - vzo);
        t = to + fr * dt;
        kind = (int)p->pl_kind[ip];
        face = ip;
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
    if (x / mm > (double)(nx - 1)) {
      kind = 1;
      break;
    }
  }
  if (rec && kind != -2)
    write_row(rec + nrec * p->rec_cols, p->flags, tob + t, x, y, z, vx, vy, vz, ez * fscale,
              er * fscale, ncol, m_ion, path, eat, ket);
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
  out->label[ii] = label;
  out->ncol[ii] = ncol;
  out->steps[ii] = step;
  out->nrec[ii] = nrec;
}

int or_rz_fly(const or_rz_fields *f, const or_params *p, const or_ions *ions,
              or_out *out, int n_threads) {
  if (!f || !p || !ions || !out || f->nx < 2 || f->nu < 2 || p->rec_every < 1) return -1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(n_threads > 0 ? n_threads : 1)
  for (int64_t i = 0; i < ions->n; ++i) fly_one(f, p, ions, out, i);
  return 0;
}
