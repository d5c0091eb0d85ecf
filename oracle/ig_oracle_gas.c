/*
 * Buffer-gas models of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * Hard sphere: ion_gym/physics/collision3d.py:265-374 (_c_bar_rel, _mfp_mm,
 * _collide, hs_collision_step).  SDS: ion_gym/physics/sds.py:134-169
 * (_icdf_steps_at, icdf_jump_steps, isotropic_vector).  Expressions keep
 * the reference's evaluation order (Python evaluates a*b*c as (a*b)*c).
 */
#include <math.h>

#include "ig_oracle_gas.h"

#define KB 1.380649e-23 /* ion_gym/physics/constants.py:19 */

double or_c_bar_rel(double speed, double c_star, double c_bar) {
  /* collision3d.py:265-272 */
  double s = speed / c_star;
  if (s < 1e-12) return c_bar;
  return c_bar * ((s + 1.0 / (2.0 * s)) * 0.5 * sqrt(M_PI) * erf(s) +
                  0.5 * exp(-s * s));
}

static double mfp_mm(double speed, double T_k, double P_pa, double sigma,
                     double c_star, double c_bar) {
  /* collision3d.py:275-278 */
  double cbr = or_c_bar_rel(speed, c_star, c_bar);
  return 1000.0 * KB * T_k * (speed / cbr) / (P_pa * sigma);
}

/* collision3d.py:281-330.  Returns 0, or -1 when no partner is accepted in
   MAX_PARTNER_TRIES draws (the reference raises ValueError there). */
int or_collide(or_rng *rng, double *vx, double *vy, double *vz, double m_ion,
               double m_gas, double sig1d, double speed_rel) {
  double scale = speed_rel + sig1d * 5.196152422706632;
  double rx = 0, ry = 0, rz = 0, rlen = 0;
  int accepted = 0;
  for (int tr = 0; tr < 10000; ++tr) {
    double gx = or_normal(rng) * sig1d;
    double gy = or_normal(rng) * sig1d;
    double gz = or_normal(rng) * sig1d;
    rx = *vx - gx;
    ry = *vy - gy;
    rz = *vz - gz;
    rlen = sqrt(rx * rx + ry * ry + rz * rz);
    if (or_uniform(rng) < rlen / scale) {
      accepted = 1;
      break;
    }
  }
  if (!accepted) return -1;
  if (rlen < 1e-12) return 0;
  double alpha = asin(sqrt(0.999999999 * or_uniform(rng)));
  double theta = 2.0 * M_PI * or_uniform(rng);
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
  *vx = *vx - f * rhx;
  *vy = *vy - f * rhy;
  *vz = *vz - f * rhz;
  return 0;
}

/* collision3d.py:333-374.  Returns 1 on a collision, 0 without, -1 when
   the partner draw was refused. */
int or_hs_step(or_rng *rng, double *vx, double *vy, double *vz, double gfx,
               double gfy, double gfz, double dt, double dt_to_us, double T_k,
               double P_pa, double sigma, double c_star, double c_bar,
               double m_ion, double m_gas, double sig1d) {
  double wx = *vx - gfx, wy = *vy - gfy, wz = *vz - gfz;
  double sp = sqrt(wx * wx + wy * wy + wz * wz);
  if (sp < 1e-7) sp = 1e-7;
  double lam = mfp_mm(sp, T_k, P_pa, sigma, c_star, c_bar);
  double u = or_uniform(rng);
  if (u < 1.0 - exp(-sp * dt * dt_to_us / lam)) {
    if (or_collide(rng, &wx, &wy, &wz, m_ion, m_gas, sig1d, sp) < 0) return -1;
    *vx = wx + gfx;
    *vy = wy + gfy;
    *vz = wz + gfz;
    return 1;
  }
  return 0;
}

double or_icdf_steps_at(const double *stats, double log_mr, double u) {
  /* sds.py:134-151; the table is (5, 1002) */
  double x = u * (1002 - 2);
  int64_t i = (int64_t)x;
  double w = x - (double)i;
  int64_t k = (int64_t)floor(log_mr);
  if (k < 0)
    k = 0;
  else if (k > 5 - 2)
    k = 5 - 2;
  double t = log_mr - (double)k;
  const double *a = stats + k * 1002, *b = stats + (k + 1) * 1002;
  double lo = a[i] + w * (a[i + 1] - a[i]);
  double hi = b[i] + w * (b[i + 1] - b[i]);
  double llo = log10(lo);
  return pow(10.0, llo + t * (log10(hi) - llo));
}

void or_sds_jump(or_rng *rng, const double *stats, double log_mr, double V,
                 double mfp, double dt, double *jx, double *jy, double *jz) {
  /* tracer_rz.py:264-267 with sds.py:154-169 */
  double dsteps = or_icdf_steps_at(stats, log_mr, or_uniform(rng));
  double nc = V / mfp * dt;
  double rj = sqrt(nc / 100000.0) * dsteps * mfp;
  double c = 2.0 * or_uniform(rng) - 1.0;
  double ph = 2.0 * M_PI * or_uniform(rng);
  double one_m = 1.0 - c * c;
  double s = rj * sqrt(one_m > 0.0 ? one_m : 0.0);
  *jx = s * cos(ph);
  *jy = s * sin(ph);
  *jz = rj * c;
}

void or_collide_run(uint32_t seed, int64_t n, double vx, double vy, double vz,
                    double m_ion, double m_gas, double sig1d, double *out) {
  or_rng r;
  or_mt_seed(&r, seed);
  for (int64_t i = 0; i < n; ++i) {
    double sp = sqrt(vx * vx + vy * vy + vz * vz);
    or_collide(&r, &vx, &vy, &vz, m_ion, m_gas, sig1d, sp);
    out[3 * i] = vx;
    out[3 * i + 1] = vy;
    out[3 * i + 2] = vz;
  }
}
