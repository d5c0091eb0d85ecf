/* CPU restatement of ion_gym's 3-D Laplace basis solvers -- test
 * infrastructure only (the checker of the engine's ig_solve3d).
 *
 *   or_sweep3       solver3d._sweep            (solver3d.py:131-198)
 *   or_solve3d      solver3d.solve3d           (solver3d.py:250-293)
 *   or_residual3    multigrid3d._residual      (multigrid3d.py:305-354)
 *   or_sweep_rhs    multigrid3d._sweep_rhs     (multigrid3d.py:357-391)
 *   or_residual_rhs multigrid3d._residual_rhs  (multigrid3d.py:394-417)
 *   or_restrict_fw  multigrid3d._restrict_fw_njit (multigrid3d.py:420-455)
 *   or_prolong_add  multigrid3d._prolong_add_njit (multigrid3d.py:458-499)
 *   or_coarsen_any  multigrid3d._coarsen_fixed_any (multigrid3d.py:507-526)
 *   or_optimal_omega solver3d.optimal_omega    (solver3d.py:201-226), unit spacing
 *   or_solve_err    multigrid3d._solve_err     (multigrid3d.py:553-599)
 *   or_solve3d_vmg  multigrid3d.solve3d_vmg    (multigrid3d.py:602-731)
 *   or_coarsen_mask multigrid3d._coarsen_mask  (multigrid3d.py:84-135)
 *   or_prolongate   multigrid3d._prolongate    (multigrid3d.py:138-172)
 *   or_solve3d_mg   multigrid3d.solve3d_mg     (multigrid3d.py:175-230)
 *
 * Arrays are C order (nx, ny, nz), z fastest; the leg fractions theta are
 * (6, nx, ny, nz) float32 in the order -x, +x, -y, +y, -z, +z, or NULL for
 * all ones.  Arithmetic follows the reference expression by expression in
 * IEEE double (float32 where the reference multiplies float32 fractions),
 * no contraction, so a correct device restatement matches it bit for bit:
 * a colour's nodes read only the other colour, so any visiting order of
 * one colour gives the same field, and every reduction is a maximum.
 * Convergence (parity with the reference itself): unpinned bitwise, since
 * the reference's numba build cannot run here; pinned by the reference's
 * own gates restated in tests (exact quadratic harmonic, second order on a
 * smooth harmonic, folded against full domain, fractional legs against
 * the coaxial closed form, V-cycle and cascade against SOR at the same
 * tolerance, the size-1 axis limits; the cascade's transfers against the
 * reference's numpy forms and affine fields).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IX(i, j, k) (((int64_t)(i) * ny + (j)) * nz + (k))

/* neighbour values with even reflection at every outer face and the node
 * itself along a size-1 axis (solver3d.py:139-171) */
static inline void nbrs(const double *p, int64_t nx, int64_t ny, int64_t nz, int64_t i,
                        int64_t j, int64_t k, double v[6]) {
  const double c = p[IX(i, j, k)];
  v[0] = i > 0 ? p[IX(i - 1, j, k)] : (nx > 1 ? p[IX(1, j, k)] : c);
  v[1] = i < nx - 1 ? p[IX(i + 1, j, k)] : (nx > 1 ? p[IX(nx - 2, j, k)] : c);
  v[2] = j > 0 ? p[IX(i, j - 1, k)] : (ny > 1 ? p[IX(i, 1, k)] : c);
  v[3] = j < ny - 1 ? p[IX(i, j + 1, k)] : (ny > 1 ? p[IX(i, ny - 2, k)] : c);
  v[4] = k > 0 ? p[IX(i, j, k - 1)] : (nz > 1 ? p[IX(i, j, 1)] : c);
  v[5] = k < nz - 1 ? p[IX(i, j, k + 1)] : (nz > 1 ? p[IX(i, j, nz - 2)] : c);
}

/* the stencil's numerator and denominator (solver3d.py:173-192) */
static inline void stencil(const double v[6], const float *th, int64_t n, int64_t q, int gl,
                           double *num, double *den) {
  float t[6];
  for (int d = 0; d < 6; ++d) t[d] = th ? th[(int64_t)d * n + q] : 1.0f;
  if (gl) {
    const double wxm = 1.0 / (double)t[0], wxp = 1.0 / (double)t[1];
    const double wym = 1.0 / (double)t[2], wyp = 1.0 / (double)t[3];
    const double wzm = 1.0 / (double)t[4], wzp = 1.0 / (double)t[5];
    *num = v[0] * wxm + v[1] * wxp + v[2] * wym + v[3] * wyp + v[4] * wzm + v[5] * wzp;
    *den = wxm + wxp + wym + wyp + wzm + wzp;
  } else {
    const float sx = t[0] + t[1], sy = t[2] + t[3], sz = t[4] + t[5];
    const float ax = t[0] * sx, bx = t[1] * sx, ay = t[2] * sy, by = t[3] * sy;
    const float az = t[4] * sz, bz = t[5] * sz;
    *num = v[0] / (double)ax + v[1] / (double)bx + v[2] / (double)ay + v[3] / (double)by +
           v[4] / (double)az + v[5] / (double)bz;
    const float px = t[0] * t[1], py = t[2] * t[3], pz = t[4] * t[5];
    *den = 1.0 / (double)px + 1.0 / (double)py + 1.0 / (double)pz;
  }
}

double or_sweep3(double *phi, const uint8_t *fixed, const float *th, int64_t nx, int64_t ny,
                 int64_t nz, double omega, int colour, int gl) {
  const int64_t n = nx * ny * nz;
  double dmax = 0.0;
#pragma omp parallel for reduction(max : dmax) schedule(static)
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = (i + j + colour) & 1; k < nz; k += 2) {
        const int64_t q = IX(i, j, k);
        if (fixed[q]) continue;
        double v[6], num, den;
        nbrs(phi, nx, ny, nz, i, j, k, v);
        stencil(v, th, n, q, gl, &num, &den);
        const double old = phi[q];
        const double nw = old + omega * (num / den - old);
        const double d = fabs(nw - old);
        if (d > dmax) dmax = d;
        phi[q] = nw;
      }
  return dmax;
}

/* solve3d: phi holds the initial guess; fixed nodes are set to val first.
 * Returns 0 when converged, 1 when max_sweeps ran out. */
int or_solve3d(double *phi, const uint8_t *fixed, const double *val, const float *th,
               int64_t nx, int64_t ny, int64_t nz, double tol, int64_t max_sweeps,
               double omega, int gl, int64_t *sweeps, double *delta) {
  const int64_t n = nx * ny * nz;
  for (int64_t q = 0; q < n; ++q)
    if (fixed[q]) phi[q] = val[q];
  double d = INFINITY;
  for (int64_t s = 0; s < max_sweeps; ++s) {
    const double d1 = or_sweep3(phi, fixed, th, nx, ny, nz, omega, 0, gl);
    const double d2 = or_sweep3(phi, fixed, th, nx, ny, nz, omega, 1, gl);
    d = d1 > d2 ? d1 : d2;
    if (d < tol) {
      *sweeps = s + 1;
      *delta = d;
      return 0;
    }
  }
  *sweeps = max_sweeps;
  *delta = d;
  return 1;
}

void or_residual3(const double *phi, const uint8_t *fixed, const float *th, int64_t nx,
                  int64_t ny, int64_t nz, int gl, double *out) {
  const int64_t n = nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = 0; k < nz; ++k) {
        const int64_t q = IX(i, j, k);
        if (fixed[q]) {
          out[q] = 0.0;
          continue;
        }
        double v[6], num, den;
        nbrs(phi, nx, ny, nz, i, j, k, v);
        stencil(v, th, n, q, gl, &num, &den);
        out[q] = num - den * phi[q];
      }
}

double or_sweep_rhs(double *e, const uint8_t *fixed, const double *f, int64_t nx, int64_t ny,
                    int64_t nz, double omega, int colour) {
  double dmax = 0.0;
#pragma omp parallel for reduction(max : dmax) schedule(static)
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = (i + j + colour) & 1; k < nz; k += 2) {
        const int64_t q = IX(i, j, k);
        if (fixed[q]) continue;
        double v[6];
        nbrs(e, nx, ny, nz, i, j, k, v);
        const double nv = (v[0] + v[1] + v[2] + v[3] + v[4] + v[5] - f[q]) / 6.0;
        const double old = e[q];
        const double nw = old + omega * (nv - old);
        const double d = fabs(nw - old);
        if (d > dmax) dmax = d;
        e[q] = nw;
      }
  return dmax;
}

void or_residual_rhs(const double *e, const uint8_t *fixed, const double *f, int64_t nx,
                     int64_t ny, int64_t nz, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t k = 0; k < nz; ++k) {
        const int64_t q = IX(i, j, k);
        if (fixed[q]) {
          out[q] = 0.0;
          continue;
        }
        double v[6];
        nbrs(e, nx, ny, nz, i, j, k, v);
        out[q] = (v[0] + v[1] + v[2] + v[3] + v[4] + v[5] - 6.0 * e[q]) - f[q];
      }
}

/* full weighting onto the (n+1)/2 grid, edges clamped */
void or_restrict_fw(const double *fine, int64_t nx, int64_t ny, int64_t nz, double *out,
                    int64_t cx, int64_t cy, int64_t cz) {
#pragma omp parallel for schedule(static)
  for (int64_t ic = 0; ic < cx; ++ic)
    for (int64_t jc = 0; jc < cy; ++jc)
      for (int64_t kc = 0; kc < cz; ++kc) {
        const int64_t i = 2 * ic, j = 2 * jc, k = 2 * kc;
        double acc = 0.0;
        for (int di = -1; di <= 1; ++di) {
          int64_t ii = i + di;
          if (ii < 0) ii = 0;
          if (ii > nx - 1) ii = nx - 1;
          const double wi = di == 0 ? 1.0 : 0.5;
          for (int dj = -1; dj <= 1; ++dj) {
            int64_t jj = j + dj;
            if (jj < 0) jj = 0;
            if (jj > ny - 1) jj = ny - 1;
            const double wj = dj == 0 ? 1.0 : 0.5;
            for (int dk = -1; dk <= 1; ++dk) {
              int64_t kk = k + dk;
              if (kk < 0) kk = 0;
              if (kk > nz - 1) kk = nz - 1;
              const double wk = dk == 0 ? 1.0 : 0.5;
              acc += wi * wj * wk * fine[IX(ii, jj, kk)];
            }
          }
        }
        out[(ic * cy + jc) * cz + kc] = acc / 8.0;
      }
}

/* one axis of the trilinear prolongation: the coarse pair and fraction of
 * fine index i (multigrid3d.py:464-472) */
static inline void pro_axis(int64_t i, int64_t c, int64_t *a0, int64_t *a1, double *t) {
  double xf = i * 0.5;
  if (xf > c - 1) xf = c - 1.0;
  int64_t x0 = (int64_t)xf;
  if (x0 > c - 2) x0 = c > 1 ? c - 2 : 0;
  *a0 = x0;
  *a1 = c > 1 ? x0 + 1 : x0;
  *t = xf - x0;
}

void or_prolong_add(const double *co, int64_t cx, int64_t cy, int64_t cz, double *out,
                    const uint8_t *fixed, int64_t nx, int64_t ny, int64_t nz) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nx; ++i) {
    int64_t x0, x1;
    double tx;
    pro_axis(i, cx, &x0, &x1, &tx);
    for (int64_t j = 0; j < ny; ++j) {
      int64_t y0, y1;
      double ty;
      pro_axis(j, cy, &y0, &y1, &ty);
      for (int64_t k = 0; k < nz; ++k) {
        const int64_t q = IX(i, j, k);
        if (fixed[q]) continue;
        int64_t z0, z1;
        double tz;
        pro_axis(k, cz, &z0, &z1, &tz);
#define C(a, b, c) co[((a) * cy + (b)) * cz + (c)]
        const double c00 = C(x0, y0, z0) * (1 - tx) + C(x1, y0, z0) * tx;
        const double c10 = C(x0, y1, z0) * (1 - tx) + C(x1, y1, z0) * tx;
        const double c01 = C(x0, y0, z1) * (1 - tx) + C(x1, y0, z1) * tx;
        const double c11 = C(x0, y1, z1) * (1 - tx) + C(x1, y1, z1) * tx;
#undef C
        const double c0 = c00 * (1 - ty) + c10 * ty;
        const double c1 = c01 * (1 - ty) + c11 * ty;
        out[q] += c0 * (1 - tz) + c1 * tz;
      }
    }
  }
}

/* coarse node fixed iff any fine node of its 2 x 2 x 2 block is */
void or_coarsen_any(const uint8_t *fixed, int64_t nx, int64_t ny, int64_t nz, uint8_t *out,
                    int64_t cx, int64_t cy, int64_t cz) {
  for (int64_t ic = 0; ic < cx; ++ic)
    for (int64_t jc = 0; jc < cy; ++jc)
      for (int64_t kc = 0; kc < cz; ++kc) {
        uint8_t a = 0;
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
              const int64_t i = 2 * ic + di, j = 2 * jc + dj, k = 2 * kc + dk;
              if (i < nx && j < ny && k < nz && fixed[IX(i, j, k)]) a = 1;
            }
        out[(ic * cy + jc) * cz + kc] = a;
      }
}

/* optimal_omega (solver3d.py:201-226) for unit spacing */
double or_optimal_omega(int64_t nx, int64_t ny, int64_t nz) {
  const int64_t s[3] = {nx, ny, nz};
  double num = 0.0;
  int cnt = 0;
  for (int a = 0; a < 3; ++a)
    if (s[a] >= 2) {
      num = num + 1.0 * cos(M_PI / (double)(s[a] + 1));
      cnt += 1;
    }
  if (cnt == 0) return 1.0;
  double rho = num / (double)cnt;
  if (rho > 1.0 - 1e-15) rho = 1.0 - 1e-15;
  return 2.0 / (1.0 + sqrt(1.0 - rho * rho));
}

static double maxabs(const double *a, int64_t n) {
  double m = 0.0;
  for (int64_t q = 0; q < n; ++q) {
    const double v = fabs(a[q]);
    if (v > m) m = v;
  }
  return m;
}

/* _solve_err: solve the plain 7-point error equation L e = f to a relative
 * tolerance (multigrid3d.py:553-599); e is zeroed first */
void or_solve_err(const uint8_t *fixed, const double *f, int64_t nx, int64_t ny, int64_t nz,
                  double rel_tol, int max_cycles, int nu1, int nu2, int64_t min_coarse,
                  double *e) {
  const int64_t n = nx * ny * nz;
  memset(e, 0, (size_t)n * sizeof(double));
  double scale = maxabs(f, n);
  if (scale < 1e-300) scale = 1e-300;
  const double tol = rel_tol * scale / 6.0;
  int64_t mn = nx < ny ? nx : ny;
  if (nz < mn) mn = nz;
  if (mn <= min_coarse) {
    const double om = or_optimal_omega(nx, ny, nz);
    for (int it = 0; it < 20000; ++it) {
      const double d1 = or_sweep_rhs(e, fixed, f, nx, ny, nz, om, 0);
      const double d2 = or_sweep_rhs(e, fixed, f, nx, ny, nz, om, 1);
      if ((d1 > d2 ? d1 : d2) < 0.02 * tol) break;
    }
    return;
  }
  const int64_t cx = (nx + 1) / 2, cy = (ny + 1) / 2, cz = (nz + 1) / 2, nc = cx * cy * cz;
  uint8_t *fc = malloc((size_t)nc);
  double *r = malloc((size_t)n * sizeof(double));
  double *rc = malloc((size_t)nc * sizeof(double));
  double *ec = malloc((size_t)nc * sizeof(double));
  or_coarsen_any(fixed, nx, ny, nz, fc, cx, cy, cz);
  for (int cyc = 0; cyc < max_cycles; ++cyc) {
    for (int s = 0; s < nu1; ++s) {
      or_sweep_rhs(e, fixed, f, nx, ny, nz, 1.0, 0);
      or_sweep_rhs(e, fixed, f, nx, ny, nz, 1.0, 1);
    }
    or_residual_rhs(e, fixed, f, nx, ny, nz, r);
    or_restrict_fw(r, nx, ny, nz, rc, cx, cy, cz);
    for (int64_t q = 0; q < nc; ++q) rc[q] = -(rc[q] * 4.0);
    or_solve_err(fc, rc, cx, cy, cz, rel_tol, max_cycles, nu1, nu2, min_coarse, ec);
    or_prolong_add(ec, cx, cy, cz, e, fixed, nx, ny, nz);
    double d = 0.0;
    for (int s = 0; s < nu2; ++s) {
      const double d1 = or_sweep_rhs(e, fixed, f, nx, ny, nz, 1.0, 0);
      const double d2 = or_sweep_rhs(e, fixed, f, nx, ny, nz, 1.0, 1);
      d = d1 > d2 ? d1 : d2;
    }
    if (d < tol) break;
  }
  free(fc);
  free(r);
  free(rc);
  free(ec);
}

/* solve3d_vmg: phi holds the initial guess.  sym_axes bit a set: project
 * onto the mirror-symmetric subspace along axis a after each cycle (the
 * caller has verified the symmetry).  Returns 0 converged, 1 cycles
 * exhausted, 2 stagnated (phi then holds the stalled iterate). */
int or_solve3d_vmg(double *phi, const uint8_t *fixed, const double *val, const float *th,
                   int64_t nx, int64_t ny, int64_t nz, double tol, int max_cycles, int nu1,
                   int nu2, int gl, int sym_axes, int64_t *sweeps, double *delta) {
  const int64_t n = nx * ny * nz;
  for (int64_t q = 0; q < n; ++q)
    if (fixed[q]) phi[q] = val[q];
  const int64_t cx = (nx + 1) / 2, cy = (ny + 1) / 2, cz = (nz + 1) / 2, nc = cx * cy * cz;
  double *c0 = malloc((size_t)n * sizeof(double));
  double *r = malloc((size_t)n * sizeof(double));
  double *rc = malloc((size_t)nc * sizeof(double));
  double *ec = malloc((size_t)nc * sizeof(double));
  uint8_t *fcm = malloc((size_t)nc);
  or_coarsen_any(fixed, nx, ny, nz, fcm, cx, cy, cz);
  const int64_t dims[3] = {nx, ny, nz};
  double prev = INFINITY, cd = INFINITY;
  int64_t sw = 0;
  int rc_status = 1;
  for (int cyc = 0; cyc < max_cycles; ++cyc) {
    memcpy(c0, phi, (size_t)n * sizeof(double));
    for (int s = 0; s < nu1; ++s) {
      or_sweep3(phi, fixed, th, nx, ny, nz, 1.0, 0, gl);
      or_sweep3(phi, fixed, th, nx, ny, nz, 1.0, 1, gl);
      sw += 1;
    }
    or_residual3(phi, fixed, th, nx, ny, nz, gl, r);
    or_restrict_fw(r, nx, ny, nz, rc, cx, cy, cz);
    for (int64_t q = 0; q < nc; ++q) rc[q] = -(rc[q] * 4.0);
    or_solve_err(fcm, rc, cx, cy, cz, 0.05, 30, nu1, nu2, 9, ec);
    or_prolong_add(ec, cx, cy, cz, phi, fixed, nx, ny, nz);
    for (int s = 0; s < nu2; ++s) {
      or_sweep3(phi, fixed, th, nx, ny, nz, 1.0, 0, gl);
      or_sweep3(phi, fixed, th, nx, ny, nz, 1.0, 1, gl);
      sw += 1;
    }
    for (int a = 0; a < 3; ++a) {
      if (!((sym_axes >> a) & 1)) continue;
      for (int64_t i = 0; i < nx; ++i)
        for (int64_t j = 0; j < ny; ++j)
          for (int64_t k = 0; k < nz; ++k) {
            int64_t m[3] = {i, j, k};
            m[a] = dims[a] - 1 - m[a];
            const int64_t q = IX(i, j, k), p = IX(m[0], m[1], m[2]);
            if (p < q) continue; /* each pair once, from its lower index */
            const double s1 = (phi[q] + phi[p]) * 0.5, s2 = (phi[p] + phi[q]) * 0.5;
            phi[q] = s1;
            phi[p] = s2;
          }
    }
    cd = 0.0;
    for (int64_t q = 0; q < n; ++q) {
      const double v = fabs(phi[q] - c0[q]);
      if (v > cd) cd = v;
    }
    if (cd < tol) {
      rc_status = 0;
      break;
    }
    if (cd >= prev) {
      rc_status = 2;
      break;
    }
    prev = cd;
  }
  *sweeps = sw;
  *delta = cd;
  free(c0);
  free(r);
  free(rc);
  free(ec);
  free(fcm);
  return rc_status;
}

/* _coarsen_mask (multigrid3d.py:84-135): coarse node c takes the fine node
 * 2c by injection; a coarse node whose injection point is vacuum becomes
 * fixed when any node of the fine 2 x 2 x 2 block at 2c is, with the value
 * of the block's fixed node of largest |val| (the first such in the order
 * di, dj, dk with dk fastest, as argmax takes it).  Free coarse nodes keep
 * the injected value, which no solve reads. */
void or_coarsen_mask(const uint8_t *fixed, const double *val, int64_t nx, int64_t ny,
                     int64_t nz, uint8_t *fc, double *vc, int64_t cx, int64_t cy, int64_t cz) {
  for (int64_t ic = 0; ic < cx; ++ic)
    for (int64_t jc = 0; jc < cy; ++jc)
      for (int64_t kc = 0; kc < cz; ++kc) {
        const int64_t c = (ic * cy + jc) * cz + kc;
        const int64_t q = IX(2 * ic, 2 * jc, 2 * kc);
        fc[c] = fixed[q];
        vc[c] = val[q];
        if (fixed[q]) continue;
        double best = -1.0, v = 0.0;
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
              const int64_t i = 2 * ic + di, j = 2 * jc + dj, k = 2 * kc + dk;
              if (i >= nx || j >= ny || k >= nz || !fixed[IX(i, j, k)]) continue;
              const double m = fabs(val[IX(i, j, k)]);
              if (m > best) {
                best = m;
                v = val[IX(i, j, k)];
              }
            }
        if (best >= 0.0) {
          fc[c] = 1;
          vc[c] = v;
        }
      }
}

/* one axis of _prolongate (multigrid3d.py:144-156): fine index i sits at
 * coarse coordinate min(i / 2, c - 1) */
static inline void cas_axis(int64_t i, int64_t c, int64_t *a0, int64_t *a1, double *t) {
  double xf = i / 2.0;
  if (xf > (double)(c - 1)) xf = (double)(c - 1);
  const int64_t x0 = (int64_t)floor(xf);
  *a0 = x0;
  *a1 = x0 + 1 < c - 1 ? x0 + 1 : c - 1;
  *t = xf - (double)x0;
}

/* _prolongate (multigrid3d.py:138-172): the trilinear interpolant of the
 * coarse field at every fine node, x first, then y, then z */
void or_prolongate(const double *co, int64_t cx, int64_t cy, int64_t cz, double *out,
                   int64_t nx, int64_t ny, int64_t nz) {
  for (int64_t i = 0; i < nx; ++i) {
    int64_t x0, x1;
    double tx;
    cas_axis(i, cx, &x0, &x1, &tx);
    for (int64_t j = 0; j < ny; ++j) {
      int64_t y0, y1;
      double ty;
      cas_axis(j, cy, &y0, &y1, &ty);
      for (int64_t k = 0; k < nz; ++k) {
        int64_t z0, z1;
        double tz;
        cas_axis(k, cz, &z0, &z1, &tz);
#define C(a, b, c) co[((a) * cy + (b)) * cz + (c)]
        const double c00 = C(x0, y0, z0) * (1 - tx) + C(x1, y0, z0) * tx;
        const double c10 = C(x0, y1, z0) * (1 - tx) + C(x1, y1, z0) * tx;
        const double c01 = C(x0, y0, z1) * (1 - tx) + C(x1, y0, z1) * tx;
        const double c11 = C(x0, y1, z1) * (1 - tx) + C(x1, y1, z1) * tx;
#undef C
        const double c0 = c00 * (1 - ty) + c10 * ty;
        const double c1 = c01 * (1 - ty) + c11 * ty;
        out[IX(i, j, k)] = c0 * (1 - tz) + c1 * tz;
      }
    }
  }
}

/* solve3d_mg (multigrid3d.py:175-230), the cascade: coarsen (fixed, val)
 * while every axis exceeds min_coarse and halves to at least 3 nodes (at
 * most max_levels grids), solve the coarsest from zeros to tol, each
 * intermediate grid from the prolongated coarser solution to
 * tol * level_tol_factor (all with plain unit fractions and the grid's own
 * optimal omega), then the fine grid with its fractions, omega_fine and
 * tol.  Returns 0 converged, or 1 with *level the grid (0 = fine) whose
 * SOR ran out of sweeps; *sweeps and *delta are the fine solve's (or the
 * failing grid's). */
int or_solve3d_mg(double *phi, const uint8_t *fixed, const double *val, const float *th,
                  int64_t nx, int64_t ny, int64_t nz, double tol, int64_t max_sweeps,
                  double omega_fine, int gl, int min_coarse, int max_levels,
                  double level_tol_factor, int64_t *sweeps, double *delta, int *level) {
  enum { kMaxGrids = 64 };
  int64_t dims[kMaxGrids][3];
  uint8_t *fl[kMaxGrids];
  double *vl[kMaxGrids];
  int nl = 1;
  dims[0][0] = nx;
  dims[0][1] = ny;
  dims[0][2] = nz;
  fl[0] = (uint8_t *)fixed;
  vl[0] = (double *)val;
  while (nl < max_levels && nl < kMaxGrids) {
    const int64_t *s = dims[nl - 1];
    int64_t mn = s[0] < s[1] ? s[0] : s[1];
    mn = mn < s[2] ? mn : s[2];
    if (mn <= min_coarse || (mn + 1) / 2 < 3) break;
    for (int a = 0; a < 3; ++a) dims[nl][a] = (s[a] + 1) / 2;
    const int64_t m = dims[nl][0] * dims[nl][1] * dims[nl][2];
    fl[nl] = (uint8_t *)malloc(m);
    vl[nl] = (double *)malloc(m * sizeof(double));
    or_coarsen_mask(fl[nl - 1], vl[nl - 1], s[0], s[1], s[2], fl[nl], vl[nl], dims[nl][0],
                    dims[nl][1], dims[nl][2]);
    ++nl;
  }
  int st = 0;
  double *pc = NULL;
  for (int l = nl - 1; l >= 0 && !st; --l) {
    const int64_t *s = dims[l];
    const int64_t m = s[0] * s[1] * s[2];
    double *p = l == 0 ? phi : (double *)malloc(m * sizeof(double));
    if (pc) {
      const int64_t *c = dims[l + 1];
      or_prolongate(pc, c[0], c[1], c[2], p, s[0], s[1], s[2]);
      free(pc);
      pc = NULL;
    } else {
      memset(p, 0, m * sizeof(double));
    }
    const double lt = (l == nl - 1 || l == 0) ? tol : tol * level_tol_factor;
    const double om = l == 0 ? omega_fine : or_optimal_omega(s[0], s[1], s[2]);
    int64_t sw = 0;
    double d = INFINITY;
    st = or_solve3d(p, fl[l], vl[l], l == 0 ? th : NULL, s[0], s[1], s[2], lt, max_sweeps, om,
                    gl, &sw, &d);
    *sweeps = sw;
    *delta = d;
    *level = l;
    if (l > 0) pc = p;
  }
  if (pc) free(pc);
  for (int l = 1; l < nl; ++l) {
    free(fl[l]);
    free(vl[l]);
  }
  return st;
}
