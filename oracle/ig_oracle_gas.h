/* Buffer-gas models of the CPU oracle (test infrastructure only). */
#ifndef IG_ORACLE_GAS_H
#define IG_ORACLE_GAS_H
#include "ig_oracle_rng.h"

int or_collide(or_rng *rng, double *vx, double *vy, double *vz, double m_ion,
               double m_gas, double sig1d, double speed_rel);
int or_hs_step(or_rng *rng, double *vx, double *vy, double *vz, double gfx,
               double gfy, double gfz, double dt, double dt_to_us, double T_k,
               double P_pa, double sigma, double c_star, double c_bar,
               double m_ion, double m_gas, double sig1d);
void or_sds_jump(or_rng *rng, const double *stats, double log_mr, double V,
                 double mfp, double dt, double *jx, double *jy, double *jz);
#endif
