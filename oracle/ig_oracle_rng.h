/* Random streams of the CPU oracle (test infrastructure only). */
#ifndef IG_ORACLE_RNG_H
#define IG_ORACLE_RNG_H
#include <stdint.h>

#include "ig_oracle.h"

typedef struct {
  int kind;
  /* MT19937 */
  uint32_t mt[624];
  int pos;
  /* Philox */
  uint32_t key[2];
  uint64_t n;
  /* polar-method cache */
  int has_gauss;
  double gauss;
} or_rng;

void or_mt_seed(or_rng *r, uint32_t s);
void or_philox_seed(or_rng *r, uint64_t seed, uint64_t run_key);
double or_uniform(or_rng *r);
double or_normal(or_rng *r);
#endif
