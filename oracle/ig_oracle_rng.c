/*
 * Random streams of the CPU oracle.  TEST INFRASTRUCTURE ONLY.
 *
 * OR_RNG_MT restates the stream numba gives the reference kernels after
 * np.random.seed(seed) (ion_gym/physics/tracer_rz.py:147,
 * collision3d.py:295-300, sds.py:160-169): MT19937 seeded with
 * init_genrand, a double from two 32-bit words (27 + 26 bits), and
 * normals by the polar method that caches the second value of each pair.
 * The test tests/test_oracle_golden.py pins it against the values the
 * reference pins in tests/test_collide_bound.py:21-23.
 *
 * OR_RNG_PHILOX is the engine's stream: Philox4x32-10 (Salmon et al.,
 * "Parallel random numbers: as easy as 1, 2, 3", SC11) keyed by the ion's
 * seed XOR the run key; draw n uses block n/2, words (0,1) or (2,3) for
 * n even or odd, turned into a double the same way.  Normals use the same
 * polar method.  Pinned by the published Philox4x32-10 known answers.
 */
#include "ig_oracle_rng.h"

#include <math.h>

#define MT_N 624
#define MT_M 397

void or_mt_seed(or_rng *r, uint32_t s) {
  r->mt[0] = s;
  for (int i = 1; i < MT_N; ++i)
    r->mt[i] = 1812433253u * (r->mt[i - 1] ^ (r->mt[i - 1] >> 30)) + (uint32_t)i;
  r->pos = MT_N;
  r->has_gauss = 0;
  r->gauss = 0.0;
  r->kind = OR_RNG_MT;
}

static uint32_t mt_next(or_rng *r) {
  if (r->pos >= MT_N) {
    for (int k = 0; k < MT_N; ++k) {
      uint32_t y = (r->mt[k] & 0x80000000u) | (r->mt[(k + 1) % MT_N] & 0x7fffffffu);
      uint32_t v = r->mt[(k + MT_M) % MT_N] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      r->mt[k] = v;
    }
    r->pos = 0;
  }
  uint32_t y = r->mt[r->pos++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

static inline uint32_t mulhilo(uint32_t a, uint32_t b, uint32_t *hi) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  return (uint32_t)p;
}

void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, hi1;
    uint32_t lo0 = mulhilo(0xD2511F53u, c0, &hi0);
    uint32_t lo1 = mulhilo(0xCD9E8D57u, c2, &hi1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void or_philox_seed(or_rng *r, uint64_t seed, uint64_t run_key) {
  uint64_t k = seed ^ run_key;
  r->key[0] = (uint32_t)k;
  r->key[1] = (uint32_t)(k >> 32);
  r->n = 0;
  r->has_gauss = 0;
  r->gauss = 0.0;
  r->kind = OR_RNG_PHILOX;
}

static inline double words_to_double(uint32_t a, uint32_t b) {
  a >>= 5;
  b >>= 6;
  return ((double)a * 67108864.0 + (double)b) / 9007199254740992.0;
}

double or_uniform(or_rng *r) {
  if (r->kind == OR_RNG_MT) {
    uint32_t a = mt_next(r);
    uint32_t b = mt_next(r);
    return words_to_double(a, b);
  }
  uint64_t blk = r->n >> 1;
  int half = (int)(r->n & 1u);
  r->n += 1;
  uint32_t ctr[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), 0u, 0u};
  uint32_t o[4];
  or_philox4x32_10(ctr, r->key, o);
  return half ? words_to_double(o[2], o[3]) : words_to_double(o[0], o[1]);
}

double or_normal(or_rng *r) {
  if (r->has_gauss) {
    r->has_gauss = 0;
    return r->gauss;
  }
  double x1, x2, r2;
  do {
    x1 = 2.0 * or_uniform(r) - 1.0;
    x2 = 2.0 * or_uniform(r) - 1.0;
    r2 = x1 * x1 + x2 * x2;
  } while (r2 >= 1.0 || r2 == 0.0);
  double f = sqrt(-2.0 * log(r2) / r2);
  r->gauss = f * x1;
  r->has_gauss = 1;
  return f * x2;
}

void or_mt_draws(uint32_t seed, int kind, int64_t n, double *out) {
  or_rng r;
  or_mt_seed(&r, seed);
  for (int64_t i = 0; i < n; ++i) out[i] = kind ? or_normal(&r) : or_uniform(&r);
}
