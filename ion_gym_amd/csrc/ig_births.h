// Births on the device: generate_births (ion_gym/physics/sim_build.py:82-277)
// for a whole source in parallel, drawing from the same numpy PCG64 stream
// in the same order as the reference's sequential loop.
//
// The reference draws ion after ion from one np.random.default_rng(seed):
// ion i's first draw follows ion i-1's last.  Most draws take one 64-bit
// word of the stream (uniform, random), but a standard normal takes one
// word 99% of the time and more when numpy's ziggurat rejects, and a
// truncated-Gaussian position repeats normals until one falls inside, so
// where ion i starts is only known after every earlier ion is drawn.  The
// device resolves that dependency without a sequential pass:
//
//   1. words   : the stream's first nW 64-bit outputs, each thread jumping
//                ahead to its chunk (Brown's arbitrary-stride LCG jump,
//                F. B. Brown, Trans. Am. Nucl. Soc. 71 (1994) 202) and
//                stepping PCG64 (XSL-RR 128/64, O'Neill 2014) from there;
//   2. next    : for EVERY word position p, where an ion that starts at p
//                would end (p + the words its draws take) -- a function of
//                p alone, since no draw count depends on the ion index;
//   3. jumps   : binary lifting, J_{k+1}[p] = J_k[J_k[p]];
//   4. ions    : ion i starts at next^i(0), composed from the bits of i,
//                and draws its values from there.
//
// Positions past the words generated are absorbing sentinels, so an
// undersized stream is detected (and the host grows it) instead of read
// out of bounds.  The standard normal is numpy's: the Marsaglia-Tsang
// ziggurat (J. Stat. Softw. 5 (2000) 8) with 256 layers as numpy lays it
// out (layer index in the low 8 bits of the word, the sign in bit 8, a
// 52-bit magnitude above; the Marsaglia tail beyond r = 3.6541528853610088
// from two log1p draws), with numpy's own layer tables passed in.
#pragma once
#pragma clang fp contract(off)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ion_gym_amd.h"

namespace ig {

// the source as the kernels read it (host pointers replaced by device ones)
struct BirthsK {
  int32_t dist, axis;
  int64_t n_per, n;
  int32_t n_mz;
  int64_t grid_g; // int(ceil(sqrt(n_per))) for the grid layout
  const double *mz;
  double p0[3], r_mm, len_mm, box[3], gsig[3], gtrunc[3];
  int32_t gtries;
  int32_t has_ke, has_dir, has_dv, has_drift;
  double ke_lo, ke_hi, dir[3], temp_k, dvs[3], vd[3], tob_span;
  const uint64_t *zk;
  const double *zw, *zf;
};

// destination of the seven birth columns: element (i, c) at
// col[c] + i * stride (row-major (N, 7): base + c, stride 7; an
// ensemble's SoA arrays: stride 1)
struct BirthsOut {
  double *col[7];
  int64_t stride;
};

// launch the four passes on `stream`; W holds nW words, J holds
// levels * (P + 2) int32 entries; flag (int64 x 2) is cleared first and
// reads back {status bits (1 = stream too short, 2 = truncation tries
// exhausted), lowest failing ion}
int births_launch(const BirthsK &b, uint64_t s_hi, uint64_t s_lo, uint64_t i_hi, uint64_t i_lo,
                  uint64_t *W, int64_t nW, int32_t *J, int64_t P, int32_t levels,
                  const BirthsOut &out, int64_t *flag, hipStream_t stream);

} // namespace ig
