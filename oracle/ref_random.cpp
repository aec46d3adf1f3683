// ref_random.cpp -- drives the reference's OWN RandomGen (libs/lib/random.cpp,
// compiled from /root/reference by oracle/Makefile into
// oracle/_ref/libcrava_ref_random.so) as FFTGrid::fillInComplexNoise uses it.
//
// TEST INFRASTRUCTURE ONLY (the checker): the product never links or loads
// it.  The noise-fill loop restates src/fftgrid.cpp:1214-1263, because
// src/fftgrid.cpp needs the rest of the reference; every random number comes
// from the reference's own RandomGen::rnorm01 / unif01.
#include <cmath>

#include "lib/random.h"

extern "C" {

// n draws of RandomGen::rnorm01 from a generator seeded with `seed`
// (RandomGen::RandomGen(unsigned int), libs/lib/random.cpp:13-17).
void ref_rnorm01_stream(unsigned seed, long long n, double *out) {
  RandomGen g(seed);
  for (long long i = 0; i < n; ++i) out[i] = RandomGen::rnorm01();
}

// n draws of RandomGen::unif01 from `seed`.
void ref_unif01_stream(unsigned seed, long long n, double *out) {
  RandomGen g(seed);
  for (long long i = 0; i < n; ++i) out[i] = RandomGen::unif01();
}

// FFTGrid::fillInComplexNoise (src/fftgrid.cpp:1214-1263) on a half-complex
// float grid [nzp][nyp][nxp/2+1] {re, im}, generator seeded with `seed`, then
// `extra` more rnorm01 draws from the same stream into `after` (they show
// that the fill advanced the stream as the reference does).
void ref_fill_complex_noise(float *grid, int nxp, int nyp, int nzp, unsigned seed, int extra, double *after) {
  RandomGen g(seed);
  const int cnxp = nxp / 2 + 1;
  const long long csize = (long long)cnxp * nyp * nzp;
  const float std_ = float(1 / sqrt(2.0));
  for (long long i = 0; i < csize; i++) {
    float *c = grid + 2 * i;
    if (((i % cnxp) == 0) || (((i % cnxp) == cnxp - 1) && ((i % 2) == 1))) {
      const long long xshift = i % cnxp;
      const long long jkind = (i - xshift) / cnxp;
      const long long jind = jkind % nyp;
      const long long kind = jkind / nyp;
      const long long jccind = jind == 0 ? 0 : nyp - jind;
      const long long kccind = kind == 0 ? 0 : nzp - kind;
      const long long jkccind = jccind + kccind * nyp;
      if (jkccind == jkind) {
        c[0] = float(RandomGen::rnorm01());
        c[1] = 0;
      } else if (jkccind > jkind) {
        c[0] = float(std_ * RandomGen::rnorm01());
        c[1] = float(std_ * RandomGen::rnorm01());
      } else {
        const float *cc = grid + 2 * (jkccind * cnxp + xshift);
        c[0] = cc[0];
        c[1] = -cc[1];
      }
    } else {
      c[0] = float(std_ * RandomGen::rnorm01());
      c[1] = float(std_ * RandomGen::rnorm01());
    }
  }
  for (int i = 0; i < extra; ++i) after[i] = RandomGen::rnorm01();
}

}  // extern "C"
