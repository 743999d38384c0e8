/*
 * CPU oracle for the MISO MHD time step -- TEST INFRASTRUCTURE ONLY.
 * See mhd_oracle.h.  The body (mhd_oracle_body.inc) is instantiated for
 * Real = double and Real = float, mirroring MISO's templates.
 */
#define _GNU_SOURCE
#include "mhd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* util::sqrt/exp/fabs/copysign resolve to the std:: overload of the
 * argument type on the host (miso/include/miso/utility.hpp:31-41). */
static inline double msqrt_f64(double x) { return sqrt(x); }
static inline double mexp_f64(double x) { return exp(x); }
static inline double mfabs_f64(double x) { return fabs(x); }
static inline double mcopysign_f64(double x, double y) { return copysign(x, y); }
static inline float msqrt_f32(float x) { return sqrtf(x); }
static inline float mexp_f32(float x) { return expf(x); }
static inline float mfabs_f32(float x) { return fabsf(x); }
static inline float mcopysign_f32(float x, float y) { return copysignf(x, y); }

#define R double
#define S f64
#include "mhd_oracle_body.inc"
#undef R
#undef S

#define R float
#define S f32
#include "mhd_oracle_body.inc"
#undef R
#undef S

size_t og_work_elems(const og_grid *g) {
  return (size_t)24 * g->i_total * g->j_total * g->k_total;
}

void og_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int og_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
