/*
 * TEST INFRASTRUCTURE ONLY — not part of the product (see srf_oracle_impl.h).
 *
 * Builds the FP64 oracle (suffix _f64, double outputs) and the FP32 CPU port of
 * the reference projectors (suffix _f32, float outputs) from one restatement.
 * OpenMP parallelises over events; scatter ops use per-thread image replicas.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define REAL double
#define REAL_MAX DBL_MAX
#define SUFFIX _f64
#include "srf_oracle_impl.h"
#undef REAL
#undef REAL_MAX
#undef SUFFIX

#define REAL float
#define REAL_MAX FLT_MAX
#define SUFFIX _f32
#include "srf_oracle_impl.h"
#undef REAL
#undef REAL_MAX
#undef SUFFIX

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
