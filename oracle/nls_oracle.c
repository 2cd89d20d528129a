/*
 * nls_oracle.c -- CPU oracle for the time-stepping core of nls.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of nls/nls.f90 (bands
 * :61-295, 2-D blocks :297-527, gbmv call :530-541, reservoir and RHS
 * :570-650 and :829-870, RK4 :705-734 and :873-901, chemical potential
 * :921-971), extended with periodic boundaries and a dynamic reservoir.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * it, and only as a checker; the product never calls it.
 *
 * Built twice from nls_oracle_impl.h: REAL = double (suffix _f64) and
 * REAL = float (suffix _f32, single precision like the reference).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REAL double
#define SFX _f64
#include "nls_oracle_impl.h"
#undef REAL
#undef SFX

#define REAL float
#define SFX _f32
#include "nls_oracle_impl.h"
#undef REAL
#undef SFX

/* Threads used by the 2-D Laplacian loops (OpenMP); 1 = serial. */
int oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void oracle_version(int *major, int *minor, int *patch) { *major = 0; *minor = 2; *patch = 0; }
