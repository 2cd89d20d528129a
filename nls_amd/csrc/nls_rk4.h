// nls_rk4.h -- launch interface of the plain RK4 stage kernels.
#pragma once
#include "nls_core.h"

namespace nls {

template <typename T> struct StageArgs {
    int n;                       // points per axis
    int stage;                   // 1..4
    T dt, scale;                 // stage input y = u + (k_prev*dt)*scale
    Coeffs<T> c;
    T w0, w1, w2, w3;            // Cartesian second-derivative weights s2[R+d]
    const T *wtab;               // 1-D radial rows (n x order), ref 1-D only
    const cplx_t<T> *u;          // state psi
    const T *un;                 // state n (dynamic reservoir)
    const cplx_t<T> *kin;        // previous stage derivative
    const T *kin_n;
    cplx_t<T> *kout;             // this stage's derivative
    T *kout_n;
    cplx_t<T> *acc;              // RK4 accumulator; after stage 4 the new state
    T *acc_n;
    const T *P;                  // pumping
};

template <typename T>
void launch_rk4_stage(int dim, int order, int bc, bool dyn, const StageArgs<T> &a, hipStream_t s);

template <typename T>
void launch_radial_table(int n, int order, double h, T *wtab, hipStream_t s);

}  // namespace nls
