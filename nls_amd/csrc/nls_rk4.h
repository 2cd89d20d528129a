// nls_rk4.h -- launch interface of the single-instance RK4 stage kernels.
#pragma once
#include "nls_stage.h"

namespace nls {

template <typename T>
void launch_rk4_stage(int dim, int order, int bc, bool dyn, const StageArgs<T> &a, hipStream_t s);

template <typename T>
void launch_radial_table(int n, int order, double h, T *wtab, hipStream_t s);

// Tile height (4, 8, 16 or 32 rows) of the single-instance 2-D stage kernels, per
// dtype (0 = f32, 1 = f64).
int set_stage_tile(int dtype, int ty);
int get_stage_tile(int dtype);

}  // namespace nls
