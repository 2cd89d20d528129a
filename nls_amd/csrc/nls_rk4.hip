// nls_rk4.hip -- the plain RK4 path for one instance: one fused kernel per
// RK4 stage over the finite-difference Laplacian (bodies in nls_stage.h), in
// float or double, for the reference's boundary families (1-D radial / 2-D
// zero-outside) and periodic grids, with the algebraic or the dynamic
// reservoir.  2-D: one 64 x 16 tile per workgroup; 1-D: 256 points.
#include "nls_rk4.h"

namespace nls {

template <typename T, int ORDER, int BC, bool DYN>
__global__ void __launch_bounds__(256) k_rk4_stage_2d(const StageArgs<T> a) {
    constexpr int R = (ORDER - 1) / 2;
    __shared__ cplx_t<T> tile[(TILE2_Y + 2 * R) * (TILE2_X + 2 * R)];
    stage_tile_2d<T, ORDER, BC, DYN>(a, blockIdx.x * TILE2_X, blockIdx.y * TILE2_Y, tile);
}

template <typename T, int ORDER, int BC, bool DYN>
__global__ void __launch_bounds__(256) k_rk4_stage_1d(const StageArgs<T> a) {
    constexpr int R = (ORDER - 1) / 2;
    __shared__ cplx_t<T> tile[TILE1 + 2 * R];
    stage_tile_1d<T, ORDER, BC, DYN>(a, blockIdx.x * TILE1, tile);
}

// Rows of the 1-D radial operator, built on the device from the closed form
// in nls_stencil.h (double precision, rounded once to T).
template <typename T>
__global__ void k_radial_table(int n, int order, double h, T *wtab) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < order; ++k) wtab[(size_t)i * order + k] = (T)radial_weight(i, k, n, order, h);
}

template <typename T, int ORDER, int BC, bool DYN>
static void launch_stage(int dim, const StageArgs<T> &a, hipStream_t s) {
    if (dim == 2) {
        dim3 grid((a.n + TILE2_X - 1) / TILE2_X, (a.n + TILE2_Y - 1) / TILE2_Y);
        hipLaunchKernelGGL((k_rk4_stage_2d<T, ORDER, BC, DYN>), grid, dim3(256), 0, s, a);
    } else {
        dim3 grid((a.n + TILE1 - 1) / TILE1);
        hipLaunchKernelGGL((k_rk4_stage_1d<T, ORDER, BC, DYN>), grid, dim3(TILE1), 0, s, a);
    }
}

template <typename T>
void launch_rk4_stage(int dim, int order, int bc, bool dyn, const StageArgs<T> &a, hipStream_t s) {
#define NLS_CASE(O, B, D) \
    if (order == O && bc == B && dyn == D) return launch_stage<T, O, B, D>(dim, a, s);
    NLS_CASE(3, 0, false) NLS_CASE(3, 0, true) NLS_CASE(3, 1, false) NLS_CASE(3, 1, true)
    NLS_CASE(5, 0, false) NLS_CASE(5, 0, true) NLS_CASE(5, 1, false) NLS_CASE(5, 1, true)
    NLS_CASE(7, 0, false) NLS_CASE(7, 0, true) NLS_CASE(7, 1, false) NLS_CASE(7, 1, true)
#undef NLS_CASE
}

template <typename T>
void launch_radial_table(int n, int order, double h, T *wtab, hipStream_t s) {
    hipLaunchKernelGGL((k_radial_table<T>), dim3((n + 255) / 256), dim3(256), 0, s, n, order, h, wtab);
}

template void launch_rk4_stage<float>(int, int, int, bool, const StageArgs<float> &, hipStream_t);
template void launch_rk4_stage<double>(int, int, int, bool, const StageArgs<double> &, hipStream_t);
template void launch_radial_table<float>(int, int, double, float *, hipStream_t);
template void launch_radial_table<double>(int, int, double, double *, hipStream_t);

}  // namespace nls
