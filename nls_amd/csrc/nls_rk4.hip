// nls_rk4.hip -- the plain RK4 path for one instance: one fused kernel per
// RK4 stage over the finite-difference Laplacian (bodies in nls_stage.h), in
// float or double, for the reference's boundary families (1-D radial / 2-D
// zero-outside) and periodic grids, with the algebraic or the dynamic
// reservoir.  2-D: one 64 x 16 tile per workgroup; 1-D: 256 points.
#include <atomic>

#include "nls_rk4.h"

namespace nls {

// Tiles are numbered row-major over the grid; workgroups reach the 8 XCDs
// round-robin, so block b takes tile (b % 8) * chunk + b / 8 and every XCD
// sweeps a contiguous band of tile rows (its halo rows stay in its L2).
template <typename T, int ORDER, int BC, bool DYN, int TY, bool NT>
__global__ void __launch_bounds__(256) k_rk4_stage_2d(const StageArgs<T> a, int ntx, int ntiles, int chunk) {
    constexpr int R = (ORDER - 1) / 2;
    __shared__ cplx_t<T> tile[(TY + 2 * R) * (TILE2_X + 2 * R)];
    const int b = blockIdx.x;
    const int t = (b & 7) * chunk + (b >> 3);
    if (t >= ntiles) return;
    const int ty = t / ntx, tx = t - ty * ntx;
    stage_tile_2d<T, ORDER, BC, DYN, TY, NT>(a, tx * TILE2_X, ty * TY, tile);
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

// tile height of the 2-D stage kernels per dtype (a process-wide knob, set
// by nls_tune_stage_tile from any thread)
static std::atomic<int> g_stage_ty[2] = {{16}, {16}};

// Non-temporal stores of k and the accumulator pay once a stage's arrays
// (state, pumping, two k buffers and the accumulator) exceed the 256 MB
// last-level cache: 2048^2 FP64 dynamic (436 MB) 294.4 -> 262.0 us per step;
// below it the next stage finds them in that cache, and the hint costs
// (2048^2 FP32, 218 MB: 123.5 -> 136.6 us).
constexpr double NT_STAGE_BYTES = 256.0 * 1024 * 1024;

template <typename T, int ORDER, int BC, bool DYN, int TY>
static void launch_stage_2d(const StageArgs<T> &a, int ntx, int ntiles, int chunk, hipStream_t s) {
    const double pts = (double)a.n * a.n;
    const double bytes = pts * (4.0 * sizeof(cplx_t<T>) + (DYN ? 5.0 : 1.0) * sizeof(T));
    if (bytes > NT_STAGE_BYTES)
        hipLaunchKernelGGL((k_rk4_stage_2d<T, ORDER, BC, DYN, TY, true>), dim3(8 * chunk), dim3(256), 0, s, a, ntx,
                           ntiles, chunk);
    else
        hipLaunchKernelGGL((k_rk4_stage_2d<T, ORDER, BC, DYN, TY, false>), dim3(8 * chunk), dim3(256), 0, s, a, ntx,
                           ntiles, chunk);
}

template <typename T, int ORDER, int BC, bool DYN>
static void launch_stage(int dim, const StageArgs<T> &a, hipStream_t s) {
    if (dim == 2) {
        const int ty = g_stage_ty[sizeof(T) == 8 ? 1 : 0];
        const int ntx = (a.n + TILE2_X - 1) / TILE2_X;
        const int nty = (a.n + ty - 1) / ty;
        const int ntiles = ntx * nty, chunk = (ntiles + 7) / 8;
        if (ty == 4) launch_stage_2d<T, ORDER, BC, DYN, 4>(a, ntx, ntiles, chunk, s);
        else if (ty == 8) launch_stage_2d<T, ORDER, BC, DYN, 8>(a, ntx, ntiles, chunk, s);
        else if (ty == 16) launch_stage_2d<T, ORDER, BC, DYN, 16>(a, ntx, ntiles, chunk, s);
        else launch_stage_2d<T, ORDER, BC, DYN, 32>(a, ntx, ntiles, chunk, s);
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

int set_stage_tile(int dtype, int ty) {
    if (ty != 4 && ty != 8 && ty != 16 && ty != 32) return -1;
    g_stage_ty[dtype ? 1 : 0] = ty;
    return 0;
}

int get_stage_tile(int dtype) { return g_stage_ty[dtype ? 1 : 0]; }

template <typename T>
void launch_radial_table(int n, int order, double h, T *wtab, hipStream_t s) {
    hipLaunchKernelGGL((k_radial_table<T>), dim3((n + 255) / 256), dim3(256), 0, s, n, order, h, wtab);
}

template void launch_rk4_stage<float>(int, int, int, bool, const StageArgs<float> &, hipStream_t);
template void launch_rk4_stage<double>(int, int, int, bool, const StageArgs<double> &, hipStream_t);
template void launch_radial_table<float>(int, int, double, float *, hipStream_t);
template void launch_radial_table<double>(int, int, double, double *, hipStream_t);

}  // namespace nls
