// nls_rk4.hip -- the plain RK4 path: one fused kernel per RK4 stage over the
// finite-difference Laplacian, in float or double, for the reference's
// boundary families (1-D radial / 2-D zero-outside) and periodic grids, with
// the algebraic or the dynamic reservoir.
//
// Reference loop (runge_kutta, nls/nls.f90:725-733; runge_kutta_2d,
// :892-900):  k1 = F(u), k2 = F(u + k1*dt/2), k3 = F(u + k2*dt/2),
// k4 = F(u + k3*dt), u = u + (k1 + 2*k2 + 2*k3 + k4)*dt/6.
//
// Each stage kernel forms its stage input y = u + (k_prev*dt)*a on the fly
// (a = 0, 1/2, 1/2, 1), so only k_prev, u and the accumulator cross HBM:
//   stage 1: read u, P               -> write k1            (kA)
//   stage 2: read u, k1, P           -> write k2, acc=k1+2k2 (kB, acc)
//   stage 3: read u, k2, acc, P      -> write k3, acc+=2k3   (kA, acc)
//   stage 4: read u, k3, acc, P      -> write u + (acc+k4)*dt/6 into acc,
//            which becomes u of the next step (buffers swap roles).
// Per grid point and step that is 15 S + 4 r bytes (S = bytes of the state:
// one complex, plus one real for the dynamic reservoir; r = one real).
//
// 2-D tiles are 64 x 16 points per 256-thread workgroup: the stage input of
// the tile plus its halo of R = (order-1)/2 is built in LDS, then each
// thread applies the stencil to 4 points of one column.  1-D tiles are 256
// points.
#include "nls_rk4.h"

namespace nls {

template <typename T, bool DYN>
__device__ __forceinline__ void stage_epilogue(const StageArgs<T> &a, size_t gi, cplx_t<T> u,
                                               cplx_t<T> kin, T kin_n, cplx_t<T> k, T kn, T un) {
    const int st = a.stage;
    if (st == 1) {
        a.kout[gi] = k;
        if (DYN) a.kout_n[gi] = kn;
    } else if (st == 2) {
        a.kout[gi] = k;
        a.acc[gi] = mk<T>(kin.x + (T)2 * k.x, kin.y + (T)2 * k.y);
        if (DYN) { a.kout_n[gi] = kn; a.acc_n[gi] = kin_n + (T)2 * kn; }
    } else if (st == 3) {
        a.kout[gi] = k;
        const cplx_t<T> s = a.acc[gi];
        a.acc[gi] = mk<T>(s.x + (T)2 * k.x, s.y + (T)2 * k.y);
        if (DYN) { a.kout_n[gi] = kn; a.acc_n[gi] = a.acc_n[gi] + (T)2 * kn; }
    } else {
        const cplx_t<T> s = a.acc[gi];
        const T sr = s.x + k.x, si = s.y + k.y;
        a.acc[gi] = mk<T>(u.x + (sr * a.dt) / (T)6, u.y + (si * a.dt) / (T)6);
        if (DYN) { const T sn = a.acc_n[gi] + kn; a.acc_n[gi] = un + (sn * a.dt) / (T)6; }
    }
}

// Stage input at a global point: y = u + (k_prev*dt)*a.
template <typename T>
__device__ __forceinline__ cplx_t<T> stage_input(const StageArgs<T> &a, size_t gi) {
    cplx_t<T> u = a.u[gi];
    if (a.stage != 1) {
        const cplx_t<T> k = a.kin[gi];
        u.x = u.x + (k.x * a.dt) * a.scale;
        u.y = u.y + (k.y * a.dt) * a.scale;
    }
    return u;
}

template <typename T, int ORDER, int BC, bool DYN>
__global__ void __launch_bounds__(256) k_rk4_stage_2d(const StageArgs<T> a) {
    constexpr int R = (ORDER - 1) / 2;
    constexpr int TX = 64, TY = 16, LW = TX + 2 * R, LH = TY + 2 * R;
    __shared__ cplx_t<T> tile[LH * LW];
    const int n = a.n;
    const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
    const int tid = threadIdx.x;

    for (int idx = tid; idx < LW * LH; idx += 256) {
        const int ly = idx / LW, lx = idx - ly * LW;
        int gx = x0 + lx - R, gy = y0 + ly - R;
        cplx_t<T> y = mk<T>(0, 0);
        if (BC == BC_PERIODIC) {
            gx = wrap_index(gx, n);
            gy = wrap_index(gy, n);
            y = stage_input(a, (size_t)gy * n + gx);
        } else if (gx >= 0 && gx < n && gy >= 0 && gy < n) {
            y = stage_input(a, (size_t)gy * n + gx);
        }
        tile[idx] = y;
    }
    __syncthreads();

    const int tx = tid & 63, ty = tid >> 6;
    const int gx = x0 + tx;
    if (gx >= n) return;
    const T w[4] = {a.w0, a.w1, a.w2, a.w3};
#pragma unroll
    for (int j = 0; j < TY / 4; ++j) {
        const int gy = y0 + ty + 4 * j;
        if (gy >= n) break;
        const int l = (ty + 4 * j + R) * LW + tx + R;
        const cplx_t<T> c = tile[l];
        T lr = (T)2 * w[0] * c.x, li = (T)2 * w[0] * c.y;
#pragma unroll
        for (int d = 1; d <= R; ++d) {
            const cplx_t<T> a0 = tile[l - d], a1 = tile[l + d];
            const cplx_t<T> b0 = tile[l - d * LW], b1 = tile[l + d * LW];
            lr += w[d] * ((a0.x + a1.x) + (b0.x + b1.x));
            li += w[d] * ((a0.y + a1.y) + (b0.y + b1.y));
        }
        const size_t gi = (size_t)gy * n + gx;
        const cplx_t<T> u = a.u[gi];
        cplx_t<T> kin = mk<T>(0, 0);
        T kin_n = 0, un = 0, yn = 0;
        if (a.stage != 1) kin = a.kin[gi];
        if (DYN) {
            un = a.un[gi];
            yn = un;
            if (a.stage != 1) { kin_n = a.kin_n[gi]; yn = un + (kin_n * a.dt) * a.scale; }
        }
        cplx_t<T> k;
        T kn = 0;
        rhs_point<T, DYN>(a.c, a.P[gi], c, yn, mk<T>(lr, li), k, kn);
        stage_epilogue<T, DYN>(a, gi, u, kin, kin_n, k, kn, un);
    }
}

template <typename T, int ORDER, int BC, bool DYN>
__global__ void __launch_bounds__(256) k_rk4_stage_1d(const StageArgs<T> a) {
    constexpr int R = (ORDER - 1) / 2;
    constexpr int TX = 256, LW = TX + 2 * R;
    __shared__ cplx_t<T> tile[LW];
    const int n = a.n;
    const int x0 = blockIdx.x * TX;
    for (int idx = threadIdx.x; idx < LW; idx += 256) {
        int gx = x0 + idx - R;
        cplx_t<T> y = mk<T>(0, 0);
        if (BC == BC_PERIODIC) {
            gx = wrap_index(gx, n);
            y = stage_input(a, (size_t)gx);
        } else if (gx >= 0 && gx < n) {
            y = stage_input(a, (size_t)gx);
        }
        tile[idx] = y;
    }
    __syncthreads();
    const int gx = x0 + threadIdx.x;
    if (gx >= n) return;
    const int l = threadIdx.x + R;
    const cplx_t<T> c = tile[l];
    T lr = 0, li = 0;
    if (BC == BC_PERIODIC) {
        const T w[4] = {a.w0, a.w1, a.w2, a.w3};
        lr = w[0] * c.x;
        li = w[0] * c.y;
#pragma unroll
        for (int d = 1; d <= R; ++d) {
            const cplx_t<T> a0 = tile[l - d], a1 = tile[l + d];
            lr += w[d] * (a0.x + a1.x);
            li += w[d] * (a0.y + a1.y);
        }
    } else {
        const T *wr = a.wtab + (size_t)gx * ORDER;
#pragma unroll
        for (int d = -R; d <= R; ++d) {
            const T wd = wr[d + R];
            const cplx_t<T> v = tile[l + d];
            lr += wd * v.x;
            li += wd * v.y;
        }
    }
    const size_t gi = (size_t)gx;
    const cplx_t<T> u = a.u[gi];
    cplx_t<T> kin = mk<T>(0, 0);
    T kin_n = 0, un = 0, yn = 0;
    if (a.stage != 1) kin = a.kin[gi];
    if (DYN) {
        un = a.un[gi];
        yn = un;
        if (a.stage != 1) { kin_n = a.kin_n[gi]; yn = un + (kin_n * a.dt) * a.scale; }
    }
    cplx_t<T> k;
    T kn = 0;
    rhs_point<T, DYN>(a.c, a.P[gi], c, yn, mk<T>(lr, li), k, kn);
    stage_epilogue<T, DYN>(a, gi, u, kin, kin_n, k, kn, un);
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
        dim3 grid((a.n + 63) / 64, (a.n + 15) / 16);
        hipLaunchKernelGGL((k_rk4_stage_2d<T, ORDER, BC, DYN>), grid, dim3(256), 0, s, a);
    } else {
        dim3 grid((a.n + 255) / 256);
        hipLaunchKernelGGL((k_rk4_stage_1d<T, ORDER, BC, DYN>), grid, dim3(256), 0, s, a);
    }
}

template <typename T, int ORDER, int BC>
static void launch_stage_r(int dim, bool dyn, const StageArgs<T> &a, hipStream_t s) {
    if (dyn) launch_stage<T, ORDER, BC, true>(dim, a, s);
    else launch_stage<T, ORDER, BC, false>(dim, a, s);
}

template <typename T, int ORDER>
static void launch_stage_b(int dim, int bc, bool dyn, const StageArgs<T> &a, hipStream_t s) {
    if (bc == BC_PERIODIC) launch_stage_r<T, ORDER, BC_PERIODIC>(dim, dyn, a, s);
    else launch_stage_r<T, ORDER, BC_REF>(dim, dyn, a, s);
}

template <typename T>
void launch_rk4_stage(int dim, int order, int bc, bool dyn, const StageArgs<T> &a, hipStream_t s) {
    switch (order) {
        case 3: launch_stage_b<T, 3>(dim, bc, dyn, a, s); break;
        case 5: launch_stage_b<T, 5>(dim, bc, dyn, a, s); break;
        default: launch_stage_b<T, 7>(dim, bc, dyn, a, s); break;
    }
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
