// nls_stage.h -- device bodies of the per-stage RK4 kernels, shared by the
// single-instance launches (nls_rk4.hip) and the batched launches over a
// ragged list of tiles (nls_batch.hip).
//
// Reference loop (runge_kutta, nls/nls.f90:725-733; runge_kutta_2d,
// :892-900):  k1 = F(u), k2 = F(u + k1*dt/2), k3 = F(u + k2*dt/2),
// k4 = F(u + k3*dt), u = u + (k1 + 2*k2 + 2*k3 + k4)*dt/6.
//
// A stage forms its input y = u + (k_prev*dt)*a on the fly (a = 0, 1/2,
// 1/2, 1), so only k_prev, u and the accumulator cross HBM:
//   stage 1: read u, P               -> write k1             (kA)
//   stage 2: read u, k1, P           -> write k2, acc=k1+2k2 (kB, acc)
//   stage 3: read u, k2, acc, P      -> write k3, acc+=2k3   (kA, acc)
//   stage 4: read u, k3, acc, P      -> write u + (acc+k4)*dt/6 into acc,
//            which becomes u of the next step (buffers swap roles).
// Per grid point and step: 15 S + 4 r bytes (S = bytes of the state: one
// complex, plus one real for the dynamic reservoir; r = one real).
#pragma once
#include "nls_core.h"

namespace nls {

template <typename T> struct StageArgs {
    int n;                       // points per axis
    int stage;                   // 1..4
    T dt, scale;                 // stage input y = u + (k_prev*dt)*scale
    Coeffs<T> c;
    T w0, w1, w2, w3;            // Cartesian second-derivative weights s2(d)
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

// 2-D tile: 64 x 16 points per 256-thread workgroup.
constexpr int TILE2_X = 64, TILE2_Y = 16;
// 1-D tile: 256 points per 256-thread workgroup.
constexpr int TILE1 = 256;

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

// Stage input at a point: y = u + (k_prev*dt)*a.
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

// Pointwise part of a stage at one point, given y and its Laplacian.
template <typename T, bool DYN>
__device__ __forceinline__ void stage_point(const StageArgs<T> &a, size_t gi, cplx_t<T> y, T lr, T li) {
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
    rhs_point<T, DYN>(a.c, a.P[gi], y, yn, mk<T>(lr, li), k, kn);
    stage_epilogue<T, DYN>(a, gi, u, kin, kin_n, k, kn, un);
}

// One 64 x 16 tile of a 2-D stage with origin (x0, y0); called by all 256
// threads of the workgroup.  tile_lds holds (16+2R) x (64+2R) values.
template <typename T, int ORDER, int BC, bool DYN>
__device__ __forceinline__ void stage_tile_2d(const StageArgs<T> &a, int x0, int y0, cplx_t<T> *tile) {
    constexpr int R = (ORDER - 1) / 2;
    constexpr int LW = TILE2_X + 2 * R, LH = TILE2_Y + 2 * R;
    const int n = a.n;
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
    for (int j = 0; j < TILE2_Y / 4; ++j) {
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
        stage_point<T, DYN>(a, (size_t)gy * n + gx, c, lr, li);
    }
}

// One 256-point tile of a 1-D stage starting at x0.
template <typename T, int ORDER, int BC, bool DYN>
__device__ __forceinline__ void stage_tile_1d(const StageArgs<T> &a, int x0, cplx_t<T> *tile) {
    constexpr int R = (ORDER - 1) / 2;
    constexpr int LW = TILE1 + 2 * R;
    const int n = a.n;
    for (int idx = threadIdx.x; idx < LW; idx += TILE1) {
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
    stage_point<T, DYN>(a, (size_t)gx, c, lr, li);
}

}  // namespace nls
