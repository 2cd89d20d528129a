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

// 2-D tiles: 64 x TY points per 256-thread workgroup (TY = 16 or 32).
constexpr int TILE2_X = 64;
// 1-D tile: 256 points per 256-thread workgroup.
constexpr int TILE1 = 256;

// Stage bookkeeping at one point.  The results are formed as values first
// and written through exactly two store sites with fixed base pointers (k of
// this stage unless it is the last one; the accumulator unless it is the
// first); an if-chain with a store per stage let the compiler merge the
// accumulator stores behind a base-address phi that it left undefined on one
// path.  NT: non-temporal stores, for grids whose per-stage arrays exceed the
// last-level cache (the next stage reads them from HBM in any case; the hint
// keeps them from evicting the lines the next tiles' halos re-read).
template <typename T, bool DYN, bool NT = false>
__device__ __forceinline__ void stage_epilogue(const StageArgs<T> &a, size_t gi, cplx_t<T> u,
                                               cplx_t<T> kin, T kin_n, cplx_t<T> acc, T accn, cplx_t<T> k, T kn,
                                               T un) {
    const int st = a.stage;
    // acc := k1 + 2 k2 (stage 2), acc + 2 k3 (stage 3), u + (acc + k4) dt/6 (stage 4)
    const cplx_t<T> base = st == 2 ? kin : acc;
    const T basen = st == 2 ? kin_n : accn;
    const T wk = st == 4 ? (T)1 : (T)2;
    const T sr = base.x + wk * k.x, si = base.y + wk * k.y, sn = basen + wk * kn;
    cplx_t<T> av;
    T avn;
    if (st == 4) {
        av = mk<T>(u.x + (sr * a.dt) / (T)6, u.y + (si * a.dt) / (T)6);
        avn = un + (sn * a.dt) / (T)6;
    } else {
        av = mk<T>(sr, si);
        avn = sn;
    }
    if constexpr (NT) {
        using V2 = T __attribute__((ext_vector_type(2)));
        if (st != 4) {
            __builtin_nontemporal_store(V2{k.x, k.y}, reinterpret_cast<V2 *>(&a.kout[gi]));
            if (DYN) __builtin_nontemporal_store(kn, &a.kout_n[gi]);
        }
        if (st != 1) {
            __builtin_nontemporal_store(V2{av.x, av.y}, reinterpret_cast<V2 *>(&a.acc[gi]));
            if (DYN) __builtin_nontemporal_store(avn, &a.acc_n[gi]);
        }
    } else {
        if (st != 4) {
            a.kout[gi] = k;
            if (DYN) a.kout_n[gi] = kn;
        }
        if (st != 1) {
            a.acc[gi] = av;
            if (DYN) a.acc_n[gi] = avn;
        }
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
    cplx_t<T> acc = mk<T>(0, 0);
    T accn = 0;
    if (a.stage >= 3) {
        acc = a.acc[gi];
        if (DYN) accn = a.acc_n[gi];
    }
    cplx_t<T> k;
    T kn = 0;
    rhs_point<T, DYN>(a.c, a.P[gi], y, yn, mk<T>(lr, li), k, kn);
    stage_epilogue<T, DYN>(a, gi, u, kin, kin_n, acc, accn, k, kn, un);
}

// One 64 x TY tile of a 2-D stage with origin (x0, y0); called by all 256
// threads of the workgroup.  Thread (tx, ty) = (tid % 64, tid / 64) owns the
// TY/4 points (x0 + tx, y0 + ty + 4j).  Every global load of a thread's own
// points is issued before the LDS barrier, and the own points' stage inputs
// double as the interior of the LDS tile, so u and k_prev are read once; only
// the R-wide halo ring is loaded separately.  tile holds (TY+2R) x (64+2R).
template <typename T, int ORDER, int BC, bool DYN, int TY, bool NT = false>
__device__ __forceinline__ void stage_tile_2d(const StageArgs<T> &a, int x0, int y0, cplx_t<T> *tile) {
    using C = cplx_t<T>;
    constexpr int R = (ORDER - 1) / 2;
    constexpr int LW = TILE2_X + 2 * R, LH = TY + 2 * R;
    constexpr int J = TY / 4;
    constexpr int HALO = 2 * R * TILE2_X + 2 * R * LH;   // top/bottom rows + left/right columns
    const int n = a.n;
    const int tid = threadIdx.x;
    const int tx = tid & 63, ty = tid >> 6;
    const bool first = a.stage == 1;

    C u[J], kin[J], acc[J];
    T un[J], kn_in[J], pp[J], accn[J];
    size_t gi[J];
    bool own[J];
    int gx = x0 + tx;
    const bool xin = gx < n;
    if (BC == BC_PERIODIC) gx = wrap_index(gx, n);
#pragma unroll
    for (int j = 0; j < J; ++j) {
        int gy = y0 + ty + 4 * j;
        own[j] = xin && gy < n;
        bool load = own[j];
        if (BC == BC_PERIODIC) { gy = wrap_index(gy, n); load = true; }
        else load = load && gy < n;
        gi[j] = load ? (size_t)gy * n + gx : 0;
        u[j] = load ? a.u[gi[j]] : mk<T>(0, 0);
        kin[j] = (load && !first) ? a.kin[gi[j]] : mk<T>(0, 0);
        un[j] = (DYN && load) ? a.un[gi[j]] : (T)0;
        kn_in[j] = (DYN && load && !first) ? a.kin_n[gi[j]] : (T)0;
        pp[j] = own[j] ? a.P[gi[j]] : (T)0;
        acc[j] = (own[j] && a.stage >= 3) ? a.acc[gi[j]] : mk<T>(0, 0);
        accn[j] = (DYN && own[j] && a.stage >= 3) ? a.acc_n[gi[j]] : (T)0;
    }
    // halo ring: every halo point's loads of this thread are issued before
    // any LDS store, so the thread waits out the memory latency once
    constexpr int HITER = (HALO + 255) / 256;
    C hv[HITER], hk[HITER];
    int hl[HITER];
#pragma unroll
    for (int q0 = 0; q0 < HITER; ++q0) {
        const int h = tid + 256 * q0;
        const bool mine = h < HALO;
        const int hh = mine ? h : 0;
        int lx, ly;
        if (hh < 2 * R * TILE2_X) {                 // rows above and below the tile
            const int r = hh / TILE2_X;
            lx = R + (hh - r * TILE2_X);
            ly = r < R ? r : TY + r;                 // r in [0,R): top; [R,2R): bottom
        } else {                                     // columns left and right, full height
            const int q = hh - 2 * R * TILE2_X;
            const int c = q / LH;
            ly = q - c * LH;
            lx = c < R ? c : TILE2_X + c;            // c in [0,R): left; [R,2R): right
        }
        int hx = x0 + lx - R, hy = y0 + ly - R;
        bool ok = mine;
        if (BC == BC_PERIODIC) { hx = wrap_index(hx, n); hy = wrap_index(hy, n); }
        else ok = ok && hx >= 0 && hx < n && hy >= 0 && hy < n;
        // unconditional loads (index 0 when the point is not used) keep every
        // halo load of the thread in one group; the stage input is formed at
        // the store, as stage_input does
        const size_t hi = ok ? (size_t)hy * n + hx : (size_t)0;
        const C v = a.u[hi];
        const C k = first ? mk<T>(0, 0) : a.kin[hi];
        hv[q0] = ok ? v : mk<T>(0, 0);
        hk[q0] = ok ? k : mk<T>(0, 0);
        hl[q0] = mine ? ly * LW + lx : -1;
    }
#pragma unroll
    for (int q0 = 0; q0 < HITER; ++q0) {
        C y = hv[q0];
        if (!first) {
            y.x = y.x + (hk[q0].x * a.dt) * a.scale;
            y.y = y.y + (hk[q0].y * a.dt) * a.scale;
        }
        if (hl[q0] >= 0) tile[hl[q0]] = y;
    }
    // interior: the own points' stage inputs
#pragma unroll
    for (int j = 0; j < J; ++j) {
        C y = u[j];
        if (!first) { y.x = y.x + (kin[j].x * a.dt) * a.scale; y.y = y.y + (kin[j].y * a.dt) * a.scale; }
        tile[(ty + 4 * j + R) * LW + tx + R] = y;
    }
    __syncthreads();
    const T w[4] = {a.w0, a.w1, a.w2, a.w3};
#pragma unroll
    for (int j = 0; j < J; ++j) {
        if (!own[j]) continue;
        const int l = (ty + 4 * j + R) * LW + tx + R;
        const C c = tile[l];
        T lr = lap_centre<T>((T)2 * w[0], c.x), li = lap_centre<T>((T)2 * w[0], c.y);
#pragma unroll
        for (int d = 1; d <= R; ++d) {
            const C a0 = tile[l - d], a1 = tile[l + d];
            const C b0 = tile[l - d * LW], b1 = tile[l + d * LW];
            lr += lap_quad<T>(w[d], c.x, a0.x, a1.x, b0.x, b1.x, d);
            li += lap_quad<T>(w[d], c.y, a0.y, a1.y, b0.y, b1.y, d);
        }
        T yn = 0;
        if (DYN) yn = first ? un[j] : un[j] + (kn_in[j] * a.dt) * a.scale;
        C k;
        T kn = 0;
        rhs_point<T, DYN>(a.c, pp[j], c, yn, mk<T>(lr, li), k, kn);
        stage_epilogue<T, DYN, NT>(a, gi[j], u[j], kin[j], kn_in[j], acc[j], accn[j], k, kn, un[j]);
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
        lr = lap_centre<T>(w[0], c.x);
        li = lap_centre<T>(w[0], c.y);
#pragma unroll
        for (int d = 1; d <= R; ++d) {
            const cplx_t<T> a0 = tile[l - d], a1 = tile[l + d];
            lr += lap_pair<T>(w[d], c.x, a0.x, a1.x, d);
            li += lap_pair<T>(w[d], c.y, a0.y, a1.y, d);
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
