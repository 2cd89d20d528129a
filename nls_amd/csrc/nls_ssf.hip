// nls_ssf.hip -- split-step Fourier scheme on periodic grids (no counterpart
// in the reference, which only has RK4 over finite differences).
//
// One step (Strang splitting):  N(dt/2) -> L(dt) -> N(dt/2)
//   L(dt): psi_hat(k) *= exp(-i |k|^2 dt), the exact flow of psi_t = i lap psi
//          on the spectral wavenumbers of the grid;
//   N(t):  one classical RK4 step of size t of the pointwise part of the
//          reference's right-hand side (nls/nls.f90:570-581, :637-644) plus
//          the reservoir equation when c11 != 0.
//
// 2-D: the state between passes is kept transformed along x ("row space"):
//   row pass  : inverse x-FFT -> N(dt/2) [end of step] -> N(dt/2) [start of
//               the next] -> forward x-FFT           (reads/writes psi, n, P)
//   column pass: forward y-FFT -> phase exp(-i (kx^2 + ky^2) dt)/n^2 ->
//               inverse y-FFT                          (reads/writes psi)
// so a step moves 4 c + 3 r bytes per point (c complex, r real).  The first
// row pass has no inverse transform and the last one no forward transform.
// 1-D: a workgroup holds a whole line (n <= 4096) in LDS and runs all steps
// on chip.
//
// Precision: the fields sit in HBM in the problem's dtype; every kernel
// computes in double precision (FFT stages, pointwise RK4, twiddle and phase
// tables), with a single-precision state converted on load and rounded once
// on store.  On MI355X double-precision vector arithmetic issues at the
// single-precision rate (packed FP32 aside), so single precision only saves
// HBM bytes; rounding the state once per pass instead of at every FFT stage
// keeps its error far below that of single-precision transforms.
//
// FFTs: in-place Stockham stages, radix-8 while at least three factors of two
// remain, then radix-4 and a last radix-2 as needed, on the complex values a
// workgroup holds in LDS (2048 in the 2-D passes, 4096 for a 1-D line); each
// stage stages its butterfly inputs and twiddles in registers, synchronises,
// then writes the outputs in place.  Twiddles exp(-2 pi i m / n) and the
// phase table are generated in double precision.
#include <cstring>

#include "nls_core.h"
#include "nls_internal.h"

namespace nls {

// Points and threads per workgroup.  Row and column passes hold one
// 2048-point line (or several shorter lines) per workgroup: 32 KB of LDS in
// double precision, so up to five workgroups share a CU.  The 1-D kernel
// holds a whole line of up to 4096 points in one workgroup (1024 threads in
// single precision; 512 in double precision, within 256 registers).
constexpr int ROW_PTS = 2048, ROW_LOG_PTS = 11, ROW_NTH = 256;
constexpr int LINE_PTS = 4096;
template <typename T> struct LineShape { static constexpr int NTH = sizeof(T) == 8 ? 512 : 1024; };
// Column pass: 2^LOG_L adjacent columns per workgroup (fewer when n is
// smaller), so each row segment it reads and writes is 32 bytes: two FP64
// or four FP32 columns.  More columns per workgroup would only leave fewer
// workgroups for small grids.  WIDE: the 2048-point column pass (stages in
// registers), four columns per workgroup of 1024 threads, 136 KB of LDS, one
// workgroup per CU: 64-byte row segments (2048^2 per step, same box, two
// rounds: FP64 99.6 -> 96.6 us, FP32 storage 89.9 -> 88.1 us; one column
// per workgroup, 16-byte segments: 123 / 108 us).
template <typename T, bool WIDE = false> struct ColShape;
template <> struct ColShape<double, false> { static constexpr int LOG_L = 1, PTS = 4096, NTH = 512; };
template <> struct ColShape<double, true> { static constexpr int LOG_L = 2, PTS = 8192, NTH = 1024; };

// LDS layout of the transforms: element i at i + (i >> SHIFT), one complex of
// padding per 2^SHIFT, so that the strided writes of the first radix-8
// stages (stride 8, then groups of 8 at stride 64) spread over the LDS banks
// instead of piling on a few; SHIFT = 31 is the plain layout.  Each kernel
// takes its own shift (the padding costs index arithmetic where the
// conflicts it removes are few).
#ifndef NLS_SSF_ROW_PAD
#define NLS_SSF_ROW_PAD 4
#endif
#ifndef NLS_SSF_COL_PAD
#define NLS_SSF_COL_PAD 4
#endif
// The 2048-point column pass (four columns, stages in registers) is the
// faster unpadded: 2048^2 per step, same box, two rounds each, FP64 95.2 /
// 94.8 -> 93.8 / 93.7 us, FP32 storage 87.8 / 87.6 -> 87.0 / 87.1 us (one pad
// per 8 or per 32 complex: in between).
#ifndef NLS_SSF_COL2048_PAD
#define NLS_SSF_COL2048_PAD 31
#endif
#ifndef NLS_SSF_LINE_PAD
#define NLS_SSF_LINE_PAD 4
#endif
constexpr int lds_padded(int pts, int shift) { return pts + (shift < 31 ? pts >> shift : 0); }
template <typename T, int SHIFT> struct Lds {
    cplx_t<T> *p;
    __device__ __forceinline__ cplx_t<T> &operator[](int i) const { return p[SHIFT < 31 ? i + (i >> SHIFT) : i]; }
};

template <typename T>
__device__ __forceinline__ cplx_t<T> cmul(cplx_t<T> a, cplx_t<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Element m of line l lives at buf[l*lstride + m*estride].  LINE_FAST maps
// consecutive threads to consecutive lines (column pass: lines are adjacent
// in LDS) instead of consecutive elements (row pass).  A stage reads its
// butterfly inputs from LDS and its twiddles from the table (global memory,
// cached) before the barrier, so both are in flight together; after the
// barrier it computes and writes the outputs in place.
template <typename T, bool INV, bool LINE_FAST, int PTS, int NTH, typename LB>
__device__ __forceinline__ void fft_radix4_stage(LB buf, int logN, int logL, int lstride, int estride,
                                                 int logNs, const cplx_t<T> *tw) {
    using C = cplx_t<T>;
    constexpr int QB = (PTS / 4 + NTH - 1) / NTH;    // butterflies per thread
    const int lognb = logN - 2;
    const int total = 1 << (logL + lognb);           // butterflies in the workgroup
    const int Ns = 1 << logNs;
    const int sh = logN - logNs - 2;
    C v[QB][4], w[QB][3];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 1; r < 4; ++r) w[q][r - 1] = tw[(r * k) << sh];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[q][r] = buf[l * lstride + (j + (r << lognb)) * estride];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 1; r < 4; ++r) {
                C wr = w[q][r - 1];
                if (INV) wr.y = -wr.y;
                v[q][r] = cmul<T>(v[q][r], wr);
            }
            const C t0 = mk<T>(v[q][0].x + v[q][2].x, v[q][0].y + v[q][2].y);
            const C t1 = mk<T>(v[q][0].x - v[q][2].x, v[q][0].y - v[q][2].y);
            const C t2 = mk<T>(v[q][1].x + v[q][3].x, v[q][1].y + v[q][3].y);
            const C d13 = mk<T>(v[q][1].x - v[q][3].x, v[q][1].y - v[q][3].y);
            // forward: (a1 - a3) * (-i); inverse: * (+i)
            const C t3 = INV ? mk<T>(-d13.y, d13.x) : mk<T>(d13.y, -d13.x);
            const int d = ((j >> logNs) << (logNs + 2)) + k;
            const int o = l * lstride;
            buf[o + (d) * estride] = mk<T>(t0.x + t2.x, t0.y + t2.y);
            buf[o + (d + Ns) * estride] = mk<T>(t1.x + t3.x, t1.y + t3.y);
            buf[o + (d + 2 * Ns) * estride] = mk<T>(t0.x - t2.x, t0.y - t2.y);
            buf[o + (d + 3 * Ns) * estride] = mk<T>(t1.x - t3.x, t1.y - t3.y);
        }
    }
    __syncthreads();
}

// Radix-8 stage: the 8-point DFT of each butterfly in registers, as two
// 4-point DFTs (even and odd inputs) joined by the eighth roots of unity.
template <typename T, bool INV>
__device__ __forceinline__ void dft4(cplx_t<T> &a0, cplx_t<T> &a1, cplx_t<T> &a2, cplx_t<T> &a3) {
    using C = cplx_t<T>;
    const C t0 = mk<T>(a0.x + a2.x, a0.y + a2.y);
    const C t1 = mk<T>(a0.x - a2.x, a0.y - a2.y);
    const C t2 = mk<T>(a1.x + a3.x, a1.y + a3.y);
    const C d13 = mk<T>(a1.x - a3.x, a1.y - a3.y);
    const C t3 = INV ? mk<T>(-d13.y, d13.x) : mk<T>(d13.y, -d13.x);
    a0 = mk<T>(t0.x + t2.x, t0.y + t2.y);
    a1 = mk<T>(t1.x + t3.x, t1.y + t3.y);
    a2 = mk<T>(t0.x - t2.x, t0.y - t2.y);
    a3 = mk<T>(t1.x - t3.x, t1.y - t3.y);
}

// w8 = exp(-2 pi i / 8), the twiddle table's entry N/8.
template <typename T, bool INV, bool LINE_FAST, int PTS, int NTH, typename LB>
__device__ __forceinline__ void fft_radix8_stage(LB buf, int logN, int logL, int lstride, int estride,
                                                 int logNs, const cplx_t<T> *tw, cplx_t<T> w8) {
    using C = cplx_t<T>;
    constexpr int QB = (PTS / 8 + NTH - 1) / NTH;
    const int lognb = logN - 3;
    const int total = 1 << (logL + lognb);
    const int Ns = 1 << logNs;
    const int sh = logN - logNs - 3;
    C v[QB][8], w[QB][7];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 1; r < 8; ++r) w[q][r - 1] = tw[(r * k) << sh];
#pragma unroll
            for (int r = 0; r < 8; ++r) v[q][r] = buf[l * lstride + (j + (r << lognb)) * estride];
        }
    }
    __syncthreads();
    if (INV) w8.y = -w8.y;
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 1; r < 8; ++r) {
                C wr = w[q][r - 1];
                if (INV) wr.y = -wr.y;
                v[q][r] = cmul<T>(v[q][r], wr);
            }
            C e0 = v[q][0], e1 = v[q][2], e2 = v[q][4], e3 = v[q][6];
            C o0 = v[q][1], o1 = v[q][3], o2 = v[q][5], o3 = v[q][7];
            dft4<T, INV>(e0, e1, e2, e3);
            dft4<T, INV>(o0, o1, o2, o3);
            // o_r *= w8^r (inverse: conjugate), r = 1, 2, 3; w8^2 = -i, w8^3 = -i w8
            o1 = cmul<T>(o1, w8);
            o2 = INV ? mk<T>(-o2.y, o2.x) : mk<T>(o2.y, -o2.x);
            o3 = cmul<T>(o3, w8);
            o3 = INV ? mk<T>(-o3.y, o3.x) : mk<T>(o3.y, -o3.x);
            const int d = ((j >> logNs) << (logNs + 3)) + k;
            const int o = l * lstride;
            buf[o + d * estride] = mk<T>(e0.x + o0.x, e0.y + o0.y);
            buf[o + (d + Ns) * estride] = mk<T>(e1.x + o1.x, e1.y + o1.y);
            buf[o + (d + 2 * Ns) * estride] = mk<T>(e2.x + o2.x, e2.y + o2.y);
            buf[o + (d + 3 * Ns) * estride] = mk<T>(e3.x + o3.x, e3.y + o3.y);
            buf[o + (d + 4 * Ns) * estride] = mk<T>(e0.x - o0.x, e0.y - o0.y);
            buf[o + (d + 5 * Ns) * estride] = mk<T>(e1.x - o1.x, e1.y - o1.y);
            buf[o + (d + 6 * Ns) * estride] = mk<T>(e2.x - o2.x, e2.y - o2.y);
            buf[o + (d + 7 * Ns) * estride] = mk<T>(e3.x - o3.x, e3.y - o3.y);
        }
    }
    __syncthreads();
}

template <typename T, bool INV, bool LINE_FAST, int PTS, int NTH, typename LB>
__device__ __forceinline__ void fft_radix2_stage(LB buf, int logN, int logL, int lstride, int estride,
                                                 int logNs, const cplx_t<T> *tw) {
    using C = cplx_t<T>;
    constexpr int QB = (PTS / 2 + NTH - 1) / NTH;
    const int lognb = logN - 1;
    const int total = 1 << (logL + lognb);
    const int Ns = 1 << logNs;
    const int sh = logN - logNs - 1;
    C v[QB][2], w[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            w[q] = tw[(j & (Ns - 1)) << sh];
            v[q][0] = buf[l * lstride + j * estride];
            v[q][1] = buf[l * lstride + (j + (1 << lognb)) * estride];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
            C wq = w[q];
            if (INV) wq.y = -wq.y;
            const C a1 = cmul<T>(v[q][1], wq);
            const int d = ((j >> logNs) << (logNs + 1)) + k;
            const int o = l * lstride;
            buf[o + d * estride] = mk<T>(v[q][0].x + a1.x, v[q][0].y + a1.y);
            buf[o + (d + Ns) * estride] = mk<T>(v[q][0].x - a1.x, v[q][0].y - a1.y);
        }
    }
    __syncthreads();
}

// Stages whose inputs come from ld(l, e) and whose outputs go to st(l, e, v)
// (element e of line l; q, r: the thread's butterfly and its output, known
// at compile time once unrolled) instead of the LDS buffer: the first stage
// of a transform can read its inputs from HBM, the last one write to HBM or
// apply a pointwise factor on the way.  SYNC: synchronise between the reads and
// the writes (outputs written in place over inputs other threads read); the
// caller synchronises after the writes where they go to LDS.  TW: multiply by
// the twiddles (the first stage's are all 1 and are skipped).  The butterfly
// arithmetic is that of fft_radix8_stage / fft_radix4_stage.
template <typename T, bool INV, bool LINE_FAST, int PTS, int NTH, bool SYNC, bool TW, typename LD, typename ST>
__device__ __forceinline__ void fft_radix8_io(int logN, int logL, int logNs, const cplx_t<T> *tw, cplx_t<T> w8,
                                              LD ld, ST st) {
    using C = cplx_t<T>;
    constexpr int QB = (PTS / 8 + NTH - 1) / NTH;
    const int lognb = logN - 3;
    const int total = 1 << (logL + lognb);
    const int Ns = 1 << logNs;
    const int sh = logN - logNs - 3;
    C v[QB][8], w[QB][7];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
            if (TW) {
#pragma unroll
                for (int r = 1; r < 8; ++r) w[q][r - 1] = tw[(r * k) << sh];
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) v[q][r] = ld(l, j + (r << lognb));
        }
    }
    if (SYNC) __syncthreads();
    if (INV) w8.y = -w8.y;
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
            if (TW) {
#pragma unroll
                for (int r = 1; r < 8; ++r) {
                    C wr = w[q][r - 1];
                    if (INV) wr.y = -wr.y;
                    v[q][r] = cmul<T>(v[q][r], wr);
                }
            }
            C e0 = v[q][0], e1 = v[q][2], e2 = v[q][4], e3 = v[q][6];
            C o0 = v[q][1], o1 = v[q][3], o2 = v[q][5], o3 = v[q][7];
            dft4<T, INV>(e0, e1, e2, e3);
            dft4<T, INV>(o0, o1, o2, o3);
            o1 = cmul<T>(o1, w8);
            o2 = INV ? mk<T>(-o2.y, o2.x) : mk<T>(o2.y, -o2.x);
            o3 = cmul<T>(o3, w8);
            o3 = INV ? mk<T>(-o3.y, o3.x) : mk<T>(o3.y, -o3.x);
            const int d = ((j >> logNs) << (logNs + 3)) + k;
            st(l, d, mk<T>(e0.x + o0.x, e0.y + o0.y), q, 0);
            st(l, d + Ns, mk<T>(e1.x + o1.x, e1.y + o1.y), q, 1);
            st(l, d + 2 * Ns, mk<T>(e2.x + o2.x, e2.y + o2.y), q, 2);
            st(l, d + 3 * Ns, mk<T>(e3.x + o3.x, e3.y + o3.y), q, 3);
            st(l, d + 4 * Ns, mk<T>(e0.x - o0.x, e0.y - o0.y), q, 4);
            st(l, d + 5 * Ns, mk<T>(e1.x - o1.x, e1.y - o1.y), q, 5);
            st(l, d + 6 * Ns, mk<T>(e2.x - o2.x, e2.y - o2.y), q, 6);
            st(l, d + 7 * Ns, mk<T>(e3.x - o3.x, e3.y - o3.y), q, 7);
        }
    }
}

template <typename T, bool INV, bool LINE_FAST, int PTS, int NTH, bool SYNC, typename LD, typename ST>
__device__ __forceinline__ void fft_radix4_io(int logN, int logL, int logNs, const cplx_t<T> *tw, LD ld, ST st) {
    using C = cplx_t<T>;
    constexpr int QB = (PTS / 4 + NTH - 1) / NTH;
    const int lognb = logN - 2;
    const int total = 1 << (logL + lognb);
    const int Ns = 1 << logNs;
    const int sh = logN - logNs - 2;
    C v[QB][4], w[QB][3];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 1; r < 4; ++r) w[q][r - 1] = tw[(r * k) << sh];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[q][r] = ld(l, j + (r << lognb));
        }
    }
    if (SYNC) __syncthreads();
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int b = threadIdx.x + NTH * q;
        if (b < total) {
            const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> lognb);
            const int j = LINE_FAST ? (b >> logL) : (b & ((1 << lognb) - 1));
            const int k = j & (Ns - 1);
#pragma unroll
            for (int r = 1; r < 4; ++r) {
                C wr = w[q][r - 1];
                if (INV) wr.y = -wr.y;
                v[q][r] = cmul<T>(v[q][r], wr);
            }
            const C t0 = mk<T>(v[q][0].x + v[q][2].x, v[q][0].y + v[q][2].y);
            const C t1 = mk<T>(v[q][0].x - v[q][2].x, v[q][0].y - v[q][2].y);
            const C t2 = mk<T>(v[q][1].x + v[q][3].x, v[q][1].y + v[q][3].y);
            const C d13 = mk<T>(v[q][1].x - v[q][3].x, v[q][1].y - v[q][3].y);
            const C t3 = INV ? mk<T>(-d13.y, d13.x) : mk<T>(d13.y, -d13.x);
            const int d = ((j >> logNs) << (logNs + 2)) + k;
            st(l, d, mk<T>(t0.x + t2.x, t0.y + t2.y), q, 0);
            st(l, d + Ns, mk<T>(t1.x + t3.x, t1.y + t3.y), q, 1);
            st(l, d + 2 * Ns, mk<T>(t0.x - t2.x, t0.y - t2.y), q, 2);
            st(l, d + 3 * Ns, mk<T>(t1.x - t3.x, t1.y - t3.y), q, 3);
        }
    }
}

// The last stage of one 2048-point transform (radix 4, A) and the first of
// the next (radix 8, B) without LDS in between.  A thread's two radix-4
// butterflies of A produce the elements j0 + 256 s (s = 0 .. 7) of its line,
// and those are exactly the inputs of its radix-8 butterfly of B; so the
// outputs of A go through the pointwise map g(l, e, s, v) in registers (the
// point's fields are the thread's slot s) straight into B, whose outputs go to
// st(l, e, v, q, r).  One barrier between A's LDS reads and B's LDS writes
// (B writes elements other threads read in A).  Arithmetic as in
// fft_radix4_stage and fft_radix8_stage (B's twiddles are all 1, skipped).
template <typename T, bool INV_A, bool INV_B, bool LINE_FAST, int PTS, int NTH, typename LD, typename G, typename ST>
__device__ __forceinline__ void fft_2048_turn(int logL, const cplx_t<T> *tw, cplx_t<T> w8, LD ld, G g, ST st) {
    using C = cplx_t<T>;
    constexpr int LN = 11;
    static_assert((PTS / 4 + NTH - 1) / NTH == 2 && (PTS / 8 + NTH - 1) / NTH == 1 && PTS == NTH * 8,
                  "a thread holds two radix-4 and one radix-8 butterflies of its line");
    C v[2][4], w[2][3];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int b = threadIdx.x + NTH * q;
        const int l = LINE_FAST ? (b & ((1 << logL) - 1)) : (b >> (LN - 2));
        const int j = LINE_FAST ? (b >> logL) : (b & ((1 << (LN - 2)) - 1));
#pragma unroll
        for (int r = 1; r < 4; ++r) w[q][r - 1] = tw[r * j];       // k = j: the last stage, Ns = 512
#pragma unroll
        for (int r = 0; r < 4; ++r) v[q][r] = ld(l, j + (r << (LN - 2)));
    }
    __syncthreads();
    const int bb = threadIdx.x;
    const int l = LINE_FAST ? (bb & ((1 << logL) - 1)) : (bb >> (LN - 3));
    const int j0 = LINE_FAST ? (bb >> logL) : (bb & ((1 << (LN - 3)) - 1));
    C x[8];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
#pragma unroll
        for (int r = 1; r < 4; ++r) {
            C wr = w[q][r - 1];
            if (INV_A) wr.y = -wr.y;
            v[q][r] = cmul<T>(v[q][r], wr);
        }
        const C t0 = mk<T>(v[q][0].x + v[q][2].x, v[q][0].y + v[q][2].y);
        const C t1 = mk<T>(v[q][0].x - v[q][2].x, v[q][0].y - v[q][2].y);
        const C t2 = mk<T>(v[q][1].x + v[q][3].x, v[q][1].y + v[q][3].y);
        const C d13 = mk<T>(v[q][1].x - v[q][3].x, v[q][1].y - v[q][3].y);
        const C t3 = INV_A ? mk<T>(-d13.y, d13.x) : mk<T>(d13.y, -d13.x);
        // element j0 + 256 q + 512 r of the line: slot s = q + 2 r
        x[q] = mk<T>(t0.x + t2.x, t0.y + t2.y);
        x[q + 2] = mk<T>(t1.x + t3.x, t1.y + t3.y);
        x[q + 4] = mk<T>(t0.x - t2.x, t0.y - t2.y);
        x[q + 6] = mk<T>(t1.x - t3.x, t1.y - t3.y);
    }
    // the pointwise map one point at a time, after both butterflies (with
    // the maps' temporaries interleaved, a SIMD would hold half the waves)
#pragma unroll
    for (int sl = 0; sl < 8; ++sl) {
        x[sl] = g(l, j0 + 256 * sl, sl, x[sl]);
        __builtin_amdgcn_sched_barrier(0);
    }
    if (INV_B) w8.y = -w8.y;
    C e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    C o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<T, INV_B>(e0, e1, e2, e3);
    dft4<T, INV_B>(o0, o1, o2, o3);
    o1 = cmul<T>(o1, w8);
    o2 = INV_B ? mk<T>(-o2.y, o2.x) : mk<T>(o2.y, -o2.x);
    o3 = cmul<T>(o3, w8);
    o3 = INV_B ? mk<T>(-o3.y, o3.x) : mk<T>(o3.y, -o3.x);
    const int d = j0 << 3;                                          // Ns = 1
    st(l, d, mk<T>(e0.x + o0.x, e0.y + o0.y), 0, 0);
    st(l, d + 1, mk<T>(e1.x + o1.x, e1.y + o1.y), 0, 1);
    st(l, d + 2, mk<T>(e2.x + o2.x, e2.y + o2.y), 0, 2);
    st(l, d + 3, mk<T>(e3.x + o3.x, e3.y + o3.y), 0, 3);
    st(l, d + 4, mk<T>(e0.x - o0.x, e0.y - o0.y), 0, 4);
    st(l, d + 5, mk<T>(e1.x - o1.x, e1.y - o1.y), 0, 5);
    st(l, d + 6, mk<T>(e2.x - o2.x, e2.y - o2.y), 0, 6);
    st(l, d + 7, mk<T>(e3.x - o3.x, e3.y - o3.y), 0, 7);
}

// In-place FFT of 2^logL lines of length 2^logN (2^(logL+logN) <= PTS) in LDS.
template <typename T, bool INV, bool LINE_FAST, int PTS, int NTH, typename LB>
__device__ __forceinline__ void lds_fft(LB buf, int logN, int logL, int lstride, int estride,
                                        const cplx_t<T> *tw) {
    int logNs = 0;
    if (sizeof(T) == 8 && logN >= 3) {
        // radix-8 stages (2048 = 8 8 8 4: four passes through LDS instead of six)
        const cplx_t<T> w8 = tw[(1 << logN) >> 3];
        do {
            fft_radix8_stage<T, INV, LINE_FAST, PTS, NTH>(buf, logN, logL, lstride, estride, logNs, tw, w8);
            logNs += 3;
        } while (logN - logNs >= 3);
    }
    while (logN - logNs >= 2) {
        fft_radix4_stage<T, INV, LINE_FAST, PTS, NTH>(buf, logN, logL, lstride, estride, logNs, tw);
        logNs += 2;
    }
    if (logN - logNs == 1) fft_radix2_stage<T, INV, LINE_FAST, PTS, NTH>(buf, logN, logL, lstride, estride, logNs, tw);
}

// One RK4 step of size tau of the pointwise system at one point.
template <typename T, bool DYN>
__device__ __forceinline__ void pointwise_rk4(const Coeffs<T> &c, T inv_c11, T P, T tau, cplx_t<T> &psi, T &n) {
    auto f = [&](cplx_t<T> y, T yn, cplx_t<T> &k, T &kn) {
        const T s = y.x * y.x + y.y * y.y;
        const T r = DYN ? yn : c.c12 * P / (c.c13 + c.c14 * s);
        const T g = c.c3 * r - c.c4, q = c.c5 * s + c.c6 * r;
        k = mk<T>(g * y.x + q * y.y, g * y.y - q * y.x);
        kn = DYN ? (c.c12 * P - (c.c13 + c.c14 * s) * yn) * inv_c11 : (T)0;
    };
    cplx_t<T> k1, k2, k3, k4;
    T m1, m2, m3, m4;
    const T h = tau * (T)0.5;
    f(psi, n, k1, m1);
    f(mk<T>(psi.x + k1.x * h, psi.y + k1.y * h), n + m1 * h, k2, m2);
    f(mk<T>(psi.x + k2.x * h, psi.y + k2.y * h), n + m2 * h, k3, m3);
    f(mk<T>(psi.x + k3.x * tau, psi.y + k3.y * tau), n + m3 * tau, k4, m4);
    const T t6 = tau / (T)6;
    psi = mk<T>(psi.x + (k1.x + (T)2 * k2.x + (T)2 * k3.x + k4.x) * t6,
                psi.y + (k1.y + (T)2 * k2.y + (T)2 * k3.y + k4.y) * t6);
    if (DYN) n = n + (m1 + (T)2 * m2 + (T)2 * m3 + m4) * t6;
}

// S: the precision of the fields in HBM (the problem's dtype); T: the
// arithmetic, LDS and tables, double in both cases -- on this GPU double
// precision vector arithmetic runs at the single-precision rate (no packed
// FP32), so a single-precision state only halves the HBM bytes and is
// rounded once per pass, not at every FFT stage.
template <typename S, typename T> struct SsfArgs {
    int logN;
    T half_dt, inv_c11;
    Coeffs<T> c;
    cplx_t<S> *psi;
    S *nres;
    const S *P;
    const cplx_t<T> *tw;      // exp(-2 pi i m / N)
    const cplx_t<T> *ph;      // exp(-i k_m^2 dt) / N
    int iters;                // 1-D kernel only
};

// A stored complex value in the arithmetic precision, and back.
template <typename T, typename S> __device__ __forceinline__ cplx_t<T> widen(cplx_t<S> v) {
    return mk<T>((T)v.x, (T)v.y);
}
template <typename S, typename T> __device__ __forceinline__ cplx_t<S> narrow(cplx_t<T> v) {
    return mk<S>((S)v.x, (S)v.y);
}

enum { ROWS_FIRST = 0, ROWS_MIDDLE = 1, ROWS_LAST = 2 };

// Lines of length 2^logN per workgroup of 2^logP points (at most n lines).
__host__ __device__ constexpr int lines_per_wg(int logN, int logP) {
    return logP - logN < logN ? logP - logN : logN;
}

// Row pass: 2^(11-logN) rows per workgroup (at most n).
// ROW2048: the middle row pass of a grid of 2048-point rows, a kernel of its
// own, held to 128 registers (four waves per SIMD; in one kernel with the
// general path, or left to the compiler with double-precision storage, it
// took 131-136 and three).
template <typename S, typename T, bool DYN, int MODE, bool ROW2048 = false>
__global__ void __launch_bounds__(ROW_NTH, ROW2048 ? 4 : 1) k_ssf_rows(const SsfArgs<S, T> a) {
    using C = cplx_t<T>;
    constexpr int PPT = ROW_PTS / ROW_NTH;
    __shared__ C buf_mem[lds_padded(ROW_PTS, NLS_SSF_ROW_PAD)];
    const Lds<T, NLS_SSF_ROW_PAD> buf{buf_mem};
    const int logN = a.logN, N = 1 << logN;
    const int logL = lines_per_wg(logN, ROW_LOG_PTS);
    const int pts = 1 << (logL + logN);
    const size_t base = (size_t)blockIdx.x << (logL + logN);    // rows are contiguous
    T pp[PPT], nn[PPT];
    if constexpr (ROW2048) {
        static_assert(MODE == ROWS_MIDDLE && sizeof(T) == 8, "the 2048-point middle pass");
        // one 2048-point row (8 x 8 x 8 x 4, compile-time sizes): the first
        // inverse stage reads HBM; the last inverse stage, the two pointwise
        // half steps and the first forward stage run in registers
        // (fft_2048_turn: a thread's points are the elements tid + 256 s, its
        // slots s of pp and nn); the last forward stage writes HBM -- six
        // passes through LDS fewer and about half the barriers
        constexpr int LN = ROW_LOG_PTS, NN = 1 << LN;
        static_assert(PPT == 8, "a thread holds eight points of the row");
        // the fields stay in the storage precision until used (exact
        // conversion; single-precision storage then holds them in half the
        // registers)
        S ps[PPT], ns[PPT];
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int idx = threadIdx.x + ROW_NTH * q;
            ps[q] = a.P[base + idx];
            ns[q] = DYN ? a.nres[base + idx] : (S)0;
        }
        auto lds_ld = [&](int, int e) { return buf[e]; };
        auto lds_st = [&](int, int e, C v, int, int) { buf[e] = v; };
        const C w8 = a.tw[NN >> 3];
        fft_radix8_io<T, true, false, ROW_PTS, ROW_NTH, false, false>(
            LN, 0, 0, a.tw, w8, [&](int, int e) { return widen<T, S>(a.psi[base + e]); }, lds_st);
        __syncthreads();
        fft_radix8_stage<T, true, false, ROW_PTS, ROW_NTH>(buf, LN, 0, NN, 1, 3, a.tw, w8);
        fft_radix8_stage<T, true, false, ROW_PTS, ROW_NTH>(buf, LN, 0, NN, 1, 6, a.tw, w8);
        fft_2048_turn<T, true, false, false, ROW_PTS, ROW_NTH>(
            0, a.tw, w8, lds_ld,
            [&](int, int e, int slot, C v) {
                T nv = (T)ns[slot];
                const T pv = (T)ps[slot];
                pointwise_rk4<T, DYN>(a.c, a.inv_c11, pv, a.half_dt, v, nv);
                pointwise_rk4<T, DYN>(a.c, a.inv_c11, pv, a.half_dt, v, nv);
                // the reservoir is final here (the transforms act on psi only)
                if (DYN) a.nres[base + e] = (S)nv;
                return v;
            },
            lds_st);
        __syncthreads();
        fft_radix8_stage<T, false, false, ROW_PTS, ROW_NTH>(buf, LN, 0, NN, 1, 3, a.tw, w8);
        fft_radix8_stage<T, false, false, ROW_PTS, ROW_NTH>(buf, LN, 0, NN, 1, 6, a.tw, w8);
        fft_radix4_io<T, false, false, ROW_PTS, ROW_NTH, false>(
            LN, 0, 9, a.tw, lds_ld, [&](int, int e, C v, int, int) { a.psi[base + e] = narrow<S, T>(v); });
        return;
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + ROW_NTH * q;
        pp[q] = 0; nn[q] = 0;
        if (idx < pts) {
            buf[idx] = widen<T, S>(a.psi[base + idx]);
            pp[q] = (T)a.P[base + idx];
            if (DYN) nn[q] = (T)a.nres[base + idx];
        }
    }
    __syncthreads();
    // a grid of 2048-point rows (one a workgroup) takes a copy of the
    // transform with its sizes known at compile time: every stage's index
    // arithmetic folds to constants (6% of the step at 2048^2)
    if (MODE != ROWS_FIRST) {
        if (logN == ROW_LOG_PTS) lds_fft<T, true, false, ROW_PTS, ROW_NTH>(buf, ROW_LOG_PTS, 0, ROW_PTS, 1, a.tw);
        else lds_fft<T, true, false, ROW_PTS, ROW_NTH>(buf, logN, logL, N, 1, a.tw);
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + ROW_NTH * q;
        if (idx < pts) {
            C v = buf[idx];
            T nv = nn[q];
            pointwise_rk4<T, DYN>(a.c, a.inv_c11, pp[q], a.half_dt, v, nv);
            if (MODE == ROWS_MIDDLE) pointwise_rk4<T, DYN>(a.c, a.inv_c11, pp[q], a.half_dt, v, nv);
            buf[idx] = v;
            nn[q] = nv;
        }
    }
    __syncthreads();
    if (MODE != ROWS_LAST) {
        if (logN == ROW_LOG_PTS) lds_fft<T, false, false, ROW_PTS, ROW_NTH>(buf, ROW_LOG_PTS, 0, ROW_PTS, 1, a.tw);
        else lds_fft<T, false, false, ROW_PTS, ROW_NTH>(buf, logN, logL, N, 1, a.tw);
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + ROW_NTH * q;
        if (idx < pts) {
            a.psi[base + idx] = narrow<S, T>(buf[idx]);
            if (DYN) a.nres[base + idx] = (S)nn[q];
        }
    }
}

// Column pass: min(2^LOG_L, n) adjacent columns per workgroup, LDS
// [row][col].  Workgroups of one XCD take adjacent column blocks
// (round-robin dispatch).
template <typename S, typename T, bool COL2048 = false>
__global__ void __launch_bounds__((ColShape<T, COL2048>::NTH)) k_ssf_cols(const SsfArgs<S, T> a, int nblocks,
                                                                           int chunk) {
    using C = cplx_t<T>;
    using CS = ColShape<T, COL2048>;
    constexpr int COL_PTS = CS::PTS, COL_NTH = CS::NTH;
    constexpr int PPT = COL_PTS / COL_NTH;
    static_assert(COL_PTS == 2048 << CS::LOG_L, "column block of 2048-point columns");
    constexpr int PAD = COL2048 ? NLS_SSF_COL2048_PAD : NLS_SSF_COL_PAD;
    __shared__ C buf_mem[lds_padded(COL_PTS, PAD)];
    const Lds<T, PAD> buf{buf_mem};
    const int b = blockIdx.x;
    const int blk = (b & 7) * chunk + (b >> 3);
    if (blk >= nblocks) return;
    const int logN = a.logN, N = 1 << logN;
    const int logL = CS::LOG_L < logN ? CS::LOG_L : logN;
    const int L = 1 << logL, pts = L << logN;
    const int col0 = blk << logL;
    constexpr int LL = CS::LOG_L;
    if constexpr (COL2048) {
        static_assert(sizeof(T) == 8, "the 2048-point column pass");
        // 2048-point columns (8 x 8 x 8 x 4, compile-time sizes): the first
        // forward stage reads HBM, the last forward stage, the phase and the
        // first inverse stage run in registers (fft_2048_turn), and the last
        // inverse stage writes HBM -- six passes through LDS (reads and
        // writes) fewer and half the barriers of load / transform / phase /
        // transform / store
        constexpr int LN = ROW_LOG_PTS, NN = 1 << LN, LW = 1 << LL;
        auto lds_ld = [&](int l, int e) { return buf[l + e * LW]; };
        auto lds_st = [&](int l, int e, C v, int, int) { buf[l + e * LW] = v; };
        const C w8 = a.tw[NN >> 3];
        fft_radix8_io<T, false, true, COL_PTS, COL_NTH, false, false>(
            LN, LL, 0, a.tw, w8, [&](int l, int e) { return widen<T, S>(a.psi[(size_t)e * NN + col0 + l]); }, lds_st);
        __syncthreads();
        fft_radix8_stage<T, false, true, COL_PTS, COL_NTH>(buf, LN, LL, 1, LW, 3, a.tw, w8);
        fft_radix8_stage<T, false, true, COL_PTS, COL_NTH>(buf, LN, LL, 1, LW, 6, a.tw, w8);
        fft_2048_turn<T, false, true, true, COL_PTS, COL_NTH>(
            LL, a.tw, w8, lds_ld, [&](int l, int e, int, C v) { return cmul<T>(v, cmul<T>(a.ph[e], a.ph[col0 + l])); },
            lds_st);
        __syncthreads();
        fft_radix8_stage<T, true, true, COL_PTS, COL_NTH>(buf, LN, LL, 1, LW, 3, a.tw, w8);
        fft_radix8_stage<T, true, true, COL_PTS, COL_NTH>(buf, LN, LL, 1, LW, 6, a.tw, w8);
        fft_radix4_io<T, true, true, COL_PTS, COL_NTH, false>(
            LN, LL, 9, a.tw, lds_ld,
            [&](int l, int e, C v, int, int) { a.psi[(size_t)e * NN + col0 + l] = narrow<S, T>(v); });
        return;
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + COL_NTH * q;
        if (idx < pts) {
            const int m = idx >> logL, cc = idx & (L - 1);
            buf[idx] = widen<T, S>(a.psi[(size_t)m * N + col0 + cc]);
        }
    }
    __syncthreads();
    // 2048-point columns: the transform with compile-time sizes, as in the row pass
    if (logN == ROW_LOG_PTS) lds_fft<T, false, true, COL_PTS, COL_NTH>(buf, ROW_LOG_PTS, LL, 1, 1 << LL, a.tw);
    else lds_fft<T, false, true, COL_PTS, COL_NTH>(buf, logN, logL, 1, L, a.tw);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + COL_NTH * q;
        if (idx < pts) {
            const int m = idx >> logL, cc = idx & (L - 1);
            buf[idx] = cmul<T>(buf[idx], cmul<T>(a.ph[m], a.ph[col0 + cc]));
        }
    }
    __syncthreads();
    if (logN == ROW_LOG_PTS) lds_fft<T, true, true, COL_PTS, COL_NTH>(buf, ROW_LOG_PTS, LL, 1, 1 << LL, a.tw);
    else lds_fft<T, true, true, COL_PTS, COL_NTH>(buf, logN, logL, 1, L, a.tw);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + COL_NTH * q;
        if (idx < pts) {
            const int m = idx >> logL, cc = idx & (L - 1);
            a.psi[(size_t)m * N + col0 + cc] = narrow<S, T>(buf[idx]);
        }
    }
}

// 1-D: one workgroup holds the whole line in LDS and runs every step; the
// pumping and reservoir of the thread's points stay in registers.
template <typename S, typename T, bool DYN>
__global__ void __launch_bounds__(LineShape<T>::NTH) k_ssf_line(const SsfArgs<S, T> a) {
    using C = cplx_t<T>;
    constexpr int LINE_NTH = LineShape<T>::NTH;
    constexpr int PPT = LINE_PTS / LINE_NTH;
    __shared__ C buf_mem[lds_padded(LINE_PTS, NLS_SSF_LINE_PAD)];
    const Lds<T, NLS_SSF_LINE_PAD> buf{buf_mem};
    const int logN = a.logN, N = 1 << logN;
    T pp[PPT], nn[PPT];
    C ph[PPT];                                       // the phase of each point, for every step
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + LINE_NTH * q;
        pp[q] = 0; nn[q] = 0; ph[q] = mk<T>(0, 0);
        if (idx < N) {
            ph[q] = a.ph[idx];
            pp[q] = (T)a.P[idx];
            if (DYN) nn[q] = (T)a.nres[idx];
        }
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + LINE_NTH * q;
        if (idx < N) buf[idx] = widen<T, S>(a.psi[idx]);
    }
    for (int it = 0; it < a.iters; ++it) {
        // each thread only touches its own points here, so no barrier is
        // needed before the first FFT stage (which starts with one)
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int idx = threadIdx.x + LINE_NTH * q;
            if (idx < N) {
                C v = buf[idx];
                pointwise_rk4<T, DYN>(a.c, a.inv_c11, pp[q], a.half_dt, v, nn[q]);
                buf[idx] = v;
            }
        }
        __syncthreads();
        lds_fft<T, false, false, LINE_PTS, LINE_NTH>(buf, logN, 0, N, 1, a.tw);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int idx = threadIdx.x + LINE_NTH * q;
            if (idx < N) buf[idx] = cmul<T>(buf[idx], ph[q]);
        }
        __syncthreads();
        lds_fft<T, true, false, LINE_PTS, LINE_NTH>(buf, logN, 0, N, 1, a.tw);
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int idx = threadIdx.x + LINE_NTH * q;
            if (idx < N) {
                C v = buf[idx];
                pointwise_rk4<T, DYN>(a.c, a.inv_c11, pp[q], a.half_dt, v, nn[q]);
                buf[idx] = v;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + LINE_NTH * q;
        if (idx < N) {
            a.psi[idx] = narrow<S, T>(buf[idx]);
            if (DYN) a.nres[idx] = (S)nn[q];
        }
    }
}

// Radix-8 butterfly of Stockham stage logNs of a 2^LN-point transform on
// the inputs x[0..7] (elements j + r 2^(LN-3) of the line), in place: x[r]
// becomes the output element d + r Ns, d = ((j >> logNs) << (logNs + 3)) +
// (j & (Ns - 1)).  The arithmetic of fft_radix8_stage; TW = false skips the
// twiddles of the first stage, which are all 1.
template <typename T, bool INV, bool TW>
__device__ __forceinline__ void bfly8(cplx_t<T> x[8], int j, int logNs, int LN, const cplx_t<T> *tw, cplx_t<T> w8) {
    using C = cplx_t<T>;
    if (TW) {
        const int k = j & ((1 << logNs) - 1), sh = LN - logNs - 3;
#pragma unroll
        for (int r = 1; r < 8; ++r) {
            C wr = tw[(r * k) << sh];
            if (INV) wr.y = -wr.y;
            x[r] = cmul<T>(x[r], wr);
        }
    }
    if (INV) w8.y = -w8.y;
    C e0 = x[0], e1 = x[2], e2 = x[4], e3 = x[6];
    C o0 = x[1], o1 = x[3], o2 = x[5], o3 = x[7];
    dft4<T, INV>(e0, e1, e2, e3);
    dft4<T, INV>(o0, o1, o2, o3);
    o1 = cmul<T>(o1, w8);
    o2 = INV ? mk<T>(-o2.y, o2.x) : mk<T>(o2.y, -o2.x);
    o3 = cmul<T>(o3, w8);
    o3 = INV ? mk<T>(-o3.y, o3.x) : mk<T>(o3.y, -o3.x);
    x[0] = mk<T>(e0.x + o0.x, e0.y + o0.y);
    x[1] = mk<T>(e1.x + o1.x, e1.y + o1.y);
    x[2] = mk<T>(e2.x + o2.x, e2.y + o2.y);
    x[3] = mk<T>(e3.x + o3.x, e3.y + o3.y);
    x[4] = mk<T>(e0.x - o0.x, e0.y - o0.y);
    x[5] = mk<T>(e1.x - o1.x, e1.y - o1.y);
    x[6] = mk<T>(e2.x - o2.x, e2.y - o2.y);
    x[7] = mk<T>(e3.x - o3.x, e3.y - o3.y);
}

// 1-D, n = 4096 = 8^4 (the longest line): the line's values stay in
// registers.  Every radix-8 stage of this transform reads the elements
// j + 512 r of thread j's butterfly, and the last stage writes them back to
// the same elements, so a thread's eight points (j + 512 s) are its inputs of
// the first stage and its outputs of the last: the pointwise half steps and
// the phase run in registers, and only the three stage-to-stage exchanges of
// a transform go through LDS, alternating between two buffers (one barrier
// each instead of two).  Six barriers a step instead of about nineteen.
template <typename S, typename T, bool DYN>
__global__ void __launch_bounds__(512) k_ssf_line4096(const SsfArgs<S, T> a) {
    using C = cplx_t<T>;
    constexpr int LN = 12, NN = 1 << LN, NTH = NN / 8;
    static_assert(sizeof(T) == 8, "double-precision transforms");
    __shared__ C mem[2][lds_padded(NN, NLS_SSF_LINE_PAD)];
    const Lds<T, NLS_SSF_LINE_PAD> A{mem[0]}, B{mem[1]};
    const int j = threadIdx.x;
    C x[8], ph[8];
    T pp[8], nn[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int e = j + NTH * r;
        x[r] = widen<T, S>(a.psi[e]);
        ph[r] = a.ph[e];
        pp[r] = (T)a.P[e];
        nn[r] = DYN ? (T)a.nres[e] : (T)0;
    }
    const C w8 = a.tw[NN >> 3];
    // outputs of stage logNs to buffer L (elements d + r Ns), then the
    // thread's inputs of the next stage (elements j + 512 r)
    auto exchange = [&](const Lds<T, NLS_SSF_LINE_PAD> &L, int logNs) {
        const int Ns = 1 << logNs;
        const int d = ((j >> logNs) << (logNs + 3)) + (j & (Ns - 1));
#pragma unroll
        for (int r = 0; r < 8; ++r) L[d + r * Ns] = x[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) x[r] = L[j + NTH * r];
    };
    for (int it = 0; it < a.iters; ++it) {
#pragma unroll
        for (int r = 0; r < 8; ++r) pointwise_rk4<T, DYN>(a.c, a.inv_c11, pp[r], a.half_dt, x[r], nn[r]);
        bfly8<T, false, false>(x, j, 0, LN, a.tw, w8);
        exchange(A, 0);
        bfly8<T, false, true>(x, j, 3, LN, a.tw, w8);
        exchange(B, 3);
        bfly8<T, false, true>(x, j, 6, LN, a.tw, w8);
        exchange(A, 6);
        bfly8<T, false, true>(x, j, 9, LN, a.tw, w8);     // outputs j + 512 r: the thread's own points
#pragma unroll
        for (int r = 0; r < 8; ++r) x[r] = cmul<T>(x[r], ph[r]);
        bfly8<T, true, false>(x, j, 0, LN, a.tw, w8);
        exchange(B, 0);
        bfly8<T, true, true>(x, j, 3, LN, a.tw, w8);
        exchange(A, 3);
        bfly8<T, true, true>(x, j, 6, LN, a.tw, w8);
        exchange(B, 6);
        bfly8<T, true, true>(x, j, 9, LN, a.tw, w8);
#pragma unroll
        for (int r = 0; r < 8; ++r) pointwise_rk4<T, DYN>(a.c, a.inv_c11, pp[r], a.half_dt, x[r], nn[r]);
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int e = j + NTH * r;
        a.psi[e] = narrow<S, T>(x[r]);
        if (DYN) a.nres[e] = (S)nn[r];
    }
}

// The unit complex number (c, s) in T.  In single precision the rounded
// pair has |.|^2 = 1 + O(6e-8), a fixed gain or loss of the mode it
// multiplies on every use; of the pairs within one ulp of (c, s) the one
// closest to the unit circle is taken (|.|^2 - 1 ~ 1e-9), so thousands of
// transforms and phase steps do not drift the mode amplitudes.
template <typename T> __device__ __forceinline__ cplx_t<T> unit_round(double c, double s) {
    if (sizeof(T) == 8) return mk<T>((T)c, (T)s);
    const float c0 = (float)c, s0 = (float)s;
    float bc = c0, bs = s0;
    double best = fabs((double)c0 * c0 + (double)s0 * s0 - 1.0);
    const float cs[3] = {nextafterf(c0, -2.0f), c0, nextafterf(c0, 2.0f)};
    const float ss[3] = {nextafterf(s0, -2.0f), s0, nextafterf(s0, 2.0f)};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double e = fabs((double)cs[i] * cs[i] + (double)ss[j] * ss[j] - 1.0);
            if (e < best) { best = e; bc = cs[i]; bs = ss[j]; }
        }
    return mk<T>((T)bc, (T)bs);
}

// Twiddles and phase, double precision, rounded to T on the unit circle
// (unit_round); the phase carries the transform's 1/N (a power of two).
template <typename T>
__global__ void k_ssf_tables(int N, double dx, double dt, double scale, cplx_t<T> *tw, cplx_t<T> *ph) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= N) return;
    double s, c;
    sincospi(2.0 * (double)m / (double)N, &s, &c);
    tw[m] = unit_round<T>(c, -s);
    const int f = m < N / 2 ? m : m - N;                 // FFT frequency order
    const double k = 2.0 * M_PI * (double)f / ((double)N * dx);
    double ps, pc;
    sincos(k * k * dt, &ps, &pc);
    const cplx_t<T> u = unit_round<T>(pc, -ps);
    ph[m] = mk<T>(u.x * (T)scale, u.y * (T)scale);
}

static size_t align256s(size_t b) { return (b + 255) & ~(size_t)255; }

static int log2_exact(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return (1 << l) == n ? l : -1;
}

template <typename S, bool DYN>
static int ssf_run(const nls_problem *p, int iters, const S *P, cplx_t<S> *psi, S *nres, char *ws, hipStream_t s) {
    using T = double;                 // arithmetic, LDS and tables
    using C = cplx_t<T>;
    const int N = p->n, logN = log2_exact(N);
    C *tw = (C *)ws, *ph = (C *)(ws + align256s(N * sizeof(C)));
    // 1/N per inverse transform: once in 1-D, once per axis (folded into the
    // product ph[ky] * ph[kx]) in 2-D
    hipLaunchKernelGGL(k_ssf_tables<T>, dim3((N + 255) / 256), dim3(256), 0, s, N, p->dx, p->dt, 1.0 / N, tw, ph);
    SsfArgs<S, T> a;
    memset(&a, 0, sizeof a);
    a.logN = logN;
    a.half_dt = (T)(0.5 * p->dt);
    a.c = make_coeffs<T>(p->coeffs);
    a.inv_c11 = DYN ? (T)(1.0 / p->coeffs[10]) : (T)0;
    a.psi = psi; a.nres = nres; a.P = P; a.tw = tw; a.ph = ph;
    if (p->dim == 1) {
        a.iters = iters;
        if (logN == 12) hipLaunchKernelGGL((k_ssf_line4096<S, T, DYN>), dim3(1), dim3(512), 0, s, a);
        else hipLaunchKernelGGL((k_ssf_line<S, T, DYN>), dim3(1), dim3(LineShape<T>::NTH), 0, s, a);
        return hip_check(hipGetLastError(), "ssf line launch");
    }
    const int rblocks = N >> lines_per_wg(logN, ROW_LOG_PTS);
    const bool wide = logN == ROW_LOG_PTS;
    const int clog = wide ? ColShape<T, true>::LOG_L : ColShape<T>::LOG_L;
    const int cblocks = N >> (clog < logN ? clog : logN);
    const int chunk = (cblocks + 7) / 8;
    hipLaunchKernelGGL((k_ssf_rows<S, T, DYN, ROWS_FIRST>), dim3(rblocks), dim3(ROW_NTH), 0, s, a);
    for (int it = 0; it < iters; ++it) {
        if (wide)
            hipLaunchKernelGGL((k_ssf_cols<S, T, true>), dim3(8 * chunk), dim3(ColShape<T, true>::NTH), 0, s, a,
                               cblocks, chunk);
        else
            hipLaunchKernelGGL((k_ssf_cols<S, T>), dim3(8 * chunk), dim3(ColShape<T>::NTH), 0, s, a, cblocks, chunk);
        if (it + 1 < iters)
        {
            if (logN == ROW_LOG_PTS)
                hipLaunchKernelGGL((k_ssf_rows<S, T, DYN, ROWS_MIDDLE, true>), dim3(rblocks), dim3(ROW_NTH), 0, s, a);
            else
                hipLaunchKernelGGL((k_ssf_rows<S, T, DYN, ROWS_MIDDLE>), dim3(rblocks), dim3(ROW_NTH), 0, s, a);
        }
        else
            hipLaunchKernelGGL((k_ssf_rows<S, T, DYN, ROWS_LAST>), dim3(rblocks), dim3(ROW_NTH), 0, s, a);
    }
    return hip_check(hipGetLastError(), "ssf launch");
}

// ---------------------------------------------------------------------------
// Kinetic energy with the scheme's own (spectral) operator:
//   -Re<psi, L psi> = (1 / N^d) sum_k |k|^2 |psi_hat(k)|^2
// (psi_hat the unnormalised DFT, N^d grid points).  In 2-D it is taken from
// 1-D transforms only: by Parseval along the other axis,
//   sum_{kx,ky} kx^2 |psi_hat|^2 = N sum_rows sum_kx kx^2 |DFT_x(row)|^2
// and likewise for ky over the columns, so
//   -Re<psi, L psi> = (1/N) (sum over rows + sum over columns)
// with no 2-D transform and no intermediate field.  Each workgroup reduces
// its lines' sum of k^2 |DFT|^2 in a fixed tree into one partial; one
// workgroup adds the partials in order.  Double precision throughout.

// sum over the workgroup's threads of v, in a fixed order, into *out
// (thread 0); sh: NTH doubles of LDS, free on entry
template <int NTH> __device__ __forceinline__ void fixed_block_sum(double *sh, double v, double *out) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
#pragma unroll
    for (int s = NTH / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    if (t == 0) *out = sh[0];
}

// k_m^2 |v|^2 for FFT output index m of a length-N transform (frequency m,
// or m - N past N/2, as the phase table), k_m = k0 * frequency
__device__ __forceinline__ double k2_abs2(int m, int N, double k0, cplx_t<double> v) {
    const int f = m < N / 2 ? m : m - N;
    const double k = k0 * (double)f;
    return k * k * (v.x * v.x + v.y * v.y);
}

template <typename S>
__global__ void __launch_bounds__(ROW_NTH) k_ssf_kin_rows(const cplx_t<S> *psi, int logN, const cplx_t<double> *tw,
                                                          double k0, double *part) {
    using C = cplx_t<double>;
    constexpr int PPT = ROW_PTS / ROW_NTH;
    __shared__ C buf_mem[lds_padded(ROW_PTS, NLS_SSF_ROW_PAD)];
    const Lds<double, NLS_SSF_ROW_PAD> buf{buf_mem};
    const int N = 1 << logN;
    const int logL = lines_per_wg(logN, ROW_LOG_PTS);
    const int pts = 1 << (logL + logN);
    const size_t base = (size_t)blockIdx.x << (logL + logN);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + ROW_NTH * q;
        if (idx < pts) buf[idx] = widen<double, S>(psi[base + idx]);
    }
    __syncthreads();
    lds_fft<double, false, false, ROW_PTS, ROW_NTH>(buf, logN, logL, N, 1, tw);
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + ROW_NTH * q;
        if (idx < pts) v += k2_abs2(idx & (N - 1), N, k0, buf[idx]);
    }
    __syncthreads();
    fixed_block_sum<ROW_NTH>((double *)buf_mem, v, part + blockIdx.x);
}

template <typename S>
__global__ void __launch_bounds__(ColShape<double>::NTH) k_ssf_kin_cols(const cplx_t<S> *psi, int logN,
                                                                         const cplx_t<double> *tw, double k0,
                                                                         double *part) {
    using C = cplx_t<double>;
    constexpr int COL_PTS = ColShape<double>::PTS, COL_NTH = ColShape<double>::NTH;
    constexpr int PPT = COL_PTS / COL_NTH;
    __shared__ C buf_mem[lds_padded(COL_PTS, NLS_SSF_COL_PAD)];
    const Lds<double, NLS_SSF_COL_PAD> buf{buf_mem};
    const int N = 1 << logN;
    const int logL = ColShape<double>::LOG_L < logN ? ColShape<double>::LOG_L : logN;
    const int L = 1 << logL, pts = L << logN;
    const int col0 = blockIdx.x << logL;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + COL_NTH * q;
        if (idx < pts) buf[idx] = widen<double, S>(psi[(size_t)(idx >> logL) * N + col0 + (idx & (L - 1))]);
    }
    __syncthreads();
    lds_fft<double, false, true, COL_PTS, COL_NTH>(buf, logN, logL, 1, L, tw);
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + COL_NTH * q;
        if (idx < pts) v += k2_abs2(idx >> logL, N, k0, buf[idx]);
    }
    __syncthreads();
    fixed_block_sum<COL_NTH>((double *)buf_mem, v, part + blockIdx.x);
}

// 1-D: one workgroup, the whole line (n <= 4096)
template <typename S>
__global__ void __launch_bounds__(LineShape<double>::NTH) k_ssf_kin_line(const cplx_t<S> *psi, int logN,
                                                                          const cplx_t<double> *tw, double k0,
                                                                          double *part) {
    using C = cplx_t<double>;
    constexpr int LINE_NTH = LineShape<double>::NTH;
    constexpr int PPT = LINE_PTS / LINE_NTH;
    __shared__ C buf_mem[lds_padded(LINE_PTS, NLS_SSF_LINE_PAD)];
    const Lds<double, NLS_SSF_LINE_PAD> buf{buf_mem};
    const int N = 1 << logN;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + LINE_NTH * q;
        if (idx < N) buf[idx] = widen<double, S>(psi[idx]);
    }
    __syncthreads();
    lds_fft<double, false, false, LINE_PTS, LINE_NTH>(buf, logN, 0, N, 1, tw);
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = threadIdx.x + LINE_NTH * q;
        if (idx < N) v += k2_abs2(idx, N, k0, buf[idx]);
    }
    __syncthreads();
    fixed_block_sum<LINE_NTH>((double *)buf_mem, v, part);
}

// partials 0 .. nb-1 added in order by one workgroup of 256 threads: thread
// t adds t, t + 256, ...; then a fixed tree
__global__ void __launch_bounds__(256) k_ssf_kin_sum(int nb, const double *part, double *out) {
    __shared__ double sh[256];
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += part[b];
    fixed_block_sum<256>(sh, v, out);
}

size_t ssf_kinetic_scratch_bytes(const nls_problem *p) {
    // twiddles and phases (the table kernel writes both), the partials, the sum
    return 2 * align256s((size_t)p->n * 16) + align256s(((size_t)2 * p->n + 1) * sizeof(double));
}

int ssf_kinetic(const nls_problem *p, const void *psi, void *scratch, double *dev_sum, hipStream_t s) {
    int rc = ssf_check(p);
    if (rc) return rc;
    using C = cplx_t<double>;
    const int N = p->n, logN = log2_exact(N);
    char *ws = (char *)scratch;
    C *tw = (C *)ws, *ph = (C *)(ws + align256s((size_t)N * sizeof(C)));
    double *part = (double *)(ws + 2 * align256s((size_t)N * sizeof(C)));
    hipLaunchKernelGGL(k_ssf_tables<double>, dim3((N + 255) / 256), dim3(256), 0, s, N, p->dx, p->dt, 1.0, tw, ph);
    const double k0 = 2.0 * M_PI / ((double)N * p->dx);
    const bool f64 = p->dtype == NLS_F64;
    int nb;
    if (p->dim == 1) {
        nb = 1;
        if (f64) hipLaunchKernelGGL(k_ssf_kin_line<double>, dim3(1), dim3(LineShape<double>::NTH), 0, s,
                                    (const double2 *)psi, logN, tw, k0, part);
        else hipLaunchKernelGGL(k_ssf_kin_line<float>, dim3(1), dim3(LineShape<double>::NTH), 0, s,
                                (const float2 *)psi, logN, tw, k0, part);
    } else {
        const int rblocks = N >> lines_per_wg(logN, ROW_LOG_PTS);
        const int cblocks = N >> (ColShape<double>::LOG_L < logN ? ColShape<double>::LOG_L : logN);
        nb = rblocks + cblocks;
        if (f64) {
            hipLaunchKernelGGL(k_ssf_kin_rows<double>, dim3(rblocks), dim3(ROW_NTH), 0, s, (const double2 *)psi,
                               logN, tw, k0, part);
            hipLaunchKernelGGL(k_ssf_kin_cols<double>, dim3(cblocks), dim3(ColShape<double>::NTH), 0, s,
                               (const double2 *)psi, logN, tw, k0, part + rblocks);
        } else {
            hipLaunchKernelGGL(k_ssf_kin_rows<float>, dim3(rblocks), dim3(ROW_NTH), 0, s, (const float2 *)psi,
                               logN, tw, k0, part);
            hipLaunchKernelGGL(k_ssf_kin_cols<float>, dim3(cblocks), dim3(ColShape<double>::NTH), 0, s,
                               (const float2 *)psi, logN, tw, k0, part + rblocks);
        }
    }
    hipLaunchKernelGGL(k_ssf_kin_sum, dim3(1), dim3(256), 0, s, nb, part, dev_sum);
    return hip_check(hipGetLastError(), "ssf kinetic energy launch");
}

size_t ssf_workspace_bytes(const nls_problem *p) {
    return 2 * align256s((size_t)p->n * 16);          // twiddles and phases, double precision
}

int ssf_check(const nls_problem *p) {
    const int l = log2_exact(p->n);
    const int maxn = p->dim == 1 ? 4096 : 2048;
    if (l < 0 || p->n < 16 || p->n > maxn)
        return fail(NLS_EUNSUPPORTED, "split-step Fourier needs n a power of two in [16, %d] (got %d)", maxn, p->n);
    return NLS_OK;
}

int ssf_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *ws,
                hipStream_t s) {
    int rc = ssf_check(p);
    if (rc) return rc;
    const bool dyn = p->coeffs[10] != 0.0;
    if (p->dtype == NLS_F64) {
        if (dyn) return ssf_run<double, true>(p, iters, (const double *)pumping, (double2 *)psi, (double *)nres, (char *)ws, s);
        return ssf_run<double, false>(p, iters, (const double *)pumping, (double2 *)psi, nullptr, (char *)ws, s);
    }
    if (dyn) return ssf_run<float, true>(p, iters, (const float *)pumping, (float2 *)psi, (float *)nres, (char *)ws, s);
    return ssf_run<float, false>(p, iters, (const float *)pumping, (float2 *)psi, nullptr, (char *)ws, s);
}

}  // namespace nls
