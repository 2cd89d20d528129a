// nls_batch.hip -- batched solves: many independent instances (members) of
// mixed sizes packed into shared device arrays and advanced together, one
// launch covering the tiles of every member.
//
// Members share dim, order, boundary family, reservoir mode and dtype; each
// has its own n, dt, dx, coefficients and step budget.  A member whose
// budget is used up passes its state through unchanged, so a batch can hold
// solves of different lengths (the serving loop relies on this).
//
// Engines:
//   NLS_ENGINE_STAGE    one launch per RK4 stage (bodies in nls_stage.h), 15 S
//                       + 4 r bytes per point and step; any dim/boundary/order.
//   NLS_ENGINE_FUSED    one launch per RK4 step, 2-D only.  A workgroup loads
//                       its output tile plus a halo of 4R points (R = stencil
//                       radius) into registers and LDS once, runs the four
//                       stages on the shrinking region on chip, and writes the
//                       new state of the tile: 2 S + r bytes per point and
//                       step from HBM plus the halo re-reads, at the price of
//                       recomputing the halo's stages.
//   NLS_ENGINE_RESIDENT a workgroup per member holds the member on chip and
//                       runs every step of a call (1-D n <= 4096, 2-D n <= 64).
//   NLS_ENGINE_HYBRID   the members that fit the resident engine on it, on a
//                       stream of the batch, beside the others on the fused
//                       (2-D) or per-stage (1-D) engine.
#include <algorithm>
#include <atomic>
#include <type_traits>
#include <cstring>
#include <vector>

#include "nls_amd.h"
#include "nls_internal.h"
#include "nls_rk4.h"

namespace nls {

constexpr int BATCH_STAGE_TY = 16;    // tile height of the STAGE engine's 2-D tiles

template <typename T> struct MemberDev {
    int n, steps;
    long long off;          // first point of the member in the packed arrays
    long long tab;          // first entry of its radial table (1-D ref), else -1
    int id, pad;            // the member's index in the caller's batch (observables go to out[4 id ..])
    double dV;              // volume element of its observables
    T dt;
    T w0, w1, w2, w3;       // Cartesian second-derivative weights
    Coeffs<T> c;
};

struct TileDesc {
    int m, x0, y0, pad;
};

template <typename T> struct BatchArgs {
    const MemberDev<T> *mem;
    const TileDesc *tiles;
    int ntiles, chunk;      // XCD-aware mapping: blocks of one XCD take a contiguous tile range
    int it;                 // step index within this advance call
    int stage;
    T scale;
    const T *wtab;
    const cplx_t<T> *u; const T *un;
    const cplx_t<T> *kin; const T *kin_n;
    cplx_t<T> *kout; T *kout_n;
    cplx_t<T> *acc; T *acc_n;
    const T *P;
};

// Workgroups are dispatched to the 8 XCDs round-robin; give each XCD a
// contiguous range of tiles so neighbouring tiles share its L2.
__device__ __forceinline__ int xcd_tile(int ntiles, int chunk) {
    const int b = blockIdx.x;
    const int t = (b & 7) * chunk + (b >> 3);
    return t < ntiles ? t : -1;
}

template <typename T>
__device__ __forceinline__ StageArgs<T> member_view(const BatchArgs<T> &b, const MemberDev<T> &m) {
    StageArgs<T> a;
    a.n = m.n; a.stage = b.stage; a.dt = m.dt; a.scale = b.scale; a.c = m.c;
    a.w0 = m.w0; a.w1 = m.w1; a.w2 = m.w2; a.w3 = m.w3;
    a.wtab = m.tab >= 0 ? b.wtab + m.tab : nullptr;
    const long long o = m.off;
    a.u = b.u + o; a.un = b.un ? b.un + o : nullptr;
    a.kin = b.kin ? b.kin + o : nullptr; a.kin_n = b.kin_n ? b.kin_n + o : nullptr;
    a.kout = b.kout ? b.kout + o : nullptr; a.kout_n = b.kout_n ? b.kout_n + o : nullptr;
    a.acc = b.acc + o; a.acc_n = b.acc_n ? b.acc_n + o : nullptr;
    a.P = b.P + o;
    return a;
}

// ---------------------------------------------------------------------------
// Engine STAGE
// ---------------------------------------------------------------------------

template <typename T, int ORDER, int BC, bool DYN, int DIM>
__global__ void __launch_bounds__(256) k_batch_stage(const BatchArgs<T> b) {
    constexpr int R = (ORDER - 1) / 2;
    constexpr int LDS = DIM == 2 ? (BATCH_STAGE_TY + 2 * R) * (TILE2_X + 2 * R) : TILE1 + 2 * R;
    __shared__ cplx_t<T> tile[LDS];
    const int t = xcd_tile(b.ntiles, b.chunk);
    if (t < 0) return;
    const TileDesc td = b.tiles[t];
    const MemberDev<T> m = b.mem[td.m];
    const StageArgs<T> a = member_view(b, m);
    if (b.it >= m.steps) {
        // budget used up: carry the state into the buffer that becomes u
        if (b.stage == 4) {
            const int n = m.n;
            if (DIM == 2) {
                const int gx = td.x0 + (threadIdx.x & 63);
                for (int j = 0; j < BATCH_STAGE_TY / 4 && gx < n; ++j) {
                    const int gy = td.y0 + (threadIdx.x >> 6) + 4 * j;
                    if (gy >= n) break;
                    const size_t gi = (size_t)gy * n + gx;
                    a.acc[gi] = a.u[gi];
                    if (DYN) a.acc_n[gi] = a.un[gi];
                }
            } else {
                const int gx = td.x0 + threadIdx.x;
                if (gx < n) { a.acc[gx] = a.u[gx]; if (DYN) a.acc_n[gx] = a.un[gx]; }
            }
        }
        return;
    }
    if (DIM == 2) stage_tile_2d<T, ORDER, BC, DYN, BATCH_STAGE_TY>(a, td.x0, td.y0, tile);
    else stage_tile_1d<T, ORDER, BC, DYN>(a, td.x0, tile);
}

// ---------------------------------------------------------------------------
// Engine FUSED (2-D): one RK4 step per launch, halo recomputation.
// Output tile 64 x TY, halo H = 4R, region W x HH = (64 + 2H) x (TY + 2H).
// Stage s (1..4) computes the rectangle of points with at least sR halo rings
// around them.  The region's points are listed in ring order (RingMap): the
// output tile row-major, then the ring only stages 1..3 compute, then stage
// 1..2's ring, stage 1's ring and the outermost ring that is only loaded.
// Every stage's rectangle is then a prefix of the list; with 256 threads
// owning list entries t + 256 j, a stage's active lanes fill whole
// wavefronts but for one, and inactive wavefronts skip the stage entirely.
// Per owned point the state u, reservoir, pumping (as its pump_term), the
// reservoir's stage input and the RK4 accumulators stay in registers; the
// psi stage input sits in LDS so the stencil can read the neighbours.
// ---------------------------------------------------------------------------

// Buffer loads at a 32-bit byte offset from a buffer resource (base address,
// size in bytes); offsets at or past the size read zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void *base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ cplx_t<T> buffer_load_c(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ float2 buffer_load_c<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <> __device__ __forceinline__ double2 buffer_load_c<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}
template <typename T> __device__ __forceinline__ T buffer_load_r(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ float buffer_load_r<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <> __device__ __forceinline__ double buffer_load_r<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

// Grids (members, row slabs) the fused kernel holds.  In single precision it
// addresses a member's arrays with 32-bit byte offsets from buffer
// resources: complex arrays smaller than 4 GiB.  In double precision it takes
// 64-bit addresses from a 32-bit point index: fewer than 2^32 points.
bool fused_holds(long long points, int dtype) {
    return dtype == NLS_F64 ? points <= 0xffffffffLL : points * 8LL <= 0xffffffffLL;
}

template <typename T> struct FusedArgs {
    const MemberDev<T> *mem;
    const TileDesc *tiles;
    int ntiles, chunk;
    int it;
    const cplx_t<T> *u; const T *un;     // state at the start of the step
    cplx_t<T> *uo; T *uno;               // state at its end
    const T *P;
    // One grid instead of a member table: grid_ntx > 0 tiles per row of the
    // grid described by `grid`; tile t sits at (t % grid_ntx, t / grid_ntx).
    int grid_ntx;
    MemberDev<T> grid;
    // BC_SLAB: buffer rows 0..4R-1 and slab_end..slab_end+4R-1 are ghost
    // rows; rows 4R..slab_end-1 are the slab's own (the only ones written)
    int slab_end;
};

// Tile shape and launch shape of the fused kernel.  TY: output rows.  NT:
// threads per workgroup; each thread owns ceil(region / NT) points, so NT
// trades registers per thread against waves per SIMD (the kernel is bound by
// latency, and waves per SIMD hide it).  DB: the psi stage input is
// double-buffered in LDS (stage s reads one buffer and writes the next
// stage's input into the other: one barrier per stage and no registers
// holding the next input).  WAVES: waves per SIMD the register budget is
// sized for.
//
// FORM, chosen per launch by fused_form():
//  - FORM_TALL: 64 x 32 tiles of 1024 threads (64 x 16 at order 7), for
//    launches of several rounds, where the tall tile's smaller halo per
//    output point pays, or whose tall tiles fit one per compute unit.
//  - FORM_SHORT: 64 x 16 tiles of 1024 threads, for launches with no more
//    such tiles than the device has compute units (a lone grid, a few small
//    members): the step is then one workgroup's critical path, which grows
//    with the region points per thread (2.5 instead of 3.75).
//  - FORM_HALF (double precision, periodic grids and row slabs): 64 x 16
//    tiles of 512 threads with a single stage buffer, 48 KB of LDS, so a CU
//    holds two workgroups (the 1024-thread 64 x 32 tile needs 137 KB and a
//    CU holds one) and one workgroup's loads overlap the other's arithmetic,
//    which outweighs the larger halo when the launch runs more than two
//    rounds.
//    With reflecting boundaries this form would spill registers.
//  - FORM_TALL512 (single precision, order 5, periodic: the pair kernel):
//    64 x 32 tiles of 512 threads, four pairs a thread.  A stage's pairs then
//    fill 87% of the lanes of its passes instead of 75% (the stages step
//    1672, 1440, 1224 and 1024 pairs), at half the waves per SIMD; chosen by
//    the autotuner (batches of all-fused 2-D members: 92.8 -> 90.8 us on
//    sweep512; a lone 2048^2 grid: 33.1 -> 35.2 us).
// (A 64 x 64 single-precision form, 1.20 region points per output point
// instead of 1.31 but 110 KB of LDS and so one workgroup per CU, measured
// 111 us against 81 us per step on the sweep512 batch: with one workgroup a
// CU idles at every stage barrier.)
enum { FORM_TALL = 0, FORM_SHORT = 1, FORM_HALF = 2, FORM_TALL512 = 3 };
template <typename T, int ORDER, int FORM> struct FusedTile;
template <int ORDER, int FORM> struct FusedTile<float, ORDER, FORM> {
    static_assert(FORM != FORM_HALF, "single precision has no half-size form");
    static constexpr int TY = ORDER == 7 || FORM == FORM_SHORT ? 16 : 32;
    static constexpr int NT = FORM == FORM_TALL512 ? 512 : 1024;
    static constexpr bool DB = true;
    static constexpr int WAVES = FORM == FORM_TALL512 ? 4 : 8;
};
template <int ORDER, int FORM> struct FusedTile<double, ORDER, FORM> {
    static constexpr bool HALF = FORM == FORM_HALF;
    static constexpr int TY = ORDER == 7 || FORM != FORM_TALL ? 16 : 32;
    static constexpr int NT = HALF ? 512 : 1024;
    static constexpr bool DB = !HALF;
    static constexpr int WAVES = 4;
};
// LDS row stride of the region (W + pad), chosen so that the lanes of the
// ring order's side bands (2R points per row over consecutive rows) fall in
// distinct banks.
constexpr int fused_lds_stride(int W, int R) { return W + (R == 1 ? 2 : R == 2 ? 4 : 14); }
static int fused_ty(int order, int form) {
    return order == 7 || (form != FORM_TALL && form != FORM_TALL512) ? 16 : 32;
}

// Region points in ring order, built at compile time: entry k holds the
// point's LDS index y WS + x in bits 0..15, x in bits 16..23, y in 24..31.
template <int W, int HH, int R, int WS> struct RingMap {
    unsigned v[W * HH];
    constexpr RingMap() : v() {
        int k = 0;
        for (int s = 4; s >= 0; --s) {           // rectangle s minus rectangle s + 1
            const int a = s * R, b = (s + 1) * R;
            for (int y = a; y < HH - a; ++y)
                for (int x = a; x < W - a; ++x) {
                    const bool inner = s < 4 && y >= b && y < HH - b && x >= b && x < W - b;
                    if (!inner) v[k++] = (unsigned)(y * WS + x) | (unsigned)x << 16 | (unsigned)y << 24;
                }
        }
    }
};
static_assert(RingMap<80, 48, 2, 84>().v[0] == (8 * 84 + 8 | 8 << 16 | 8 << 24), "ring order starts at the output tile");
static_assert(RingMap<80, 48, 2, 84>().v[2048] == (6 * 84 + 6 | 6 << 16 | 6 << 24), "stage-3 ring follows the 64 x 32 tile");
static_assert(RingMap<80, 32, 2, 84>().v[1024] == (6 * 84 + 6 | 6 << 16 | 6 << 24), "stage-3 ring follows the 64 x 16 tile");
template <int W, int HH, int R, int WS> __device__ const RingMap<W, HH, R, WS> ring_map{};

// S32 (double precision only): the fields sit in HBM in single precision --
// state, reservoir and pumping are widened on load and the new state rounded
// once on store -- and everything else is the double-precision step: FP64
// stage arithmetic and LDS at FP32 storage bytes.  The FusedArgs pointers
// then hold the single-precision arrays.
template <typename T, int ORDER, int BC, bool DYN, int FORM, bool S32 = false>
__global__ void __launch_bounds__((FusedTile<T, ORDER, FORM>::NT), (FusedTile<T, ORDER, FORM>::WAVES))
k_fused_step_2d(const FusedArgs<T> f) {
    static_assert(!S32 || sizeof(T) == 8, "single-precision storage under double-precision arithmetic");
    using C = cplx_t<T>;
    using FT = FusedTile<T, ORDER, FORM>;
    constexpr int R = (ORDER - 1) / 2, H = 4 * R;
    constexpr int TX = 64, TY = FT::TY, NT = FT::NT;
    constexpr bool DB = FT::DB;
    constexpr int W = TX + 2 * H, HH = TY + 2 * H, NP = W * HH;
    constexpr int WS = fused_lds_stride(W, R);
    constexpr int J = (NP + NT - 1) / NT;
    constexpr int JOUT = TX * TY / NT;    // list entries t + NT j, j < JOUT, are the output tile
    static_assert(TX * TY % NT == 0 && W < 256 && HH < 256, "fused tile shape");
    // psi stage input, point (x, y) of the region at ys_mem[b][PAD + y WS + x]:
    // the front pad makes every stencil read a non-negative constant offset
    // from the point's own index, so each read is one LDS instruction with an
    // immediate offset from a single per-point base register
    constexpr int PAD = R * WS + R;
    __shared__ C ys_mem[DB ? 2 : 1][HH * WS + 2 * PAD];
    C *const ya = ys_mem[0] + PAD;            // input of stages 1 and 3 (and 2, 4 unless DB)
    C *const yb = ys_mem[DB ? 1 : 0] + PAD;   // input of stages 2 and 4

    const int t = xcd_tile(f.ntiles, f.chunk);
    if (t < 0) return;
    TileDesc td;
    MemberDev<T> m;
    if (f.grid_ntx > 0) {
        td.m = 0; td.x0 = (t % f.grid_ntx) * TX; td.y0 = (t / f.grid_ntx) * TY + (BC == BC_SLAB ? H : 0);
        m = f.grid;
    } else {
        td = f.tiles[t];
        m = f.mem[td.m];
    }
    const int n = m.n;
    const int ylim = BC == BC_SLAB ? f.slab_end : n;   // output rows end here
    const long long off = m.off;
    using SC = std::conditional_t<S32, float2, C>;      // storage types
    using SR = std::conditional_t<S32, float, T>;
    const SC *u = (const SC *)f.u + off;
    const SR *un = DYN ? (const SR *)f.un + off : nullptr;
    const SR *P = (const SR *)f.P + off;
    SC *uo = (SC *)f.uo + off;
    SR *uno = DYN ? (SR *)f.uno + off : nullptr;
    const int tid = threadIdx.x;
    // single precision: the region loads take 32-bit byte offsets from
    // buffer resources over the member's arrays (fused_holds: under 4 GiB)
    const unsigned cbytes = (unsigned)((unsigned long long)(BC == BC_SLAB ? f.slab_end + H : n) * n * sizeof(C));
    const __amdgpu_buffer_rsrc_t ru = buffer_of(u, cbytes), rp = buffer_of(P, cbytes / 2),
                                 rn = buffer_of(DYN ? un : P, cbytes / 2);

    if (f.it >= m.steps) {            // budget used up: pass the state through
#pragma unroll
        for (int j = 0; j < JOUT; ++j) {
            const int k = tid + NT * j;
            const int gx = td.x0 + (k & (TX - 1)), gy = td.y0 + k / TX;
            if (gx >= n || gy >= ylim) continue;
            const size_t gi = (size_t)gy * n + gx;
            uo[gi] = u[gi];
            if (DYN) uno[gi] = un[gi];
        }
        return;
    }

    const T dt = m.dt;
    const T dt2 = dt * (T)0.5;
    const T dt6 = dt / (T)6;
    Coeffs<T> c = m.c;
    if constexpr (sizeof(T) == 4) {
        // coefficients that meet another coefficient in one fused
        // multiply-add of the right-hand side (c3 r - c4, d13 + d14 s,
        // c13 + c14 s): a VALU instruction reads one scalar register, so
        // held in scalar registers each would be copied to a vector register
        // per point and stage; kept in vector registers for the whole kernel
        asm volatile("" : "+v"(c.c4), "+v"(c.d13), "+v"(c.c13));
    }
    const T w0 = (T)2 * m.w0;          // centre weight of the 2-D Laplacian
    const T w[4] = {m.w0, m.w1, m.w2, m.w3};
    T wc = w0;                         // meets c6 in one fused multiply-add: held in a vector register
    if constexpr (sizeof(T) == 4) asm volatile("" : "+v"(wc));
    const unsigned *map = ring_map<W, HH, R, WS>.v;

    C uu[J], acc[J], ynext[DB ? 1 : J];
    T cp[J], nn[J], yn[J], accn[J];
    int li[J];                         // LDS index of the point
    unsigned long long inside = 0;     // bit j: point j lies on the grid
    static_assert(J <= 64, "inside mask");

    const bool near = n >= TX + H;     // region offsets stay within one period of the grid

    // load the state, reservoir and pumping of the whole region once.  In
    // single precision the ring-order table entries of all this thread's
    // points are loaded first, then the fields of all its points, then the
    // LDS stores follow, so that each group of loads is in flight together
    // and waited for once (point by point, each thread waits out the memory
    // latency once per point: 2.5% slower).  In double precision the 512-thread
    // form, two workgroups per CU, groups its loads too (4% at 2048^2); with
    // one 1024-thread workgroup per CU point by point is 7% faster and is kept.
    constexpr bool GROUPED = sizeof(T) == 4 || FORM == FORM_HALF;
    unsigned ent[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int k = tid + NT * j;
        // the first TX*TY list entries are the output tile, row-major: their
        // coordinates need no table entry
        ent[j] = (GROUPED && j >= JOUT && (NP % NT == 0 || j < J - 1 || k < NP)) ? map[k] : 0u;
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int k = tid + NT * j;
        uu[j] = mk<T>(0, 0); cp[j] = 0; nn[j] = 0;
        acc[j] = mk<T>(0, 0); accn[j] = 0;
        if (!DB) ynext[j] = mk<T>(0, 0);
        li[j] = 0;
        if (NP % NT == 0 || j < J - 1 || k < NP) {
            int lx, ly;
            if (j < JOUT) {
                lx = H + (k & (TX - 1));
                ly = H + k / TX;
                li[j] = ly * WS + lx;
            } else {
                const unsigned e = GROUPED ? ent[j] : map[k];
                li[j] = e & 0xffff;
                lx = (int)((e >> 16) & 255);
                ly = (int)(e >> 24);
            }
            int gx = td.x0 - H + lx, gy = td.y0 - H + ly;
            bool ok = true;
            if (BC == BC_PERIODIC) {
                if (near) {
                    if constexpr (sizeof(T) == 4) { gx = wrap_near3(gx, n); gy = wrap_near3(gy, n); }
                    else { gx = wrap_near(gx, n); gy = wrap_near(gy, n); }
                } else {
                    gx = wrap_index(gx, n);
                    gy = wrap_index(gy, n);
                }
            } else if (BC == BC_SLAB) {
                // rows from 0 (the tile rows start at 4R); rows past the
                // bottom ghost rows reach no written point within 4R rows
                gx = near ? (sizeof(T) == 4 ? wrap_near3(gx, n) : wrap_near(gx, n)) : wrap_index(gx, n);
                ok = gy < ylim + H;
            } else {
                ok = gx >= 0 && gx < n && gy >= 0 && gy < n;
            }
            if (ok) {
                const unsigned gi = grid_index(gy, n, gx);
                inside |= 1ull << j;
                if constexpr (sizeof(T) == 4) {
                    uu[j] = buffer_load_c<T>(ru, gi * (unsigned)sizeof(C));
                    cp[j] = buffer_load_r<T>(rp, gi * (unsigned)sizeof(T));   // the pump term is formed at the store
                    if (DYN) nn[j] = buffer_load_r<T>(rn, gi * (unsigned)sizeof(T));
                } else {
                    // double precision keeps 64-bit addresses: its 512-thread
                    // form has no registers to spare for the buffer resources
                    const SC v = u[gi];
                    uu[j] = mk<T>((T)v.x, (T)v.y);
                    cp[j] = (T)P[gi];
                    if (DYN) nn[j] = (T)un[gi];
                }
            }
            if (!GROUPED) ya[li[j]] = uu[j];
        }
        if (!GROUPED) {
            cp[j] = pump_term<T, DYN>(c, cp[j]);
            yn[j] = nn[j];
        }
    }
    if (GROUPED) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (NP % NT == 0 || j < J - 1 || tid + NT * j < NP) ya[li[j]] = uu[j];
            cp[j] = pump_term<T, DYN>(c, cp[j]);
            yn[j] = nn[j];
        }
    }
    __syncthreads();

    // Laplacian and right-hand side of the stage input yin at point j
    typedef T V2 __attribute__((ext_vector_type(2)));
    using LV = __attribute__((address_space(3))) const V2;
    auto stage_rhs = [&](const C *yin, int j, C &k, T &kn) {
        // LDS byte address of the stencil's lowest point (R rows up), kept
        // opaque to the compiler: every read is then an immediate offset of
        // at most 2 R WS + R elements from this one register.  Folded into
        // the offsets, the buffer's own LDS address pushes the paired reads
        // out of their 8-bit offset range and costs a base add per pair.
        unsigned a = (unsigned)(size_t)((LV *)yin + (li[j] - R * WS));
        asm volatile("" : "+v"(a));
        LV *ys = (LV *)(size_t)a;
        auto rd = [&](int i) { const V2 v = ys[i]; return mk<T>(v.x, v.y); };
        const int idx = R * WS;
        const C y = rd(idx);
        // the stencil sum as written (neighbour sums weighted; the
        // difference form of lap_quad costs 8 more operations per point and
        // stage, and this kernel is bound by its arithmetic).  In single
        // precision the centre term joins the point's own factor in the
        // right-hand side (rhs_point_centre: two products fewer); double
        // precision adds it here.
        T lr, lm;
        if constexpr (sizeof(T) == 8) { lr = w0 * y.x; lm = w0 * y.y; }
#pragma unroll
        for (int d = 1; d <= R; ++d) {
            const C a0 = rd(idx - d), a1 = rd(idx + d);
            const C b0 = rd(idx - d * WS), b1 = rd(idx + d * WS);
            const T sr = (a0.x + a1.x) + (b0.x + b1.x), sm = (a0.y + a1.y) + (b0.y + b1.y);
            if (sizeof(T) == 4 && d == 1) {
                lr = w[1] * sr;
                lm = w[1] * sm;
            } else {
                lr += w[d] * sr;
                lm += w[d] * sm;
            }
        }
        kn = 0;
        if constexpr (sizeof(T) == 4) rhs_point_centre<T, DYN>(c, cp[j], wc, y, yn[j], mk<T>(lr, lm), k, kn);
        else rhs_point_cp<T, DYN>(c, cp[j], y, yn[j], mk<T>(lr, lm), k, kn);
    };

    // stages 1..3 on the shrinking rectangles.  Stage s computes the points
    // with at least sR halo rings around them: the first cnt list entries.
    // The test against cnt is made opaque to the compiler so that each
    // point's body stays a block of its own; resolved at compile time, the
    // stage's bodies would merge and all their LDS reads be hoisted together.
#pragma unroll
    for (int s = 1; s <= 3; ++s) {
        const int cnt = (W - 2 * s * R) * (HH - 2 * s * R);
        int lim = cnt;
        asm volatile("" : "+s"(lim));
        const T dta = (s == 3) ? dt : dt2;           // y_{s+1} = u + k_s * dt * (1/2, 1/2, 1)
        const C *yin = (s == 2) ? yb : ya;
        C *yout = (s == 2) ? ya : yb;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (NT * j >= cnt) break;
            if (tid + NT * j >= lim) continue;
            C k;
            T kn;
            stage_rhs(yin, j, k, kn);
            if (s == 1) {
                acc[j] = k;
                if (DYN) accn[j] = kn;
            } else {
                acc[j].x = acc[j].x + (T)2 * k.x;
                acc[j].y = acc[j].y + (T)2 * k.y;
                if (DYN) accn[j] = accn[j] + (T)2 * kn;
            }
            C yv = mk<T>(uu[j].x + k.x * dta, uu[j].y + k.y * dta);
            if (BC != BC_PERIODIC && !((inside >> j) & 1ull)) yv = mk<T>(0, 0);
            if (DB) yout[li[j]] = yv;
            else ynext[j] = yv;
            if (DYN) yn[j] = nn[j] + kn * dta;
        }
        __syncthreads();
        if (!DB) {
            // stage s+1 reads its input on the stage-s rectangle
#pragma unroll
            for (int j = 0; j < J; ++j) {
                if (NT * j >= cnt) break;
                if (tid + NT * j >= lim) continue;
                ya[li[j]] = ynext[DB ? 0 : j];
            }
            __syncthreads();
        }
    }

    // stage 4 on the output tile, where it lies on the grid, and
    // u_new = u + (k1 + 2 k2 + 2 k3 + k4) dt/6
#pragma unroll
    for (int j = 0; j < JOUT; ++j) {
        const int k0 = tid + NT * j;
        const int gx = td.x0 + (k0 & (TX - 1)), gy = td.y0 + k0 / TX;
        if (gx >= n || gy >= ylim) continue;
        C k;
        T kn;
        stage_rhs(yb, j, k, kn);
        const size_t gi = (size_t)gy * n + gx;
        const C v = mk<T>(uu[j].x + (acc[j].x + k.x) * dt6, uu[j].y + (acc[j].y + k.y) * dt6);
        uo[gi] = SC{(SR)v.x, (SR)v.y};
        if (DYN) uno[gi] = (SR)(nn[j] + (accn[j] + kn) * dt6);
    }
}

// Single precision, order 5, periodic grids: the same step with each thread
// owning vertical pairs of points, (x, y) and (x, y + 1).  Every ring of the region is two
// rows tall and every stage's rectangle starts on an even row, so pairs tile
// all of them; listed in ring order (PairMap) each stage's rectangle is still
// a prefix.  The two points of a pair share the column of six values their
// stencils read (rows y - 2 .. y + 3): 14 LDS reads a pair instead of 18, and
// one address.  A point's arithmetic is the same expression as in
// k_fused_step_2d, so results are the same bit for bit.
#ifndef NLS_FUSED_PAIRS
#define NLS_FUSED_PAIRS 1
#endif
// LDS row stride of the pair kernel's region: W + 2.  In ring order a
// 16-lane group of the side bands spans four row pairs; a row pair is
// 2 WS complex values = 4 WS banks apart, which at W + 2 = 82 puts the four
// in distinct banks (at W + 4 every second one collides: a model of the
// kernel's reads and writes gives 0.07 against 0.40 conflict cycles per LDS
// instruction; sweep512 on the hybrid engine 81.7 -> 81.4 us per step).
#ifndef NLS_PAIR_WS_PAD
#define NLS_PAIR_WS_PAD 2
#endif
constexpr int pair_lds_stride(int W) { return W + NLS_PAIR_WS_PAD; }

template <int W, int HH, int WS> struct PairMap {
    unsigned v[W * HH / 2];
    constexpr PairMap() : v() {
        int k = 0;
        constexpr int R = 2;
        for (int s = 4; s >= 0; --s) {
            const int a = s * R, b = (s + 1) * R;
            for (int y = a; y < HH - a; y += 2)
                for (int x = a; x < W - a; ++x) {
                    const bool inner = s < 4 && y >= b && y < HH - b && x >= b && x < W - b;
                    if (!inner) v[k++] = (unsigned)(y * WS + x) | (unsigned)x << 16 | (unsigned)y << 24;
                }
        }
    }
};
static_assert(PairMap<80, 48, 82>().v[0] == (8 * 82 + 8 | 8 << 16 | 8 << 24), "pair order starts at the output tile");
static_assert(PairMap<80, 48, 82>().v[1024] == (6 * 82 + 6 | 6 << 16 | 6 << 24), "stage-3 ring follows the tile");
static_assert(PairMap<80, 32, 82>().v[512] == (6 * 82 + 6 | 6 << 16 | 6 << 24), "stage-3 ring follows the short tile");
template <int W, int HH, int WS> __device__ const PairMap<W, HH, WS> pair_map{};

template <int BC, bool DYN, int FORM>
__global__ void __launch_bounds__((FusedTile<float, 5, FORM>::NT), (FusedTile<float, 5, FORM>::WAVES))
k_fused_pairs_2d(const FusedArgs<float> f) {
    using T = float;
    using C = float2;
    using FT = FusedTile<float, 5, FORM>;
    constexpr int R = 2, H = 4 * R;
    constexpr int TX = 64, TY = FT::TY, NT = FT::NT;
    constexpr int W = TX + 2 * H, HH = TY + 2 * H, NPP = W * HH / 2;
    constexpr int WS = pair_lds_stride(W);
    constexpr int J = (NPP + NT - 1) / NT;          // pairs per thread
    constexpr int NOUT = TX * TY / 2;              // output pairs: the first NOUT entries
    static_assert(TY % 2 == 0 && HH % 2 == 0 && W < 256 && HH < 256, "pair tile shape");
    constexpr int PAD = R * WS + R;
    __shared__ C ys_mem[2][HH * WS + 2 * PAD];
    C *const ya = ys_mem[0] + PAD;
    C *const yb = ys_mem[1] + PAD;

    const int t = xcd_tile(f.ntiles, f.chunk);
    if (t < 0) return;
    TileDesc td;
    MemberDev<T> m;
    if (f.grid_ntx > 0) {
        td.m = 0; td.x0 = (t % f.grid_ntx) * TX; td.y0 = (t / f.grid_ntx) * TY + (BC == BC_SLAB ? H : 0);
        m = f.grid;
    } else {
        td = f.tiles[t];
        m = f.mem[td.m];
    }
    const int n = m.n;
    const int ylim = BC == BC_SLAB ? f.slab_end : n;
    const long long off = m.off;
    const C *u = f.u + off;
    const T *un = DYN ? f.un + off : nullptr;
    const T *P = f.P + off;
    C *uo = f.uo + off;
    T *uno = DYN ? f.uno + off : nullptr;
    const int tid = threadIdx.x;
    const unsigned cbytes = (unsigned)((unsigned long long)(BC == BC_SLAB ? f.slab_end + H : n) * n * sizeof(C));
    const __amdgpu_buffer_rsrc_t ru = buffer_of(u, cbytes), rp = buffer_of(P, cbytes / 2),
                                 rn = buffer_of(DYN ? un : P, cbytes / 2);

    if (f.it >= m.steps) {            // budget used up: pass the state through
        for (int k = tid; k < TX * TY; k += NT) {
            const int gx = td.x0 + (k & (TX - 1)), gy = td.y0 + k / TX;
            if (gx >= n || gy >= ylim) continue;
            const size_t gi = (size_t)gy * n + gx;
            uo[gi] = u[gi];
            if (DYN) uno[gi] = un[gi];
        }
        return;
    }

    const T dt = m.dt;
    const T dt2 = dt * (T)0.5;
    const T dt6 = dt / (T)6;
    Coeffs<T> c = m.c;
    asm volatile("" : "+v"(c.c4), "+v"(c.d13), "+v"(c.c13));
    T wc = (T)2 * m.w0;                // centre weight of the 2-D Laplacian, joins the point's factor
    asm volatile("" : "+v"(wc));
    const T w1 = m.w1, w2 = m.w2;
    const unsigned *map = pair_map<W, HH, WS>.v;

    C uu[J][2], acc[J][2];
    T cp[J][2], nn[J][2], yn[J][2], accn[J][2];
    int li[J];                         // LDS index of the pair's upper point
    unsigned long long inside = 0;     // bit 2 j + q: point q of pair j lies on the grid
    static_assert(2 * J <= 64, "inside mask");
    const bool near = n >= TX + H;

    unsigned ent[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int k = tid + NT * j;
        ent[j] = (k >= NOUT && (NPP % NT == 0 || j < J - 1 || k < NPP)) ? map[k] : 0u;
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int k = tid + NT * j;
        li[j] = 0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            uu[j][q] = mk<T>(0, 0); cp[j][q] = 0; nn[j][q] = 0; acc[j][q] = mk<T>(0, 0); accn[j][q] = 0;
        }
        if (NPP % NT == 0 || j < J - 1 || k < NPP) {
            int lx, ly;
            if (k < NOUT) {
                lx = H + (k & (TX - 1));
                ly = H + 2 * (k / TX);
                li[j] = ly * WS + lx;
            } else {
                const unsigned e = ent[j];
                li[j] = e & 0xffff;
                lx = (int)((e >> 16) & 255);
                ly = (int)(e >> 24);
            }
            int gx = td.x0 - H + lx;
            if (BC == BC_PERIODIC || BC == BC_SLAB) gx = near ? wrap_near3(gx, n) : wrap_index(gx, n);
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                int gy = td.y0 - H + ly + q;
                bool ok = true;
                if (BC == BC_PERIODIC) gy = near ? wrap_near3(gy, n) : wrap_index(gy, n);
                else if (BC == BC_SLAB) ok = gy < ylim + H;
                else ok = gx >= 0 && gx < n && gy >= 0 && gy < n;
                if (ok) {
                    const unsigned gi = grid_index(gy, n, gx);
                    inside |= 1ull << (2 * j + q);
                    uu[j][q] = buffer_load_c<T>(ru, gi * (unsigned)sizeof(C));
                    cp[j][q] = buffer_load_r<T>(rp, gi * (unsigned)sizeof(T));
                    if (DYN) nn[j][q] = buffer_load_r<T>(rn, gi * (unsigned)sizeof(T));
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < J; ++j) {
        if (NPP % NT == 0 || j < J - 1 || tid + NT * j < NPP) {
            ya[li[j]] = uu[j][0];
            ya[li[j] + WS] = uu[j][1];
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            cp[j][q] = pump_term<T, DYN>(c, cp[j][q]);
            yn[j][q] = nn[j][q];
        }
    }
    __syncthreads();

    typedef T V2 __attribute__((ext_vector_type(2)));
    using LV = __attribute__((address_space(3))) const V2;
    // right-hand sides of the pair's two points from stage input yin
    auto pair_rhs = [&](const C *yin, int j, C k[2], T kn[2]) {
        // LDS byte address of the column's top (R rows above the upper point),
        // kept opaque: every read is an immediate offset from it
        unsigned a = (unsigned)(size_t)((LV *)yin + (li[j] - R * WS));
        asm volatile("" : "+v"(a));
        LV *ys = (LV *)(size_t)a;
        auto rd = [&](int i) { const V2 v = ys[i]; return mk<T>(v.x, v.y); };
        const C c0 = rd(0), c1 = rd(WS), yA = rd(2 * WS), yB = rd(3 * WS), c4 = rd(4 * WS), c5 = rd(5 * WS);
        const C a1m = rd(2 * WS - 1), a1p = rd(2 * WS + 1), a2m = rd(2 * WS - 2), a2p = rd(2 * WS + 2);
        // the stencil sum as in k_fused_step_2d: (horizontal pair + vertical pair) per distance
        T lr = w1 * ((a1m.x + a1p.x) + (c1.x + yB.x));
        T lm = w1 * ((a1m.y + a1p.y) + (c1.y + yB.y));
        lr += w2 * ((a2m.x + a2p.x) + (c0.x + c4.x));
        lm += w2 * ((a2m.y + a2p.y) + (c0.y + c4.y));
        kn[0] = 0;
        rhs_point_centre<T, DYN>(c, cp[j][0], wc, yA, yn[j][0], mk<T>(lr, lm), k[0], kn[0]);
        // the lower point's row after the upper point is done, so that its
        // reads do not all live in registers at once
        __builtin_amdgcn_sched_barrier(0);
        const C b1m = rd(3 * WS - 1), b1p = rd(3 * WS + 1), b2m = rd(3 * WS - 2), b2p = rd(3 * WS + 2);
        lr = w1 * ((b1m.x + b1p.x) + (yA.x + c4.x));
        lm = w1 * ((b1m.y + b1p.y) + (yA.y + c4.y));
        lr += w2 * ((b2m.x + b2p.x) + (c1.x + c5.x));
        lm += w2 * ((b2m.y + b2p.y) + (c1.y + c5.y));
        kn[1] = 0;
        rhs_point_centre<T, DYN>(c, cp[j][1], wc, yB, yn[j][1], mk<T>(lr, lm), k[1], kn[1]);
    };

#pragma unroll
    for (int s = 1; s <= 3; ++s) {
        const int cnt = (W - 2 * s * R) * (HH - 2 * s * R) / 2;
        int lim = cnt;
        asm volatile("" : "+s"(lim));
        const T dta = (s == 3) ? dt : dt2;
        const C *yin = (s == 2) ? yb : ya;
        C *yout = (s == 2) ? ya : yb;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            if (NT * j >= cnt) break;
            if (tid + NT * j >= lim) continue;
            C k[2];
            T kn[2];
            pair_rhs(yin, j, k, kn);
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                if (s == 1) {
                    acc[j][q] = k[q];
                    if (DYN) accn[j][q] = kn[q];
                } else {
                    acc[j][q].x = acc[j][q].x + (T)2 * k[q].x;
                    acc[j][q].y = acc[j][q].y + (T)2 * k[q].y;
                    if (DYN) accn[j][q] = accn[j][q] + (T)2 * kn[q];
                }
                C yv = mk<T>(uu[j][q].x + k[q].x * dta, uu[j][q].y + k[q].y * dta);
                if (BC != BC_PERIODIC && !((inside >> (2 * j + q)) & 1ull)) yv = mk<T>(0, 0);
                yout[li[j] + q * WS] = yv;
                if (DYN) yn[j][q] = nn[j][q] + kn[q] * dta;
            }
        }
        __syncthreads();
    }

    // stage 4 on the output tile and u_new = u + (k1 + 2 k2 + 2 k3 + k4) dt/6
#pragma unroll
    for (int j = 0; j < (NOUT + NT - 1) / NT; ++j) {
        const int k0 = tid + NT * j;
        if (k0 >= NOUT) break;
        const int gx = td.x0 + (k0 & (TX - 1)), gy = td.y0 + 2 * (k0 / TX);
        if (gx >= n || gy >= ylim) continue;
        C k[2];
        T kn[2];
        pair_rhs(yb, j, k, kn);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            if (gy + q >= ylim) break;
            const size_t gi = (size_t)(gy + q) * n + gx;
            uo[gi] = mk<T>(uu[j][q].x + (acc[j][q].x + k[q].x) * dt6, uu[j][q].y + (acc[j][q].y + k[q].y) * dt6);
            if (DYN) uno[gi] = nn[j][q] + (accn[j][q] + kn[q]) * dt6;
        }
    }
}

// ---------------------------------------------------------------------------
// Engine RESIDENT: one workgroup per member holds the member's
// whole grid on chip (1-D n <= 4096, 2-D n <= 64) and runs all of the call's
// steps in one launch -- a persistent kernel per step block.  LDS holds the
// psi state at the start of the step (U) and the psi stage input (Y);
// registers hold each thread's pumping, reservoir, reservoir stage input and
// accumulators.  HBM is touched only to load the state
// at the start and store it at the end.
// ---------------------------------------------------------------------------

constexpr int RES_MAXP = 4096;
// threads per workgroup: 1024 in single precision (4 points each); 512 in
// double precision (8 points each) so the register budget stays at 256
template <typename T> struct ResShape;
template <> struct ResShape<float> { static constexpr int THREADS = 1024; };
template <> struct ResShape<double> { static constexpr int THREADS = 512; };

template <typename T> struct ResidentArgs {
    const MemberDev<T> *mem;
    int iters;
    const T *wtab;
    cplx_t<T> *psi;
    T *nres;
    const T *P;
};

template <typename T, int ORDER, int BC, int DIM>
__device__ __forceinline__ cplx_t<T> resident_lap(const cplx_t<T> *Y, int idx, int n, const T *w,
                                                  const T *wrow) {
    constexpr int R = (ORDER - 1) / 2;
    const cplx_t<T> c = Y[idx];
    if (DIM == 1) {
        T lr = 0, li = 0;
        if (BC == BC_PERIODIC) {
            lr = lap_centre<T>(w[0], c.x); li = lap_centre<T>(w[0], c.y);
#pragma unroll
            for (int d = 1; d <= R; ++d) {
                int im = idx - d, ip = idx + d;
                if (im < 0) im += n;
                if (ip >= n) ip -= n;
                lr += lap_pair<T>(w[d], c.x, Y[im].x, Y[ip].x, d);
                li += lap_pair<T>(w[d], c.y, Y[im].y, Y[ip].y, d);
            }
        } else {
#pragma unroll
            for (int d = -R; d <= R; ++d) {
                const int j = idx + d;
                if (j < 0 || j >= n) continue;
                lr += wrow[d + R] * Y[j].x;
                li += wrow[d + R] * Y[j].y;
            }
        }
        return mk<T>(lr, li);
    }
    const int row = idx / n, col = idx - row * n;
    T lr = lap_centre<T>((T)2 * w[0], c.x), li = lap_centre<T>((T)2 * w[0], c.y);
#pragma unroll
    for (int d = 1; d <= R; ++d) {
        int cm = col - d, cp = col + d, rm = row - d, rp = row + d;
        cplx_t<T> a0 = mk<T>(0, 0), a1 = mk<T>(0, 0), b0 = mk<T>(0, 0), b1 = mk<T>(0, 0);
        if (BC == BC_PERIODIC) {
            if (cm < 0) cm += n;
            if (cp >= n) cp -= n;
            if (rm < 0) rm += n;
            if (rp >= n) rp -= n;
            a0 = Y[row * n + cm]; a1 = Y[row * n + cp]; b0 = Y[rm * n + col]; b1 = Y[rp * n + col];
        } else {
            if (cm >= 0) a0 = Y[row * n + cm];
            if (cp < n) a1 = Y[row * n + cp];
            if (rm >= 0) b0 = Y[rm * n + col];
            if (rp < n) b1 = Y[rp * n + col];
        }
        lr += lap_quad<T>(w[d], c.x, a0.x, a1.x, b0.x, b1.x, d);
        li += lap_quad<T>(w[d], c.y, a0.y, a1.y, b0.y, b1.y, d);
    }
    return mk<T>(lr, li);
}

template <typename T, int ORDER, int BC, bool DYN, int DIM>
__global__ void __launch_bounds__(ResShape<T>::THREADS) k_resident(const ResidentArgs<T> r) {
    using C = cplx_t<T>;
    constexpr int RES_THREADS = ResShape<T>::THREADS, RES_PPT = RES_MAXP / RES_THREADS;
    // Single precision: Y[b] is the psi stage input, double-buffered -- stage
    // s reads Y[b] and writes the next stage's input into Y[b ^ 1], so one
    // barrier per stage suffices (a thread writes Y[b] again only after the
    // next barrier, which every thread reaches after its reads of Y[b]); the
    // state at the start of the step is read at the thread's own points only
    // and stays in registers.  Double precision does the same on 1-D lines of
    // order 3 and 5; elsewhere it has no registers to spare for that state (it
    // would spill): Y[0] is the stage input, Y[1] the state, and each stage
    // waits for all reads of Y[0] before writing it.
    constexpr bool PINGPONG = sizeof(T) == 4 || (DIM == 1 && ORDER < 7);
    __shared__ C Y[2][RES_MAXP];
    const MemberDev<T> m = r.mem[blockIdx.x];
    const int n = m.n;
    const int N = DIM == 1 ? n : n * n;
    const int steps = r.iters < m.steps ? r.iters : m.steps;
    if (N > RES_MAXP || steps <= 0) return;        // the host never plans such members
    C *psi = r.psi + m.off;
    T *nres = DYN ? r.nres + m.off : nullptr;
    const T *P = r.P + m.off;
    const T *wtab = (DIM == 1 && BC == BC_REF) ? r.wtab + m.tab : nullptr;
    const T w[4] = {m.w0, m.w1, m.w2, m.w3};
    const T dt = m.dt;
    const int tid = threadIdx.x;
    T pp[RES_PPT], un[RES_PPT], yn[RES_PPT], accn[RES_PPT];
    C acc[RES_PPT], uu[PINGPONG ? RES_PPT : 1];
#pragma unroll
    for (int q = 0; q < RES_PPT; ++q) {
        const int idx = tid + RES_THREADS * q;
        pp[q] = 0; un[q] = 0; yn[q] = 0; accn[q] = 0; acc[q] = mk<T>(0, 0);
        if (PINGPONG) uu[PINGPONG ? q : 0] = mk<T>(0, 0);
        if (idx < N) {
            const C v = psi[idx];
            if (PINGPONG) uu[PINGPONG ? q : 0] = v;
            else Y[1][idx] = v;
            pp[q] = P[idx];
            if (DYN) { un[q] = nres[idx]; yn[q] = un[q]; }
        }
    }
#pragma unroll
    for (int q = 0; q < RES_PPT; ++q) {
        const int idx = tid + RES_THREADS * q;
        if (idx < N) Y[0][idx] = PINGPONG ? uu[PINGPONG ? q : 0] : Y[1][idx];
    }
    __syncthreads();
    int b = 0;
    for (int it = 0; it < steps; ++it) {
#pragma unroll
        for (int s = 1; s <= 4; ++s) {
            const T wacc = (s == 1 || s == 4) ? (T)1 : (T)2;
            const T anext = (s == 3) ? (T)1 : (T)0.5;
            // acc += wacc k; the next stage input u + k dt anext, or at stage
            // 4 the new state u + acc dt / 6
            auto advance_point = [&](int q, int idx, const C &kk, T kn, const C &u, C *Yout) {
                if (s == 1) { acc[q] = kk; accn[q] = kn; }
                else {
                    acc[q] = mk<T>(acc[q].x + wacc * kk.x, acc[q].y + wacc * kk.y);
                    accn[q] = accn[q] + wacc * kn;
                }
                if (s < 4) {
                    Yout[idx] = mk<T>(u.x + (kk.x * dt) * anext, u.y + (kk.y * dt) * anext);
                    if (DYN) yn[q] = un[q] + (kn * dt) * anext;
                    return u;
                }
                const C v = mk<T>(u.x + (acc[q].x * dt) / (T)6, u.y + (acc[q].y * dt) / (T)6);
                Yout[idx] = v;
                if (DYN) { un[q] = un[q] + (accn[q] * dt) / (T)6; yn[q] = un[q]; }
                return v;
            };
            if constexpr (PINGPONG) {
                const C *Yin = Y[b];
                C *Yout = Y[b ^ 1];
#pragma unroll
                for (int q = 0; q < RES_PPT; ++q) {
                    const int idx = tid + RES_THREADS * q;
                    if (idx >= N) continue;
                    C kk;
                    T kn = 0;
                    const C lap = resident_lap<T, ORDER, BC, DIM>(Yin, idx, n, w, wtab ? wtab + (size_t)idx * ORDER : nullptr);
                    rhs_point<T, DYN>(m.c, pp[q], Yin[idx], yn[q], lap, kk, kn);
                    uu[PINGPONG ? q : 0] = advance_point(q, idx, kk, kn, uu[PINGPONG ? q : 0], Yout);
                }
                __syncthreads();
                b ^= 1;
            } else {
                C kk[RES_PPT];
                T kn[RES_PPT];
#pragma unroll
                for (int q = 0; q < RES_PPT; ++q) {
                    const int idx = tid + RES_THREADS * q;
                    kk[q] = mk<T>(0, 0); kn[q] = 0;
                    if (idx < N) {
                        const C lap = resident_lap<T, ORDER, BC, DIM>(Y[0], idx, n, w, wtab ? wtab + (size_t)idx * ORDER : nullptr);
                        rhs_point<T, DYN>(m.c, pp[q], Y[0][idx], yn[q], lap, kk[q], kn[q]);
                    }
                }
                __syncthreads();                     // every read of Y[0] for this stage is done
#pragma unroll
                for (int q = 0; q < RES_PPT; ++q) {
                    const int idx = tid + RES_THREADS * q;
                    if (idx < N) Y[1][idx] = advance_point(q, idx, kk[q], kn[q], Y[1][idx], Y[0]);
                }
                __syncthreads();
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RES_PPT; ++q) {
        const int idx = tid + RES_THREADS * q;
        if (idx < N) {
            psi[idx] = PINGPONG ? uu[PINGPONG ? q : 0] : Y[1][idx];
            if (DYN) nres[idx] = un[q];
        }
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

template <typename T, int ORDER, int BC, bool DYN>
static void launch_batch_stage_t(int dim, const BatchArgs<T> &b, hipStream_t s) {
    const dim3 grid(8 * b.chunk);
    if (dim == 2) hipLaunchKernelGGL((k_batch_stage<T, ORDER, BC, DYN, 2>), grid, dim3(256), 0, s, b);
    else hipLaunchKernelGGL((k_batch_stage<T, ORDER, BC, DYN, 1>), grid, dim3(256), 0, s, b);
}

template <typename T>
static void launch_batch_stage(int dim, int order, int bc, bool dyn, const BatchArgs<T> &b, hipStream_t s) {
#define NLS_CASE(O, B, D) \
    if (order == O && bc == B && dyn == D) return launch_batch_stage_t<T, O, B, D>(dim, b, s);
    NLS_CASE(3, 0, false) NLS_CASE(3, 0, true) NLS_CASE(3, 1, false) NLS_CASE(3, 1, true)
    NLS_CASE(5, 0, false) NLS_CASE(5, 0, true) NLS_CASE(5, 1, false) NLS_CASE(5, 1, true)
    NLS_CASE(7, 0, false) NLS_CASE(7, 0, true) NLS_CASE(7, 1, false) NLS_CASE(7, 1, true)
#undef NLS_CASE
}

// One fused step launch in the given form; forms a configuration does not
// build (order 7 has one form; FORM_HALF is double precision without
// reflecting boundaries; FORM_TALL512 is the single-precision pair kernel)
// fall back to FORM_TALL, which fused_form never asks otherwise.
template <typename T, int O, int B, bool D, bool S32 = false>
static void launch_fused_form(int form, const FusedArgs<T> &f, hipStream_t s) {
    const dim3 grid(8 * f.chunk);
    // periodic grids (the variants with reflecting boundaries or row slabs
    // would spill registers in pairs and keep one point per entry)
    if constexpr (sizeof(T) == 4 && O == 5 && B == BC_PERIODIC && NLS_FUSED_PAIRS) {
        if (form == FORM_SHORT)
            hipLaunchKernelGGL((k_fused_pairs_2d<B, D, FORM_SHORT>), grid, dim3(1024), 0, s, f);
        else if (form == FORM_TALL512)
            hipLaunchKernelGGL((k_fused_pairs_2d<B, D, FORM_TALL512>), grid, dim3(512), 0, s, f);
        else
            hipLaunchKernelGGL((k_fused_pairs_2d<B, D, FORM_TALL>), grid, dim3(1024), 0, s, f);
        return;
    }
    if constexpr (O != 7) {
        if (form == FORM_SHORT) {
            hipLaunchKernelGGL((k_fused_step_2d<T, O, B, D, FORM_SHORT, S32>), grid,
                               dim3(FusedTile<T, O, FORM_SHORT>::NT), 0, s, f);
            return;
        }
        if constexpr (sizeof(T) == 8 && B != BC_REF)
            if (form == FORM_HALF) {
                hipLaunchKernelGGL((k_fused_step_2d<T, O, B, D, FORM_HALF, S32>), grid,
                                   dim3(FusedTile<T, O, FORM_HALF>::NT), 0, s, f);
                return;
            }
    }
    hipLaunchKernelGGL((k_fused_step_2d<T, O, B, D, FORM_TALL, S32>), grid, dim3(FusedTile<T, O, FORM_TALL>::NT), 0,
                       s, f);
}

// Double-precision arithmetic on single-precision fields (NLS_FLAG_F64_ARITH):
// periodic grids, the double-precision forms.
static void launch_fused_s32(int order, bool dyn, int form, const FusedArgs<double> &f, hipStream_t s) {
#define NLS_CASE(O, D) \
    if (order == O && dyn == D) return launch_fused_form<double, O, BC_PERIODIC, D, true>(form, f, s);
    NLS_CASE(3, false) NLS_CASE(3, true) NLS_CASE(5, false) NLS_CASE(5, true) NLS_CASE(7, false) NLS_CASE(7, true)
#undef NLS_CASE
}

template <typename T>
static void launch_fused(int order, int bc, bool dyn, int form, const FusedArgs<T> &f, hipStream_t s) {
#define NLS_CASE(O, B, D) \
    if (order == O && bc == B && dyn == D) return launch_fused_form<T, O, B, D>(form, f, s);
    NLS_CASE(3, 0, false) NLS_CASE(3, 0, true) NLS_CASE(3, 1, false) NLS_CASE(3, 1, true)
    NLS_CASE(5, 0, false) NLS_CASE(5, 0, true) NLS_CASE(5, 1, false) NLS_CASE(5, 1, true)
    NLS_CASE(7, 0, false) NLS_CASE(7, 0, true) NLS_CASE(7, 1, false) NLS_CASE(7, 1, true)
    NLS_CASE(3, 2, false) NLS_CASE(3, 2, true) NLS_CASE(5, 2, false) NLS_CASE(5, 2, true)
    NLS_CASE(7, 2, false) NLS_CASE(7, 2, true)
#undef NLS_CASE
}

// Compute units of the current device, read once per device.
static int device_cus() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    cache[dev].store(v, std::memory_order_relaxed);
    return v;
}

// The fused kernel's form for a launch over grids (or row slabs) of the
// given sizes, from the count of its 64 x 16 tiles (tiles16) and 64 x 32
// tiles (tiles32) against the compute units.  Measured on MI355X, order 5:
//  - tiles16 <= CUs, FORM_SHORT: a lone 256^2 grid 7.1 -> 5.2 us per step in
//    single precision and 8.8 -> 6.2 us in double; 512^2 (256 tiles)
//    7.3 -> 5.7 and 9.4 -> 7.1 us.
//  - otherwise FORM_TALL: where CUs would hold two short tiles the tall
//    ones are as fast or faster (a 2048 x 256 row slab: 9.7 us short, 9.5 us
//    tall), and launches of several rounds are faster tall (a batch of
//    512^2-class grids: 94.6 us tall, 112.4 us short);
//  - except in double precision without reflecting boundaries when the tall
//    tiles take more than two rounds of the CUs, FORM_HALF (alternating runs
//    on one device): 2048^2 (8 rounds) 89 -> 82 us, the sweep512 batch
//    196 -> 190 us; at two rounds or fewer the tall tiles are as fast or
//    faster (1024^2: 41 us tall, 43 us half; a 2048 x 256 slab, one round:
//    13.1 us tall, 14.9 us half).
// A process-wide choice of form per dtype (nls_tune_fused_form, set by the
// autotuner), -1: the rule below.
static std::atomic<int> g_fused_form[2] = {{-1}, {-1}};

int set_fused_form(int dtype, int form) {
    if (form < -1 || form > FORM_TALL512) return -1;
    g_fused_form[dtype == NLS_F64 ? 1 : 0].store(form);
    return 0;
}
int get_fused_form(int dtype) { return g_fused_form[dtype == NLS_F64 ? 1 : 0].load(); }

static int fused_form(int dtype, int order, int bc, long long tiles16, long long tiles32) {
    const int forced = get_fused_form(dtype);
    if (forced >= 0) {
        // forms a configuration does not build run tall (launch_fused_form)
        if (order == 7 || (forced == FORM_HALF && (dtype != NLS_F64 || bc == BC_REF))) return FORM_TALL;
        if (forced == FORM_TALL512 && !(dtype == NLS_F32 && order == 5 && bc == BC_PERIODIC && NLS_FUSED_PAIRS))
            return FORM_TALL;
        return forced;
    }
    if (order == 7) return FORM_TALL;
    const int cus = device_cus();
    if (cus <= 0) return FORM_TALL;
    if (tiles16 <= cus) return FORM_SHORT;
    if (dtype == NLS_F64 && bc != BC_REF && tiles32 > 2LL * cus) return FORM_HALF;
    return FORM_TALL;
}
static long long tiles_of(int n, int rows, int ty) { return (long long)((n + 63) / 64) * ((rows + ty - 1) / ty); }

// The double-precision 2-D order-7 variant would spill registers to scratch;
// it is not built and such batches use the fused or stage engine.
template <typename T> constexpr bool resident_2d_ok(int order) { return !(sizeof(T) == 8 && order == 7); }

template <typename T>
static void launch_resident(int dim, int order, int bc, bool dyn, int count, const ResidentArgs<T> &r, hipStream_t s) {
#define NLS_CASE(O, B, D)                                                                                     \
    if (order == O && bc == B && dyn == D) {                                                                 \
        if (dim == 1)                                                                                        \
            hipLaunchKernelGGL((k_resident<T, O, B, D, 1>), dim3(count), dim3(ResShape<T>::THREADS), 0, s, r); \
        else if (resident_2d_ok<T>(O))                                                                       \
            hipLaunchKernelGGL((k_resident<T, O, B, D, (resident_2d_ok<T>(O) ? 2 : 1)>), dim3(count),         \
                               dim3(ResShape<T>::THREADS), 0, s, r);                                         \
        return;                                                                                               \
    }
    NLS_CASE(3, 0, false) NLS_CASE(3, 0, true) NLS_CASE(3, 1, false) NLS_CASE(3, 1, true)
    NLS_CASE(5, 0, false) NLS_CASE(5, 0, true) NLS_CASE(5, 1, false) NLS_CASE(5, 1, true)
    NLS_CASE(7, 0, false) NLS_CASE(7, 0, true) NLS_CASE(7, 1, false) NLS_CASE(7, 1, true)
#undef NLS_CASE
}

template <typename T>
__global__ void k_batch_radial_tables(const MemberDev<T> *mem, int count, int order, const double *dx, T *wtab) {
    // one workgroup per member, rows strided over its threads
    const int mi = blockIdx.x;
    if (mi >= count) return;
    const MemberDev<T> m = mem[mi];
    if (m.tab < 0) return;
    for (int i = threadIdx.x; i < m.n; i += blockDim.x)
        for (int k = 0; k < order; ++k)
            wtab[m.tab + (long long)i * order + k] = (T)radial_weight(i, k, m.n, order, dx[mi]);
}

static size_t align256b(size_t b) { return (b + 255) & ~(size_t)255; }

// Plain 2-D RK4 solve of one n x n grid on the fused kernel.  The member
// descriptor and the tile grid travel in the kernel arguments; the workspace
// holds the other half of the state ping-pong.
size_t fused_grid_workspace_bytes(const nls_problem *p) {
    const size_t N = (size_t)p->n * p->n, rs = p->dtype == NLS_F64 ? 8 : 4;
    return align256b(N * 2 * rs) + (p->coeffs[10] != 0.0 ? align256b(N * rs) : 0);
}

// Member descriptor of a plain (single-instance) solve.
template <typename T> static MemberDev<T> plain_member(const nls_problem *p) {
    MemberDev<T> d;
    memset(&d, 0, sizeof d);
    d.n = p->n;
    d.steps = 0x7fffffff;
    d.tab = -1;
    d.dt = (T)p->dt;
    d.c = make_coeffs<T>(p->coeffs);
    T w[4] = {0, 0, 0, 0};
    for (int q = 0; q <= (p->order - 1) / 2; ++q) w[q] = (T)stencil_s2(p->order, q, p->dx);
    d.w0 = w[0]; d.w1 = w[1]; d.w2 = w[2]; d.w3 = w[3];
    return d;
}

template <typename T>
static int fused_grid_run(const nls_problem *p, int iters, const T *P, cplx_t<T> *psi, T *nres, char *ws,
                          hipStream_t s) {
    using C = cplx_t<T>;
    const size_t N = (size_t)p->n * p->n;
    const bool dyn = p->coeffs[10] != 0.0;
    const int form = fused_form(p->dtype, p->order, p->bc, tiles_of(p->n, p->n, 16), tiles_of(p->n, p->n, 32));
    const int ty = fused_ty(p->order, form);
    FusedArgs<T> f;
    memset(&f, 0, sizeof f);
    f.grid_ntx = (p->n + 63) / 64;
    f.ntiles = f.grid_ntx * ((p->n + ty - 1) / ty);
    f.chunk = (f.ntiles + 7) / 8;
    f.grid = plain_member<T>(p);
    f.P = P;
    C *cur = psi, *alt = (C *)ws;
    T *curn = nres, *altn = dyn ? (T *)(ws + align256b(N * sizeof(C))) : nullptr;
    for (int it = 0; it < iters; ++it) {
        f.it = it;
        f.u = cur; f.un = curn; f.uo = alt; f.uno = altn;
        launch_fused<T>(p->order, p->bc, dyn, form, f, s);
        std::swap(cur, alt);
        std::swap(curn, altn);
    }
    int rc = hip_check(hipGetLastError(), "fused step launch");
    if (rc) return rc;
    if (cur != psi) {
        rc = hip_check(hipMemcpyAsync(psi, cur, N * sizeof(C), hipMemcpyDeviceToDevice, s), "copy state");
        if (rc) return rc;
        if (dyn) rc = hip_check(hipMemcpyAsync(nres, curn, N * sizeof(T), hipMemcpyDeviceToDevice, s), "copy reservoir");
    }
    return rc;
}

// One RK4 step of one row slab of a grid periodic along its rows (BC_SLAB):
// buffers of (rows + 8R) x n points, the slab's rows preceded and followed by
// 4R ghost rows; the step writes rows 4R .. 4R + rows - 1 of the output.
// The 4R-deep halo is exactly what the fused kernel recomputes on chip, so
// one exchange of ghost rows per step suffices.
template <typename T>
static int fused_slab_run(const nls_problem *p, int rows, const T *P, const cplx_t<T> *psi_in, const T *nres_in,
                          cplx_t<T> *psi_out, T *nres_out, hipStream_t s) {
    const int H = 2 * (p->order - 1);
    const int form = fused_form(p->dtype, p->order, BC_SLAB, tiles_of(p->n, rows, 16), tiles_of(p->n, rows, 32));
    const int ty = fused_ty(p->order, form);
    FusedArgs<T> f;
    memset(&f, 0, sizeof f);
    f.grid_ntx = (p->n + 63) / 64;
    f.ntiles = f.grid_ntx * ((rows + ty - 1) / ty);
    f.chunk = (f.ntiles + 7) / 8;
    f.grid = plain_member<T>(p);
    f.slab_end = H + rows;
    f.P = P;
    f.u = psi_in; f.un = nres_in; f.uo = psi_out; f.uno = nres_out;
    launch_fused<T>(p->order, BC_SLAB, p->coeffs[10] != 0.0, form, f, s);
    return hip_check(hipGetLastError(), "slab step launch");
}

int fused_slab_step(const nls_problem *p, int rows, const void *P, const void *psi_in, const void *nres_in,
                    void *psi_out, void *nres_out, hipStream_t s) {
    if (!fused_holds((long long)(rows + 4 * (p->order - 1)) * p->n, p->dtype))
        return fail(NLS_EUNSUPPORTED, "a slab of %d rows x %d points is too large for the fused kernel (single "
                    "precision: complex array of 4 GiB or more; double: 2^32 points or more)", rows, p->n);
    if (p->dtype == NLS_F64)
        return fused_slab_run<double>(p, rows, (const double *)P, (const cplx_t<double> *)psi_in,
                                      (const double *)nres_in, (cplx_t<double> *)psi_out, (double *)nres_out, s);
    return fused_slab_run<float>(p, rows, (const float *)P, (const cplx_t<float> *)psi_in, (const float *)nres_in,
                                 (cplx_t<float> *)psi_out, (float *)nres_out, s);
}

// Plain solve of one line (n <= 4096) or small grid (n <= 64) on the
// resident kernel: one workgroup holds the instance on chip and runs a block
// of steps per launch.  Workspace: the radial row table of a 1-D ref solve.
bool resident_fits(const nls_problem *p) {
    if (p->scheme != NLS_SCHEME_RK4) return false;
    if (p->dim == 1) return p->n <= RES_MAXP;
    return (size_t)p->n * p->n <= (size_t)RES_MAXP && !(p->dtype == NLS_F64 && p->order == 7);
}

size_t resident_workspace_bytes(const nls_problem *p) {
    const size_t rs = p->dtype == NLS_F64 ? 8 : 4;
    return 256 + (p->dim == 1 && p->bc == NLS_BC_REF ? align256b((size_t)p->n * p->order * rs) : 0);
}

// Writes a by-value member descriptor into device memory (the one-entry
// member table of a plain resident solve), so no host-to-device copy is needed.
template <typename T> __global__ void k_put_member(MemberDev<T> *dst, const MemberDev<T> v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = v;
}

template <typename T>
static int resident_run(const nls_problem *p, int iters, const T *P, cplx_t<T> *psi, T *nres, char *ws,
                        hipStream_t s) {
    static_assert(sizeof(MemberDev<T>) <= 256, "member descriptor fits its workspace slot");
    MemberDev<T> one = plain_member<T>(p);
    ResidentArgs<T> r;
    memset(&r, 0, sizeof r);
    r.psi = psi;
    r.nres = nres;
    r.P = P;
    if (p->dim == 1 && p->bc == NLS_BC_REF) {
        launch_radial_table<T>(p->n, p->order, p->dx, (T *)(ws + 256), s);
        r.wtab = (const T *)(ws + 256);
        one.tab = 0;
    }
    hipLaunchKernelGGL((k_put_member<T>), dim3(1), dim3(64), 0, s, (MemberDev<T> *)ws, one);
    r.mem = (const MemberDev<T> *)ws;
    // bounded launches: each runs at most 4096 steps
    for (int done = 0; done < iters;) {
        r.iters = iters - done < 4096 ? iters - done : 4096;
        launch_resident<T>(p->dim, p->order, p->bc, p->coeffs[10] != 0.0, 1, r, s);
        done += r.iters;
    }
    return hip_check(hipGetLastError(), "resident launch");
}

int resident_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *ws,
                     hipStream_t s) {
    if (p->dtype == NLS_F64)
        return resident_run<double>(p, iters, (const double *)pumping, (double2 *)psi, (double *)nres, (char *)ws, s);
    return resident_run<float>(p, iters, (const float *)pumping, (float2 *)psi, (float *)nres, (char *)ws, s);
}

// NLS_FLAG_F64_ARITH: the double-precision step (forms, member descriptor,
// arithmetic) on single-precision arrays.  The FusedArgs pointers carry the
// float arrays; the kernel (S32) reads and writes them as such.
static int fused_grid_run_s32(const nls_problem *p, int iters, const float *P, float2 *psi, float *nres, char *ws,
                              hipStream_t s) {
    const size_t N = (size_t)p->n * p->n;
    const bool dyn = p->coeffs[10] != 0.0;
    const int form = fused_form(NLS_F64, p->order, p->bc, tiles_of(p->n, p->n, 16), tiles_of(p->n, p->n, 32));
    const int ty = fused_ty(p->order, form);
    FusedArgs<double> f;
    memset(&f, 0, sizeof f);
    f.grid_ntx = (p->n + 63) / 64;
    f.ntiles = f.grid_ntx * ((p->n + ty - 1) / ty);
    f.chunk = (f.ntiles + 7) / 8;
    f.grid = plain_member<double>(p);
    f.P = (const double *)P;
    float2 *cur = psi, *alt = (float2 *)ws;
    float *curn = nres, *altn = dyn ? (float *)(ws + align256b(N * sizeof(float2))) : nullptr;
    for (int it = 0; it < iters; ++it) {
        f.it = it;
        f.u = (const double2 *)cur; f.un = (const double *)curn; f.uo = (double2 *)alt; f.uno = (double *)altn;
        launch_fused_s32(p->order, dyn, form, f, s);
        std::swap(cur, alt);
        std::swap(curn, altn);
    }
    int rc = hip_check(hipGetLastError(), "fused step launch");
    if (rc) return rc;
    if (cur != psi) {
        rc = hip_check(hipMemcpyAsync(psi, cur, N * sizeof(float2), hipMemcpyDeviceToDevice, s), "copy state");
        if (rc) return rc;
        if (dyn) rc = hip_check(hipMemcpyAsync(nres, curn, N * sizeof(float), hipMemcpyDeviceToDevice, s), "copy reservoir");
    }
    return rc;
}

int fused_grid_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *ws,
                       hipStream_t s) {
    if (p->flags & NLS_FLAG_F64_ARITH) {
        // double-precision addressing: grids of fewer than 2^32 points
        if (!fused_holds((long long)p->n * p->n, NLS_F64))
            return fail(NLS_EUNSUPPORTED, "an %d x %d grid is too large for the fused kernel", p->n, p->n);
        return fused_grid_run_s32(p, iters, (const float *)pumping, (float2 *)psi, (float *)nres, (char *)ws, s);
    }
    if (!fused_holds((long long)p->n * p->n, p->dtype))
        return fail(NLS_EUNSUPPORTED, "an %d x %d grid is too large for the fused kernel (single precision: "
                    "complex array of 4 GiB or more; double: 2^32 points or more); the per-stage engine holds it",
                    p->n, p->n);
    if (p->dtype == NLS_F64)
        return fused_grid_run<double>(p, iters, (const double *)pumping, (double2 *)psi, (double *)nres, (char *)ws, s);
    return fused_grid_run<float>(p, iters, (const float *)pumping, (float2 *)psi, (float *)nres, (char *)ws, s);
}

}  // namespace nls

using namespace nls;

struct nls_batch {
    int dim, order, bc, dtype, engine;
    bool dyn;
    int count;
    size_t points;                 // length of the packed arrays
    int ntiles, chunk;
    // device tables (member table, tile list, member spacings, radial
    // tables) and their capacities in bytes; they grow when an update needs
    // more and are otherwise reused
    void *d_mem = nullptr, *d_tiles = nullptr, *d_dx = nullptr, *d_wtab = nullptr;
    size_t cap_mem = 0, cap_tiles = 0, cap_dx = 0, cap_wtab = 0;
    void *d_obs = nullptr;              // observables of every member (4 doubles each)
    size_t cap_obs = 0;
    // pinned host staging for the asynchronous table uploads: two slots used
    // in turn, each free again once its copy has run (event)
    struct Stage {
        void *host = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
    } stage[2];
    int stage_next = 0;
    std::vector<unsigned char> h_mem;   // host image of the member table
    std::vector<TileDesc> h_tiles;      // host image of the tile list
    std::vector<double> h_dx;           // grid spacing of each member
    size_t mem_stride = 0;
    long long tab_total = 0;            // radial table entries (1-D reference grids)
    int min_steps = 0x7fffffff;         // smallest member step budget
    int fform = FORM_TALL;              // fused engine: tile form (fused_form)
    int flags = 0;                      // NLS_FLAG_EAGER: no step-block graphs
    // the captured step block and the buffers it was captured on
    hipGraphExec_t graph = nullptr;
    const void *g_P = nullptr, *g_psi = nullptr, *g_nres = nullptr, *g_ws = nullptr;
    hipStream_t capture = nullptr;
    // NLS_ENGINE_HYBRID: the members the resident engine holds (part 0) and
    // the others (part 1: fused in 2-D, per-stage in 1-D), each a batch of
    // its own over the same packed arrays (member offsets unchanged); ids[k]
    // lists part k's members by index in this batch, which is also the id in
    // the part's member table.  Part 0 runs on `side`, forked from the
    // caller's stream and joined back to it.
    nls_batch *part[2] = {nullptr, nullptr};
    std::vector<int> ids[2];
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
};
static bool is_hybrid(const nls_batch *b) { return b->part[1] != nullptr; }

// Step-block graphs of a batch (fused and per-stage engines): a block of
// BATCH_GRAPH_BLOCK steps (even: the state ping-pong ends in the caller's
// buffers) is captured once per set of buffers and replayed as one graph
// launch.  A captured block bakes in the step index of each launch, which
// decides where a member whose budget ends inside the call starts passing
// its state through; blocks are therefore replayed only when every member's
// budget covers the whole call, and other calls issue their launches
// directly.
constexpr int BATCH_GRAPH_BLOCK = 64;

// Host image of the member table.  ids: each member's index in the caller's
// batch (NULL: its position here).
template <typename T>
static void fill_members(nls_batch *b, const nls_member *ms, const int *ids) {
    b->mem_stride = sizeof(MemberDev<T>);
    b->h_mem.assign(sizeof(MemberDev<T>) * b->count, 0);
    MemberDev<T> *md = (MemberDev<T> *)b->h_mem.data();
    const bool radial = b->dim == 1 && b->bc == NLS_BC_REF;
    long long tab = 0;
    for (int i = 0; i < b->count; ++i) {
        const nls_member &m = ms[i];
        MemberDev<T> d;
        memset(&d, 0, sizeof d);
        d.n = m.n;
        d.steps = 0x7fffffff;
        d.off = m.offset;
        d.id = ids ? ids[i] : i;
        // volume element of the observables: 2 pi r dr on the radial grid
        // (the sums carry the weight i, r = i dx), dx on a line, dx^2 on a square
        d.dV = radial ? 2.0 * M_PI * m.dx * m.dx : (b->dim == 1 ? m.dx : m.dx * m.dx);
        d.dt = (T)m.dt;
        d.c = make_coeffs<T>(m.coeffs);
        const int R = (b->order - 1) / 2;
        T w[4] = {0, 0, 0, 0};
        for (int q = 0; q <= R; ++q) w[q] = (T)stencil_s2(b->order, q, m.dx);
        d.w0 = w[0]; d.w1 = w[1]; d.w2 = w[2]; d.w3 = w[3];
        d.tab = -1;
        if (radial) { d.tab = tab; tab += (long long)m.n * b->order; }
        md[i] = d;
    }
    b->tab_total = tab;
}

template <typename T>
static void set_member_steps(nls_batch *b, int i, int steps) {
    ((MemberDev<T> *)b->h_mem.data())[i].steps = steps;
}

// Copies the points of the batch's members from src to dst, a workgroup per
// entry of the tile list (tiles of tx x ty points): a call whose state ends in
// the workspace moves it back into the caller's buffers without touching the
// gaps between members, which in a hybrid batch hold the resident part's
// members (stepped concurrently on another stream).
template <typename T>
__global__ void __launch_bounds__(256) k_batch_carry(const MemberDev<T> *mem, const TileDesc *tiles, int ntiles,
                                                     int dim, int tx, int ty, const cplx_t<T> *src, const T *srcn,
                                                     cplx_t<T> *dst, T *dstn) {
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const TileDesc td = tiles[t];
    const MemberDev<T> m = mem[td.m];
    const int n = m.n;
    const long long o = m.off;
    for (int k = threadIdx.x; k < tx * ty; k += 256) {
        const int gx = td.x0 + k % tx, gy = td.y0 + k / tx;
        if (gx >= n || gy >= (dim == 2 ? n : 1)) continue;
        const size_t gi = (size_t)o + (dim == 2 ? (size_t)gy * n + gx : (size_t)gx);
        dst[gi] = src[gi];
        if (dstn) dstn[gi] = srcn[gi];
    }
}

// Steps it0 .. it0 + iters - 1 of the fused or per-stage engine, each
// launch issued on s.
template <typename T>
static int batch_steps(nls_batch *b, int it0, int iters, const T *P, cplx_t<T> *psi, T *nres, char *ws,
                       hipStream_t s) {
    using C = cplx_t<T>;
    const size_t cb = align256b(b->points * sizeof(C)), rb = align256b(b->points * sizeof(T));
    const bool dyn = b->dyn;
    C *cur = psi;
    T *curn = nres;
    if (b->engine == NLS_ENGINE_FUSED) {
        C *alt = (C *)ws;
        T *altn = dyn ? (T *)(ws + cb) : nullptr;
        FusedArgs<T> f;
        memset(&f, 0, sizeof f);
        f.mem = (const MemberDev<T> *)b->d_mem;
        f.tiles = (const TileDesc *)b->d_tiles;
        f.ntiles = b->ntiles;
        f.chunk = b->chunk;
        f.P = P;
        for (int it = 0; it < iters; ++it) {
            f.it = it0 + it;
            f.u = cur; f.un = curn; f.uo = alt; f.uno = altn;
            launch_fused<T>(b->order, b->bc, dyn, b->fform, f, s);
            std::swap(cur, alt);
            std::swap(curn, altn);
        }
    } else {
        C *kA = (C *)ws, *kB = (C *)(ws + cb), *acc = (C *)(ws + 2 * cb);
        T *kAn = nullptr, *kBn = nullptr, *accn = nullptr;
        if (dyn) { kAn = (T *)(ws + 3 * cb); kBn = (T *)(ws + 3 * cb + rb); accn = (T *)(ws + 3 * cb + 2 * rb); }
        C *alt = acc;
        T *altn = accn;
        BatchArgs<T> a;
        memset(&a, 0, sizeof a);
        a.mem = (const MemberDev<T> *)b->d_mem;
        a.tiles = (const TileDesc *)b->d_tiles;
        a.ntiles = b->ntiles;
        a.chunk = b->chunk;
        a.wtab = (const T *)b->d_wtab;
        a.P = P;
        for (int it = 0; it < iters; ++it) {
            a.it = it0 + it;
            a.u = cur; a.un = curn; a.acc = alt; a.acc_n = altn;
            a.stage = 1; a.scale = 0; a.kin = nullptr; a.kin_n = nullptr; a.kout = kA; a.kout_n = kAn;
            launch_batch_stage<T>(b->dim, b->order, b->bc, dyn, a, s);
            a.stage = 2; a.scale = (T)0.5; a.kin = kA; a.kin_n = kAn; a.kout = kB; a.kout_n = kBn;
            launch_batch_stage<T>(b->dim, b->order, b->bc, dyn, a, s);
            a.stage = 3; a.scale = (T)0.5; a.kin = kB; a.kin_n = kBn; a.kout = kA; a.kout_n = kAn;
            launch_batch_stage<T>(b->dim, b->order, b->bc, dyn, a, s);
            a.stage = 4; a.scale = (T)1; a.kin = kA; a.kin_n = kAn; a.kout = nullptr; a.kout_n = nullptr;
            launch_batch_stage<T>(b->dim, b->order, b->bc, dyn, a, s);
            std::swap(cur, alt);
            std::swap(curn, altn);
        }
    }
    int rc = hip_check(hipGetLastError(), "batch launch");
    if (rc) return rc;
    if (cur != psi) {
        // an odd step count ends in the workspace: carry the members' points
        // (only theirs) back
        const int tx = b->dim == 1 ? TILE1 : 64;
        const int ty = b->dim == 1 ? 1 : b->engine == NLS_ENGINE_FUSED ? fused_ty(b->order, b->fform) : BATCH_STAGE_TY;
        hipLaunchKernelGGL((k_batch_carry<T>), dim3(b->ntiles), dim3(256), 0, s, (const MemberDev<T> *)b->d_mem,
                           (const TileDesc *)b->d_tiles, b->ntiles, b->dim, tx, ty, (const C *)cur,
                           dyn ? (const T *)curn : nullptr, psi, dyn ? nres : nullptr);
        rc = hip_check(hipGetLastError(), "carry launch");
    }
    return rc;
}

// Per-member observables: one workgroup per member sums |psi|^2 (times the
// point index on the 1-D radial grid, whose weight is r = i dx) and the
// reservoir, takes the peak density and counts non-finite points, in double
// precision.  Each thread accumulates a fixed strided slice; the wavefront
// reduces with lane shuffles (xor 32 .. 1) and the four wavefronts' sums are
// added in order, so a member's values do not depend on the rest of the batch.
// Member i's values go to out[4 id ..], the norm and population times its
// volume element dV.
template <typename T>
__global__ void __launch_bounds__(256) k_batch_obs(const MemberDev<T> *mem, int dim, int radial,
                                                   const cplx_t<T> *psi, const T *nres, double *out) {
    __shared__ double part[4][4];
    const MemberDev<T> m = mem[blockIdx.x];
    const long long N = dim == 1 ? (long long)m.n : (long long)m.n * m.n;
    const cplx_t<T> *u = psi + m.off;
    const T *nn = nres ? nres + m.off : nullptr;
    double v[4] = {0, 0, 0, 0};
    for (long long i = threadIdx.x; i < N; i += 256) {
        const double w = radial ? (double)i : 1.0;
        const double pr = u[i].x, pi = u[i].y;
        const double sq = pr * pr + pi * pi;
        const double r = nn ? (double)nn[i] : 0.0;
        if (!isfinite(sq) || !isfinite(r)) { v[3] += 1.0; continue; }
        v[0] += sq * w;
        v[1] += r * w;
        v[2] = sq > v[2] ? sq : v[2];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v[0] += __shfl_xor(v[0], off, 64);
        v[1] += __shfl_xor(v[1], off, 64);
        v[2] = fmax(v[2], __shfl_xor(v[2], off, 64));
        v[3] += __shfl_xor(v[3], off, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 4; ++q) part[wave][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[4] = {part[0][0], part[0][1], part[0][2], part[0][3]};
        for (int w = 1; w < 4; ++w) {
            o[0] += part[w][0]; o[1] += part[w][1]; o[2] = fmax(o[2], part[w][2]); o[3] += part[w][3];
        }
        double *dst = out + (size_t)m.id * 4;
        dst[0] = __dmul_rn(o[0], m.dV);
        dst[1] = __dmul_rn(o[1], m.dV);
        dst[2] = o[2];
        dst[3] = o[3];
    }
}

// Energy terms of every member (nls_energy per member): a workgroup per
// member, the Laplacian of each point from its neighbours in the packed
// array (the resident kernel's stencil: the member's weights, its radial
// table on the 1-D reference grid), FP64 accumulators, wavefront shuffles and
// the four wavefronts added in order.  out[4 i + 1..3] = kinetic
// -Re(conj psi L psi) w, (c5/2) |psi|^4 w, c6 r |psi|^2 w summed (w = i on the
// radial grid, else 1), each times the volume element dV, and out[4 id] their
// sum (products and sums rounded one by one, no fused multiply-add).
template <typename T, int ORDER, int BC, int DIM>
__global__ void __launch_bounds__(256) k_batch_energy(const MemberDev<T> *mem, const T *wtab,
                                                      const cplx_t<T> *psi, const T *P, const T *nres,
                                                      double *out) {
    __shared__ double part[4][3];
    const MemberDev<T> m = mem[blockIdx.x];
    const int n = m.n;
    const long long N = DIM == 1 ? (long long)n : (long long)n * n;
    const cplx_t<T> *u = psi + m.off;
    const T *pp = P + m.off;
    const T *nn = nres ? nres + m.off : nullptr;
    const T *wt = (DIM == 1 && BC == BC_REF) ? wtab + m.tab : nullptr;
    const T w[4] = {m.w0, m.w1, m.w2, m.w3};
    const double c12 = m.c.c12, c13 = m.c.c13, c14 = m.c.c14;
    double v[3] = {0, 0, 0};
    for (long long i = threadIdx.x; i < N; i += 256) {
        const double wi = (DIM == 1 && BC == BC_REF) ? (double)i : 1.0;
        const cplx_t<T> lap = resident_lap<T, ORDER, BC, DIM>(u, (int)i, n, w, wt ? wt + (size_t)i * ORDER : nullptr);
        const double pr = u[i].x, pi = u[i].y;
        const double sq = pr * pr + pi * pi;
        const double r = nn ? (double)nn[i] : c12 * (double)pp[i] / (c13 + c14 * sq);
        v[0] -= (pr * (double)lap.x + pi * (double)lap.y) * wi;
        v[1] += sq * sq * wi;
        v[2] += r * sq * wi;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 3; ++q) v[q] += __shfl_xor(v[q], off, 64);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < 3; ++q) part[wave][q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[3] = {part[0][0], part[0][1], part[0][2]};
        for (int k = 1; k < 4; ++k)
            for (int q = 0; q < 3; ++q) o[q] += part[k][q];
        double *dst = out + (size_t)m.id * 4;
        const double e1 = __dmul_rn(o[0], m.dV);
        const double e2 = __dmul_rn(__dmul_rn(__dmul_rn(0.5, (double)m.c.c5), o[1]), m.dV);
        const double e3 = __dmul_rn(__dmul_rn((double)m.c.c6, o[2]), m.dV);
        dst[0] = __dadd_rn(__dadd_rn(e1, e2), e3);
        dst[1] = e1;
        dst[2] = e2;
        dst[3] = e3;
    }
}

template <typename T>
static void launch_batch_energy(const nls_batch *b, const T *P, const cplx_t<T> *psi, const T *nres, double *out,
                                hipStream_t s) {
    const MemberDev<T> *mem = (const MemberDev<T> *)b->d_mem;
    const T *wt = (const T *)b->d_wtab;
    const dim3 grid(b->count), blk(256);
#define NLS_CASE(O, B, D)                                                                                     \
    if (b->order == O && b->bc == B && b->dim == D) {                                                          \
        hipLaunchKernelGGL((k_batch_energy<T, O, B, D>), grid, blk, 0, s, mem, wt, psi, P, nres, out);         \
        return;                                                                                               \
    }
    NLS_CASE(3, 0, 1) NLS_CASE(3, 1, 1) NLS_CASE(3, 0, 2) NLS_CASE(3, 1, 2)
    NLS_CASE(5, 0, 1) NLS_CASE(5, 1, 1) NLS_CASE(5, 0, 2) NLS_CASE(5, 1, 2)
    NLS_CASE(7, 0, 1) NLS_CASE(7, 1, 1) NLS_CASE(7, 0, 2) NLS_CASE(7, 1, 2)
#undef NLS_CASE
}

static int drop_graph(nls_batch *b) {
    if (!b->graph) return NLS_OK;
    // an earlier launch of it may still be in flight on any stream
    int rc = hip_check(hipDeviceSynchronize(), "release batch graph");
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    return rc;
}

// The batch's captured block of BATCH_GRAPH_BLOCK steps on these buffers,
// captured on the batch's private stream at first use.
template <typename T>
static int batch_graph(nls_batch *b, const T *P, cplx_t<T> *psi, T *nres, char *ws) {
    if (b->graph && b->g_P == P && b->g_psi == psi && b->g_nres == nres && b->g_ws == ws) return NLS_OK;
    int rc = drop_graph(b);
    if (rc) return rc;
    if (!b->capture && (rc = hip_check(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking), "capture stream")))
        return rc;
    rc = hip_check(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    rc = batch_steps<T>(b, 0, BATCH_GRAPH_BLOCK, P, psi, nres, ws, b->capture);
    hipGraph_t g = nullptr;
    const int rc_end = hip_check(hipStreamEndCapture(b->capture, &g), "end capture");
    if (rc || rc_end) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : rc_end;
    }
    rc = hip_check(hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0), "instantiate graph");
    (void)hipGraphDestroy(g);
    if (rc) {
        b->graph = nullptr;
        return rc;
    }
    b->g_P = P; b->g_psi = psi; b->g_nres = nres; b->g_ws = ws;
    return NLS_OK;
}

static bool batch_graphs_apply(const nls_batch *b, int iters) {
    return b->engine != NLS_ENGINE_RESIDENT && !(b->flags & NLS_FLAG_EAGER) && iters >= 2 * BATCH_GRAPH_BLOCK &&
           b->min_steps >= iters;
}

template <typename T>
static int batch_advance(nls_batch *b, int iters, const T *P, cplx_t<T> *psi, T *nres, char *ws, hipStream_t s) {
    if (b->engine == NLS_ENGINE_RESIDENT) {
        // a persistent launch per block of steps already
        ResidentArgs<T> r;
        memset(&r, 0, sizeof r);
        r.mem = (const MemberDev<T> *)b->d_mem;
        r.iters = iters;
        r.wtab = (const T *)b->d_wtab;
        r.psi = psi;
        r.nres = nres;
        r.P = P;
        launch_resident<T>(b->dim, b->order, b->bc, b->dyn, b->count, r, s);
        return hip_check(hipGetLastError(), "resident launch");
    }
    if (!batch_graphs_apply(b, iters)) return batch_steps<T>(b, 0, iters, P, psi, nres, ws, s);
    int rc = batch_graph<T>(b, P, psi, nres, ws);
    if (rc) return rc;
    const int blocks = iters / BATCH_GRAPH_BLOCK, rest = iters % BATCH_GRAPH_BLOCK;
    for (int k = 0; k < blocks; ++k)
        if ((rc = hip_check(hipGraphLaunch(b->graph, s), "graph launch"))) return rc;
    return rest ? batch_steps<T>(b, iters - rest, rest, P, psi, nres, ws, s) : NLS_OK;
}

// ---- batch tables: planning and asynchronous upload -----------------------

// Member checks shared by create and update: sizes, spacings, offsets, one
// reservoir mode, no overlap in the packed arrays; *points = the packed
// length the members need.
static int check_members(int dim, int order, bool dyn, const nls_member *members, int count, size_t *points) {
    size_t pts = 0;
    std::vector<std::pair<long long, long long>> iv;
    for (int i = 0; i < count; ++i) {
        const nls_member &m = members[i];
        if (m.n < order) return fail(NLS_EINVAL, "member %d: n=%d smaller than the stencil width", i, m.n);
        if (dim == 2 && m.n > 46340) return fail(NLS_EINVAL, "member %d: n too large", i);
        if (!(m.dx > 0.0)) return fail(NLS_EINVAL, "member %d: dx must be positive", i);
        if (m.offset < 0) return fail(NLS_EINVAL, "member %d: negative offset", i);
        if ((m.coeffs[10] != 0.0) != dyn) return fail(NLS_EINVAL, "members mix reservoir modes");
        const long long np = dim == 1 ? (long long)m.n : (long long)m.n * m.n;
        pts = std::max(pts, (size_t)(m.offset + np));
        iv.push_back({m.offset, m.offset + np});
    }
    std::sort(iv.begin(), iv.end());
    for (size_t i = 1; i < iv.size(); ++i)
        if (iv[i].first < iv[i - 1].second) return fail(NLS_EINVAL, "members overlap in the packed arrays");
    *points = pts;
    return NLS_OK;
}

// Host images of the member table, the tile list and the spacings of
// `members` (ids: their indices in the caller's batch, NULL for 0..count-1).
static void plan(nls_batch *b, const nls_member *members, int count, const int *ids, size_t points) {
    b->count = count;
    b->points = points;
    b->h_dx.assign(count, 0.0);
    for (int i = 0; i < count; ++i) b->h_dx[i] = members[i].dx;
    if (b->dtype == NLS_F64) fill_members<double>(b, members, ids);
    else fill_members<float>(b, members, ids);
    b->min_steps = 0x7fffffff;
    // tiles, member by member, row-major inside a member
    int tx = 64, ty = BATCH_STAGE_TY;
    if (b->engine == NLS_ENGINE_FUSED) {
        long long tiles16 = 0, tiles32 = 0;
        for (int i = 0; i < count; ++i) {
            tiles16 += tiles_of(members[i].n, members[i].n, 16);
            tiles32 += tiles_of(members[i].n, members[i].n, 32);
        }
        b->fform = fused_form(b->dtype, b->order, b->bc, tiles16, tiles32);
        ty = fused_ty(b->order, b->fform);
    }
    if (b->dim == 1) { tx = TILE1; ty = 1; }
    b->h_tiles.clear();
    if (b->engine != NLS_ENGINE_RESIDENT)
        for (int i = 0; i < count; ++i) {
            const int n = members[i].n;
            const int nyt = b->dim == 1 ? 1 : (n + ty - 1) / ty, nxt = (n + tx - 1) / tx;
            for (int y = 0; y < nyt; ++y)
                for (int x = 0; x < nxt; ++x) b->h_tiles.push_back({i, x * tx, b->dim == 1 ? 0 : y * ty, 0});
        }
    b->ntiles = (int)b->h_tiles.size();
    b->chunk = (b->ntiles + 7) / 8;
}

// Device buffer of at least `need` bytes; a larger one replaces it after the
// work queued on s (the only stream that reads it) has run.
static int grow(void **p, size_t *cap, size_t need, hipStream_t s, const char *what) {
    if (need <= *cap) return NLS_OK;
    int rc = NLS_OK;
    if (*p) {
        rc = hip_check(hipStreamSynchronize(s), what);
        (void)hipFree(*p);
        *p = nullptr;
    }
    const size_t c = std::max(need, 2 * *cap);
    *cap = 0;
    if (!rc) rc = hip_check(hipMalloc(p, c), what);
    if (!rc) *cap = c;
    return rc;
}

// Copies the member table -- and with `tables` the tile list and spacings,
// then rebuilds the radial tables -- to the device, enqueued on s through the
// next pinned staging slot (the host waits only if that slot's previous copy
// has not run yet).
static int upload(nls_batch *b, hipStream_t s, bool tables) {
    const size_t nm = b->h_mem.size();
    if (nm == 0) return NLS_OK;
    const size_t nt = tables ? b->h_tiles.size() * sizeof(TileDesc) : 0;
    const size_t nd = tables ? b->h_dx.size() * sizeof(double) : 0;
    const size_t rs = b->dtype == NLS_F64 ? 8 : 4;
    int rc = grow(&b->d_mem, &b->cap_mem, nm, s, "batch members");
    if (!rc && nt) rc = grow(&b->d_tiles, &b->cap_tiles, nt, s, "batch tiles");
    if (!rc && nd) rc = grow(&b->d_dx, &b->cap_dx, nd, s, "batch spacings");
    if (!rc && tables && b->tab_total > 0) rc = grow(&b->d_wtab, &b->cap_wtab, b->tab_total * rs, s, "radial tables");
    if (rc) return rc;
    nls_batch::Stage &st = b->stage[b->stage_next];
    b->stage_next ^= 1;
    if (!st.done) rc = hip_check(hipEventCreateWithFlags(&st.done, hipEventDisableTiming), "staging event");
    else rc = hip_check(hipEventSynchronize(st.done), "staging slot");
    if (rc) return rc;
    const size_t om = 0, ot = align256b(nm), od = ot + align256b(nt), need = od + nd;
    if (need > st.cap) {
        if (st.host) (void)hipHostFree(st.host);
        st.host = nullptr;
        st.cap = 0;
        const size_t c = std::max(need, 2 * st.cap);
        if ((rc = hip_check(hipHostMalloc(&st.host, c, hipHostMallocDefault), "staging buffer"))) return rc;
        st.cap = c;
    }
    char *h = (char *)st.host;
    memcpy(h + om, b->h_mem.data(), nm);
    if (nt) memcpy(h + ot, b->h_tiles.data(), nt);
    if (nd) memcpy(h + od, b->h_dx.data(), nd);
    rc = hip_check(hipMemcpyAsync(b->d_mem, h + om, nm, hipMemcpyHostToDevice, s), "upload members");
    if (!rc && nt) rc = hip_check(hipMemcpyAsync(b->d_tiles, h + ot, nt, hipMemcpyHostToDevice, s), "upload tiles");
    if (!rc && nd) rc = hip_check(hipMemcpyAsync(b->d_dx, h + od, nd, hipMemcpyHostToDevice, s), "upload spacings");
    if (!rc) rc = hip_check(hipEventRecord(st.done, s), "staging event");
    if (!rc && tables && b->tab_total > 0) {
        if (b->dtype == NLS_F64)
            hipLaunchKernelGGL(k_batch_radial_tables<double>, dim3(b->count), dim3(256), 0, s,
                               (const MemberDev<double> *)b->d_mem, b->count, b->order, (const double *)b->d_dx,
                               (double *)b->d_wtab);
        else
            hipLaunchKernelGGL(k_batch_radial_tables<float>, dim3(b->count), dim3(256), 0, s,
                               (const MemberDev<float> *)b->d_mem, b->count, b->order, (const double *)b->d_dx,
                               (float *)b->d_wtab);
        rc = hip_check(hipGetLastError(), "radial tables");
    }
    return rc;
}

// Budgets into the host image of the member table.
static int set_budgets(nls_batch *b, const int32_t *steps) {
    for (int i = 0; i < b->count; ++i)
        if (steps[i] < 0) return fail(NLS_EINVAL, "member %d: negative step budget", i);
    b->min_steps = 0x7fffffff;
    for (int i = 0; i < b->count; ++i) {
        if (b->dtype == NLS_F64) set_member_steps<double>(b, i, steps[i]);
        else set_member_steps<float>(b, i, steps[i]);
        b->min_steps = std::min(b->min_steps, (int)steps[i]);
    }
    return NLS_OK;
}

// A batch of one engine (not hybrid) over checked members; ids as plan().
static int create_single(int dim, int order, int bc, int dtype, int engine, bool dyn, const nls_member *members,
                         int count, const int *ids, size_t points, nls_batch **out) {
    nls_batch *b = new nls_batch();
    b->dim = dim; b->order = order; b->bc = bc; b->dtype = dtype; b->engine = engine; b->dyn = dyn;
    plan(b, members, count, ids, points);
    int rc = upload(b, nullptr, true);
    if (!rc) rc = hip_check(hipStreamSynchronize(nullptr), "batch tables");
    if (rc) { nls_batch_destroy(b); return rc; }
    *out = b;
    return NLS_OK;
}

// Observables / energy of every member into device memory out (4 doubles a
// member at 4 id), enqueued on s.
template <typename T>
static void launch_obs_t(nls_batch *b, const void *psi, const void *nres, double *out, hipStream_t s) {
    const int radial = (b->dim == 1 && b->bc == NLS_BC_REF) ? 1 : 0;
    hipLaunchKernelGGL((k_batch_obs<T>), dim3(b->count), dim3(256), 0, s, (const MemberDev<T> *)b->d_mem, b->dim,
                       radial, (const cplx_t<T> *)psi, b->dyn ? (const T *)nres : nullptr, out);
}

static int launch_obs(nls_batch *b, const void *psi, const void *nres, double *out, hipStream_t s) {
    if (is_hybrid(b)) {
        for (nls_batch *q : b->part) {
            const int rc = launch_obs(q, psi, nres, out, s);
            if (rc) return rc;
        }
        return NLS_OK;
    }
    if (b->count == 0) return NLS_OK;
    if (b->dtype == NLS_F64) launch_obs_t<double>(b, psi, nres, out, s);
    else launch_obs_t<float>(b, psi, nres, out, s);
    return hip_check(hipGetLastError(), "observables launch");
}

static int launch_energy(nls_batch *b, const void *P, const void *psi, const void *nres, double *out, hipStream_t s) {
    if (is_hybrid(b)) {
        for (nls_batch *q : b->part) {
            const int rc = launch_energy(q, P, psi, nres, out, s);
            if (rc) return rc;
        }
        return NLS_OK;
    }
    if (b->count == 0) return NLS_OK;
    if (b->dtype == NLS_F64)
        launch_batch_energy<double>(b, (const double *)P, (const double2 *)psi, b->dyn ? (const double *)nres : nullptr,
                                    out, s);
    else
        launch_batch_energy<float>(b, (const float *)P, (const float2 *)psi, b->dyn ? (const float *)nres : nullptr,
                                   out, s);
    return hip_check(hipGetLastError(), "energy launch");
}

// Runs launch (into the batch's device buffer) and downloads the 4 count
// doubles to the host array out; blocks until done.
template <typename F> static int obs_blocking(nls_batch *b, double *out, hipStream_t s, F launch) {
    const size_t bytes = (size_t)b->count * 4 * sizeof(double);
    if (bytes == 0) return NLS_OK;
    int rc = grow(&b->d_obs, &b->cap_obs, bytes, s, "observables buffer");
    if (!rc) rc = launch((double *)b->d_obs);
    if (!rc) rc = hip_check(hipMemcpyAsync(out, b->d_obs, bytes, hipMemcpyDeviceToHost, s), "observables download");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "observables sync");
    return rc;
}

extern "C" {

int nls_batch_create(int dim, int order, int bc, int dtype, int engine, const nls_member *members, int count,
                     nls_batch **out) {
    if (!out || !members || count <= 0) return fail(NLS_EINVAL, "batch needs members and an output handle");
    *out = nullptr;
    if (dim != 1 && dim != 2) return fail(NLS_EINVAL, "dim must be 1 or 2");
    if (order != 3 && order != 5 && order != 7) return fail(NLS_EINVAL, "order must be 3, 5 or 7");
    if (bc != NLS_BC_REF && bc != NLS_BC_PERIODIC) return fail(NLS_EINVAL, "bad bc");
    if (dtype != NLS_F32 && dtype != NLS_F64) return fail(NLS_EINVAL, "bad dtype");
    const int res_max = dim == 1 ? 4096 : 64;
    const bool res_built = !(dim == 2 && dtype == NLS_F64 && order == 7);
    int fit = 0;
    for (int i = 0; i < count; ++i) fit += res_built && members[i].n <= res_max;
    bool holds = true;                 // every member fits the fused kernel's 32-bit offsets
    for (int i = 0; i < count; ++i) holds = holds && fused_holds((long long)members[i].n * members[i].n, dtype);
    const int rest_engine = dim == 2 && holds ? NLS_ENGINE_FUSED : NLS_ENGINE_STAGE;
    if (engine == NLS_ENGINE_AUTO) engine = fit == count ? NLS_ENGINE_RESIDENT : fit > 0 ? NLS_ENGINE_HYBRID : rest_engine;
    if (engine == NLS_ENGINE_HYBRID && (fit == 0 || fit == count))      // one part only
        engine = fit == count ? NLS_ENGINE_RESIDENT : rest_engine;
    if (engine != NLS_ENGINE_STAGE && engine != NLS_ENGINE_FUSED && engine != NLS_ENGINE_RESIDENT &&
        engine != NLS_ENGINE_HYBRID)
        return fail(NLS_EINVAL, "bad engine");
    if (engine == NLS_ENGINE_RESIDENT && !res_built)
        return fail(NLS_EUNSUPPORTED, "the resident engine has no 2-D order-7 double-precision variant");
    if (engine == NLS_ENGINE_RESIDENT)
        for (int i = 0; i < count; ++i)
            if (members[i].n > res_max)
                return fail(NLS_EUNSUPPORTED, "member %d: n=%d does not fit one workgroup (resident engine: n <= %d)",
                            i, members[i].n, res_max);
    if (engine == NLS_ENGINE_FUSED && dim != 2) return fail(NLS_EUNSUPPORTED, "the fused engine is 2-D only");
    if (engine == NLS_ENGINE_FUSED && !holds)
        return fail(NLS_EUNSUPPORTED, "the fused engine holds members whose complex array is under 4 GiB");
    const bool dyn = members[0].coeffs[10] != 0.0;
    size_t points = 0;
    int rc = check_members(dim, order, dyn, members, count, &points);
    if (rc) return rc;
    if (engine != NLS_ENGINE_HYBRID)
        return create_single(dim, order, bc, dtype, engine, dyn, members, count, nullptr, points, out);
    nls_batch *b = new nls_batch();
    b->dim = dim; b->order = order; b->bc = bc; b->dtype = dtype; b->engine = engine; b->dyn = dyn;
    b->count = count; b->points = points;
    std::vector<nls_member> ms[2];
    for (int i = 0; i < count; ++i) {
        b->h_dx.push_back(members[i].dx);
        const int k = res_built && members[i].n <= res_max ? 0 : 1;
        b->ids[k].push_back(i);
        ms[k].push_back(members[i]);
    }
    size_t pts[2] = {0, 0};
    for (int k = 0; k < 2 && !rc; ++k) rc = check_members(dim, order, dyn, ms[k].data(), (int)ms[k].size(), &pts[k]);
    if (!rc)
        rc = create_single(dim, order, bc, dtype, NLS_ENGINE_RESIDENT, dyn, ms[0].data(), (int)ms[0].size(),
                           b->ids[0].data(), pts[0], &b->part[0]);
    if (!rc)
        rc = create_single(dim, order, bc, dtype, rest_engine, dyn, ms[1].data(), (int)ms[1].size(), b->ids[1].data(),
                           pts[1], &b->part[1]);
    if (!rc) rc = hip_check(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking), "hybrid stream");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming), "hybrid event");
    if (!rc) rc = hip_check(hipEventCreateWithFlags(&b->join_ev, hipEventDisableTiming), "hybrid event");
    if (rc) { nls_batch_destroy(b); return rc; }
    b->ntiles = b->part[1]->ntiles;
    *out = b;
    return NLS_OK;
}

int nls_batch_destroy(nls_batch *b) {
    if (!b) return NLS_OK;
    for (nls_batch *q : b->part) nls_batch_destroy(q);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    if (b->join_ev) (void)hipEventDestroy(b->join_ev);
    (void)drop_graph(b);
    if (b->capture) (void)hipStreamDestroy(b->capture);
    for (nls_batch::Stage &st : b->stage) {
        if (st.done) {
            (void)hipEventSynchronize(st.done);
            (void)hipEventDestroy(st.done);
        }
        if (st.host) (void)hipHostFree(st.host);
    }
    for (void *p : {b->d_mem, b->d_tiles, b->d_dx, b->d_wtab, b->d_obs})
        if (p) (void)hipFree(p);
    delete b;
    return NLS_OK;
}

size_t nls_batch_points(const nls_batch *b) { return b ? b->points : 0; }

int nls_batch_tiles(const nls_batch *b) { return b ? b->ntiles : 0; }

size_t nls_batch_workspace_bytes(const nls_batch *b) {
    if (!b) return 0;
    if (is_hybrid(b)) return nls_batch_workspace_bytes(b->part[1]);     // the resident part needs none
    const size_t rs = b->dtype == NLS_F64 ? 8 : 4;
    const size_t cb = align256b(b->points * 2 * rs), rb = align256b(b->points * rs);
    if (b->engine == NLS_ENGINE_RESIDENT) return 0;
    if (b->engine == NLS_ENGINE_FUSED) return cb + (b->dyn ? rb : 0);
    return 3 * cb + (b->dyn ? 3 * rb : 0);
}

int nls_batch_update(nls_batch *b, const nls_member *members, int count, const int32_t *steps, void *stream) {
    if (!b) return fail(NLS_EINVAL, "NULL batch");
    if (count < 0 || (count > 0 && !members)) return fail(NLS_EINVAL, "bad member list");
    if (is_hybrid(b))
        return fail(NLS_EUNSUPPORTED, "a hybrid batch splits its members between engines at creation; update "
                    "batches of one engine");
    if (b->engine == NLS_ENGINE_RESIDENT)
        for (int i = 0; i < count; ++i)
            if (members[i].n > (b->dim == 1 ? 4096 : 64))
                return fail(NLS_EUNSUPPORTED, "member %d: n=%d does not fit one workgroup", i, members[i].n);
    if (b->engine == NLS_ENGINE_FUSED)
        for (int i = 0; i < count; ++i)
            if (!fused_holds((long long)members[i].n * members[i].n, b->dtype))
                return fail(NLS_EUNSUPPORTED, "member %d is too large for the fused engine", i);
    size_t points = 0;
    int rc = check_members(b->dim, b->order, b->dyn, members, count, &points);
    if (rc) return rc;
    if (steps)
        for (int i = 0; i < count; ++i)
            if (steps[i] < 0) return fail(NLS_EINVAL, "member %d: negative step budget", i);
    hipStream_t s = (hipStream_t)stream;
    if (b->graph) {
        // a captured block holds the old launch shapes; it may be in flight on s
        if ((rc = hip_check(hipStreamSynchronize(s), "update"))) return rc;
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    plan(b, members, count, nullptr, points);
    if (steps && (rc = set_budgets(b, steps))) return rc;
    return upload(b, s, true);
}

int nls_batch_set_steps(nls_batch *b, const int32_t *steps, void *stream) {
    if (!b || !steps) return fail(NLS_EINVAL, "NULL argument");
    for (int i = 0; i < b->count; ++i)
        if (steps[i] < 0) return fail(NLS_EINVAL, "member %d: negative step budget", i);
    if (is_hybrid(b)) {
        // uploads on the caller's stream, which the side stream follows
        for (int k = 0; k < 2; ++k) {
            std::vector<int32_t> st;
            for (int i : b->ids[k]) st.push_back(steps[i]);
            const int rc = nls_batch_set_steps(b->part[k], st.data(), stream);
            if (rc) return rc;
        }
        return NLS_OK;
    }
    const int rc = set_budgets(b, steps);
    return rc ? rc : upload(b, (hipStream_t)stream, false);
}

int nls_batch_observables(nls_batch *b, const void *psi, const void *nres, double *out, void *stream) {
    if (!b || !psi || !out) return fail(NLS_EINVAL, "NULL argument");
    if (b->dyn && !nres) return fail(NLS_EINVAL, "NULL reservoir buffer");
    hipStream_t s = (hipStream_t)stream;
    return obs_blocking(b, out, s, [&](double *d) { return launch_obs(b, psi, nres, d, s); });
}

int nls_batch_observables_async(nls_batch *b, const void *psi, const void *nres, double *out, void *stream) {
    if (!b || !psi || !out) return fail(NLS_EINVAL, "NULL argument");
    if (b->dyn && !nres) return fail(NLS_EINVAL, "NULL reservoir buffer");
    return launch_obs(b, psi, nres, out, (hipStream_t)stream);
}

int nls_batch_energy(nls_batch *b, const void *pumping, const void *psi, const void *nres, double *out,
                     void *stream) {
    if (!b || !pumping || !psi || !out) return fail(NLS_EINVAL, "NULL argument");
    if (b->dyn && !nres) return fail(NLS_EINVAL, "NULL reservoir buffer");
    hipStream_t s = (hipStream_t)stream;
    return obs_blocking(b, out, s, [&](double *d) { return launch_energy(b, pumping, psi, nres, d, s); });
}

int nls_batch_energy_async(nls_batch *b, const void *pumping, const void *psi, const void *nres, double *out,
                           void *stream) {
    if (!b || !pumping || !psi || !out) return fail(NLS_EINVAL, "NULL argument");
    if (b->dyn && !nres) return fail(NLS_EINVAL, "NULL reservoir buffer");
    return launch_energy(b, pumping, psi, nres, out, (hipStream_t)stream);
}

int nls_batch_prepare(nls_batch *b, int iters, const void *pumping, void *psi, void *nres, void *workspace,
                      size_t workspace_bytes) {
    if (!b) return fail(NLS_EINVAL, "NULL batch");
    if (!pumping || !psi || (b->dyn && !nres)) return fail(NLS_EINVAL, "NULL device buffer");
    const size_t need = nls_batch_workspace_bytes(b);
    if (workspace_bytes < need || (need && !workspace))
        return fail(NLS_EINVAL, "workspace of %zu bytes is smaller than the %zu required", workspace_bytes, need);
    if (is_hybrid(b)) return nls_batch_prepare(b->part[1], iters, pumping, psi, nres, workspace, workspace_bytes);
    if (b->count == 0 || !batch_graphs_apply(b, iters)) return NLS_OK;
    if (b->dtype == NLS_F64)
        return batch_graph<double>(b, (const double *)pumping, (double2 *)psi, (double *)nres, (char *)workspace);
    return batch_graph<float>(b, (const float *)pumping, (float2 *)psi, (float *)nres, (char *)workspace);
}

int nls_batch_set_flags(nls_batch *b, int flags) {
    if (!b) return fail(NLS_EINVAL, "NULL batch");
    if (flags & ~NLS_FLAG_EAGER) return fail(NLS_EINVAL, "unknown flags 0x%x", flags);
    b->flags = flags;
    for (nls_batch *q : b->part)
        if (q) q->flags = flags;
    return NLS_OK;
}

int nls_batch_advance(nls_batch *b, int iters, const void *pumping, void *psi, void *nres, void *workspace,
                      size_t workspace_bytes, void *stream) {
    if (!b) return fail(NLS_EINVAL, "NULL batch");
    if (iters < 0) return fail(NLS_EINVAL, "iters must be >= 0");
    if (!pumping || !psi || (b->dyn && !nres)) return fail(NLS_EINVAL, "NULL device buffer");
    const size_t need = nls_batch_workspace_bytes(b);
    if (workspace_bytes < need || (need && !workspace))
        return fail(NLS_EINVAL, "workspace of %zu bytes is smaller than the %zu required", workspace_bytes, need);
    if (iters == 0 || b->count == 0) return NLS_OK;
    hipStream_t s = (hipStream_t)stream;
    if (is_hybrid(b)) {
        // resident part on the side stream, the rest on the caller's, joined
        int rc = hip_check(hipEventRecord(b->fork_ev, s), "hybrid fork");
        if (!rc) rc = hip_check(hipStreamWaitEvent(b->side, b->fork_ev, 0), "hybrid fork");
        if (!rc) rc = nls_batch_advance(b->part[0], iters, pumping, psi, nres, workspace, workspace_bytes, b->side);
        const int rc1 = rc ? rc : nls_batch_advance(b->part[1], iters, pumping, psi, nres, workspace, workspace_bytes, s);
        // join even after a failed launch, so the caller's stream never runs ahead of the side stream
        int rj = hip_check(hipEventRecord(b->join_ev, b->side), "hybrid join");
        if (!rj) rj = hip_check(hipStreamWaitEvent(s, b->join_ev, 0), "hybrid join");
        return rc1 ? rc1 : rj;
    }
    if (b->dtype == NLS_F64)
        return batch_advance<double>(b, iters, (const double *)pumping, (double2 *)psi, (double *)nres,
                                     (char *)workspace, s);
    return batch_advance<float>(b, iters, (const float *)pumping, (float2 *)psi, (float *)nres, (char *)workspace, s);
}

}  // extern "C"
