// nls_capi.hip -- extern "C" entry points of libnls_amd.so (declared in
// include/nls_amd.h): argument checking, workspace layout, the RK4 step
// driver and the host-memory drop-ins of the reference's Fortran
// procedures.
#include <cstdarg>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "nls_amd.h"
#include "nls_internal.h"
#include "nls_rk4.h"

namespace nls {

int check_problem(const nls_problem *p) {
    if (!p) return fail(NLS_EINVAL, "problem is NULL");
    if (p->dim != 1 && p->dim != 2) return fail(NLS_EINVAL, "dim must be 1 or 2 (got %d)", p->dim);
    if (p->order != 3 && p->order != 5 && p->order != 7)
        return fail(NLS_EINVAL, "order must be 3, 5 or 7 (got %d)", p->order);
    if (p->n < p->order) return fail(NLS_EINVAL, "n=%d is smaller than the stencil width %d", p->n, p->order);
    if (p->dim == 2 && p->n > 46340) return fail(NLS_EINVAL, "n=%d too large for a 2-D grid", p->n);
    if (p->bc != NLS_BC_REF && p->bc != NLS_BC_PERIODIC) return fail(NLS_EINVAL, "bc must be 0 or 1");
    if (p->dtype != NLS_F32 && p->dtype != NLS_F64) return fail(NLS_EINVAL, "dtype must be 0 or 1");
    if (p->scheme != NLS_SCHEME_RK4 && p->scheme != NLS_SCHEME_SSF)
        return fail(NLS_EINVAL, "scheme must be 0 or 1");
    if (!(p->dx > 0.0)) return fail(NLS_EINVAL, "dx must be positive");
    if (p->scheme == NLS_SCHEME_SSF && p->bc != NLS_BC_PERIODIC)
        return fail(NLS_EUNSUPPORTED, "the split-step Fourier scheme needs periodic boundaries");
    if ((p->flags & NLS_FLAG_F64_ARITH) &&
        (p->dtype != NLS_F32 || p->scheme != NLS_SCHEME_RK4 || p->dim != 2 || p->bc != NLS_BC_PERIODIC ||
         (p->engine != NLS_ENGINE_AUTO && p->engine != NLS_ENGINE_FUSED)))
        return fail(NLS_EUNSUPPORTED, "NLS_FLAG_F64_ARITH: single-precision 2-D periodic RK4 problems on the fused "
                                      "engine");
    if (p->scheme == NLS_SCHEME_SSF) return ssf_check(p);
    if (p->engine == NLS_ENGINE_HYBRID)
        return fail(NLS_EINVAL, "NLS_ENGINE_HYBRID splits a batch between engines; a single instance takes one engine");
    if (p->engine < NLS_ENGINE_AUTO || p->engine > NLS_ENGINE_RESIDENT)
        return fail(NLS_EINVAL, "engine must be one of NLS_ENGINE_* (got %d)", p->engine);
    if (p->engine == NLS_ENGINE_FUSED && p->dim != 2)
        return fail(NLS_EUNSUPPORTED, "the fused engine is 2-D only");

    if (p->engine == NLS_ENGINE_RESIDENT && p->scheme == NLS_SCHEME_RK4 && !resident_fits(p))
        return fail(NLS_EUNSUPPORTED, "the resident engine holds 1-D n <= 4096 or 2-D n <= 64 "
                                      "(not double-precision 2-D order 7)");
    return NLS_OK;
}

// RK4 engine of a plain solve.  AUTO: the resident kernel when the instance
// fits on chip (one launch runs a block of steps) -- every 1-D line up to
// 4096 points, 2-D grids up to 64 x 64 in single precision (in double
// precision one workgroup computing the whole grid is slower than the fused
// kernel's tiles) -- else the fused one-launch-per-step kernel for 2-D grids
// and the per-stage kernels for 1-D lines.
int plain_engine(const nls_problem *p) {
    if (p->flags & NLS_FLAG_F64_ARITH) return NLS_ENGINE_FUSED;
    if (p->engine != NLS_ENGINE_AUTO) return p->engine;
    if (resident_fits(p) && (p->dim == 1 || p->dtype == NLS_F32)) return NLS_ENGINE_RESIDENT;
    return p->dim == 2 && fused_holds((long long)p->n * p->n, p->dtype) ? NLS_ENGINE_FUSED : NLS_ENGINE_STAGE;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t points_of(const nls_problem *p) { return p->dim == 1 ? (size_t)p->n : (size_t)p->n * p->n; }

// RK4 workspace: kA, kB, acc (each one complex per point, plus one real for
// the dynamic reservoir) and, for the 1-D radial operator, its row table.
static size_t rk4_workspace(const nls_problem *p, size_t *off_tab) {
    const size_t N = points_of(p), rs = p->dtype == NLS_F64 ? 8 : 4;
    const bool dyn = p->coeffs[10] != 0.0;
    size_t b = 3 * align256(N * 2 * rs);
    if (dyn) b += 3 * align256(N * rs);
    if (off_tab) *off_tab = b;
    if (p->dim == 1 && p->bc == NLS_BC_REF) b += align256((size_t)p->n * p->order * rs);
    return b;
}

template <typename T>
static void fill_weights(const nls_problem *p, StageArgs<T> &a) {
    const int R = (p->order - 1) / 2;
    T w[4] = {0, 0, 0, 0};
    for (int d = 0; d <= R; ++d) w[d] = (T)stencil_s2(p->order, d, p->dx);
    a.w0 = w[0]; a.w1 = w[1]; a.w2 = w[2]; a.w3 = w[3];
}

// A caller-given operator in place of the one built from (order, dx): the
// row table of a 1-D band (n x order, row i holding the weights of columns
// i - R .. i + R) or the weights w0 .. wR of a 2-D Cartesian stencil.
struct OperatorGiven {
    const double *tab = nullptr;
    const double *w = nullptr;
};

// Operator of the stage kernels: weights (and, for the 1-D radial operator,
// its row table in the workspace slot `tab`).  Returns an error code.
template <typename T>
static int set_operator(const nls_problem *p, StageArgs<T> &a, T *tab, const OperatorGiven *og, hipStream_t s) {
    fill_weights<T>(p, a);
    if (og && og->w) {
        a.w0 = (T)og->w[0]; a.w1 = (T)og->w[1]; a.w2 = (T)og->w[2]; a.w3 = (T)og->w[3];
    }
    if (p->dim == 1 && p->bc == NLS_BC_REF) {
        if (og && og->tab) {
            std::vector<T> t((size_t)p->n * p->order);
            for (size_t i = 0; i < t.size(); ++i) t[i] = (T)og->tab[i];
            int rc = hip_check(hipMemcpy(tab, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice), "upload operator");
            if (rc) return rc;
        } else {
            launch_radial_table<T>(p->n, p->order, p->dx, tab, s);
        }
        a.wtab = tab;
    }
    return NLS_OK;
}

template <typename T>
static int rk4_advance(const nls_problem *p, int iters, const T *P, cplx_t<T> *psi, T *nres, char *ws,
                       hipStream_t s, const OperatorGiven *og = nullptr) {
    using C = cplx_t<T>;
    const size_t N = points_of(p);
    const bool dyn = p->coeffs[10] != 0.0;
    size_t off_tab = 0;
    rk4_workspace(p, &off_tab);
    const size_t cb = align256(N * sizeof(C)), rb = align256(N * sizeof(T));
    C *kA = (C *)ws, *kB = (C *)(ws + cb), *acc = (C *)(ws + 2 * cb);
    T *kAn = nullptr, *kBn = nullptr, *accn = nullptr;
    if (dyn) {
        kAn = (T *)(ws + 3 * cb);
        kBn = (T *)(ws + 3 * cb + rb);
        accn = (T *)(ws + 3 * cb + 2 * rb);
    }
    StageArgs<T> a;
    memset(&a, 0, sizeof a);
    a.n = p->n;
    a.dt = (T)p->dt;
    a.c = make_coeffs<T>(p->coeffs);
    a.P = P;
    int rc = set_operator<T>(p, a, (T *)(ws + off_tab), og, s);
    if (rc) return rc;
    C *cur = psi, *alt = acc;
    T *curn = nres, *altn = accn;
    for (int it = 0; it < iters; ++it) {
        a.u = cur; a.un = curn; a.acc = alt; a.acc_n = altn;
        a.stage = 1; a.scale = 0; a.kin = nullptr; a.kin_n = nullptr; a.kout = kA; a.kout_n = kAn;
        launch_rk4_stage<T>(p->dim, p->order, p->bc, dyn, a, s);
        a.stage = 2; a.scale = (T)0.5; a.kin = kA; a.kin_n = kAn; a.kout = kB; a.kout_n = kBn;
        launch_rk4_stage<T>(p->dim, p->order, p->bc, dyn, a, s);
        a.stage = 3; a.scale = (T)0.5; a.kin = kB; a.kin_n = kBn; a.kout = kA; a.kout_n = kAn;
        launch_rk4_stage<T>(p->dim, p->order, p->bc, dyn, a, s);
        a.stage = 4; a.scale = (T)1; a.kin = kA; a.kin_n = kAn; a.kout = nullptr; a.kout_n = nullptr;
        launch_rk4_stage<T>(p->dim, p->order, p->bc, dyn, a, s);
        std::swap(cur, alt);
        std::swap(curn, altn);
    }
    rc = hip_check(hipGetLastError(), "RK4 stage launch");
    if (rc) return rc;
    if (cur != psi) {
        rc = hip_check(hipMemcpyAsync(psi, cur, N * sizeof(C), hipMemcpyDeviceToDevice, s), "copy state");
        if (rc) return rc;
        if (dyn) rc = hip_check(hipMemcpyAsync(nres, curn, N * sizeof(T), hipMemcpyDeviceToDevice, s), "copy reservoir");
    }
    return rc;
}

template <typename T>
static int rhs_once(const nls_problem *p, const T *P, const cplx_t<T> *psi, const T *nres, cplx_t<T> *dpsi,
                    T *dn, char *ws, hipStream_t s, const OperatorGiven *og = nullptr) {
    const bool dyn = p->coeffs[10] != 0.0;
    StageArgs<T> a;
    memset(&a, 0, sizeof a);
    a.n = p->n;
    a.dt = (T)p->dt;
    a.c = make_coeffs<T>(p->coeffs);
    a.P = P;
    int rc = set_operator<T>(p, a, (T *)ws, og, s);
    if (rc) return rc;
    a.stage = 1; a.u = psi; a.un = nres; a.kout = dpsi; a.kout_n = dn;
    launch_rk4_stage<T>(p->dim, p->order, p->bc, dyn, a, s);
    return hip_check(hipGetLastError(), "RHS launch");
}

// Device scratch owned for the duration of one host-memory call.
struct DevBuf {
    void *ptr = nullptr;
    ~DevBuf() { if (ptr) (void)hipFree(ptr); }
    int alloc(size_t bytes, const char *what) {
        return hip_check(hipMalloc(&ptr, bytes ? bytes : 1), what);
    }
};

}  // namespace nls

using namespace nls;

extern "C" {

size_t nls_workspace_bytes(const nls_problem *p) {
    if (check_problem(p)) return 0;
    if (p->scheme == NLS_SCHEME_SSF) return ssf_workspace_bytes(p);
    switch (plain_engine(p)) {
    case NLS_ENGINE_FUSED: return fused_grid_workspace_bytes(p);
    case NLS_ENGINE_RESIDENT: return resident_workspace_bytes(p);
    default: return rk4_workspace(p, nullptr);
    }
}

// Enqueues iters steps as plain launches.
static int advance_direct(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *ws,
                          hipStream_t s) {
    if (p->scheme == NLS_SCHEME_SSF) return ssf_advance(p, iters, pumping, psi, nres, ws, s);
    const int engine = plain_engine(p);
    if (engine == NLS_ENGINE_FUSED) return fused_grid_advance(p, iters, pumping, psi, nres, ws, s);
    if (engine == NLS_ENGINE_RESIDENT) return resident_advance(p, iters, pumping, psi, nres, ws, s);
    if (p->dtype == NLS_F64)
        return rk4_advance<double>(p, iters, (const double *)pumping, (double2 *)psi, (double *)nres, (char *)ws, s);
    return rk4_advance<float>(p, iters, (const float *)pumping, (float2 *)psi, (float *)nres, (char *)ws, s);
}

// ---- step-block graphs -------------------------------------------------------
// A block of GRAPH_BLOCK steps (even, so the state ping-pong ends where it
// started and the block holds kernel nodes only) is captured once on a
// private stream of the calling thread and replayed as one graph launch.
// Entries are keyed by device, problem and buffer addresses; the cache holds
// GRAPH_CACHE entries per thread (least recently used evicted).  Capture
// streams have the default priority.
constexpr int GRAPH_BLOCK = 64;
constexpr size_t GRAPH_CACHE = 8;

struct GraphEntry {
    int device;
    int stage_ty;          // tile height of the per-stage kernels when captured
    int fform;             // fused kernel's tile form setting when captured
    nls_problem p;
    const void *P;
    void *psi, *nres, *ws;
    hipGraphExec_t exec;
    unsigned long long used;
};

struct GraphCache {
    std::vector<GraphEntry> entries;
    std::vector<hipStream_t> capture;   // per device, created on first use
    unsigned long long clock = 0;
};

// Never destroyed at thread exit (the HIP runtime may be gone by then);
// nls_release() frees it explicitly.
static thread_local GraphCache *g_graphs = nullptr;

static bool graphs_apply(const nls_problem *p, int iters) {
    if ((p->flags & NLS_FLAG_EAGER) || iters < 2 * GRAPH_BLOCK || p->scheme != NLS_SCHEME_RK4) return false;
    const int engine = plain_engine(p);
    return engine == NLS_ENGINE_FUSED || engine == NLS_ENGINE_STAGE;
}

static int block_graph(const nls_problem *p, const void *P, void *psi, void *nres, void *ws, hipGraphExec_t *out) {
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "get device");
    if (rc) return rc;
    if (!g_graphs) g_graphs = new GraphCache;
    GraphCache &gc = *g_graphs;
    // a graph holds the kernels of the launch configuration it was captured
    // with: the per-stage tile height (nls_tune_stage_tile) is part of the key
    const int ty = get_stage_tile(p->dtype),
              ff = get_fused_form((p->flags & NLS_FLAG_F64_ARITH) ? NLS_F64 : p->dtype);
    for (GraphEntry &e : gc.entries)
        if (e.device == dev && e.stage_ty == ty && e.fform == ff && e.P == P && e.psi == psi && e.nres == nres &&
            e.ws == ws &&
            memcmp(&e.p, p, sizeof *p) == 0) {
            e.used = ++gc.clock;
            *out = e.exec;
            return NLS_OK;
        }
    if ((int)gc.capture.size() <= dev) gc.capture.resize(dev + 1, nullptr);
    if (!gc.capture[dev]) {
        rc = hip_check(hipStreamCreateWithFlags(&gc.capture[dev], hipStreamNonBlocking), "capture stream");
        if (rc) return rc;
    }
    hipStream_t cs = gc.capture[dev];
    rc = hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    rc = advance_direct(p, GRAPH_BLOCK, P, psi, nres, ws, cs);
    hipGraph_t g = nullptr;
    const int rc_end = hip_check(hipStreamEndCapture(cs, &g), "end capture");
    if (rc || rc_end) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : rc_end;
    }
    hipGraphExec_t exec = nullptr;
    rc = hip_check(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "instantiate graph");
    (void)hipGraphDestroy(g);
    if (rc) return rc;
    if (gc.entries.size() >= GRAPH_CACHE) {
        // evict the least recently used entry; an earlier launch of it may
        // still be in flight on any stream
        size_t v = 0;
        for (size_t i = 1; i < gc.entries.size(); ++i)
            if (gc.entries[i].used < gc.entries[v].used) v = i;
        rc = hip_check(hipDeviceSynchronize(), "evict graph");
        if (rc) { (void)hipGraphExecDestroy(exec); return rc; }
        (void)hipGraphExecDestroy(gc.entries[v].exec);
        gc.entries.erase(gc.entries.begin() + v);
    }
    GraphEntry e;
    e.device = dev;
    e.stage_ty = ty;
    e.fform = ff;
    e.p = *p;
    e.P = P; e.psi = psi; e.nres = nres; e.ws = ws;
    e.exec = exec;
    e.used = ++gc.clock;
    gc.entries.push_back(e);
    *out = exec;
    return NLS_OK;
}

int nls_release(void) {
    if (!g_graphs) return NLS_OK;
    int rc = hip_check(hipDeviceSynchronize(), "release");
    for (GraphEntry &e : g_graphs->entries) (void)hipGraphExecDestroy(e.exec);
    for (hipStream_t cs : g_graphs->capture)
        if (cs) (void)hipStreamDestroy(cs);
    delete g_graphs;
    g_graphs = nullptr;
    return rc;
}

int nls_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *workspace,
                size_t workspace_bytes, void *stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (iters < 0) return fail(NLS_EINVAL, "iters must be >= 0");
    const bool dyn = p->coeffs[10] != 0.0;
    if (!pumping || !psi || (dyn && !nres)) return fail(NLS_EINVAL, "NULL device buffer");
    const size_t need = nls_workspace_bytes(p);
    if (workspace_bytes < need || (need && !workspace))
        return fail(NLS_EINVAL, "workspace of %zu bytes is smaller than the %zu required", workspace_bytes, need);
    if (iters == 0) return NLS_OK;
    hipStream_t s = (hipStream_t)stream;
    if (!graphs_apply(p, iters)) return advance_direct(p, iters, pumping, psi, nres, workspace, s);
    hipGraphExec_t exec = nullptr;
    rc = block_graph(p, pumping, psi, dyn ? nres : nullptr, workspace, &exec);
    if (rc) return rc;
    for (int b = 0; b < iters / GRAPH_BLOCK; ++b) {
        rc = hip_check(hipGraphLaunch(exec, s), "graph launch");
        if (rc) return rc;
    }
    const int rest = iters % GRAPH_BLOCK;
    return rest ? advance_direct(p, rest, pumping, psi, nres, workspace, s) : NLS_OK;
}

}  // extern "C"

// ---- advance without allocation ------------------------------------------
// The workspace of nls_advance_static: a file-scope device array, one copy on
// every GPU the library runs on, held until the process ends.  Host code never
// passes the array's name as an address: its address on the current device
// comes from hipGetSymbolAddress, taken once per device (static_workspace).
// Every call rewrites what it later reads of the array (state ping-pong,
// stage buffers, member descriptor, radial and FFT tables) before reading it,
// so nothing needs clearing; calls that overlap on one GPU share the array
// and race.
static __device__ __align__(256) unsigned char scratch[NLS_STATIC_WORKSPACE_BYTES];

static int static_workspace(void **out) {
    static std::mutex mu;
    static void *addr[64];                 // per device
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "current device");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) return fail(NLS_EUNSUPPORTED, "device %d: at most 64 devices", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *a = nullptr;
        if ((rc = hip_check(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)), "static workspace address"))) return rc;
        addr[dev] = a;
    }
    *out = addr[dev];
    return NLS_OK;
}

extern "C" {

int nls_advance_static(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (iters < 0) return fail(NLS_EINVAL, "iters must be >= 0");
    const bool dyn = p->coeffs[10] != 0.0;
    if (!pumping || !psi || (dyn && !nres)) return fail(NLS_EINVAL, "NULL device buffer");
    const size_t need = nls_workspace_bytes(p);
    if (need > (size_t)NLS_STATIC_WORKSPACE_BYTES)
        return fail(NLS_EINVAL, "this problem needs a workspace of %zu bytes; nls_advance_static holds at most %zu "
                    "(nls_advance takes a workspace of the caller's)", need, (size_t)NLS_STATIC_WORKSPACE_BYTES);
    if (iters == 0) return NLS_OK;
    void *ws = nullptr;
    if ((rc = static_workspace(&ws))) return rc;
    // direct launches only: no step-block graph is captured or cached
    return advance_direct(p, iters, pumping, psi, nres, ws, (hipStream_t)stream);
}

int nls_slab_halo_rows(int order) {
    if (order != 3 && order != 5 && order != 7) return fail(NLS_EINVAL, "order must be 3, 5 or 7 (got %d)", order);
    return 2 * (order - 1);
}

int nls_slab_step(const nls_problem *p, int rows, const void *pumping, const void *psi_in, const void *nres_in,
                  void *psi_out, void *nres_out, void *stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (p->dim != 2 || p->bc != NLS_BC_PERIODIC || p->scheme != NLS_SCHEME_RK4)
        return fail(NLS_EUNSUPPORTED, "row slabs are 2-D, periodic and RK4");
    const int h = 2 * (p->order - 1);
    if (rows < h) return fail(NLS_EINVAL, "rows=%d is smaller than the %d ghost rows of a side", rows, h);
    if ((long long)(rows + 2 * h) * p->n >= (1ll << 31))
        return fail(NLS_EINVAL, "slab of %d rows x %d points too large", rows, p->n);
    const bool dyn = p->coeffs[10] != 0.0;
    if (!pumping || !psi_in || !psi_out || (dyn && (!nres_in || !nres_out)))
        return fail(NLS_EINVAL, "NULL device buffer");
    if (psi_in == psi_out || (dyn && nres_in == nres_out))
        return fail(NLS_EINVAL, "input and output slabs must be different buffers");
    return fused_slab_step(p, rows, pumping, psi_in, dyn ? nres_in : nullptr, psi_out, dyn ? nres_out : nullptr,
                           (hipStream_t)stream);
}

int nls_rhs(const nls_problem *p, const void *pumping, const void *psi, const void *nres, void *dpsi, void *dn,
            void *stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    const bool dyn = p->coeffs[10] != 0.0;
    if (!pumping || !psi || !dpsi || (dyn && (!nres || !dn))) return fail(NLS_EINVAL, "NULL device buffer");
    hipStream_t s = (hipStream_t)stream;
    DevBuf tab;
    if (p->dim == 1 && p->bc == NLS_BC_REF) {
        rc = tab.alloc((size_t)p->n * p->order * (p->dtype == NLS_F64 ? 8 : 4), "radial table");
        if (rc) return rc;
    }
    if (p->dtype == NLS_F64)
        rc = rhs_once<double>(p, (const double *)pumping, (const double2 *)psi, (const double *)nres,
                              (double2 *)dpsi, (double *)dn, (char *)tab.ptr, s);
    else
        rc = rhs_once<float>(p, (const float *)pumping, (const float2 *)psi, (const float *)nres, (float2 *)dpsi,
                             (float *)dn, (char *)tab.ptr, s);
    if (rc) return rc;
    // the table must outlive the kernel that reads it
    return hip_check(hipStreamSynchronize(s), "RHS sync");
}

// Host-memory solve; `check` scans the downloaded state for NaN/Inf.
static int solve_host(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, bool check) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (iters < 0) return fail(NLS_EINVAL, "iters must be >= 0");
    const bool dyn = p->coeffs[10] != 0.0;
    if (!pumping || !psi || (dyn && !nres)) return fail(NLS_EINVAL, "NULL host buffer");
    const size_t N = points_of(p), rs = p->dtype == NLS_F64 ? 8 : 4;
    DevBuf dP, dpsi, dn, ws;
    const size_t wsb = nls_workspace_bytes(p);
    if ((rc = dP.alloc(N * rs, "pumping")) || (rc = dpsi.alloc(N * 2 * rs, "psi")) ||
        (rc = ws.alloc(wsb, "workspace")))
        return rc;
    if (dyn && (rc = dn.alloc(N * rs, "reservoir"))) return rc;
    if ((rc = hip_check(hipMemcpy(dP.ptr, pumping, N * rs, hipMemcpyHostToDevice), "upload pumping")) ||
        (rc = hip_check(hipMemcpy(dpsi.ptr, psi, N * 2 * rs, hipMemcpyHostToDevice), "upload psi")))
        return rc;
    if (dyn && (rc = hip_check(hipMemcpy(dn.ptr, nres, N * rs, hipMemcpyHostToDevice), "upload reservoir")))
        return rc;
    rc = nls_advance(p, iters, dP.ptr, dpsi.ptr, dn.ptr, ws.ptr, wsb, nullptr);
    if (rc) return rc;
    if ((rc = hip_check(hipDeviceSynchronize(), "solve")) ||
        (rc = hip_check(hipMemcpy(psi, dpsi.ptr, N * 2 * rs, hipMemcpyDeviceToHost), "download psi")))
        return rc;
    if (dyn && (rc = hip_check(hipMemcpy(nres, dn.ptr, N * rs, hipMemcpyDeviceToHost), "download reservoir")))
        return rc;
    if (check) {
        size_t bad = 0;
        if (p->dtype == NLS_F64) {
            const double *a = (const double *)psi, *b = (const double *)nres;
            for (size_t i = 0; i < 2 * N; ++i) bad += !std::isfinite(a[i]);
            for (size_t i = 0; dyn && i < N; ++i) bad += !std::isfinite(b[i]);
        } else {
            const float *a = (const float *)psi, *b = (const float *)nres;
            for (size_t i = 0; i < 2 * N; ++i) bad += !std::isfinite(a[i]);
            for (size_t i = 0; dyn && i < N; ++i) bad += !std::isfinite(b[i]);
        }
        if (bad)
            return fail(NLS_ENONFINITE, "the solution holds %zu non-finite values after %d steps (dt = %g above the "
                        "scheme's stability limit?); the state was written back", bad, iters, p->dt);
    }
    return NLS_OK;
}

int nls_solve_host(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres) {
    return solve_host(p, iters, pumping, psi, nres, true);
}

// ---- drop-ins of the reference's Fortran procedures ----------------------

static int dropin_solve(int dim, float dt, float dx, int n, int order, int iters, const float *pumping,
                        const float *coeffs, const void *u0, void *u) {
    if (!pumping || !coeffs || !u0 || !u) return fail(NLS_EINVAL, "NULL argument");
    nls_problem p;
    memset(&p, 0, sizeof p);
    p.dim = dim; p.n = n; p.order = order; p.bc = NLS_BC_REF; p.dtype = NLS_F32; p.scheme = NLS_SCHEME_RK4;
    p.engine = NLS_ENGINE_AUTO;
    p.dt = dt; p.dx = dx;
    for (int k = 0; k < 23; ++k) p.coeffs[k] = coeffs[k];
    p.coeffs[10] = 0.0;       // the reference never reads c11: its reservoir is algebraic (nls.f90:570-581)
    int rc = check_problem(&p);
    if (rc) return rc;
    const size_t N = points_of(&p);
    if (u != u0) memcpy(u, u0, N * 2 * sizeof(float));
    // the reference reports nothing when a solve blows up (nls.f90:725-733):
    // the drop-ins return the state as it is
    return solve_host(&p, iters, pumping, u, nullptr, false);
}

int nls_solve_nls(float dt, float dx, int n, int order, int iters, const float *pumping, const float *coeffs,
                  const void *u0, void *u) {
    return dropin_solve(1, dt, dx, n, order, iters, pumping, coeffs, u0, u);
}

int nls_solve_nls_1d(float dt, float dx, int n, int order, int iters, const float *pumping, const float *coeffs,
                     const void *u0, void *u) {
    return dropin_solve(1, dt, dx, n, order, iters, pumping, coeffs, u0, u);
}

int nls_solve_nls_2d(float dt, float dx, int n, int order, int iters, const float *pumping, const float *coeffs,
                     const void *u0, void *u) {
    return dropin_solve(2, dt, dx, n, order, iters, pumping, coeffs, u0, u);
}

// ---- the reference's runge_kutta / hamiltonian family -------------------
// Same right-hand side and RK4 as solve_nls, on an operator the caller gives
// as the reference's arrays: a BLAS band op(2 klu + 1, n) in 1-D, or the
// block-banded (blocks(n, 2m - 1), orders(m)) in 2-D.  Single precision,
// host memory, blocking; the arithmetic runs on the per-stage kernels.

// Row table of the 1-D kernels from a band: row i holds A(i, i - klu + k),
// stored at op((klu + i - j) + 1, j + 1) (BLAS band layout, nls.f90:540).
static void band_rows(int n, int klu, const float *op, std::vector<double> &tab) {
    const int m = 2 * klu + 1;
    tab.assign((size_t)n * m, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < m; ++k) {
            const int j = i + k - klu;
            if (j >= 0 && j < n) tab[(size_t)i * m + k] = op[(size_t)j * m + (m - 1 - k)];
        }
}

// Stencil weights w0 .. wR of a 2-D block-banded operator, provided it is
// one symmetric constant-coefficient Cartesian stencil with zero outside the
// grid (what make_laplacian_2d builds, nls.f90:297-403): the kernels apply
// such a stencil; other block operators are rejected.
static int block_stencil(int n, int m, const float *blocks, const int *orders, double w[4]) {
    const int R = (m - 1) / 2;
    if (n < m) return fail(NLS_EINVAL, "n=%d is smaller than the stencil width %d", n, m);
    for (int q = 0; q < m; ++q)
        if (orders[q] != (q == R ? R : 0))
            return fail(NLS_EUNSUPPORTED, "orders must be those of make_laplacian_2d (block %d has %d)", q, orders[q]);
    const float *mid = blocks + (size_t)R * n;       // band (m, n) of the diagonal block
    std::vector<float> c(m);
    for (int b = 0; b < m; ++b) c[b] = mid[(size_t)R * m + b];
    const char *msg = "the block operator is not one symmetric constant-coefficient stencil";
    for (int j = 0; j < n; ++j)
        for (int b = 0; b < m; ++b) {
            const int i = j + b - R;
            if (i >= 0 && i < n && mid[(size_t)j * m + b] != c[b]) return fail(NLS_EUNSUPPORTED, "%s", msg);
        }
    w[0] = w[1] = w[2] = w[3] = 0.0;
    for (int d = 1; d <= R; ++d) {
        const float *left = blocks + (size_t)(R - d) * n, *right = blocks + (size_t)(R + m - 1 + d) * n;
        if (c[R - d] != c[R + d]) return fail(NLS_EUNSUPPORTED, "%s", msg);
        for (int j = 0; j < n; ++j)
            if (left[j] != c[R - d] || right[j] != c[R - d]) return fail(NLS_EUNSUPPORTED, "%s", msg);
        w[d] = c[R - d];
    }
    w[0] = 0.5 * (double)c[R];      // the kernels take the 1-D centre weight; the 2-D centre is twice it
    return NLS_OK;
}

static int operator_dropin(int dim, int n, int order, float dt, int iters, const float *pumping,
                           const float *coeffs, const OperatorGiven &og, const void *u0, void *u, bool rhs_only) {
    if (!pumping || !coeffs || !u0 || !u) return fail(NLS_EINVAL, "NULL argument");
    if (iters < 0) return fail(NLS_EINVAL, "iters must be >= 0");
    nls_problem p;
    memset(&p, 0, sizeof p);
    p.dim = dim; p.n = n; p.order = order; p.bc = NLS_BC_REF; p.dtype = NLS_F32; p.scheme = NLS_SCHEME_RK4;
    p.engine = NLS_ENGINE_STAGE;
    p.dt = dt; p.dx = 1.0;              // the operator is given; dx is not used
    for (int k = 0; k < 23; ++k) p.coeffs[k] = coeffs[k];
    p.coeffs[10] = 0.0;                 // algebraic reservoir, as the reference's hamiltonian
    int rc = check_problem(&p);
    if (rc) return rc;
    const size_t N = points_of(&p);
    DevBuf dP, du, dv, ws;
    if ((rc = dP.alloc(N * 4, "pumping")) || (rc = du.alloc(N * 8, "psi")) ||
        (rc = ws.alloc(rk4_workspace(&p, nullptr), "workspace")) || (rhs_only && (rc = dv.alloc(N * 8, "result"))))
        return rc;
    if ((rc = hip_check(hipMemcpy(dP.ptr, pumping, N * 4, hipMemcpyHostToDevice), "upload pumping")) ||
        (rc = hip_check(hipMemcpy(du.ptr, u0, N * 8, hipMemcpyHostToDevice), "upload psi")))
        return rc;
    if (rhs_only)
        rc = rhs_once<float>(&p, (const float *)dP.ptr, (const float2 *)du.ptr, nullptr, (float2 *)dv.ptr, nullptr,
                             (char *)ws.ptr, nullptr, &og);
    else
        rc = rk4_advance<float>(&p, iters, (const float *)dP.ptr, (float2 *)du.ptr, nullptr, (char *)ws.ptr, nullptr,
                                &og);
    if (rc) return rc;
    if ((rc = hip_check(hipDeviceSynchronize(), "solve"))) return rc;
    return hip_check(hipMemcpy(u, rhs_only ? dv.ptr : du.ptr, N * 8, hipMemcpyDeviceToHost), "download");
}

int nls_hamiltonian(const float *pumping, const float *coeffs, const void *u, void *v, const float *op, int klu,
                    int n) {
    if (!op) return fail(NLS_EINVAL, "NULL argument");
    if (klu < 1 || klu > 3) return fail(NLS_EUNSUPPORTED, "klu must be 1, 2 or 3 (got %d)", klu);
    if (n < 2 * klu + 1) return fail(NLS_EINVAL, "n=%d is smaller than the band width %d", n, 2 * klu + 1);
    std::vector<double> tab;
    band_rows(n, klu, op, tab);
    OperatorGiven og;
    og.tab = tab.data();
    return operator_dropin(1, n, 2 * klu + 1, 0.0f, 0, pumping, coeffs, og, u, v, true);
}

int nls_runge_kutta(float dt, float t0, const void *u0, const float *op, int n, int order, int iters, void *u,
                    const float *pumping, const float *coeffs) {
    (void)t0;                           // advanced but never read by the reference (nls.f90:719-732)
    if (!op) return fail(NLS_EINVAL, "NULL argument");
    if (order != 3 && order != 5 && order != 7) return fail(NLS_EINVAL, "order must be 3, 5 or 7 (got %d)", order);
    if (n < order) return fail(NLS_EINVAL, "n=%d is smaller than the band width %d", n, order);
    std::vector<double> tab;
    band_rows(n, (order - 1) / 2, op, tab);
    OperatorGiven og;
    og.tab = tab.data();
    return operator_dropin(1, n, order, dt, iters, pumping, coeffs, og, u0, u, false);
}

int nls_hamiltonian_2d(const float *pumping, const float *coeffs, const void *u, void *v, const float *blocks,
                       const int *orders, int order, int n) {
    if (!blocks || !orders) return fail(NLS_EINVAL, "NULL argument");
    if (order != 3 && order != 5 && order != 7) return fail(NLS_EINVAL, "order must be 3, 5 or 7 (got %d)", order);
    double w[4];
    int rc = block_stencil(n, order, blocks, orders, w);
    if (rc) return rc;
    OperatorGiven og;
    og.w = w;
    return operator_dropin(2, n, order, 0.0f, 0, pumping, coeffs, og, u, v, true);
}

int nls_runge_kutta_2d(float dt, float t0, const void *u0, int n, const float *blocks, const int *orders, int order,
                       int iters, void *u, const float *pumping, const float *coeffs) {
    (void)t0;
    if (!blocks || !orders) return fail(NLS_EINVAL, "NULL argument");
    if (order != 3 && order != 5 && order != 7) return fail(NLS_EINVAL, "order must be 3, 5 or 7 (got %d)", order);
    double w[4];
    int rc = block_stencil(n, order, blocks, orders, w);
    if (rc) return rc;
    OperatorGiven og;
    og.w = w;
    return operator_dropin(2, n, order, dt, iters, pumping, coeffs, og, u0, u, false);
}

int nls_tune_fused_form(int dtype, int form) {
    if (dtype != NLS_F32 && dtype != NLS_F64) return fail(NLS_EINVAL, "bad dtype");
    if (form == -1) return get_fused_form(dtype) + 1;     // internal -1 (per launch) .. 3 -> 0 .. 4
    if (form < 0 || form > 4 || set_fused_form(dtype, form - 1))
        return fail(NLS_EINVAL, "form must be 0 (per launch), 1 (tall), 2 (short), 3 (half) or 4 (tall, 512 "
                    "threads) (got %d)", form);
    return form;
}

int nls_tune_stage_tile(int dtype, int rows) {
    if (dtype != NLS_F32 && dtype != NLS_F64) return fail(NLS_EINVAL, "bad dtype");
    if (rows == 0) return get_stage_tile(dtype);
    if (set_stage_tile(dtype, rows)) return fail(NLS_EINVAL, "tile height must be 4, 8, 16 or 32 (got %d)", rows);
    return rows;
}

}  // extern "C"
