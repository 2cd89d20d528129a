/*
 * nls_amd.h -- C ABI of libnls_amd.so, the gfx950 HIP implementation of the
 * time-stepping core of nls (Solver1D/Solver2D.solve -> solve_nls /
 * solve_nls_2d).
 *
 * Every entry point returns an int status: NLS_OK (0) or a negative NLS_E*
 * code; nls_last_error() gives a message for the calling thread.  No
 * PyTorch or other framework types appear here: device buffers are plain
 * pointers, streams are hipStream_t passed as void* (NULL = default stream).
 *
 * Data conventions.
 *   complex arrays: interleaved (re, im) pairs -- complex64 = 2 x float,
 *     complex128 = 2 x double;
 *   2-D fields: n*n values, flat index row*n + col with col contiguous.  This
 *     is byte-identical to the reference's column-major u(n, n) (first index
 *     contiguous), so arrays pass straight through;
 *   coeffs: the reference's 23-entry coefficient vector (Fortran c(k) is
 *     coeffs[k-1]); c3, c4, c5, c6, c11, c12, c13, c14 are read.  c11 == 0
 *     selects the reference's algebraic (adiabatic) reservoir
 *     r = c12 P / (c13 + c14 |psi|^2); c11 != 0 integrates the reservoir
 *     density n as a second field, c11 dn/dt = c12 P - (c13 + c14 |psi|^2) n.
 */
#ifndef NLS_AMD_H
#define NLS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define NLS_OK            0
#define NLS_EINVAL       -1   /* bad argument (shape, order, null pointer)  */
#define NLS_EUNSUPPORTED -2   /* valid but unsupported mode combination     */
#define NLS_EHIP         -3   /* HIP runtime error                          */
#define NLS_ENONFINITE   -4   /* solution became NaN/Inf                    */
#define NLS_ENOMEM       -5   /* device allocation failed                   */

/* ---- enums ------------------------------------------------------------- */
#define NLS_BC_REF       0    /* reference boundaries: 1-D radial r = k*dx with
                                 a mirror at r = 0 and zero beyond the grid;
                                 2-D zero outside the grid                  */
#define NLS_BC_PERIODIC  1    /* Cartesian stencils with wrap-around        */
#define NLS_F32          0    /* complex64 state, float arithmetic          */
#define NLS_F64          1    /* complex128 state, double arithmetic        */
#define NLS_SCHEME_RK4   0    /* classical RK4 over the FD Laplacian        */
#define NLS_SCHEME_SSF   1    /* Strang split-step Fourier (periodic only)  */
/* nls_problem.flags */
#define NLS_FLAG_EAGER    1   /* issue every launch directly: no step-block
                                 graphs (nls_advance)                        */
#define NLS_FLAG_F64_ARITH 2  /* single-precision fields (dtype NLS_F32),
                                 double-precision RK4 arithmetic: the fused
                                 engine's double-precision step on complex64 /
                                 float arrays, rounded once per step on store
                                 (2-D periodic RK4 only; not in batches)     */
/* RK4 engines (nls_problem.engine, nls_batch_create) */
#define NLS_ENGINE_AUTO  -1   /* RESIDENT if the instance (every member of a
                                 batch) fits, else FUSED for 2-D and STAGE
                                 for 1-D; a batch of which only some members
                                 fit: HYBRID                                 */
#define NLS_ENGINE_STAGE  0   /* one launch per RK4 stage                    */
#define NLS_ENGINE_FUSED  1   /* one launch per RK4 step, halo recomputed on
                                 chip (2-D only)                             */
#define NLS_ENGINE_RESIDENT 2 /* a workgroup per instance holds its grid on
                                 chip and runs a block of steps per launch
                                 (1-D n <= 4096, 2-D n <= 64; not FP64 2-D
                                 order 7)                                    */
#define NLS_ENGINE_HYBRID 3   /* batches only: the members that fit on the
                                 RESIDENT engine run there, on a stream of
                                 the batch forked from and joined to the
                                 caller's, beside the others on FUSED (2-D)
                                 or STAGE (1-D); no counterpart in the
                                 reference, which solves one instance per
                                 process                                     */

const char *nls_last_error(void);

/* version (nls.f90:29-37): 0.2.0, the reference module version. */
int nls_version(int *major, int *minor, int *patch);

/* ======================================================================
 * Drop-in host-memory entry points.  Same argument list and meaning as the
 * reference's Fortran procedures (scalars by value instead of by
 * reference); inputs and outputs are host arrays, the work runs on the
 * current HIP device.  Single precision like the reference.
 * ====================================================================== */

/* solve_nls (nls/nls.f90:797-813) and solve_nls_1d (:815-827): 1-D
 * axisymmetric solve, order 3/5/7, iters RK4 steps of dt on r_k = k*dx.
 * pumping[n] float, coeffs[23] float, u0[n] and u[n] complex64.  Like the
 * reference, the reservoir is algebraic: coeffs[10] (c11) is not read. */
int nls_solve_nls(float dt, float dx, int n, int order, int iters, const float *pumping,
                  const float *coeffs, const void *u0, void *u);
int nls_solve_nls_1d(float dt, float dx, int n, int order, int iters, const float *pumping,
                     const float *coeffs, const void *u0, void *u);

/* solve_nls_2d (nls/nls.f90:903-919): n x n grid, zero outside. */
int nls_solve_nls_2d(float dt, float dx, int n, int order, int iters, const float *pumping,
                     const float *coeffs, const void *u0, void *u);

/* chemical_potential_1d (nls.f90:921-948): mu is complex64 (2 floats). */
int nls_chemical_potential_1d(float dx, int n, const float *pumping, const float *coeffs,
                              const void *u0, void *mu);
/* chemical_potential_2d (nls.f90:950-971): real part only, like the reference. */
int nls_chemical_potential_2d(float dx, int n, const float *pumping, const float *coeffs,
                              const void *u0, float *mu);

/* make_laplacian (nls.f90:278-295): the 1-D radial operator in BLAS band
 * storage, Fortran shape (order, n) -> op[(j-1)*order + (p-1)].  Built in
 * double precision on the host. */
int nls_make_laplacian(int n, int order, double h, double *op);

/* The reference module's other public procedures (nls.f90:13-24).  Fortran
 * arrays are column-major: a(p, j) at a[(j-1)*rows + (p-1)].
 *
 * Operator builders and BLAS-2 wrappers, host code as in the reference:
 *   make_banded_matrix (:61-91): mat(m, n) from row(m);
 *   make_laplacian_2d (:297-403): blocks(n, 2m-1) and orders(m) of the 2-D
 *     Cartesian operator (the reference declares order 7 blocks(m, 2m-1),
 *     valid only at n = 7; here (n, 2m-1) at every n);
 *   rgbmv (:530-541): u += sign A x, A = op(2 klu + 1, n) in band storage;
 *   rbbmv (:408-527): y += sign B x for the block-banded B (blocks, ms);
 *   revervoir / revervoir_2d (:570-581, :829-839):
 *     r = c12 P / (c13 + c14 u_sqr) on n or n*n points. */
int nls_make_banded_matrix(int n, int m, const float *row, float *mat);
/* clear_first_row_of_derivative (:93-107), divide_derivative_on_radius
 * (:109-130): in place on a first-derivative band L1(m, n) -- the steps
 * make_laplacian_o3/o5/o7 take: zero the entries of matrix row 0, divide the
 * entries of matrix row k > 0 by k h. */
int nls_clear_first_row_of_derivative(int n, int m, float *L1);
int nls_divide_derivative_on_radius(int n, int m, float h, float *L1);
int nls_make_laplacian_2d(int n, int m, float h, float *blocks, int *orders);
int nls_rgbmv(const float *x, float *u, float sign, const float *op, int klu, int n);
int nls_rbbmv(const float *x, float *y, float sign, const float *blocks, const int *ms, int m, int n);
int nls_revervoir(const float *pumping, const float *coeffs, const float *u_sqr, float *r, int n);
int nls_revervoir_2d(const float *pumping, const float *coeffs, const float *u_sqr, float *r, int n);

/* Right-hand side and RK4 on a caller-given operator, run on the GPU (the
 * per-stage kernels), host arrays, algebraic reservoir:
 *   hamiltonian (:621-650): v = F(u) with the band op(2 klu + 1, n),
 *     klu = 1, 2 or 3 (any band of that width);
 *   runge_kutta (:705-734): iters RK4 steps of dt from u0 with op(order, n);
 *     t0 is accepted and, as in the reference, unused;
 *   hamiltonian_2d (:841-870), runge_kutta_2d (:873-901): the same on n x n
 *     with (blocks, orders); the blocks must hold one symmetric
 *     constant-coefficient Cartesian stencil with zero outside the grid, as
 *     make_laplacian_2d builds (the kernels apply such stencils); other
 *     block operators give NLS_EUNSUPPORTED. */
int nls_hamiltonian(const float *pumping, const float *coeffs, const void *u, void *v, const float *op,
                    int klu, int n);
int nls_runge_kutta(float dt, float t0, const void *u0, const float *op, int n, int order, int iters,
                    void *u, const float *pumping, const float *coeffs);
int nls_hamiltonian_2d(const float *pumping, const float *coeffs, const void *u, void *v,
                       const float *blocks, const int *orders, int order, int n);
int nls_runge_kutta_2d(float dt, float t0, const void *u0, int n, const float *blocks, const int *orders,
                       int order, int iters, void *u, const float *pumping, const float *coeffs);

/* ======================================================================
 * Extended entry points: any precision, boundary family, reservoir mode
 * and scheme; host or device buffers.
 * ====================================================================== */

typedef struct nls_problem {
    int32_t dim;        /* 1 or 2                                          */
    int32_t n;          /* points per axis                                 */
    int32_t order;      /* 3, 5 or 7 (stencil width)                       */
    int32_t bc;         /* NLS_BC_REF / NLS_BC_PERIODIC                    */
    int32_t dtype;      /* NLS_F32 / NLS_F64                               */
    int32_t scheme;     /* NLS_SCHEME_RK4 / NLS_SCHEME_SSF                 */
    int32_t engine;     /* RK4: NLS_ENGINE_* (AUTO picks by size)          */
    int32_t flags;      /* NLS_FLAG_* bits, 0 by default                   */
    double dt, dx;
    double coeffs[23];
} nls_problem;

/* Host-memory solve with the reference's semantics and the extensions:
 * advances psi (complex, dtype) and, when coeffs[10] != 0, the reservoir n
 * (real, dtype) in place by iters steps.  pumping is real (dtype).
 * Returns NLS_ENONFINITE when the final state holds NaN or Inf (a dt above
 * the scheme's stability limit); the state is written back all the same.
 * The reference reports nothing in that case (nls.f90:725-733), and neither
 * do the drop-ins above (nls_solve_nls*), which keep its behaviour. */
int nls_solve_host(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres);

/* Workspace needed by nls_advance for p (bytes). */
size_t nls_workspace_bytes(const nls_problem *p);

/* Device-memory solve on a stream: psi, nres (dynamic reservoir only) and
 * pumping are device arrays; psi/nres are advanced in place.  workspace is a
 * device buffer of at least nls_workspace_bytes(p).  Enqueues work and
 * returns; the caller synchronises the stream.
 * Solves of at least 128 steps on the fused or per-stage engine replay a
 * captured block of 64 steps as one hipGraph launch (unless
 * NLS_FLAG_EAGER): the graph is captured on a private stream of the calling
 * thread, cached per thread by device, problem and buffer addresses (up to 8
 * entries), and launched on `stream`. */
int nls_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres,
                void *workspace, size_t workspace_bytes, void *stream);

/* nls_advance without allocation: the workspace is a device array of
 * NLS_STATIC_WORKSPACE_BYTES that the library holds on each GPU it runs on
 * (from its first use there until the process ends), every launch is issued
 * directly on `stream` (no step-block graph is captured), and no device or
 * host memory is allocated.  Problems whose nls_workspace_bytes exceeds the
 * array's size are rejected with NLS_EINVAL.  Calls that overlap on one GPU
 * share the array and race: run them in stream order.  Same results as
 * nls_advance.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
#define NLS_STATIC_WORKSPACE_BYTES (64u << 20)
int nls_advance_static(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *stream);

/* Frees the calling thread's cached step-block graphs (after synchronising
 * the device).  Optional: the cache is bounded. */
int nls_release(void);

/* Row-slab decomposition of one periodic 2-D grid (one device's share of a
 * grid split along its rows; no counterpart in the reference, whose solvers
 * are serial).  A slab of `rows` rows of an n-wide grid (p->n = n) is held in
 * buffers of (rows + 2 h) x n points, h = nls_slab_halo_rows(p->order) = 4R
 * = 2 (order - 1): h ghost rows (the h rows above the slab, periodically),
 * the slab's rows, h ghost rows (the rows below it).  nls_slab_step runs one
 * RK4 step on the fused kernel and writes rows h .. h + rows - 1 of psi_out
 * and nres_out (ghost rows untouched); psi_in / nres_in / pumping are read
 * in full.  The 4R-deep halo is what one step depends on, so one exchange
 * of h rows with each neighbour per step keeps the slabs equal, bit for bit,
 * to the same rows of the whole-grid solve on the fused engine.  Needs
 * dim 2, periodic bc, the RK4 scheme and rows >= h; in and out buffers must
 * not overlap.  Enqueues on `stream` and returns. */
int nls_slab_halo_rows(int order);
int nls_slab_step(const nls_problem *p, int rows, const void *pumping, const void *psi_in,
                  const void *nres_in, void *psi_out, void *nres_out, void *stream);

/* One evaluation of the right-hand side (hamiltonian / hamiltonian_2d,
 * nls.f90:621-650, :841-870) on device arrays: dpsi = F_psi, dn = F_n
 * (dn only for the dynamic reservoir). */
int nls_rhs(const nls_problem *p, const void *pumping, const void *psi, const void *nres,
            void *dpsi, void *dn, void *stream);

/* Tile height (rows: 4, 8, 16 or 32) of the single-instance 2-D RK4 stage kernels
 * for dtype; rows = 0 only queries.  Returns the height in effect or a
 * negative NLS_E* code.  Process-wide; set by the autotuner.  No
 * counterpart in the reference. */
int nls_tune_stage_tile(int dtype, int rows);

/* Tile form of the fused RK4 kernel for dtype: 0 chooses it per launch from
 * the tile counts against the compute units (the default), 1 = 64 x 32 tiles
 * of 1024 threads (64 x 16 at order 7), 2 = 64 x 16 tiles of 1024 threads,
 * 3 = 64 x 16 tiles of 512 threads with one stage buffer (double precision
 * on periodic grids; elsewhere it runs as 1), 4 = 64 x 32 tiles of 512
 * threads (single precision, order 5, periodic grids; elsewhere it runs as
 * 1); form = -1 only queries.
 * Returns the setting in effect or a negative NLS_E* code.  Process-wide; a
 * batch takes the form when it is created or updated.  Set by the
 * autotuner.  No counterpart in the reference. */
int nls_tune_fused_form(int dtype, int form);

/* Chemical potential of a state on device, mu = i <psi, F(psi)>_w / <psi, psi>_w
 * with the weight r_k = k*dx on the 1-D radial grid and 1 otherwise
 * (chemical_potential_1d/2d, nls.f90:921-971, generalised to the problem's
 * order, boundaries and reservoir).  mu[0..1] = (re, im), accumulated in
 * double precision; blocks until done. */
int nls_chemical_potential(const nls_problem *p, const void *pumping, const void *psi, const void *nres,
                           double *mu, void *stream);

/* Observables of a device state, double-precision accumulators in a fixed
 * order: out[0] = norm  sum |psi|^2 dV, out[1] = population sum n dV,
 * out[2] = max |psi|^2, out[3] = number of non-finite points.  dV is
 * 2*pi*r*dr on the 1-D radial grid, dx on a periodic line, dx^2 in 2-D.
 * Blocks until done. */
int nls_observables(const nls_problem *p, const void *psi, const void *nres, double *out, void *stream);

/* Energy of a device state for the conservative part of the right-hand side
 * (nls.f90:637-644: the operator -L + c5|psi|^2 + c6 r), double-precision
 * accumulators in a fixed order: out[1] = kinetic -Re<psi, L psi> dV with the
 * Laplacian L the problem's scheme evolves, out[2] = interaction (c5/2) sum |psi|^4
 * dV, out[3] = reservoir coupling c6 sum r |psi|^2 dV (r = n for the dynamic
 * reservoir, else c12 P / (c13 + c14 |psi|^2)), out[0] = their sum; dV as
 * nls_observables (2*pi*r*dr with the weight r on the radial grid).  The
 * reference defines no energy; its chemical potential is the same operator's
 * expectation, Re mu * norm = out[1] + 2 out[2] + out[3] for RK4 problems
 * (nls_chemical_potential takes the finite-difference operator for every
 * scheme).  L is the order-p finite-difference stencil
 * for RK4 problems and the spectral operator for split-step (SSF) problems:
 * out[1] = (dV / N^d) sum_k |k|^2 |psi_hat(k)|^2 over the grid's
 * wavenumbers k = 2 pi f / (n dx), f = -n/2 .. n/2 - 1 (taken from the 1-D
 * transforms of every row and column), which the scheme's linear step
 * conserves exactly.  Blocks until done. */
int nls_energy(const nls_problem *p, const void *pumping, const void *psi, const void *nres, double *out,
               void *stream);

/* ======================================================================
 * Input fields generated on the device: pumping profiles and initial
 * states sampled on a grid, so inputs given by parameters need no
 * host-to-device upload.  Profiles follow the reference's functors
 * (nls/pumping.py:113-178) sampled like its models (nls/model.py:220-232);
 * values are formed in double precision in numpy's order (no fused
 * multiply-add) and rounded once to dtype.
 * ====================================================================== */

/* Grid along each axis (2-D: the same grid for x and y; element [row, col]
 * sits at (x[col], x[row]) like numpy.meshgrid(x, x)). */
#define NLS_GRID_LINSPACE  0  /* numpy.linspace(a, b, n): the reference's grids */
#define NLS_GRID_STEP      1  /* a + b*k: the periodic model grid           */
#define NLS_GRID_FRACTION  2  /* a + (b*k)/n: a period b split in n cells    */
typedef struct nls_grid {
    int32_t kind;
    int32_t reserved;
    double a, b;
} nls_grid;

#define NLS_PROFILE_CONSTANT      0  /* amplitude                                   */
#define NLS_PROFILE_GAUSSIAN      1  /* GaussianPumping1D/2D (pumping.py:113-149)    */
#define NLS_PROFILE_GAUSSIAN_RING 2  /* GaussianRingPumping1D/2D (pumping.py:152-178):
                                        Gaussians at +radius and -radius of x (1-D)
                                        or of r = |(x,y) - (x0,y0)| (2-D)          */
#define NLS_PROFILE_RING          3  /* one Gaussian at r = radius (no mirror term)  */
typedef struct nls_profile {
    int32_t kind;
    int32_t reserved;
    double amplitude;   /* power                                           */
    double x0, y0;      /* centre                                          */
    double width;       /* variation (standard deviation)                  */
    double radius;
} nls_profile;

/* out[n] (1-D) or out[n*n] (2-D), real dtype, on the device. */
int nls_fill_profile(const nls_profile *profile, int dim, int n, const nls_grid *grid, int dtype, void *out,
                     void *stream);

/* psi (complex dtype) = profile + noise*(z_re + i z_im)/sqrt(2) with
 * standard normal z from a counter-based generator keyed by seed and the
 * element index (reproducible, independent of launch shape). */
int nls_fill_state(const nls_profile *profile, double noise, uint64_t seed, int dim, int n, const nls_grid *grid,
                   int dtype, void *psi, void *stream);

/* ======================================================================
 * Batched solves: many independent instances ("members") of mixed sizes in
 * shared packed device arrays, advanced together.  Members share dim,
 * order, boundary family, reservoir mode and dtype; each has its own n, dt,
 * dx, coefficients and step budget.  No counterpart in the reference, which
 * runs one solve per worker process (schedulr/workr.py:39-107); each member
 * follows solve_nls / solve_nls_2d (nls.f90:797, :903).
 * ====================================================================== */

typedef struct nls_member {
    int32_t n;          /* points per axis                                    */
    int32_t reserved;
    int64_t offset;     /* index of the member's first point in the packed
                           psi / n / pumping arrays (elements, not bytes)     */
    double dt, dx;
    double coeffs[23];
} nls_member;

typedef struct nls_batch nls_batch;

/* Plan a batch: validates members (no overlap), builds the tile list and
 * uploads the member table (device memory owned by the batch). */
int nls_batch_create(int dim, int order, int bc, int dtype, int engine, const nls_member *members, int count,
                     nls_batch **out);
int nls_batch_destroy(nls_batch *b);
/* Length of the packed arrays (points) and number of tiles. */
size_t nls_batch_points(const nls_batch *b);
int nls_batch_tiles(const nls_batch *b);
size_t nls_batch_workspace_bytes(const nls_batch *b);
/* Step budget per member for the following advance calls (host array of
 * count entries): member i advances min(iters, steps[i]) steps and then
 * holds its state.  Default: unlimited.  The member table is uploaded
 * asynchronously, ordered on `stream` (use the stream the batch advances
 * on); the host array may be reused on return. */
int nls_batch_set_steps(nls_batch *b, const int32_t *steps, void *stream);
/* Replaces the members of a batch of one engine (not HYBRID) and, when steps
 * is not NULL, their budgets (else unlimited): the same checks as
 * nls_batch_create; count may be 0 (advances then do nothing).  The tables
 * are uploaded asynchronously, ordered on `stream`; the host waits only when
 * the tables must grow or a captured step block must be dropped.  For a
 * request loop that admits and retires members between advances without
 * re-creating its batch.  No counterpart in the reference. */
int nls_batch_update(nls_batch *b, const nls_member *members, int count, const int32_t *steps, void *stream);
/* Advance every member by up to iters steps on device arrays in place.
 * Fused and per-stage engines: a call of at least 128 steps in which every
 * member's budget covers all the steps replays a captured block of 64 steps
 * as one hipGraph launch (captured at first use on the batch's own stream,
 * per set of buffers; unless NLS_FLAG_EAGER); the resident engine runs a
 * block of steps per launch anyway.  A batch is used by one thread at a time. */
int nls_batch_advance(nls_batch *b, int iters, const void *pumping, void *psi, void *nres, void *workspace,
                      size_t workspace_bytes, void *stream);
/* NLS_FLAG_EAGER: issue every launch directly (no step-block graphs). */
int nls_batch_set_flags(nls_batch *b, int flags);
/* Capture now the step block that nls_batch_advance(b, iters, ...) on these
 * buffers would replay (nothing when it would issue its launches directly);
 * runs no step.  Optional: advance captures at first use otherwise. */
int nls_batch_prepare(nls_batch *b, int iters, const void *pumping, void *psi, void *nres, void *workspace,
                      size_t workspace_bytes);
/* Observables of every member in one launch, accumulated in double
 * precision in a fixed order (a member's values do not depend on the rest of
 * the batch; nls_observables of one instance): out[4 i + 0..3] = norm
 * sum |psi|^2 dV, reservoir population sum n dV, peak |psi|^2 and the count
 * of non-finite points of member i.  out is a host array of 4 count
 * doubles; blocks until done. */
int nls_batch_observables(nls_batch *b, const void *psi, const void *nres, double *out, void *stream);
/* Energy of every member (nls_energy of each member alone, with the member's
 * own stencil weights): out[4 i + 0..3] = energy, kinetic, interaction,
 * reservoir coupling of member i, FP64 accumulators in a fixed order per
 * member; out is a host array of 4 count doubles; blocks until done.  No
 * counterpart in the reference. */
int nls_batch_energy(nls_batch *b, const void *pumping, const void *psi, const void *nres, double *out,
                     void *stream);
/* The same two reductions written to device memory: out is a device array of
 * 4 count doubles, enqueued on `stream` and not waited for.  No counterpart
 * in the reference. */
int nls_batch_observables_async(nls_batch *b, const void *psi, const void *nres, double *out, void *stream);
int nls_batch_energy_async(nls_batch *b, const void *pumping, const void *psi, const void *nres, double *out,
                           void *stream);

#ifdef __cplusplus
}
#endif
#endif /* NLS_AMD_H */
