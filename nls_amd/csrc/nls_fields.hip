// nls_fields.hip -- input fields generated on the device: pumping profiles
// and initial states sampled on a grid, so a solve whose inputs are given by
// parameters needs no host-to-device upload.
//
// Profiles follow the reference's functors (nls/pumping.py:113-178) sampled
// on the model's grids (nls/model.py:220-232):
//   GAUSSIAN       a exp(-((x-x0)^2 + (y-y0)^2) / (2 w^2))     (1-D: y = 0)
//   GAUSSIAN_RING  a (exp(-(p-R)^2 / (2 w^2)) + exp(-(p+R)^2 / (2 w^2))),
//                  p = x (1-D, signed) or |(x,y) - (x0,y0)| (2-D)
//   RING           a exp(-(p-R)^2 / (2 w^2)), p = |x - x0| or |(x,y) - (x0,y0)|
//   CONSTANT       a
// Grids (per axis, n points):
//   LINSPACE       numpy.linspace(a, b, n): a + k (b-a)/(n-1), last point b
//   STEP           a + b k
//   FRACTION       a + (b k) / n
// Every value is formed in double precision in the order the numpy
// expressions evaluate, with no fused multiply-add, and rounded once to the
// output precision; only exp/sqrt/log/sin/cos may differ from the host's in
// the last bit.
//
// Noise: element i of a state gets (std/sqrt 2)(z_re + i z_im) with
// (z_re, z_im) a Box-Muller pair from two uniforms (0, 1] made by the
// splitmix64 finaliser of seed + 2i + 1 and seed + 2i + 2 (53-bit mantissas).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nls_amd.h"
#include "nls_core.h"
#include "nls_internal.h"

namespace nls {

struct GridDev {
    int kind, n;
    double a, b, step;
};

#pragma clang fp contract(off)
__device__ __forceinline__ double grid_at(const GridDev &g, int k) {
    if (g.kind == NLS_GRID_LINSPACE) return k == g.n - 1 ? g.b : (double)k * g.step + g.a;
    if (g.kind == NLS_GRID_STEP) return g.a + g.b * (double)k;
    return g.a + (g.b * (double)k) / (double)g.n;
}

__device__ __forceinline__ double profile_at(const nls_profile &p, int dim, double x, double y) {
    const double a = p.amplitude, w2 = 2.0 * (p.width * p.width);
    switch (p.kind) {
    case NLS_PROFILE_GAUSSIAN: {
        const double dxx = x - p.x0, dyy = (dim == 1 ? 0.0 : y) - p.y0;
        return a * exp(-(dxx * dxx + dyy * dyy) / w2);
    }
    case NLS_PROFILE_GAUSSIAN_RING: {
        double rho;
        if (dim == 1) {
            rho = x;
        } else {
            const double dxx = x - p.x0, dyy = y - p.y0;
            rho = sqrt(dxx * dxx + dyy * dyy);
        }
        // the reference sums two 1-D Gaussians centred at +R and -R, each
        // evaluated as exp(-((rho - c)^2 + (0 - 0)^2) / (2 w^2))
        const double u = rho - p.radius, v = rho - (-p.radius);
        return 1.0 * (a * exp(-(u * u + 0.0) / w2) + a * exp(-(v * v + 0.0) / w2));
    }
    case NLS_PROFILE_RING: {
        double rho;
        if (dim == 1) {
            rho = fabs(x - p.x0);
        } else {
            const double dxx = x - p.x0, dyy = y - p.y0;
            rho = sqrt(dxx * dxx + dyy * dyy);
        }
        const double u = rho - p.radius;
        return a * exp(-(u * u) / w2);
    }
    default:
        return a;
    }
}

__device__ __forceinline__ double uniform53(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 1.0) * 0x1.0p-53;      // (0, 1]
}
#pragma clang fp contract(on)

template <typename T>
__global__ void k_fill_profile(const nls_profile p, int dim, GridDev gx, GridDev gy, T *out, long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int n = gx.n;
    const int col = dim == 1 ? (int)i : (int)(i % n);
    const int row = dim == 1 ? 0 : (int)(i / n);
    // row index along the first array axis: numpy.meshgrid(x, x) gives
    // X[row, col] = x[col], Y[row, col] = x[row]
    const double x = grid_at(gx, col), y = dim == 1 ? 0.0 : grid_at(gy, row);
    out[i] = (T)profile_at(p, dim, x, y);
}

template <typename T>
__global__ void k_fill_state(const nls_profile p, int dim, GridDev gx, GridDev gy, double noise, uint64_t seed,
                             cplx_t<T> *out, long long total) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int n = gx.n;
    const int col = dim == 1 ? (int)i : (int)(i % n);
    const int row = dim == 1 ? 0 : (int)(i / n);
    const double x = grid_at(gx, col), y = dim == 1 ? 0.0 : grid_at(gy, row);
    double re = profile_at(p, dim, x, y), im = 0.0;
    if (noise != 0.0) {
        const double u1 = uniform53(seed + 2ull * (uint64_t)i + 1ull);
        const double u2 = uniform53(seed + 2ull * (uint64_t)i + 2ull);
        const double r = sqrt(-2.0 * log(u1));
        const double th = 6.283185307179586 * u2;
        const double s = noise / 1.4142135623730951;
        re = re + s * (r * cos(th));
        im = s * (r * sin(th));
    }
    out[i] = mk<T>((T)re, (T)im);
}

static int make_grid(const nls_grid *g, int n, GridDev *d) {
    if (!g) return fail(NLS_EINVAL, "grid is NULL");
    if (g->kind != NLS_GRID_LINSPACE && g->kind != NLS_GRID_STEP && g->kind != NLS_GRID_FRACTION)
        return fail(NLS_EINVAL, "grid kind must be NLS_GRID_LINSPACE, _STEP or _FRACTION (got %d)", g->kind);
    d->kind = g->kind;
    d->n = n;
    d->a = g->a;
    d->b = g->b;
    d->step = n > 1 ? (g->b - g->a) / (double)(n - 1) : 0.0;
    return NLS_OK;
}

static int check_fill(const nls_profile *p, int dim, int n, int dtype, const void *out) {
    if (!p) return fail(NLS_EINVAL, "profile is NULL");
    if (p->kind < NLS_PROFILE_CONSTANT || p->kind > NLS_PROFILE_RING)
        return fail(NLS_EINVAL, "profile kind must be one of NLS_PROFILE_* (got %d)", p->kind);
    if (p->kind != NLS_PROFILE_CONSTANT && !(p->width > 0.0))
        return fail(NLS_EINVAL, "profile width must be positive");
    if (dim != 1 && dim != 2) return fail(NLS_EINVAL, "dim must be 1 or 2 (got %d)", dim);
    if (n < 1 || (dim == 2 && n > 46340)) return fail(NLS_EINVAL, "bad grid size n=%d", n);
    if (dtype != NLS_F32 && dtype != NLS_F64) return fail(NLS_EINVAL, "dtype must be 0 or 1");
    if (!out) return fail(NLS_EINVAL, "output buffer is NULL");
    return NLS_OK;
}

}  // namespace nls

using namespace nls;

int nls_fill_profile(const nls_profile *profile, int dim, int n, const nls_grid *grid, int dtype, void *out,
                     void *stream) {
    int rc = check_fill(profile, dim, n, dtype, out);
    if (rc) return rc;
    GridDev g;
    rc = make_grid(grid, n, &g);
    if (rc) return rc;
    const long long total = dim == 1 ? (long long)n : (long long)n * n;
    const dim3 blocks((unsigned)((total + 255) / 256));
    hipStream_t s = (hipStream_t)stream;
    if (dtype == NLS_F64)
        hipLaunchKernelGGL((k_fill_profile<double>), blocks, dim3(256), 0, s, *profile, dim, g, g, (double *)out, total);
    else
        hipLaunchKernelGGL((k_fill_profile<float>), blocks, dim3(256), 0, s, *profile, dim, g, g, (float *)out, total);
    return hip_check(hipGetLastError(), "fill profile");
}

int nls_fill_state(const nls_profile *profile, double noise, uint64_t seed, int dim, int n, const nls_grid *grid,
                   int dtype, void *psi, void *stream) {
    int rc = check_fill(profile, dim, n, dtype, psi);
    if (rc) return rc;
    if (!(noise >= 0.0)) return fail(NLS_EINVAL, "noise must be >= 0");
    GridDev g;
    rc = make_grid(grid, n, &g);
    if (rc) return rc;
    const long long total = dim == 1 ? (long long)n : (long long)n * n;
    const dim3 blocks((unsigned)((total + 255) / 256));
    hipStream_t s = (hipStream_t)stream;
    if (dtype == NLS_F64)
        hipLaunchKernelGGL((k_fill_state<double>), blocks, dim3(256), 0, s, *profile, dim, g, g, noise, (uint64_t)seed,
                           (double2 *)psi, total);
    else
        hipLaunchKernelGGL((k_fill_state<float>), blocks, dim3(256), 0, s, *profile, dim, g, g, noise, (uint64_t)seed,
                           (float2 *)psi, total);
    return hip_check(hipGetLastError(), "fill state");
}
