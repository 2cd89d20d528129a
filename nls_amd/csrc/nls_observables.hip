// nls_observables.hip -- on-device reductions with double-precision
// accumulators and a fixed summation order (so a result never depends on
// what else runs on the GPU):
//   - the chemical potential mu = i <u, F(u)>_w / <u, u>_w of
//     chemical_potential_1d/2d (reference nls/nls.f90:921-971): weight
//     w_k = k*dx on the 1-D radial grid (nls.f90:941-947), 1 in 2-D;
//   - norm, reservoir population, peak density and a non-finite count;
//   - the energy of the conservative part of the right-hand side,
//     E = -Re<psi, L psi>_w + (c5/2) sum |psi|^4 w + c6 sum r |psi|^2 w
//     (times the volume element), whose variation at fixed reservoir r gives
//     the operator -L + c5|psi|^2 + c6 r of nls.f90:637-644; L is the
//     problem's finite-difference Laplacian, or for a split-step problem the
//     spectral one the scheme evolves (nls_ssf.hip, ssf_kinetic).
// Pass 1: each 256-thread workgroup reduces a grid-strided slice in a fixed
// tree; pass 2: one workgroup sums the per-workgroup partials in order.
#include <cstring>
#include <vector>

#include "nls_core.h"
#include "nls_internal.h"

namespace nls {

constexpr int RED_BLOCKS = 512;     // partials per reduction (fixed: keeps the order fixed)
constexpr int RED_VALS = 4;         // values reduced together

__device__ __forceinline__ void block_sum(double (*sh)[256], double *v, double *out, int nv) {
    const int t = threadIdx.x;
    for (int q = 0; q < nv; ++q) sh[q][t] = v[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < nv; ++q) sh[q][t] += sh[q][t + s];
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < nv; ++q) out[q] = sh[q][0];
}

// sum_i conj(a_i) b_i w_i  -> (re, im);  w_i = i*dx if radial else 1
template <typename T>
__global__ void __launch_bounds__(256) k_cdot_partial(size_t N, const cplx_t<T> *a, const cplx_t<T> *b,
                                                      int radial, double dx, double *part) {
    __shared__ double sh[RED_VALS][256];
    double v[RED_VALS] = {0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t)gridDim.x * 256) {
        const double w = radial ? (double)i * dx : 1.0;
        const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
        v[0] += (ar * br + ai * bi) * w;
        v[1] += (ar * bi - ai * br) * w;
    }
    block_sum(sh, v, part + (size_t)blockIdx.x * RED_VALS, 2);
}

// norm sum |psi|^2 w, population sum n w, max |psi|^2, non-finite count
template <typename T>
__global__ void __launch_bounds__(256) k_obs_partial(size_t N, const cplx_t<T> *psi, const T *nres,
                                                     int radial, double dx, double *part) {
    __shared__ double sh[RED_VALS][256];
    double v[RED_VALS] = {0, 0, 0, 0};
    double peak = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t)gridDim.x * 256) {
        const double w = radial ? (double)i * dx : 1.0;
        const double pr = psi[i].x, pi = psi[i].y;
        const double s = pr * pr + pi * pi;
        const double nn = nres ? (double)nres[i] : 0.0;
        if (!isfinite(s) || !isfinite(nn)) v[3] += 1.0;
        else {
            v[0] += s * w;
            v[1] += nn * w;
            peak = s > peak ? s : peak;
        }
    }
    v[2] = 0.0;
    block_sum(sh, v, part + (size_t)blockIdx.x * RED_VALS, 4);
    // peak: max-reduce separately (sum of maxima would be wrong)
    __syncthreads();
    sh[0][threadIdx.x] = peak;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * RED_VALS + 2] = sh[0][0];
}

// sum |psi|^4 w and sum r |psi|^2 w, r the reservoir: n (dynamic) or the
// algebraic c12 P / (c13 + c14 |psi|^2) (nls.f90:570-581)
template <typename T>
__global__ void __launch_bounds__(256) k_energy_partial(size_t N, const cplx_t<T> *psi, const T *P, const T *nres,
                                                        double c12, double c13, double c14, int radial, double dx,
                                                        double *part) {
    __shared__ double sh[RED_VALS][256];
    double v[RED_VALS] = {0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < N; i += (size_t)gridDim.x * 256) {
        const double w = radial ? (double)i * dx : 1.0;
        const double pr = psi[i].x, pi = psi[i].y;
        const double s = pr * pr + pi * pi;
        const double r = nres ? (double)nres[i] : c12 * (double)P[i] / (c13 + c14 * s);
        v[0] += s * s * w;
        v[1] += r * s * w;
    }
    block_sum(sh, v, part + (size_t)blockIdx.x * RED_VALS, 2);
}

__global__ void __launch_bounds__(256) k_sum_partials(int nb, const double *part, int max_slot, double *out) {
    __shared__ double sh[RED_VALS][256];
    double v[RED_VALS] = {0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += 256)
        for (int q = 0; q < RED_VALS; ++q) {
            const double x = part[(size_t)b * RED_VALS + q];
            v[q] = (q == max_slot) ? fmax(v[q], x) : v[q] + x;
        }
    const int t = threadIdx.x;
    for (int q = 0; q < RED_VALS; ++q) sh[q][t] = v[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int q = 0; q < RED_VALS; ++q)
                sh[q][t] = (q == max_slot) ? fmax(sh[q][t], sh[q][t + s]) : sh[q][t] + sh[q][t + s];
        __syncthreads();
    }
    if (t == 0)
        for (int q = 0; q < RED_VALS; ++q) out[q] = sh[q][0];
}

static int red_blocks(size_t N) {
    const size_t b = (N + 255) / 256;
    return (int)(b < (size_t)RED_BLOCKS ? (b ? b : 1) : RED_BLOCKS);
}

// <a, b>_w on device buffers; partials/out are device scratch.
template <typename T>
static void cdot(size_t N, const void *a, const void *b, int radial, double dx, double *part, double *out,
                 hipStream_t s) {
    const int nb = red_blocks(N);
    hipLaunchKernelGGL((k_cdot_partial<T>), dim3(nb), dim3(256), 0, s, N, (const cplx_t<T> *)a,
                       (const cplx_t<T> *)b, radial, dx, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, nb, part, -1, out);
}

static int chemical_potential_dev(const nls_problem *p, const void *P, const void *psi, const void *nres,
                                  double *mu, hipStream_t s) {
    const size_t N = points_of(p), rs = p->dtype == NLS_F64 ? 8 : 4;
    const bool dyn = p->coeffs[10] != 0.0;
    void *scratch = nullptr;
    const size_t fbytes = N * 2 * rs, nbytes = dyn ? N * rs : 0;
    const size_t red = (size_t)RED_BLOCKS * RED_VALS * sizeof(double) + 2 * RED_VALS * sizeof(double);
    int rc = hip_check(hipMalloc(&scratch, fbytes + nbytes + red + 512), "mu scratch");
    if (rc) return rc;
    char *base = (char *)scratch;
    void *F = base;
    void *Fn = dyn ? base + fbytes : nullptr;
    double *part = (double *)(base + ((fbytes + nbytes + 255) & ~(size_t)255));
    double *outM = part + RED_BLOCKS * RED_VALS, *outE = outM + RED_VALS;
    rc = nls_rhs(p, P, psi, nres, F, Fn, s);
    const int radial = (p->dim == 1 && p->bc == NLS_BC_REF) ? 1 : 0;
    if (!rc) {
        if (p->dtype == NLS_F64) {
            cdot<double>(N, psi, psi, radial, p->dx, part, outM, s);
            cdot<double>(N, psi, F, radial, p->dx, part, outE, s);
        } else {
            cdot<float>(N, psi, psi, radial, p->dx, part, outM, s);
            cdot<float>(N, psi, F, radial, p->dx, part, outE, s);
        }
        rc = hip_check(hipGetLastError(), "mu reduction");
    }
    double hM[RED_VALS], hE[RED_VALS];
    if (!rc) rc = hip_check(hipMemcpyAsync(hM, outM, sizeof hM, hipMemcpyDeviceToHost, s), "mu download");
    if (!rc) rc = hip_check(hipMemcpyAsync(hE, outE, sizeof hE, hipMemcpyDeviceToHost, s), "mu download");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "mu sync");
    (void)hipFree(scratch);
    if (rc) return rc;
    // E = i * (Er + i Ei) = -Ei + i Er; mu = E / M
    const double xr = -hE[1], xi = hE[0], Mr = hM[0], Mi = hM[1];
    const double den = Mr * Mr + Mi * Mi;
    mu[0] = (xr * Mr + xi * Mi) / den;
    mu[1] = (xi * Mr - xr * Mi) / den;
    return NLS_OK;
}

// Energy terms on device buffers.  The kinetic term comes from the
// right-hand side kernels with every term but the Laplacian switched off
// (c3 = c4 = c5 = c6 = c12 = c14 = 0, c13 = 1: F = i L psi), so it uses the
// solver's own discrete operator; then -Re<psi, L psi>_w = -Im<psi, F>_w.
static int energy_dev(const nls_problem *p, const void *P, const void *psi, const void *nres, double *out,
                      hipStream_t s) {
    const size_t N = points_of(p), rs = p->dtype == NLS_F64 ? 8 : 4;
    const bool dyn = p->coeffs[10] != 0.0;
    nls_problem q = *p;
    q.coeffs[2] = q.coeffs[3] = q.coeffs[4] = q.coeffs[5] = 0.0;
    q.coeffs[10] = 0.0; q.coeffs[11] = 0.0; q.coeffs[12] = 1.0; q.coeffs[13] = 0.0;
    // the kinetic term with the scheme's own Laplacian: the stencil's for
    // RK4, the spectral one for the split-step scheme
    const bool spectral = p->scheme == NLS_SCHEME_SSF;
    void *scratch = nullptr;
    const size_t fbytes = spectral ? ssf_kinetic_scratch_bytes(p) : N * 2 * rs;
    const size_t red = (size_t)RED_BLOCKS * RED_VALS * sizeof(double) + 2 * RED_VALS * sizeof(double);
    int rc = hip_check(hipMalloc(&scratch, fbytes + red + 512), "energy scratch");
    if (rc) return rc;
    char *base = (char *)scratch;
    void *F = base;
    double *part = (double *)(base + ((fbytes + 255) & ~(size_t)255));
    double *outK = part + RED_BLOCKS * RED_VALS, *outV = outK + RED_VALS;
    rc = spectral ? ssf_kinetic(p, psi, F, outK, s) : nls_rhs(&q, P, psi, nullptr, F, nullptr, s);
    const int radial = (p->dim == 1 && p->bc == NLS_BC_REF) ? 1 : 0;
    const int nb = red_blocks(N);
    if (!rc) {
        // (the spectral kinetic sum is already in outK[0])
        if (!spectral && p->dtype == NLS_F64) cdot<double>(N, psi, F, radial, p->dx, part, outK, s);
        else if (!spectral) cdot<float>(N, psi, F, radial, p->dx, part, outK, s);
        if (p->dtype == NLS_F64) {
            hipLaunchKernelGGL((k_energy_partial<double>), dim3(nb), dim3(256), 0, s, N, (const double2 *)psi,
                               (const double *)P, dyn ? (const double *)nres : nullptr, p->coeffs[11], p->coeffs[12],
                               p->coeffs[13], radial, p->dx, part);
        } else {
            hipLaunchKernelGGL((k_energy_partial<float>), dim3(nb), dim3(256), 0, s, N, (const float2 *)psi,
                               (const float *)P, dyn ? (const float *)nres : nullptr, p->coeffs[11], p->coeffs[12],
                               p->coeffs[13], radial, p->dx, part);
        }
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, nb, part, -1, outV);
        rc = hip_check(hipGetLastError(), "energy reduction");
    }
    double hK[RED_VALS], hV[RED_VALS];
    if (!rc) rc = hip_check(hipMemcpyAsync(hK, outK, sizeof hK, hipMemcpyDeviceToHost, s), "energy download");
    if (!rc) rc = hip_check(hipMemcpyAsync(hV, outV, sizeof hV, hipMemcpyDeviceToHost, s), "energy download");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "energy sync");
    (void)hipFree(scratch);
    if (rc) return rc;
    const double dV = radial ? 2.0 * M_PI * p->dx : (p->dim == 1 ? p->dx : p->dx * p->dx);
    // -Re<psi, L psi> = -Im<psi, i L psi>; spectral: (1/n) (sum over lines)
    out[1] = spectral ? hK[0] / (double)p->n * dV : -hK[1] * dV;
    out[2] = 0.5 * p->coeffs[4] * hV[0] * dV;
    out[3] = p->coeffs[5] * hV[1] * dV;
    out[0] = out[1] + out[2] + out[3];
    return NLS_OK;
}

template <typename T>
static int observables_dev(const nls_problem *p, const void *psi, const void *nres, double *out, hipStream_t s) {
    const size_t N = points_of(p);
    void *scratch = nullptr;
    int rc = hip_check(hipMalloc(&scratch, (RED_BLOCKS + 1) * RED_VALS * sizeof(double)), "obs scratch");
    if (rc) return rc;
    double *part = (double *)scratch, *dout = part + RED_BLOCKS * RED_VALS;
    const int nb = red_blocks(N);
    const int radial = (p->dim == 1 && p->bc == NLS_BC_REF) ? 1 : 0;
    hipLaunchKernelGGL((k_obs_partial<T>), dim3(nb), dim3(256), 0, s, N, (const cplx_t<T> *)psi,
                       (const T *)nres, radial, p->dx, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, nb, part, 2, dout);
    rc = hip_check(hipGetLastError(), "observables");
    double h[RED_VALS];
    if (!rc) rc = hip_check(hipMemcpyAsync(h, dout, sizeof h, hipMemcpyDeviceToHost, s), "obs download");
    if (!rc) rc = hip_check(hipStreamSynchronize(s), "obs sync");
    (void)hipFree(scratch);
    if (rc) return rc;
    // volume element: 2*pi*r*dr on the radial grid (r-weight already applied),
    // dx on a 1-D line, dx^2 on the square grid
    const double dV = radial ? 2.0 * M_PI * p->dx : (p->dim == 1 ? p->dx : p->dx * p->dx);
    out[0] = h[0] * dV;
    out[1] = h[1] * dV;
    out[2] = h[2];
    out[3] = h[3];
    return NLS_OK;
}

}  // namespace nls

using namespace nls;

extern "C" {

int nls_chemical_potential(const nls_problem *p, const void *pumping, const void *psi, const void *nres,
                           double *mu, void *stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (!pumping || !psi || !mu || (p->coeffs[10] != 0.0 && !nres)) return fail(NLS_EINVAL, "NULL argument");
    return chemical_potential_dev(p, pumping, psi, nres, mu, (hipStream_t)stream);
}

int nls_energy(const nls_problem *p, const void *pumping, const void *psi, const void *nres, double *out,
               void *stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (!pumping || !psi || !out || (p->coeffs[10] != 0.0 && !nres)) return fail(NLS_EINVAL, "NULL argument");
    if (p->scheme != NLS_SCHEME_RK4 && p->scheme != NLS_SCHEME_SSF) return fail(NLS_EINVAL, "bad scheme");
    return energy_dev(p, pumping, psi, nres, out, (hipStream_t)stream);
}

int nls_observables(const nls_problem *p, const void *psi, const void *nres, double *out, void *stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (!psi || !out) return fail(NLS_EINVAL, "NULL argument");
    if (p->dtype == NLS_F64) return observables_dev<double>(p, psi, nres, out, (hipStream_t)stream);
    return observables_dev<float>(p, psi, nres, out, (hipStream_t)stream);
}

static int dropin_mu(int dim, float dx, int n, const float *pumping, const float *coeffs, const void *u0,
                     double *mu) {
    if (!pumping || !coeffs || !u0) return fail(NLS_EINVAL, "NULL argument");
    nls_problem p;
    memset(&p, 0, sizeof p);
    p.dim = dim; p.n = n; p.order = 5; p.bc = NLS_BC_REF; p.dtype = NLS_F32; p.scheme = NLS_SCHEME_RK4;
    p.engine = NLS_ENGINE_AUTO;
    p.dt = 0.0; p.dx = dx;
    for (int k = 0; k < 23; ++k) p.coeffs[k] = coeffs[k];
    p.coeffs[10] = 0.0;                    // the reference's hamiltonian is adiabatic
    int rc = check_problem(&p);
    if (rc) return rc;
    const size_t N = points_of(&p);
    void *dP = nullptr, *du = nullptr;
    rc = hip_check(hipMalloc(&dP, N * 4), "pumping");
    if (!rc) rc = hip_check(hipMalloc(&du, N * 8), "u0");
    if (!rc) rc = hip_check(hipMemcpy(dP, pumping, N * 4, hipMemcpyHostToDevice), "upload");
    if (!rc) rc = hip_check(hipMemcpy(du, u0, N * 8, hipMemcpyHostToDevice), "upload");
    if (!rc) rc = chemical_potential_dev(&p, dP, du, nullptr, mu, nullptr);
    if (dP) (void)hipFree(dP);
    if (du) (void)hipFree(du);
    return rc;
}

int nls_chemical_potential_1d(float dx, int n, const float *pumping, const float *coeffs, const void *u0,
                              void *mu) {
    if (!mu) return fail(NLS_EINVAL, "NULL argument");
    double m[2];
    int rc = dropin_mu(1, dx, n, pumping, coeffs, u0, m);
    if (rc) return rc;
    ((float *)mu)[0] = (float)m[0];
    ((float *)mu)[1] = (float)m[1];
    return NLS_OK;
}

int nls_chemical_potential_2d(float dx, int n, const float *pumping, const float *coeffs, const void *u0,
                              float *mu) {
    if (!mu) return fail(NLS_EINVAL, "NULL argument");
    double m[2];
    int rc = dropin_mu(2, dx, n, pumping, coeffs, u0, m);
    if (rc) return rc;
    *mu = (float)m[0];
    return NLS_OK;
}

}  // extern "C"
