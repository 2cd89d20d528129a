// nls_internal.h -- helpers shared by the translation units of libnls_amd.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "nls_amd.h"

namespace nls {

int fail(int code, const char *fmt, ...);          // record message, return code
int hip_check(hipError_t e, const char *what);     // NLS_OK or NLS_EHIP/NLS_ENOMEM
int check_problem(const nls_problem *p);
size_t points_of(const nls_problem *p);

int plain_engine(const nls_problem *p);            // RK4 engine of a plain solve (NLS_ENGINE_*)

// the fused kernel holds a member (or row slab) of this many points: its
// complex array is smaller than 4 GiB (32-bit byte offsets)
bool fused_holds(long long points, int dtype);
// process-wide tile form of the fused kernel per dtype: -1 (chosen per
// launch), 0 tall, 1 short, 2 half (nls_tune_fused_form)
int set_fused_form(int dtype, int form);
int get_fused_form(int dtype);
// plain 2-D RK4 solve of one grid on the fused kernel (nls_batch.hip)
size_t fused_grid_workspace_bytes(const nls_problem *p);
int fused_grid_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *ws,
                       hipStream_t s);
// one RK4 step of a row slab with 4R ghost rows on each side (fused kernel, BC_SLAB)
int fused_slab_step(const nls_problem *p, int rows, const void *P, const void *psi_in, const void *nres_in,
                    void *psi_out, void *nres_out, hipStream_t s);
// plain RK4 solve of one line (n <= 4096) or small grid (n <= 64) on the resident kernel
bool resident_fits(const nls_problem *p);
size_t resident_workspace_bytes(const nls_problem *p);
int resident_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *ws,
                     hipStream_t s);

// split-step Fourier scheme (nls_ssf.hip)
size_t ssf_workspace_bytes(const nls_problem *p);
int ssf_check(const nls_problem *p);               // size limits of the scheme
// (1/N^d) sum_k |k|^2 |psi_hat(k)|^2 = -Re<psi, L psi> with the spectral
// Laplacian of a split-step problem, into dev_sum (device, one double);
// scratch: ssf_kinetic_scratch_bytes of device memory
size_t ssf_kinetic_scratch_bytes(const nls_problem *p);
int ssf_kinetic(const nls_problem *p, const void *psi, void *scratch, double *dev_sum, hipStream_t s);
int ssf_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *ws,
                hipStream_t s);

}  // namespace nls
