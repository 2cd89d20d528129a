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

// split-step Fourier scheme (nls_ssf.hip)
size_t ssf_workspace_bytes(const nls_problem *p);
int ssf_check(const nls_problem *p);               // size limits of the scheme
int ssf_advance(const nls_problem *p, int iters, const void *pumping, void *psi, void *nres, void *ws,
                hipStream_t s);

}  // namespace nls
