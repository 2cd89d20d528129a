// nls_status.hip -- the library's status reporting: the per-thread message of
// the last failure (nls_last_error) and the helpers every entry point uses
// to set it; the module version.  Host code only.
#include <cstdarg>
#include <cstdio>

#include "nls_amd.h"
#include "nls_internal.h"

namespace nls {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return NLS_OK;
    return fail(e == hipErrorOutOfMemory ? NLS_ENOMEM : NLS_EHIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace nls

extern "C" {

const char *nls_last_error(void) { return nls::g_err; }

// version (nls.f90:29-37): the reference module's 0.2.0.
int nls_version(int *major, int *minor, int *patch) {
    if (!major || !minor || !patch) return nls::fail(NLS_EINVAL, "NULL argument");
    *major = 0; *minor = 2; *patch = 0;
    return NLS_OK;
}

}  // extern "C"
