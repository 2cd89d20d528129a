// nls_ssf.hip -- split-step Fourier scheme (not yet implemented).
#include "nls_internal.h"

namespace nls {

size_t ssf_workspace_bytes(const nls_problem *) { return 0; }

int ssf_advance(const nls_problem *, int, const void *, void *, void *, void *, hipStream_t) {
    return fail(NLS_EUNSUPPORTED, "split-step Fourier scheme not built");
}

}  // namespace nls
