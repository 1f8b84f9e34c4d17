// flsg_cheb.hip -- local Chebyshev (placeholder until the ordered kernel lands).
#include "flsg_internal.h"

namespace flsg {
int solve_local_cheb(flsg_graph g, const flsg_params *p, int64_t s, double *d_x, double *d_r, flsg_result *res,
                     flsg_trace *tr, hipStream_t st) {
    (void)g; (void)p; (void)s; (void)d_x; (void)d_r; (void)res; (void)tr; (void)st;
    return fail(FLSG_EARG, "local Chebyshev: not built yet");
}
}  // namespace flsg
