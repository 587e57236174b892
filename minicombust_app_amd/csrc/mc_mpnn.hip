// mc_mpnn.hip -- learned pressure-correction initial guess (placeholder until
// the message-passing network kernels land: no network loaded, no guess).
#include "mc_internal.hpp"

namespace mc {
int mpnn_guess(Ctx *c, const double *, double *, int *ok)
{
    (void)c;
    *ok = 0;
    return MC_OK;
}
void mpnn_destroy(Ctx *c) { c->mpnn = nullptr; }
}  // namespace mc

extern "C" {
int mc_mpnn_load(mc_ctx *, const float *, int64_t, int, int) { return MC_ERR_STATE; }
int mc_mpnn_infer(mc_ctx *, const double *, double *) { return MC_ERR_STATE; }
}
