// nans_step.hip -- world batches: upload/download and the simulation step.
#include "nans_common.hpp"

struct nans_worlds {
    int dummy;
};

namespace nans {
int worlds_release(nans_ctx *ctx)
{
    if (ctx && ctx->worlds) {
        delete ctx->worlds;
        ctx->worlds = nullptr;
    }
    return NANS_OK;
}
}  // namespace nans
