// Learned preconditioner (placeholder until the network lands).
#include "pdx_nk.h"

void pdx_gnn_free(pdx_ctx *c) { (void)c; }

pdx_status pdx_gnn_prepare(pdx_ctx *c)
{
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the learned preconditioner is not trained");
}

pdx_status pdx_gnn_apply(pdx_ctx *c, const double *v, double *z)
{
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the learned preconditioner is not trained");
}

pdx_status pdx_gnn_record(pdx_ctx *c, const double *rhs, const double *du) { return PDX_OK; }

extern "C" pdx_status pdx_gnn_collect(pdx_ctx *c, int32_t enable)
{
    if (!c) return PDX_ERR_ARG;
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the learned preconditioner is not built yet");
}

extern "C" pdx_status pdx_gnn_train(pdx_ctx *c, int32_t epochs, double lr, uint64_t seed,
                                    double *loss_out)
{
    if (!c) return PDX_ERR_ARG;
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the learned preconditioner is not built yet");
}
