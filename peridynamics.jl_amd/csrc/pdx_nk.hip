// Newton-Krylov on the device (placeholder entry points until the solver lands).
#include "pdx_internal.h"

void pdx_nk_free(pdx_ctx *c) { (void)c; }
void pdx_gnn_free(pdx_ctx *c) { (void)c; }

extern "C" pdx_status pdx_set_displacement_bcs(pdx_ctx *c, int64_t n, const int32_t *points,
                                               const int32_t *dims, const double *values)
{
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "Newton-Krylov is not built yet");
}

extern "C" pdx_status pdx_jv(pdx_ctx *c, const double *v, double *Jv, double scale)
{
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "Newton-Krylov is not built yet");
}

extern "C" pdx_status pdx_residual(pdx_ctx *c, double *r, double *norm_out)
{
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "Newton-Krylov is not built yet");
}

extern "C" pdx_status pdx_nk_step(pdx_ctx *c, int64_t n, int64_t n_steps, const pdx_nk_opts *o,
                                  pdx_nk_stats *s)
{
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "Newton-Krylov is not built yet");
}

extern "C" pdx_status pdx_gnn_collect(pdx_ctx *c, int32_t enable)
{
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the learned preconditioner is not built yet");
}

extern "C" pdx_status pdx_gnn_train(pdx_ctx *c, int32_t epochs, double lr, uint64_t seed,
                                    double *loss_out)
{
    return pdx_fail(c, PDX_ERR_UNSUPPORTED, "the learned preconditioner is not built yet");
}
