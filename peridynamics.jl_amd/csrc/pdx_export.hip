// Export pipeline: the device side of export_results / export_reference_results
// (src/core/threads_body_data_handler.jl:197-212) and _export_results (src/auxiliary/io.jl:157-167).
//
// pdx_export_begin gathers the requested fields of the current state into a device
// snapshot in the caller's point order (one kernel per field on the compute stream, so the
// snapshot is consistent with the step that just ended), then queues the copy of that
// snapshot into pinned host memory on a second stream.  It returns at once: the steps the
// caller launches next overlap with the copy, and the host writes the files of one export
// while the device computes the next ones.  Two slots alternate, so one export can be
// written out while the next is copied.
#include <cstring>

#include "pdx_internal.h"

#define PDX_EXPORT_MAX_FIELDS 16

void pdx_export_free(pdx_ctx *c)
{
    if (c->xstream) (void)hipStreamSynchronize(c->xstream);
    for (int s = 0; s < 2; s++) {
        if (c->xdev[s]) (void)hipFree(c->xdev[s]);
        if (c->xhost[s]) (void)hipHostFree(c->xhost[s]);
        if (c->xev_snap[s]) (void)hipEventDestroy(c->xev_snap[s]);
        if (c->xev_done[s]) (void)hipEventDestroy(c->xev_done[s]);
        c->xdev[s] = c->xhost[s] = nullptr;
        c->xev_snap[s] = c->xev_done[s] = nullptr;
        c->xcap[s] = 0;
    }
    if (c->xstream) (void)hipStreamDestroy(c->xstream);
    c->xstream = nullptr;
    c->xbegun = c->xwaited = 0;
}

static pdx_status export_init(pdx_ctx *c)
{
    if (c->xstream) return PDX_OK;
    // default priority, like the compute stream
    PDX_HIP(c, hipStreamCreateWithFlags(&c->xstream, hipStreamNonBlocking));
    for (int s = 0; s < 2; s++) {
        PDX_HIP(c, hipEventCreateWithFlags(&c->xev_snap[s], hipEventDisableTiming));
        PDX_HIP(c, hipEventCreateWithFlags(&c->xev_done[s], hipEventDisableTiming));
    }
    return PDX_OK;
}

extern "C" pdx_status pdx_export_begin(pdx_ctx *c, int32_t n_fields, const int32_t *fields, int64_t tag)
{
    if (!c || n_fields < 0 || (n_fields > 0 && !fields)) return PDX_ERR_ARG;
    if (n_fields > PDX_EXPORT_MAX_FIELDS)
        return pdx_fail(c, PDX_ERR_ARG, "at most %d fields per export", PDX_EXPORT_MAX_FIELDS);
    if (!c->bonds_built) return pdx_fail(c, PDX_ERR_STATE, "pdx_build_bonds first");
    if (c->xbegun - c->xwaited >= 2)
        return pdx_fail(c, PDX_ERR_STATE, "two exports are pending: pdx_export_wait first");
    PDX_HIP(c, hipSetDevice(c->device));
    PDX_TRY(export_init(c));
    const int s = (int)(c->xbegun & 1);
    std::vector<int64_t> off(n_fields + 1, 0);
    for (int k = 0; k < n_fields; k++) {
        int comps;
        PDX_TRY(pdx_field_comps(c, fields[k], &comps));
        off[k + 1] = off[k] + c->n * comps;
    }
    const size_t need = (size_t)std::max<int64_t>(off[n_fields], 1);
    if (need > c->xcap[s]) {
        if (c->xdev[s]) (void)hipFree(c->xdev[s]);
        if (c->xhost[s]) (void)hipHostFree(c->xhost[s]);
        c->xdev[s] = c->xhost[s] = nullptr;
        c->xcap[s] = 0;
        PDX_TRY(pdx_alloc(c, &c->xdev[s], need));
        PDX_HIP(c, hipHostMalloc((void **)&c->xhost[s], sizeof(double) * need, hipHostMallocDefault));
        c->xcap[s] = need;
    }
    for (int k = 0; k < n_fields; k++) PDX_TRY(pdx_gather_field(c, fields[k], c->xdev[s] + off[k]));
    PDX_HIP(c, hipEventRecord(c->xev_snap[s], c->stream));
    PDX_HIP(c, hipStreamWaitEvent(c->xstream, c->xev_snap[s], 0));
    if (off[n_fields] > 0)
        PDX_HIP(c, hipMemcpyAsync(c->xhost[s], c->xdev[s], sizeof(double) * off[n_fields],
                                  hipMemcpyDeviceToHost, c->xstream));
    PDX_HIP(c, hipEventRecord(c->xev_done[s], c->xstream));
    c->xfields[s].assign(fields, fields + n_fields);
    c->xoff[s] = off;
    c->xtag[s] = tag;
    c->xbegun++;
    return PDX_OK;
}

extern "C" pdx_status pdx_export_wait(pdx_ctx *c, int64_t *tag, int32_t *slot, const double **field_data)
{
    if (!c) return PDX_ERR_ARG;
    if (c->xwaited >= c->xbegun) return pdx_fail(c, PDX_ERR_STATE, "no export pending");
    PDX_HIP(c, hipSetDevice(c->device));
    const int s = (int)(c->xwaited & 1);
    PDX_HIP(c, hipEventSynchronize(c->xev_done[s]));
    if (tag) *tag = c->xtag[s];
    if (slot) *slot = s;
    if (field_data)
        for (size_t k = 0; k < c->xfields[s].size(); k++) field_data[k] = c->xhost[s] + c->xoff[s][k];
    c->xwaited++;
    return PDX_OK;
}
