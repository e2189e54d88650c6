// Batched persistent launch (placeholder until the persistent kernel lands).
#include "tca_internal.hpp"

using tca::fail;

struct tca_batch {
    int n_units = 0;
};

extern "C" tca_status tca_batch_create(int n_items, const tca_batch_item *items, tca_batch **batch)
{
    (void)n_items;
    (void)items;
    if (batch) *batch = nullptr;
    return fail(TCA_ERR_UNSUPPORTED, "tca_batch_create: not implemented yet");
}
extern "C" tca_status tca_batch_run(tca_batch *batch, void *stream)
{
    (void)batch;
    (void)stream;
    return fail(TCA_ERR_UNSUPPORTED, "tca_batch_run: not implemented yet");
}
extern "C" tca_status tca_batch_destroy(tca_batch *batch)
{
    delete batch;
    return TCA_OK;
}
extern "C" tca_status tca_batch_get_info(const tca_batch *batch, int *n_units, int *n_workgroups)
{
    (void)batch;
    if (n_units) *n_units = 0;
    if (n_workgroups) *n_workgroups = 0;
    return fail(TCA_ERR_UNSUPPORTED, "tca_batch_get_info: not implemented yet");
}
