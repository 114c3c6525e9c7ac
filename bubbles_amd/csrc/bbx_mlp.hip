// bbx_mlp.hip -- learned PCISPH pressure warm start (placeholder until the
// MFMA kernels land): every entry point reports that the MLP is unavailable.
#include "bbx_common.h"

struct bbx_mlp {
    int dummy;
};

namespace bbx {
int mlp_infer_device(bbx_mlp *, const float *, int, int, float *, hipStream_t) {
    return fail(BBX_E_INVALID, "MLP not built");
}
}  // namespace bbx

using namespace bbx;

extern "C" {
int bbx_mlp_create(int, bbx_mlp **out) {
    if (out) *out = nullptr;
    return fail(BBX_E_INVALID, "MLP not built");
}
int bbx_mlp_destroy(bbx_mlp *) { return BBX_OK; }
int bbx_mlp_param_count(int *count) {
    if (count) *count = 0;
    return fail(BBX_E_INVALID, "MLP not built");
}
int bbx_mlp_get_params(bbx_mlp *, float *) { return fail(BBX_E_INVALID, "MLP not built"); }
int bbx_mlp_set_params(bbx_mlp *, const float *) { return fail(BBX_E_INVALID, "MLP not built"); }
int bbx_mlp_infer(bbx_mlp *, const float *, int, float *) { return fail(BBX_E_INVALID, "MLP not built"); }
int bbx_mlp_gradients(bbx_mlp *, const float *, const float *, int, float *, float *) {
    return fail(BBX_E_INVALID, "MLP not built");
}
int bbx_mlp_train_step(bbx_mlp *, const float *, const float *, int, float, float *) {
    return fail(BBX_E_INVALID, "MLP not built");
}
int bbx_mlp_train(bbx_mlp *, const float *, const float *, int, int, int, float, int, float *) {
    return fail(BBX_E_INVALID, "MLP not built");
}
}
