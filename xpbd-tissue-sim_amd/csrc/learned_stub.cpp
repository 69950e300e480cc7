// Temporary entry points for the learned warm start and the autotuner.
#include "capi_internal.h"

namespace xpbd {
void destroy_learned(xpbd_handle*) {}
}  // namespace xpbd

extern "C" {
int xpbd_mlp_create(xpbd_mlp**) { return XPBD_E_UNSUPPORTED; }
int xpbd_mlp_destroy(xpbd_mlp*) { return XPBD_OK; }
int xpbd_mlp_num_params(int32_t* out) { if (out) *out = 20099; return XPBD_OK; }
int xpbd_mlp_set_params(xpbd_mlp*, const float*) { return XPBD_E_UNSUPPORTED; }
int xpbd_mlp_get_params(xpbd_mlp*, float*) { return XPBD_E_UNSUPPORTED; }
int xpbd_mlp_set_scaling(xpbd_mlp*, const float*, const float*, float) { return XPBD_E_UNSUPPORTED; }
int xpbd_mlp_forward(xpbd_mlp*, int32_t, const float*, float*) { return XPBD_E_UNSUPPORTED; }
int xpbd_mlp_loss_grad(xpbd_mlp*, int32_t, const float*, const float*, double*, float*) { return XPBD_E_UNSUPPORTED; }
int xpbd_mlp_train(xpbd_mlp*, int32_t, const float*, const float*, int32_t, int32_t, float, uint64_t, double*) { return XPBD_E_UNSUPPORTED; }
int xpbd_collect_training(xpbd_handle*, int32_t, int32_t, float*, float*, int32_t*) { return XPBD_E_UNSUPPORTED; }
int xpbd_features(xpbd_handle*, float*) { return XPBD_E_UNSUPPORTED; }
int xpbd_attach_mlp(xpbd_handle*, xpbd_mlp*) { return XPBD_E_UNSUPPORTED; }
int xpbd_step_learned(xpbd_handle*, int32_t, double, double, int32_t*, double*) { return XPBD_E_UNSUPPORTED; }
int xpbd_autotune(xpbd_batch*, const char*) { return XPBD_E_UNSUPPORTED; }
}
