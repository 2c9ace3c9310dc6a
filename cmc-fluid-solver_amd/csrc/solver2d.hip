// 2D ADI (placeholder until the 2D kernels land).
#include "cmcx_internal.h"

struct cmcx2d_solver {
    int dummy;
};

extern "C" {
int cmcx2d_create(const cmcx2d_domain*, const cmcx_params*, int, void*, cmcx2d_solver** out) {
    if (out) *out = nullptr;
    cmcx::set_error("2D solver not built yet");
    return CMCX_ERR_UNSUPPORTED;
}
int cmcx2d_update_boundaries(cmcx2d_solver*) { return CMCX_ERR_UNSUPPORTED; }
int cmcx2d_step(cmcx2d_solver*, double, int, int, double*, int*) { return CMCX_ERR_UNSUPPORTED; }
int cmcx2d_get_fields(cmcx2d_solver*, void*, void*, void*, void*) { return CMCX_ERR_UNSUPPORTED; }
int cmcx2d_destroy(cmcx2d_solver*) { return CMCX_ERR_UNSUPPORTED; }
}
