// pending.hip -- entry points whose kernels are not built yet: they fail loudly.
#include "common.hpp"
using namespace tnx;
extern "C" {
int tnx_mlp_forward(const tnx_mlp_t *, const float *, int64_t, float *, void *) { set_error("not built"); return TNX_E_RANGE; }
int tnx_mlp_backward(const tnx_mlp_t *, const float *, const float *, const float *, int64_t, float *, float *, void *) { set_error("not built"); return TNX_E_RANGE; }
int tnx_mlp_adam(const tnx_mlp_t *, float *, const float *, float *, float *, int64_t, float, float, float, float, int32_t, void *) { set_error("not built"); return TNX_E_RANGE; }
}
