// nq_mlp.hip -- cost-model kernels (placeholder until the MLP lands)
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/nq.h"
extern "C" {
int nq_ctx_features(nq_ctx *, float *) { return NQ_EUNSUPPORTED; }
int nq_mlp_forward(nq_ctx *, const float *, const float *, uint64_t, float *) { return NQ_EUNSUPPORTED; }
int nq_mlp_loss_grad(nq_ctx *, const float *, const float *, const float *, uint64_t, float *, float *) { return NQ_EUNSUPPORTED; }
int nq_mlp_train(nq_ctx *, float *, const float *, const float *, uint64_t, const uint32_t *, int, int, float, float *) { return NQ_EUNSUPPORTED; }
}
