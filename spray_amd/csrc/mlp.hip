// Domain-hit predictor (learned schedule): placeholder until the MLP kernels land.
#include "ctx.h"
#include "tracer_dev.h"

namespace spray {
struct Predictor {};
void predictorRelease(Ctx* c) {
  delete c->pred;
  c->pred = nullptr;
}
__global__ void kPhaseNone(TraceParams P) {
  uint32_t n = P.ctr->q_total[0];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    P.phase_b[i] = 0;
}
void predictorPhase(Ctx* c, const TraceParams& P) { kPhaseNone<<<1024, 256, 0, c->stream>>>(P); }
}  // namespace spray

using namespace spray;
extern "C" {
int spray_predictor_log(spray_ctx*, int, long long) { return SPRAY_ERR_STATE; }
long long spray_predictor_num_samples(const spray_ctx*) { return 0; }
int spray_predictor_train(spray_ctx*, int, int, float, float*) { return SPRAY_ERR_STATE; }
int spray_predictor_eval(spray_ctx*, float*, float*) { return SPRAY_ERR_STATE; }
int spray_predictor_save(spray_ctx*, const char*) { return SPRAY_ERR_STATE; }
int spray_predictor_load(spray_ctx*, const char*) { return SPRAY_ERR_STATE; }
}
