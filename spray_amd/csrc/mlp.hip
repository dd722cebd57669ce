// Domain-hit predictor (learned schedule): placeholder until the MLP kernels land.
#include "ctx.h"

namespace spray {
struct Predictor {};
void predictorRelease(Ctx* c) {
  delete c->pred;
  c->pred = nullptr;
}
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
