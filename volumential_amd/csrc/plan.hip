// Batched list-1 evaluation plans (tile schedule + dense / split forms).
#include "internal.h"

using namespace vl1;

struct vl1_plan {
  int dummy = 0;
};

extern "C" {

int vl1_plan_create(vl1_plan** out, const vl1_payload* p, const vl1_plan_desc* d, void* stream) {
  (void)p;
  (void)d;
  (void)stream;
  if (out) *out = nullptr;
  return fail(VL1_E_UNSUPPORTED, "vl1_plan_create: not built yet");
}

int vl1_plan_destroy(vl1_plan* plan) {
  delete plan;
  return VL1_OK;
}

int vl1_plan_eval(vl1_plan* plan, int form, const double* f, int64_t ldf, int n_rhs, double* y,
                  int64_t ldy, const vl1_split_opts* opts, void* stream) {
  (void)plan; (void)form; (void)f; (void)ldf; (void)n_rhs; (void)y; (void)ldy; (void)opts;
  (void)stream;
  return fail(VL1_E_UNSUPPORTED, "vl1_plan_eval: not built yet");
}

int vl1_plan_stats(const vl1_plan* plan, int64_t* out5) {
  (void)plan;
  (void)out5;
  return fail(VL1_E_UNSUPPORTED, "vl1_plan_stats: not built yet");
}

}  // extern "C"
