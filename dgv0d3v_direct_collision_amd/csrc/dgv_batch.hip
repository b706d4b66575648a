// dgv_batch.hip -- batched evaluation of many VDFs (placeholder entry points;
// filled in below as the batched kernels land).
#include "dgv_internal.h"

extern "C" {
int dgv_batch_create(dgv_ctx* const*, int32_t, const int32_t*, int32_t, int32_t, dgv_batch** out) {
  if (out) *out = nullptr;
  dgv::set_global_error("batched evaluation not built yet");
  return DGV_ERR_UNSUPPORTED;
}
void dgv_batch_destroy(dgv_batch*) {}
int dgv_batch_layout(const dgv_batch*, int32_t, int64_t*, int64_t*, int64_t*) { return DGV_ERR_UNSUPPORTED; }
int dgv_batch_set_f(dgv_batch*, int32_t, const double*, int64_t) { return DGV_ERR_UNSUPPORTED; }
int dgv_batch_get_q(dgv_batch*, int32_t, double*, int64_t) { return DGV_ERR_UNSUPPORTED; }
int dgv_batch_eval_q(dgv_batch*) { return DGV_ERR_UNSUPPORTED; }
int dgv_batch_rk_run(dgv_batch*, double, double, int32_t, int64_t, int) { return DGV_ERR_UNSUPPORTED; }
int dgv_batch_autotune(dgv_batch*, int32_t*, double*) { return DGV_ERR_UNSUPPORTED; }
int dgv_batch_device_buffers(dgv_batch*, double**, double**) { return DGV_ERR_UNSUPPORTED; }
int dgv_batch_time_eval(dgv_batch*, int32_t, double*, double*) { return DGV_ERR_UNSUPPORTED; }
}
