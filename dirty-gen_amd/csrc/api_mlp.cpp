// C ABI of the tire sub-grid network (include/silt_mlp.h).
#include "silt_mlp.h"

#include <hip/hip_runtime.h>

#include "mlp.hpp"
#include "silt_sheet.h"

static_assert(silt::kMlpParams == SILT_MLP_PARAMS, "parameter count");
static_assert(silt::kMlpIn == SILT_MLP_IN && silt::kMlpOut == SILT_MLP_OUT, "shape");

struct silt_mlp {
  silt::Mlp m;
  float* x = nullptr;  // staging for host inputs (FP32)
  size_t cap = 0;
};

namespace {

int stage(silt_mlp* h, const float* x, const float* t, size_t n) {
  if (n > h->cap) {
    if (h->x) (void)hipFree(h->x);
    h->x = nullptr;
    h->cap = 0;
    if (hipMalloc(&h->x, n * SILT_MLP_IN * sizeof(float)) != hipSuccess) return SILT_ENOMEM;
    h->cap = n;
  }
  int rc = silt::mlp_reserve(h->m, n);
  if (rc) return rc;
  if (hipMemcpyAsync(h->x, x, n * SILT_MLP_IN * sizeof(float), hipMemcpyHostToDevice, h->m.st) != hipSuccess)
    return SILT_EDEVICE;
  if (t && hipMemcpyAsync(h->m.tg, t, n * SILT_MLP_OUT * sizeof(float), hipMemcpyHostToDevice, h->m.st) != hipSuccess)
    return SILT_EDEVICE;
  silt::mlp_load_x(h->m, h->x, n);
  return 0;
}

}  // namespace

extern "C" {

silt_mlp* silt_mlp_create(unsigned seed) {
  silt_mlp* h = new silt_mlp();
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return nullptr;
  }
  if (silt::mlp_init(h->m, 256, seed, st) != 0 || hipStreamSynchronize(st) != hipSuccess) {
    silt_mlp_destroy(h);
    return nullptr;
  }
  return h;
}

void silt_mlp_destroy(silt_mlp* h) {
  if (!h) return;
  hipStream_t st = h->m.st;
  silt::mlp_free(h->m);
  if (h->x) (void)hipFree(h->x);
  if (st) (void)hipStreamDestroy(st);
  delete h;
}

int silt_mlp_get_params(const silt_mlp* h, float* params) {
  if (!h || !params) return SILT_EINVAL;
  if (hipMemcpyAsync(params, h->m.master, SILT_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, h->m.st) !=
          hipSuccess ||
      hipStreamSynchronize(h->m.st) != hipSuccess)
    return SILT_EDEVICE;
  return SILT_OK;
}

int silt_mlp_set_params(silt_mlp* h, const float* params) {
  if (!h || !params) return SILT_EINVAL;
  if (hipMemcpyAsync(h->m.master, params, SILT_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, h->m.st) !=
      hipSuccess)
    return SILT_EDEVICE;
  silt::mlp_sync_bf16(h->m);
  return hipStreamSynchronize(h->m.st) == hipSuccess ? SILT_OK : SILT_EDEVICE;
}

int silt_mlp_forward(silt_mlp* h, const float* x, size_t n, float* y) {
  if (!h || !x || !y || n == 0) return SILT_EINVAL;
  int rc = stage(h, x, nullptr, n);
  if (rc) return rc;
  silt::mlp_forward(h->m, n);
  if (hipMemcpyAsync(y, h->m.y, n * SILT_MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, h->m.st) != hipSuccess ||
      hipStreamSynchronize(h->m.st) != hipSuccess)
    return SILT_EDEVICE;
  return SILT_OK;
}

int silt_mlp_infer(silt_mlp* h, const float* x, size_t n, float* y) {
  if (!h || !x || !y || n == 0) return SILT_EINVAL;
  int rc = stage(h, x, nullptr, n);
  if (rc) return rc;
  silt::mlp_infer(h->m, n);
  if (hipMemcpyAsync(y, h->m.y, n * SILT_MLP_OUT * sizeof(float), hipMemcpyDeviceToHost, h->m.st) != hipSuccess ||
      hipStreamSynchronize(h->m.st) != hipSuccess)
    return SILT_EDEVICE;
  return SILT_OK;
}

int silt_mlp_gradients(silt_mlp* h, const float* x, const float* t, size_t n, double* sse, float* grad) {
  if (!h || !x || !t || n == 0) return SILT_EINVAL;
  int rc = stage(h, x, t, n);
  if (rc) return rc;
  silt::mlp_forward(h->m, n);
  silt::mlp_backward(h->m, n);
  if (grad && hipMemcpyAsync(grad, h->m.grad, SILT_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, h->m.st) !=
                  hipSuccess)
    return SILT_EDEVICE;
  double s = 0.0;
  if (hipMemcpyAsync(&s, h->m.loss, sizeof s, hipMemcpyDeviceToHost, h->m.st) != hipSuccess ||
      hipStreamSynchronize(h->m.st) != hipSuccess)
    return SILT_EDEVICE;
  if (sse) *sse = s;
  return SILT_OK;
}

int silt_mlp_train_batch(silt_mlp* h, const float* x, const float* t, size_t n, float lr, double* sse) {
  if (!h || !x || !t || n == 0 || !(lr > 0.0f)) return SILT_EINVAL;
  int rc = stage(h, x, t, n);
  if (rc) return rc;
  silt::mlp_forward(h->m, n);
  silt::mlp_backward(h->m, n);
  silt::mlp_adam(h->m, lr);
  double s = 0.0;
  if (hipMemcpyAsync(&s, h->m.loss, sizeof s, hipMemcpyDeviceToHost, h->m.st) != hipSuccess ||
      hipStreamSynchronize(h->m.st) != hipSuccess)
    return SILT_EDEVICE;
  if (sse) *sse = s;
  return SILT_OK;
}

}  // extern "C"
