// Issue-rate probe: the VALU ceilings of the device the filters run on.
//
// The filter kernels (MLS, SOR, the median network) are bound by instruction
// issue, not by HBM. Their measured instruction rates are set against what a
// kernel made of nothing but fused multiply-adds issues on the same device:
// eight independent FMA chains per lane, every wave of a grid that fills the
// device, no packed instructions (packed FP32 is disabled in this build).
#include "psgc_internal.h"

namespace psgc {
namespace {

constexpr int kProbeIters = 4000;
constexpr int kProbeFmasPerIter = 32;

template <typename T>
__device__ inline T fma_t(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

template <typename T>
__global__ void __launch_bounds__(256) valu_fma_kernel(float* out, int iters, T a, T b) {
  T x0 = (T)threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5,
    x6 = x0 + 6, x7 = x0 + 7;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < kProbeFmasPerIter / 8; ++r) {
      x0 = fma_t(x0, a, b);
      x1 = fma_t(x1, a, b);
      x2 = fma_t(x2, a, b);
      x3 = fma_t(x3, a, b);
      x4 = fma_t(x4, a, b);
      x5 = fma_t(x5, a, b);
      x6 = fma_t(x6, a, b);
      x7 = fma_t(x7, a, b);
    }
  }
  // one store per thread keeps every chain live
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = (float)(x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7);
}

// Wave-instructions per second of one timed launch (after an untimed one).
template <typename T>
int time_fma(psgc_ctx* ctx, float* out, int blocks, double* ginst_s) {
  hipEvent_t a = nullptr, b = nullptr;
  PSGC_HIP(hipEventCreate(&a));
  int rc = PSGC_OK;
  float ms = 0.f;
  if (hipEventCreate(&b) != hipSuccess) rc = PSGC_E_HIP;
  for (int rep = 0; rep < 2 && rc == PSGC_OK; ++rep) {
    if (rep == 1 && hipEventRecord(a, ctx->stream) != hipSuccess) rc = PSGC_E_HIP;
    valu_fma_kernel<T><<<blocks, 256, 0, ctx->stream>>>(out, kProbeIters, (T)0.999, (T)0.001);
    if (hipGetLastError() != hipSuccess) rc = PSGC_E_HIP;
  }
  if (rc == PSGC_OK && (hipEventRecord(b, ctx->stream) != hipSuccess ||
                        hipEventSynchronize(b) != hipSuccess ||
                        hipEventElapsedTime(&ms, a, b) != hipSuccess))
    rc = PSGC_E_HIP;
  hipEventDestroy(a);
  if (b) hipEventDestroy(b);
  if (rc != PSGC_OK) PSGC_FAIL(rc, "VALU probe launch or timing failed");
  const double waves = (double)blocks * 4.0;
  *ginst_s = waves * kProbeIters * kProbeFmasPerIter / (ms * 1e-3) / 1e9;
  return PSGC_OK;
}

}  // namespace
}  // namespace psgc

using namespace psgc;

extern "C" int psgc_probe_valu_rate(psgc_ctx* ctx, double* fp32_ginst_s, double* fp64_ginst_s) {
  PSGC_CTX_ENTER(ctx);
  if (!fp32_ginst_s || !fp64_ginst_s) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  // 64 workgroups per CU: many times the resident slots, so every SIMD
  // stays full until the tail
  const int blocks = (ctx->cu_count > 0 ? ctx->cu_count : 256) * 64;
  void* out = nullptr;
  PSGC_TRY(ctx_buf(ctx, S_TMP0, (size_t)blocks * 256 * sizeof(float), &out));
  PSGC_TRY(time_fma<float>(ctx, (float*)out, blocks, fp32_ginst_s));
  PSGC_TRY(time_fma<double>(ctx, (float*)out, blocks, fp64_ginst_s));
  return PSGC_OK;
}
