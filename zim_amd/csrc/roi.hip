// roi.hip — ROI-restricted reconstruction (startCalculate_v2_roi and the
// right search window of prepareRectifiedROI,
// JSLAB_PhaseEpipolarReconstructor_V2.cpp:629-669, 693-914).
//
// The left crop left_rect = {x, y, w, h} is matched against the right crop
// right_rect = {x', y, w', h} of the same rows, read in place from the
// full-size maps: per crop row the first/last right-crop column of every
// fringe order (LDS-free atomics, as the full-image span table), then one
// thread per masked left pixel. The edge test uses right-crop columns and the
// disparity global columns, (x + j) - (x' + best) - delta, as the reference
// does. The rest of the chain runs on crop-sized buffers with the plain
// kernels, and the ROR stage keeps the reference's fallback: when ROR leaves
// fewer than 10 % of the points, the pre-ROR mask and the Gaussian points are
// used (counted and selected on the device).

#include <climits>

#include "psgc_device.h"
#include "psgc_internal.h"

namespace psgc {
namespace {

constexpr int RB = 16;

dim3 grid_of(int rows, int cols) { return dim3((cols + RB - 1) / RB, (rows + RB - 1) / RB); }

struct R1v {
  double r[9];
};

__global__ void roi_needed_kernel(const int32_t* __restrict__ lk, int cols, int lx0, int y0,
                                  int lw, int h, const uint8_t* __restrict__ mask, int pha_f,
                                  int32_t* __restrict__ needed) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= h || x >= lw || mask[(size_t)r * lw + x] == 0) return;
  const int k = lk[(size_t)(y0 + r) * cols + lx0 + x];
  if (k > 0 && k <= pha_f) needed[k] = 1;  // idempotent
}

__global__ void roi_window_kernel(const int32_t* __restrict__ rk, int rows, int cols, int y0,
                                  int h, int pha_f, const int32_t* __restrict__ needed,
                                  int32_t* __restrict__ mnmx) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  const int y = y0 + r;
  if (r >= h || x >= cols || y < 0 || y >= rows) return;
  const int k = rk[(size_t)y * cols + x];
  if (k > 0 && k <= pha_f && needed[k]) {
    atomicMin(mnmx, x);
    atomicMax(mnmx + 1, x);
  }
}

__global__ void roi_span_init_kernel(int32_t* __restrict__ span, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    span[2 * i] = INT_MAX;
    span[2 * i + 1] = -1;
  }
}

// Per crop row and order, the first and last right-crop column of that order.
__global__ void roi_span_kernel(const int32_t* __restrict__ rk, int cols, int rx0, int y0,
                                int rw, int h, int knum, int32_t* __restrict__ span) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= h || x >= rw) return;
  const int k = rk[(size_t)(y0 + r) * cols + rx0 + x];
  if (k >= 0 && k <= knum) {
    int32_t* s = span + 2 * ((size_t)r * (knum + 1) + k);
    atomicMin(s, x);
    atomicMax(s + 1, x);
  }
}

__global__ void roi_match_kernel(const double* __restrict__ lp, const double* __restrict__ rp,
                                 const int32_t* __restrict__ lk, const int32_t* __restrict__ rk,
                                 int cols, int lx0, int rx0, int y0, int lw, int rw, int h,
                                 const uint8_t* __restrict__ mask, int pha_f, double grad_thr,
                                 const int32_t* __restrict__ span, double* __restrict__ disp64,
                                 float* __restrict__ disp32) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= h || j >= lw) return;
  const size_t o = (size_t)r * lw + j;
  double d = 0.0;
  if (mask[o] != 0) {
    const size_t grow = (size_t)(y0 + r) * cols;
    const int k = lk[grow + lx0 + j];
    if (k > 0 && k <= pha_f) {
      const int32_t* s = span + 2 * ((size_t)r * (pha_f + 1) + k);
      const double* rrow = rp + grow + rx0;
      const int32_t* krow = rk + grow + rx0;
      const double lv = lp[grow + lx0 + j];
      double min_cost = 1e9;
      int best = -1;
      for (int c = s[0]; c <= s[1]; ++c) {
        if (krow[c] != k) continue;
        const double diff = fabs(rrow[c] - lv);
        if (diff < min_cost) {
          min_cost = diff;
          best = c;
        }
      }
      if (best >= 0 && min_cost < grad_thr && best >= 2 && best < rw - 1)
        d = (double)(lx0 + j) - (double)(rx0 + best) - subpixel_delta(rrow, best, lv);
    }
  }
  disp64[o] = d;
  disp32[o] = (float)d;
}

// Triangulation of the masked crop pixels, coefficients at global pixels.
__global__ void roi_poly_kernel(const double* __restrict__ disp64, const float* __restrict__ disp32,
                                const uint8_t* __restrict__ mask, int cols, int lx0, int y0,
                                int lw, int h, const float4* __restrict__ cx,
                                const float4* __restrict__ cy, const float4* __restrict__ cz,
                                R1v r1, double min_depth, double max_depth, int use_rect,
                                float* __restrict__ xyz, uint8_t* __restrict__ P) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= h || j >= lw) return;
  const size_t o = (size_t)r * lw + j;
  float out[3] = {0.f, 0.f, 0.f};
  uint8_t v = 0;
  if (mask[o] != 0) {
    const double d = disp64 ? disp64[o] : (double)disp32[o];
    const size_t g = (size_t)(y0 + r) * cols + lx0 + j;
    v = reproject_poly(cx[g], cy[g], cz[g], r1.r, d, min_depth, max_depth, use_rect, out);
  }
  xyz[o * 3 + 0] = out[0];
  xyz[o * 3 + 1] = out[1];
  xyz[o * 3 + 2] = out[2];
  P[o] = v;
}

__global__ void roi_count_kernel(const uint8_t* __restrict__ a, const uint8_t* __restrict__ b,
                                 size_t n, unsigned long long* __restrict__ cnt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool va = i < n && a[i] != 0, vb = i < n && b[i] != 0;
  const unsigned long long ma = __ballot(va), mb = __ballot(vb);
  if ((threadIdx.x & 63) == 0) {
    if (ma) atomicAdd(cnt, (unsigned long long)__popcll(ma));
    if (mb) atomicAdd(cnt + 1, (unsigned long long)__popcll(mb));
  }
}

// cnt[0] = points before ROR, cnt[1] = after; fewer than 10 %: keep the
// Gaussian points and the pre-ROR mask.
__global__ void roi_select_kernel(const unsigned long long* __restrict__ cnt,
                                  const float* __restrict__ xyz_g, const uint8_t* __restrict__ P_in,
                                  const float* __restrict__ xyz_r, const uint8_t* __restrict__ P_r,
                                  size_t n, float* __restrict__ xyz, uint8_t* __restrict__ P) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long raw = cnt[0], kept = cnt[1];
  const bool fallback = raw > 0 && (double)kept < (double)raw * 0.1;
  const float* src = fallback ? xyz_g : xyz_r;
  xyz[i * 3 + 0] = src[i * 3 + 0];
  xyz[i * 3 + 1] = src[i * 3 + 1];
  xyz[i * 3 + 2] = src[i * 3 + 2];
  P[i] = fallback ? P_in[i] : P_r[i];
}

}  // namespace

int launch_roi_right_window(const int32_t* lk, const int32_t* rk, int rows, int cols, int pha_f,
                            const int32_t* lr, const uint8_t* mask, int32_t* needed,
                            int32_t* mnmx, hipStream_t s) {
  // needed[0..pha_f] and mnmx[2] are initialised by the caller
  roi_needed_kernel<<<grid_of(lr[3], lr[2]), dim3(RB, RB), 0, s>>>(lk, cols, lr[0], lr[1], lr[2],
                                                                   lr[3], mask, pha_f, needed);
  PSGC_LAUNCHED();
  roi_window_kernel<<<grid_of(lr[3], cols), dim3(RB, RB), 0, s>>>(rk, rows, cols, lr[1], lr[3],
                                                                  pha_f, needed, mnmx);
  PSGC_LAUNCHED();
  return PSGC_OK;
}

int launch_roi_match(const double* lp, const double* rp, const int32_t* lk, const int32_t* rk,
                     int cols, const int32_t* lr, const int32_t* rr, const uint8_t* mask,
                     int pha_f, double grad_thr, int32_t* span, double* disp64, float* disp32,
                     hipStream_t s) {
  const size_t nspan = (size_t)lr[3] * (pha_f + 1);
  PSGC_PROF_BEGIN("match", s);
  roi_span_init_kernel<<<(unsigned)((nspan + 255) / 256), 256, 0, s>>>(span, nspan);
  PSGC_LAUNCHED();
  roi_span_kernel<<<grid_of(lr[3], rr[2]), dim3(RB, RB), 0, s>>>(rk, cols, rr[0], lr[1], rr[2],
                                                                 lr[3], pha_f, span);
  PSGC_LAUNCHED();
  roi_match_kernel<<<grid_of(lr[3], lr[2]), dim3(RB, RB), 0, s>>>(
      lp, rp, lk, rk, cols, lr[0], rr[0], lr[1], lr[2], rr[2], lr[3], mask, pha_f, grad_thr, span,
      disp64, disp32);
  PSGC_LAUNCHED();
  PSGC_PROF_END("match", s);
  return PSGC_OK;
}

int launch_roi_poly(const double* disp64, const float* disp32, const uint8_t* mask, int cols,
                    const int32_t* lr, const float* const* coef, const double* R1,
                    double min_depth, double max_depth, int use_rect, float* xyz, uint8_t* P,
                    hipStream_t s) {
  R1v r1v;
  for (int i = 0; i < 9; ++i) r1v.r[i] = R1[i];
  PSGC_PROF_BEGIN("reproject", s);
  roi_poly_kernel<<<grid_of(lr[3], lr[2]), dim3(RB, RB), 0, s>>>(
      disp64, disp32, mask, cols, lr[0], lr[1], lr[2], lr[3], (const float4*)coef[0],
      (const float4*)coef[1], (const float4*)coef[2], r1v, min_depth, max_depth, use_rect, xyz,
      P);
  PSGC_LAUNCHED();
  PSGC_PROF_END("reproject", s);
  return PSGC_OK;
}

int launch_roi_ror_select(const float* xyz_g, const uint8_t* P_in, const float* xyz_r,
                          const uint8_t* P_r, size_t n, unsigned long long* cnt, float* xyz,
                          uint8_t* P, hipStream_t s) {
  // cnt[0..1] are zeroed by the caller
  const unsigned g = (unsigned)((n + 255) / 256);
  roi_count_kernel<<<g, 256, 0, s>>>(P_in, P_r, n, cnt);
  PSGC_LAUNCHED();
  roi_select_kernel<<<g, 256, 0, s>>>(cnt, xyz_g, P_in, xyz_r, P_r, n, xyz, P);
  PSGC_LAUNCHED();
  return PSGC_OK;
}

}  // namespace psgc
