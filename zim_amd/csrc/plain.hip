// plain.hip — the per-pair, per-stage kernels: one thread per pixel in 16x16
// workgroups, operands read straight from HBM, one launch per stage. This is
// the structure of the reference GPU path (JSLAB_PhaseShiftDecoder_V2_GPU.cu,
// JSLAB_PhaseEpipolarReconstructor_V2_GPU.cu, JSLAB_MLS_Filter.cu) kept as
// the "plain per-pair GPU path" that the batched pipeline is measured
// against; the arithmetic follows the reference CPU path
// (JSLAB_PhaseShiftDecoder_V2.cpp, JSLAB_PhaseEpipolarReconstructor_V2.cpp,
// JSLAB_MLS_Filter.cpp) operation for operation so that integer outputs are
// bit-identical to the CPU restatement.
//
// Built with -ffp-contract=off: no mul+add pair may be fused.

#include <climits>

#include "psgc_internal.h"
#include "psgc_math.h"
#include "psgc_device.h"

namespace psgc {
namespace {

constexpr int TB = 16;

inline dim3 grid2d(int rows, int cols) {
  return dim3((cols + TB - 1) / TB, (rows + TB - 1) / TB);
}

// ------------------------------------------------------------- decode ----

__global__ void decode_plain_kernel(DecodeArgs a, double* __restrict__ phase_raw,
                                    int32_t* __restrict__ kmap,
                                    double* __restrict__ modmap,
                                    uint8_t* __restrict__ validmap) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= a.rows || x >= a.cols) return;
  const size_t p = (size_t)r * a.cols + x;

  double wrap[3] = {0, 0, 0};
  double mod0 = 0.0;
  bool ok = true;
  for (int i = 0; i < a.nlev && ok; ++i) {
    const int n = a.shift_num[i];
    const int base = a.level_base[i];
    double ssum = 0, csum = 0, gsum = 0;
    double m = 0.0;
    if (a.hdr) {
      double beta[3];
      ok = hdr_fit(a, i, r, x, beta);
      if (ok) {
        csum = beta[1] * n / 2.0;
        ssum = beta[2] * n / 2.0;
        gsum = beta[0] * n;
        ok = level_finish(ssum, csum, gsum, n, a.mod_thresh, &wrap[i], &m);
      }
    } else {
      for (int k = 0; k < n; ++k) {
        const double v = blurred_sample(a, base + k, r, x);
        ssum -= a.sin_tab[base + k] * v;
        csum += a.cos_tab[base + k] * v;
        gsum += v;
      }
      ok = level_finish(ssum, csum, gsum, n, a.mod_thresh, &wrap[i], &m);
    }
    if (i == 0) mod0 = m;
  }
  if (modmap) modmap[p] = mod0;
  double ph = 0.0;
  int32_t k = 0;
  if (ok) {
    if (a.nlev == 1) ok = gray_order(a, r, x, wrap[0], &ph, &k);
    else heterodyne(a, wrap, &ph, &k);
  }
  if (!ok) {
    // with HDR the order map starts at -1 (JSLAB_PhaseShiftDecoder_V2.cpp:
    // 199-239), otherwise at 0
    ph = 0.0;
    k = a.hdr ? -1 : 0;
  }
  phase_raw[p] = ph;
  kmap[p] = k;
  if (validmap) validmap[p] = ok ? 255 : 0;
}

// Gray-code period-jump correction against the median of the valid row
// neighbours x-2..x+2 (same rule as the CPU restatement).
__global__ void gray_fix_kernel(int rows, int cols, const double* __restrict__ ph_in,
                                double* __restrict__ ph_out,
                                const int32_t* __restrict__ k_in,
                                int32_t* __restrict__ k_out,
                                const uint8_t* __restrict__ valid) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  const size_t p = (size_t)r * cols + x;
  double ph = ph_in[p];
  int32_t k = k_in[p];
  if (valid[p]) {
    double v[5];
    int n = 0;
    for (int d = -2; d <= 2; ++d) {
      const int xx = x + d;
      if (xx < 0 || xx >= cols || !valid[(size_t)r * cols + xx]) continue;
      v[n++] = ph_in[(size_t)r * cols + xx];
    }
    const double med = select_kth(v, n, n / 2);
    const double jump = round_half_away((med - ph) / (2 * kPi));
    if (jump != 0.0) {
      ph = ph + jump * 2 * kPi;
      k = k + (int32_t)jump;
    }
  }
  ph_out[p] = ph;
  k_out[p] = k;
}

// Post-smooth: 3x3 [1/4 1/2 1/4]^2, replicated border, separable order.
// HDR post-smooth (JSLAB_PhaseShiftDecoder_V2.cpp:282-290): the [1/4,1/2,1/4]^2
// blur of the phase over the blur of the validity weights (1 for valid, 0
// otherwise, floored at 1e-12); invalid pixels become 0.
__global__ void post_smooth_valid_kernel(const double* __restrict__ in,
                                         const uint8_t* __restrict__ valid,
                                         double* __restrict__ out, int rows, int cols) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  const size_t p = (size_t)r * cols + x;
  if (!valid[p]) {
    out[p] = 0.0;
    return;
  }
  const int xl = x > 0 ? x - 1 : 0, xr = x < cols - 1 ? x + 1 : cols - 1;
  const int ru = r > 0 ? r - 1 : 0, rd = r < rows - 1 ? r + 1 : rows - 1;
  const size_t u = (size_t)ru * cols, m = (size_t)r * cols, d = (size_t)rd * cols;
  auto wv = [&](size_t q) { return valid[q] ? 1.0 : 0.0; };
  const double hu = 0.5 * in[u + x] + 0.25 * (in[u + xl] + in[u + xr]);
  const double hm = 0.5 * in[m + x] + 0.25 * (in[m + xl] + in[m + xr]);
  const double hd = 0.5 * in[d + x] + 0.25 * (in[d + xl] + in[d + xr]);
  const double ph = 0.5 * hm + 0.25 * (hu + hd);
  const double wu = 0.5 * wv(u + x) + 0.25 * (wv(u + xl) + wv(u + xr));
  const double wm = 0.5 * wv(m + x) + 0.25 * (wv(m + xl) + wv(m + xr));
  const double wd = 0.5 * wv(d + x) + 0.25 * (wv(d + xl) + wv(d + xr));
  const double w = 0.5 * wm + 0.25 * (wu + wd);
  out[p] = ph / (w > 1e-12 ? w : 1e-12);
}

__global__ void post_smooth_kernel(const double* __restrict__ in, double* __restrict__ out,
                                   int rows, int cols) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  const int xl = x > 0 ? x - 1 : 0, xr = x < cols - 1 ? x + 1 : cols - 1;
  const int ru = r > 0 ? r - 1 : 0, rd = r < rows - 1 ? r + 1 : rows - 1;
  const double* u = in + (size_t)ru * cols;
  const double* m = in + (size_t)r * cols;
  const double* d = in + (size_t)rd * cols;
  const double hu = 0.5 * u[x] + 0.25 * (u[xl] + u[xr]);
  const double hm = 0.5 * m[x] + 0.25 * (m[xl] + m[xr]);
  const double hd = 0.5 * d[x] + 0.25 * (d[xl] + d[xr]);
  out[(size_t)r * cols + x] = 0.5 * hm + 0.25 * (hu + hd);
}

// -------------------------------------------------------------- match ----

__global__ void span_init_kernel(int32_t* __restrict__ span, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    span[2 * i] = INT_MAX;
    span[2 * i + 1] = -1;
  }
}

// Per row and fringe order, first and last right column of that order.
__global__ void span_build_kernel(const int32_t* __restrict__ rk, int rows, int cols,
                                  int knum, int32_t* __restrict__ span) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  const int k = rk[(size_t)r * cols + x];
  if (k >= 0 && k <= knum) {
    int32_t* s = span + 2 * ((size_t)r * (knum + 1) + k);
    atomicMin(s, x);
    atomicMax(s + 1, x);
  }
}

__global__ void match_plain_kernel(const double* __restrict__ lp, const double* __restrict__ rp,
                                   const int32_t* __restrict__ lk,
                                   const int32_t* __restrict__ rk, int rows, int cols,
                                   int pha_f, double grad_thr,
                                   const int32_t* __restrict__ span,
                                   double* __restrict__ disp64, float* __restrict__ disp32,
                                   int32_t* __restrict__ match_col) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || j >= cols) return;
  const size_t row = (size_t)r * cols;
  double d = 0.0;
  int best = -1;
  const int k = lk[row + j];
  if (k > 0 && k <= pha_f) {
    const int32_t* s = span + 2 * ((size_t)r * (pha_f + 1) + k);
    const int lo = s[0], hi = s[1];
    const double lv = lp[row + j];
    double min_cost = 1e9;
    for (int c = lo; c <= hi; ++c) {
      if (rk[row + c] != k) continue;
      const double diff = fabs(rp[row + c] - lv);
      if (diff < min_cost) {
        min_cost = diff;
        best = c;
      }
    }
    if (best >= 0 && min_cost < grad_thr && best >= 2 && best < cols - 1) {
      d = subpixel_disparity(rp + row, best, lv, j);
    } else {
      best = -1;
    }
  }
  if (disp64) disp64[row + j] = d;
  if (disp32) disp32[row + j] = (float)d;
  if (match_col) match_col[row + j] = best;
}

// ------------------------------------------------------ median / reproject --

__global__ void median5_plain_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                     int rows, int cols) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  float v[25];
  int n = 0;
  for (int dy = -2; dy <= 2; ++dy) {
    const int yy = min(max(r + dy, 0), rows - 1);
    for (int dx = -2; dx <= 2; ++dx) {
      const int xx = min(max(x + dx, 0), cols - 1);
      v[n++] = src[(size_t)yy * cols + xx];
    }
  }
  for (int i = 0; i <= 12; ++i) {
    int mi = i;
    for (int t = i + 1; t < 25; ++t)
      if (v[t] < v[mi]) mi = t;
    const float tmp = v[i];
    v[i] = v[mi];
    v[mi] = tmp;
  }
  dst[(size_t)r * cols + x] = v[12];
}

struct QR {
  double q[16];
  double r1[9];
};

__global__ void reproject_plain_kernel(const float* __restrict__ disp32,
                                       const double* __restrict__ disp64, int rows, int cols,
                                       QR m, double min_depth, double max_depth, int use_rect,
                                       float* __restrict__ xyz, uint8_t* __restrict__ P) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= rows || j >= cols) return;
  const size_t p = (size_t)i * cols + j;
  const double d = disp64 ? disp64[p] : (double)disp32[p];
  float o[3];
  const uint8_t v = reproject_point(m.q, m.r1, i, j, d, min_depth, max_depth, use_rect, o);
  xyz[p * 3 + 0] = o[0];
  xyz[p * 3 + 1] = o[1];
  xyz[p * 3 + 2] = o[2];
  P[p] = v;
}

__global__ void reproject_poly_plain_kernel(const float* __restrict__ disp32,
                                            const double* __restrict__ disp64, int rows,
                                            int cols, const float4* __restrict__ cx,
                                            const float4* __restrict__ cy,
                                            const float4* __restrict__ cz, QR m,
                                            double min_depth, double max_depth, int use_rect,
                                            float* __restrict__ xyz, uint8_t* __restrict__ P) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= rows || j >= cols) return;
  const size_t p = (size_t)i * cols + j;
  const double d = disp64 ? disp64[p] : (double)disp32[p];
  float o[3];
  const uint8_t v =
      reproject_poly(cx[p], cy[p], cz[p], m.r1, d, min_depth, max_depth, use_rect, o);
  xyz[p * 3 + 0] = o[0];
  xyz[p * 3 + 1] = o[1];
  xyz[p * 3 + 2] = o[2];
  P[p] = v;
}

// ------------------------------------------------------------ Gaussian ----

struct GaussK {
  float w[49];
};

__global__ void gaussian_plain_kernel(const float* __restrict__ src, const uint8_t* __restrict__ P,
                                      float* __restrict__ dst, uint8_t* __restrict__ dstP,
                                      int rows, int cols, int win, GaussK k2, int semantics) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  const size_t p = (size_t)r * cols + x;
  const int half = (win - 1) / 2;
  float o0 = 0.f, o1 = 0.f, o2 = 0.f;
  uint8_t op = semantics == PSGC_SEM_CPU ? P[p] : 0;
  if (r >= half && r < rows - half && x >= half && x < cols - half && P[p] != 0) {
    const float cx = src[p * 3];
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, cnt = 0.f;
    for (int ik = -half; ik <= half; ++ik)
      for (int jk = -half; jk <= half; ++jk) {
        const size_t q = (size_t)(r + ik) * cols + (x + jk);
        if (P[q] == 0) continue;
        const float qx = src[q * 3];
        if (fabsf(qx - cx) < 3.0f) {
          const float w = k2.w[(ik + half) * win + (jk + half)];
          s0 += qx * w;
          s1 += src[q * 3 + 1] * w;
          s2 += src[q * 3 + 2] * w;
          cnt += w;
        }
      }
    if (cnt > 0.8f) {
      if (semantics == PSGC_SEM_CPU) {
        const float inv = 1.f / cnt;
        o0 = s0 * inv;
        o1 = s1 * inv;
        o2 = s2 * inv;
      } else {
        o0 = s0 / cnt;
        o1 = s1 / cnt;
        o2 = s2 / cnt;
        op = 1;
      }
    }
  }
  dst[p * 3 + 0] = o0;
  dst[p * 3 + 1] = o1;
  dst[p * 3 + 2] = o2;
  dstP[p] = op;
}

// ----------------------------------------------------------------- ROR ----

__global__ void ror_plain_kernel(const float* __restrict__ src, const uint8_t* __restrict__ P,
                                 float* __restrict__ dst, uint8_t* __restrict__ dstP, int rows,
                                 int cols, float r2, int min_nb, int half) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  const size_t p = (size_t)r * cols + x;
  bool keep = false;
  float px = 0.f, py = 0.f, pz = 0.f;
  if (P[p] != 0) {
    px = src[p * 3];
    py = src[p * 3 + 1];
    pz = src[p * 3 + 2];
    int cnt = 0;
    const int r0 = max(0, r - half), r1 = min(rows, r + half + 1);
    const int c0 = max(0, x - half), c1 = min(cols, x + half + 1);
    for (int i = r0; i < r1 && !keep; ++i)
      for (int j = c0; j < c1; ++j) {
        const size_t q = (size_t)i * cols + j;
        if (P[q] == 0 || q == p) continue;
        if (dist2(px, py, pz, src + q * 3) <= r2 && ++cnt >= min_nb) {
          keep = true;
          break;
        }
      }
  }
  dst[p * 3 + 0] = keep ? px : 0.f;
  dst[p * 3 + 1] = keep ? py : 0.f;
  dst[p * 3 + 2] = keep ? pz : 0.f;
  dstP[p] = keep ? 1 : 0;
}

// ----------------------------------------------------------------- SOR ----

constexpr int kSorMaxK = 64;

// Mean distance to the k nearest valid window neighbours; -1 when fewer
// than k candidates exist (or the point is invalid).
__global__ void sor_knn_plain_kernel(const float* __restrict__ src, const uint8_t* __restrict__ P,
                                     int rows, int cols, int k, int half,
                                     double* __restrict__ md) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  const size_t p = (size_t)r * cols + x;
  double out = -1.0;
  if (P[p] != 0) {
    const float px = src[p * 3], py = src[p * 3 + 1], pz = src[p * 3 + 2];
    float best[kSorMaxK];
    int n = 0;
    for (int i = max(0, r - half); i < min(rows, r + half + 1); ++i)
      for (int j = max(0, x - half); j < min(cols, x + half + 1); ++j) {
        const size_t q = (size_t)i * cols + j;
        if (P[q] == 0 || q == p) continue;
        const float d2 = dist2(px, py, pz, src + q * 3);
        if (n < k) {
          int t = n++;
          while (t > 0 && best[t - 1] > d2) {
            best[t] = best[t - 1];
            --t;
          }
          best[t] = d2;
        } else if (d2 < best[k - 1]) {
          int t = k - 1;
          while (t > 0 && best[t - 1] > d2) {
            best[t] = best[t - 1];
            --t;
          }
          best[t] = d2;
        }
      }
    if (n >= k) {
      double s = 0.0;
      for (int m = 0; m < k; ++m) s += sqrt((double)best[m]);
      out = s / k;
    }
  }
  md[p] = out;
}


// ----------------------------------------------------------------- MLS ----

// One thread per pixel, three passes over the (2hw+1)^2 window read from HBM
// (structure of mlsKernel, JSLAB_MLS_Filter.cu:165-391; arithmetic of
// JSLAB_MLS_Filter.cpp:176-389).
// C: the per-neighbour multiply-adds contracted (GPU semantics).
template <bool C>
__global__ void mls_plain_kernel(const float* __restrict__ src, const uint8_t* __restrict__ P,
                                 float* __restrict__ dst, int rows, int cols, int hw,
                                 float inv2s2, int order, float r2cut, int semantics) {
  __shared__ double etab[64];  // 2^(j/64) for exp_f32
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  if (tid < 64) etab[tid] = kExp2Tab64[tid];
  __syncthreads();
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y * blockDim.y + threadIdx.y;
  if (r >= rows || x >= cols) return;
  const size_t p = (size_t)r * cols + x;
  const float pi[3] = {src[p * 3], src[p * 3 + 1], src[p * 3 + 2]};
  float out[3] = {pi[0], pi[1], pi[2]};
  if (P[p] != 0) {
    const int r0 = max(0, r - hw), r1 = min(rows, r + hw + 1);
    const int c0 = max(0, x - hw), c1 = min(cols, x + hw + 1);
    MlsFit f;
    f.cx = f.cy = f.cz = 0.f;
    float wsum = 0.f;
    int nv = 0;
    for (int i = r0; i < r1; ++i)
      for (int j = c0; j < c1; ++j) {
        const size_t q = (size_t)i * cols + j;
        if (P[q] == 0) continue;
        const float* pk = src + q * 3;
        const float d2 = dist2<C>(pi[0], pi[1], pi[2], pk);
        if (d2 > r2cut) continue;
        const float w = exp_f32(-d2 * inv2s2, etab);
        mls_accum_centroid<C>(pk, w, &f, &wsum);
        nv++;
      }
    if (nv >= 3 && !(wsum < 1e-12f)) {
      const float iw = 1.0f / wsum;
      f.cx *= iw;
      f.cy *= iw;
      f.cz *= iw;
      float cov[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int i = r0; i < r1; ++i)
        for (int j = c0; j < c1; ++j) {
          const size_t q = (size_t)i * cols + j;
          if (P[q] == 0) continue;
          const float* pk = src + q * 3;
          const float d2 = dist2<C>(pi[0], pi[1], pi[2], pk);
          if (d2 > r2cut) continue;
          const float w = exp_f32(-d2 * inv2s2, etab);
          mls_accum_cov<C>(f, pk, w, cov);
        }
      mls_frame(cov, &f);
      const int cur = mls_order(order, nv, semantics);
      if (cur == 0) {
        mls_plane(f, pi, out);
      } else {
        const int nc = cur == 1 ? 3 : 6;
        float AtWA[6][6], AtWf[6];
        for (int a = 0; a < 6; ++a) {
          AtWf[a] = 0.f;
          for (int b = 0; b < 6; ++b) AtWA[a][b] = 0.f;
        }
        for (int i = r0; i < r1; ++i)
          for (int j = c0; j < c1; ++j) {
            const size_t q = (size_t)i * cols + j;
            if (P[q] == 0) continue;
            const float* pk = src + q * 3;
            const float d2 = dist2<C>(pi[0], pi[1], pi[2], pk);
            if (d2 > r2cut) continue;
            const float w = exp_f32(-d2 * inv2s2, etab);
            mls_accum_poly<C>(f, pk, w, nc, AtWA, AtWf);
          }
        mls_solve_eval(f, AtWA, AtWf, cur, pi, out);
      }
    }
  }
  dst[p * 3 + 0] = out[0];
  dst[p * 3 + 1] = out[1];
  dst[p * 3 + 2] = out[2];
}

// ------------------------------------------------------ SOR statistics ----
// Deterministic reduction: chunks of 1024 consecutive pixels tree-summed in
// LDS (s[i] += s[i + h], h = 512 .. 1), then the chunk partials tree-summed
// over the next power of two in one workgroup. The CPU restatement sums in
// the same order, so the threshold is bit-identical.

constexpr int kChunk = 1024;

__global__ void __launch_bounds__(256) sor_chunk_kernel(const double* __restrict__ md,
                                                        size_t npix, double* __restrict__ part,
                                                        size_t np2) {
  __shared__ double s[kChunk], q[kChunk], n[kChunk];
  const size_t c = blockIdx.x;
  for (int t = threadIdx.x; t < kChunk; t += blockDim.x) {
    const size_t p = c * kChunk + t;
    const double v = (p < npix) ? md[p] : -1.0;
    const bool ok = v >= 0.0;
    s[t] = ok ? v : 0.0;
    q[t] = ok ? v * v : 0.0;
    n[t] = ok ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int h = kChunk / 2; h >= 1; h /= 2) {
    for (int i = threadIdx.x; i < h; i += blockDim.x) {
      s[i] += s[i + h];
      q[i] += q[i + h];
      n[i] += n[i + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[c] = s[0];
    part[np2 + c] = q[0];
    part[2 * np2 + c] = n[0];
  }
}

__global__ void __launch_bounds__(1024) sor_final_kernel(double* __restrict__ part, size_t nchunks,
                                                         size_t np2, float std_mul,
                                                         double* __restrict__ stats) {
  double* s = part;
  double* q = part + np2;
  double* n = part + 2 * np2;
  for (size_t i = nchunks + threadIdx.x; i < np2; i += blockDim.x) {
    s[i] = 0.0;
    q[i] = 0.0;
    n[i] = 0.0;
  }
  __syncthreads();
  for (size_t h = np2 / 2; h >= 1; h /= 2) {
    for (size_t i = threadIdx.x; i < h; i += blockDim.x) {
      s[i] += s[i + h];
      q[i] += q[i + h];
      n[i] += n[i + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cnt = n[0], sum = s[0], sq = q[0];
    double mean = 0.0, sd = 0.0;
    if (cnt > 0) mean = sum / cnt;
    if (cnt > 1) {
      const double var = (sq - sum * sum / cnt) / (cnt - 1.0);
      sd = sqrt(var > 0.0 ? var : 0.0);
    }
    stats[0] = cnt;
    stats[1] = mean;
    stats[2] = sd;
    stats[3] = mean + (double)std_mul * sd;
  }
}

// All pairs' statistics in two launches: chunk partials over the
// concatenated chunks (a workgroup finds its pair), then one workgroup per
// pair for the final reduction; the same reduction tree as one pair alone.
__global__ void __launch_bounds__(256) sor_chunk_batch_kernel(const SorStatJob* __restrict__ jobs,
                                                              int njobs) {
  int lo = 0, hi = njobs - 1;  // last pair with chunk_base <= blockIdx.x
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].chunk_base <= (int64_t)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const SorStatJob& J = jobs[lo];
  __shared__ double s[kChunk], q[kChunk], n[kChunk];
  const size_t c = blockIdx.x - J.chunk_base;
  for (int t = threadIdx.x; t < kChunk; t += blockDim.x) {
    const size_t p = c * kChunk + t;
    const double v = (p < J.npix) ? J.md[p] : -1.0;
    const bool ok = v >= 0.0;
    s[t] = ok ? v : 0.0;
    q[t] = ok ? v * v : 0.0;
    n[t] = ok ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int h = kChunk / 2; h >= 1; h /= 2) {
    for (int i = threadIdx.x; i < h; i += blockDim.x) {
      s[i] += s[i + h];
      q[i] += q[i + h];
      n[i] += n[i + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    J.part[c] = s[0];
    J.part[J.np2 + c] = q[0];
    J.part[2 * J.np2 + c] = n[0];
  }
}

__global__ void __launch_bounds__(1024) sor_final_batch_kernel(const SorStatJob* __restrict__ jobs) {
  const SorStatJob& J = jobs[blockIdx.x];
  double* s = J.part;
  double* q = J.part + J.np2;
  double* n = J.part + 2 * J.np2;
  for (size_t i = J.nchunks + threadIdx.x; i < J.np2; i += blockDim.x) {
    s[i] = 0.0;
    q[i] = 0.0;
    n[i] = 0.0;
  }
  __syncthreads();
  for (size_t h = J.np2 / 2; h >= 1; h /= 2) {
    for (size_t i = threadIdx.x; i < h; i += blockDim.x) {
      s[i] += s[i + h];
      q[i] += q[i + h];
      n[i] += n[i + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double cnt = n[0], sum = s[0], sq = q[0];
    double mean = 0.0, sd = 0.0;
    if (cnt > 0) mean = sum / cnt;
    if (cnt > 1) {
      const double var = (sq - sum * sum / cnt) / (cnt - 1.0);
      sd = sqrt(var > 0.0 ? var : 0.0);
    }
    J.stats[0] = cnt;
    J.stats[1] = mean;
    J.stats[2] = sd;
    J.stats[3] = mean + (double)J.std_mul * sd;
  }
}

__global__ void sor_apply_kernel(const float* __restrict__ src, double* __restrict__ md,
                                 size_t npix, const double* __restrict__ stats,
                                 float* __restrict__ dst, uint8_t* __restrict__ dstP) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  const double thr = stats[3];
  const double v = md[p];
  const bool keep = v >= 0.0 && v <= thr;
  dst[p * 3 + 0] = keep ? src[p * 3 + 0] : 0.f;
  dst[p * 3 + 1] = keep ? src[p * 3 + 1] : 0.f;
  dst[p * 3 + 2] = keep ? src[p * 3 + 2] : 0.f;
  dstP[p] = keep ? 1 : 0;
  if (v < 0.0) md[p] = 0.0;
}

}  // namespace

// ------------------------------------------------------------ launchers ----

int launch_decode_plain(const DecodeArgs& a, double* phase_raw, int32_t* k, double* mod,
                        uint8_t* valid, hipStream_t s) {
  PSGC_PROF_BEGIN("decode", s);
  decode_plain_kernel<<<grid2d(a.rows, a.cols), dim3(TB, TB), 0, s>>>(a, phase_raw, k, mod,
                                                                       valid);
  PSGC_LAUNCHED();
  PSGC_PROF_END("decode", s);
  return PSGC_OK;
}

int launch_gray_fix(const DecodeArgs& a, const double* phase_in, double* phase_out,
                    const int32_t* k_in, int32_t* k_out, const uint8_t* valid, hipStream_t s) {
  PSGC_PROF_BEGIN("gray_fix", s);
  gray_fix_kernel<<<grid2d(a.rows, a.cols), dim3(TB, TB), 0, s>>>(a.rows, a.cols, phase_in,
                                                                   phase_out, k_in, k_out, valid);
  PSGC_LAUNCHED();
  PSGC_PROF_END("gray_fix", s);
  return PSGC_OK;
}

int launch_post_smooth_valid(const double* in, const uint8_t* valid, double* out, int rows,
                             int cols, hipStream_t s) {
  PSGC_PROF_BEGIN("post_smooth", s);
  post_smooth_valid_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(in, valid, out, rows,
                                                                       cols);
  PSGC_LAUNCHED();
  PSGC_PROF_END("post_smooth", s);
  return PSGC_OK;
}

int launch_post_smooth(const double* in, double* out, int rows, int cols, hipStream_t s) {
  PSGC_PROF_BEGIN("post_smooth", s);
  post_smooth_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(in, out, rows, cols);
  PSGC_LAUNCHED();
  PSGC_PROF_END("post_smooth", s);
  return PSGC_OK;
}

int launch_match_plain(const double* lp, const double* rp, const int32_t* lk, const int32_t* rk,
                       int rows, int cols, int pha_f, double grad_thr, int32_t* span,
                       double* disp64, float* disp32, int32_t* match_col, hipStream_t s) {
  const int knum = pha_f > 0 ? pha_f : 0;
  const size_t nspan = (size_t)rows * (knum + 1);
  PSGC_PROF_BEGIN("match", s);
  span_init_kernel<<<(unsigned)((nspan + 255) / 256), 256, 0, s>>>(span, nspan);
  PSGC_LAUNCHED();
  span_build_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(rk, rows, cols, knum, span);
  PSGC_LAUNCHED();
  match_plain_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(
      lp, rp, lk, rk, rows, cols, pha_f, grad_thr, span, disp64, disp32, match_col);
  PSGC_LAUNCHED();
  PSGC_PROF_END("match", s);
  return PSGC_OK;
}

int launch_median5(const float* src, float* dst, int rows, int cols, hipStream_t s) {
  PSGC_PROF_BEGIN("median5", s);
  median5_plain_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(src, dst, rows, cols);
  PSGC_LAUNCHED();
  PSGC_PROF_END("median5", s);
  return PSGC_OK;
}

int launch_reproject(const float* disp32, const double* disp64, int rows, int cols,
                     const double* QRh, double min_depth, double max_depth, int use_rect,
                     float* xyz, uint8_t* P, hipStream_t s) {
  QR m;
  for (int i = 0; i < 16; ++i) m.q[i] = QRh[i];
  for (int i = 0; i < 9; ++i) m.r1[i] = QRh[16 + i];
  PSGC_PROF_BEGIN("reproject", s);
  reproject_plain_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(
      disp32, disp64, rows, cols, m, min_depth, max_depth, use_rect, xyz, P);
  PSGC_LAUNCHED();
  PSGC_PROF_END("reproject", s);
  return PSGC_OK;
}

int launch_reproject_poly(const float* disp32, const double* disp64, int rows, int cols,
                          const float* const* coef, const double* R1, double min_depth,
                          double max_depth, int use_rect, float* xyz, uint8_t* P,
                          hipStream_t s) {
  QR m;
  for (int i = 0; i < 16; ++i) m.q[i] = 0.0;
  for (int i = 0; i < 9; ++i) m.r1[i] = R1[i];
  PSGC_PROF_BEGIN("reproject", s);
  reproject_poly_plain_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(
      disp32, disp64, rows, cols, (const float4*)coef[0], (const float4*)coef[1],
      (const float4*)coef[2], m, min_depth, max_depth, use_rect, xyz, P);
  PSGC_LAUNCHED();
  PSGC_PROF_END("reproject", s);
  return PSGC_OK;
}

int launch_gaussian(const float* xyz, const uint8_t* P, float* dst, uint8_t* dstP, int rows,
                    int cols, int win, const float* k2_host, int semantics, hipStream_t s) {
  GaussK k;
  for (int i = 0; i < win * win; ++i) k.w[i] = k2_host[i];
  PSGC_PROF_BEGIN("gaussian", s);
  gaussian_plain_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(xyz, P, dst, dstP, rows,
                                                                     cols, win, k, semantics);
  PSGC_LAUNCHED();
  PSGC_PROF_END("gaussian", s);
  return PSGC_OK;
}

int launch_ror(const float* xyz, const uint8_t* P, float* dst, uint8_t* dstP, int rows, int cols,
               float radius, int min_nb, int window, hipStream_t s) {
  PSGC_PROF_BEGIN("ror", s);
  ror_plain_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(
      xyz, P, dst, dstP, rows, cols, radius * radius, min_nb, window / 2);
  PSGC_LAUNCHED();
  PSGC_PROF_END("ror", s);
  return PSGC_OK;
}

int launch_sor(const float* xyz, const uint8_t* P, float* dst, uint8_t* dstP, int rows, int cols,
               int k, float std_mul, int half, double* md, double* partial, double* stats,
               hipStream_t s) {
  if (k < 1 || k > kSorMaxK) PSGC_FAIL(PSGC_E_INVALID, "sor k=%d outside [1, %d]", k, kSorMaxK);
  PSGC_PROF_BEGIN("sor_knn", s);
  sor_knn_plain_kernel<<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(xyz, P, rows, cols, k, half,
                                                                    md);
  PSGC_LAUNCHED();
  PSGC_PROF_END("sor_knn", s);
  return launch_sor_finish(xyz, md, dst, dstP, (size_t)rows * cols, std_mul, partial, stats, s);
}

int launch_sor_finish(const float* xyz, double* md, float* dst, uint8_t* dstP, size_t npix,
                      float std_mul, double* partial, double* stats, hipStream_t s) {
  const size_t nchunks = (npix + kChunk - 1) / kChunk;
  size_t np2 = 1;
  while (np2 < nchunks) np2 *= 2;
  PSGC_PROF_BEGIN("sor_stats", s);
  sor_chunk_kernel<<<(unsigned)nchunks, 256, 0, s>>>(md, npix, partial, np2);
  PSGC_LAUNCHED();
  sor_final_kernel<<<1, 1024, 0, s>>>(partial, nchunks, np2, std_mul, stats);
  PSGC_LAUNCHED();
  sor_apply_kernel<<<(unsigned)((npix + 255) / 256), 256, 0, s>>>(xyz, md, npix, stats, dst, dstP);
  PSGC_LAUNCHED();
  PSGC_PROF_END("sor_stats", s);
  return PSGC_OK;
}

int launch_sor_stats(const double* md, size_t npix, float std_mul, double* partial,
                     double* stats, hipStream_t s) {
  const size_t nchunks = (npix + kChunk - 1) / kChunk;
  size_t np2 = 1;
  while (np2 < nchunks) np2 *= 2;
  PSGC_PROF_BEGIN("sor_stats", s);
  sor_chunk_kernel<<<(unsigned)nchunks, 256, 0, s>>>(md, npix, partial, np2);
  PSGC_LAUNCHED();
  sor_final_kernel<<<1, 1024, 0, s>>>(partial, nchunks, np2, std_mul, stats);
  PSGC_LAUNCHED();
  PSGC_PROF_END("sor_stats", s);
  return PSGC_OK;
}

int launch_sor_stats_batch(const SorStatJob* jobs_dev, int njobs, int64_t total_chunks,
                           hipStream_t s) {
  if (njobs <= 0 || total_chunks <= 0) return PSGC_OK;
  PSGC_PROF_BEGIN("sor_stats", s);
  sor_chunk_batch_kernel<<<(unsigned)total_chunks, 256, 0, s>>>(jobs_dev, njobs);
  PSGC_LAUNCHED();
  sor_final_batch_kernel<<<njobs, 1024, 0, s>>>(jobs_dev);
  PSGC_LAUNCHED();
  PSGC_PROF_END("sor_stats", s);
  return PSGC_OK;
}

int launch_sor_apply(const float* xyz, double* md, float* dst, uint8_t* dstP, size_t npix,
                     const double* stats, hipStream_t s) {
  if (npix == 0) return PSGC_OK;
  sor_apply_kernel<<<(unsigned)((npix + 255) / 256), 256, 0, s>>>(xyz, md, npix, stats, dst, dstP);
  PSGC_LAUNCHED();
  return PSGC_OK;
}

size_t sor_chunks(size_t npix) { return (npix + kChunk - 1) / kChunk; }

size_t sor_partial_doubles(size_t npix) {
  const size_t nchunks = (npix + kChunk - 1) / kChunk;
  size_t np2 = 1;
  while (np2 < nchunks) np2 *= 2;
  return 3 * np2;
}

int launch_mls(const float* src, const uint8_t* P, float* dst, int rows, int cols, int hw,
               float sigma, int order, float r2cut, int semantics, hipStream_t s) {
  PSGC_PROF_BEGIN("mls", s);
  const float inv2s2 = 1.0f / (2.0f * sigma * sigma);
  if (semantics == PSGC_SEM_GPU)
    mls_plain_kernel<true><<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(
        src, P, dst, rows, cols, hw, inv2s2, order, r2cut, semantics);
  else
    mls_plain_kernel<false><<<grid2d(rows, cols), dim3(TB, TB), 0, s>>>(
        src, P, dst, rows, cols, hw, inv2s2, order, r2cut, semantics);
  PSGC_LAUNCHED();
  PSGC_PROF_END("mls", s);
  return PSGC_OK;
}

}  // namespace psgc
