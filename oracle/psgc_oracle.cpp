// psgc_oracle.cpp — TEST INFRASTRUCTURE, not product code.
//
// A plain C++17 restatement of the reference's CPU reconstruction path,
// written to be read against the reference line by line. It is the checker
// for the HIP library and the "port" CPU baseline timed beside it; only the
// tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Reference files (PhaseShiftGrayCode3D/src/):
//   decode          JSLAB_PhaseShiftDecoder_V2.cpp:6-310
//   HDR fit         JSLAB_HDRPhaseFit.h:10-61
//   match           JSLAB_PhaseEpipolarReconstructor_V2.cpp:933-1006
//   median/Q        JSLAB_PhaseEpipolarReconstructor_V2.cpp:397-461
//   Gaussian        JSLAB_PhaseEpipolarReconstructor_V2.cpp:79-124
//   ROR             JSLAB_PhaseEpipolarReconstructor_V2.cpp:126-195
//   MLS             JSLAB_MLS_Filter.cpp:11-441 (GPU fallback rule:
//                   JSLAB_MLS_Filter.cu:283-286)
// Not in the reference (textbook definitions): SOR, Gray-code unwrapping.
//
// Where the reference relies on OpenCV arithmetic the exact operation order
// is not observable without OpenCV (absent here), so it is fixed as follows
// and called "parity unpinned" at that boundary:
//   - cv::GaussianBlur 3x3 on f64: separable, row pass then column pass,
//     each out = c*x + s*(x_left + x_right), BORDER_REPLICATE;
//   - cv::getGaussianKernel: e_i = exp(-(i-h)^2/(2 sigma^2)) in f64, scaled by
//     the reciprocal of their sum, rounded to the kernel type; for ksize 3 and
//     sigma <= 0 the fixed [1/4, 1/2, 1/4];
//   - cv::medianBlur 5 on f32: exact 13th order statistic, BORDER_REPLICATE.
// atan2 is psgc::atan2 (shared with the device, within 2 ulp of libm). The
// float std::exp / std::acos / std::cos of the MLS filter are
// psgc::exp_f32 / acos_f32 / cos_f32 (shared with the device; evaluated in
// double and rounded once, checked exhaustively against the C library by
// orc_check_f32).
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -ffp-contract=off).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

#include "psgc.h"
#include "psgc_math.h"
#include "psgc_select.h"
#include "psgc_voxel.h"

namespace {

constexpr double PI = 3.14159265358979323846;

// ---- OpenCV kernel helpers ------------------------------------------------

// getGaussianKernel(n, sigma) in f64.
std::vector<double> gauss_kernel_d(int n, double sigma) {
  std::vector<double> k(n);
  if (n == 3 && sigma <= 0) {
    k = {0.25, 0.5, 0.25};
    return k;
  }
  if (sigma <= 0) sigma = ((n - 1) * 0.5 - 1) * 0.3 + 0.8;
  const double scale2 = -0.5 / (sigma * sigma);
  double sum = 0;
  for (int i = 0; i < n; ++i) {
    const double x = i - (n - 1) * 0.5;
    k[i] = std::exp(scale2 * x * x);
    sum += k[i];
  }
  sum = 1.0 / sum;
  for (int i = 0; i < n; ++i) k[i] *= sum;
  return k;
}

// 3x3 separable blur, BORDER_REPLICATE: row pass, then column pass.
void blur3_replicate(const double* src, double* dst, int rows, int cols,
                     double c, double s) {
  std::vector<double> tmp((size_t)rows * cols);
#pragma omp parallel for
  for (int r = 0; r < rows; ++r) {
    const double* in = src + (size_t)r * cols;
    double* out = tmp.data() + (size_t)r * cols;
    for (int x = 0; x < cols; ++x) {
      const double xl = in[x > 0 ? x - 1 : 0];
      const double xr = in[x < cols - 1 ? x + 1 : cols - 1];
      out[x] = c * in[x] + s * (xl + xr);
    }
  }
#pragma omp parallel for
  for (int r = 0; r < rows; ++r) {
    const double* up = tmp.data() + (size_t)(r > 0 ? r - 1 : 0) * cols;
    const double* mid = tmp.data() + (size_t)r * cols;
    const double* dn = tmp.data() + (size_t)(r < rows - 1 ? r + 1 : rows - 1) * cols;
    double* out = dst + (size_t)r * cols;
    for (int x = 0; x < cols; ++x) out[x] = c * mid[x] + s * (up[x] + dn[x]);
  }
}

double sample_at(const void* plane, int type, int64_t stride, int r, int x) {
  const char* row = (const char*)plane + (size_t)r * stride;
  switch (type) {
    case PSGC_SAMPLE_U8: return (double)((const uint8_t*)row)[x];
    case PSGC_SAMPLE_U16: return (double)((const uint16_t*)row)[x];
    default: return ((const double*)row)[x];
  }
}

int sample_size(int type) {
  return type == PSGC_SAMPLE_U8 ? 1 : (type == PSGC_SAMPLE_U16 ? 2 : 8);
}

// Frame -> f64 plane (convertTo CV_64F), JSLAB_PhaseShiftDecoder_V2.cpp:22.
std::vector<double> to_f64(const void* plane, int type, int64_t stride,
                           int rows, int cols) {
  std::vector<double> out((size_t)rows * cols);
  for (int r = 0; r < rows; ++r)
    for (int x = 0; x < cols; ++x)
      out[(size_t)r * cols + x] = sample_at(plane, type, stride, r, x);
  return out;
}

// ---- HDR fit (JSLAB_HDRPhaseFit.h:10-61) -----------------------------------
// Design A = [1, cos d, -sin d] over the usable subset; pinv(A) = (A^T A)^-1 A^T
// for the full-rank subsets that the reference keeps (its SVD rank test,
// :29-31). The last bits of OpenCV's SVD-based invert are parity unpinned.
struct HdrFit {
  int steps = 0;
  std::vector<char> valid;
  std::vector<double> w;  // [mask][k][3]
  explicit HdrFit(int n) : steps(n) {
    const size_t nm = size_t(1) << n;
    valid.assign(nm, 0);
    w.assign(nm * n * 3, 0.0);
    const double tau = 2.0 * PI;
    for (size_t mask = 0; mask < nm; ++mask) {
      std::vector<int> idx;
      for (int k = 0; k < n; ++k)
        if (mask & (size_t(1) << k)) idx.push_back(k);
      if (idx.size() < 3) continue;
      double AtA[3][3] = {};
      std::vector<double> rows(idx.size() * 3);
      for (size_t j = 0; j < idx.size(); ++j) {
        const double d = tau * idx[j] / n;
        rows[j * 3 + 0] = 1.0;
        rows[j * 3 + 1] = std::cos(d);
        rows[j * 3 + 2] = -std::sin(d);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) AtA[a][b] += rows[j * 3 + a] * rows[j * 3 + b];
      }
      // eigenvalue-free rank test: smallest singular value of A relative to
      // the largest via the 3x3 Gram determinant and trace.
      const double det =
          AtA[0][0] * (AtA[1][1] * AtA[2][2] - AtA[1][2] * AtA[2][1]) -
          AtA[0][1] * (AtA[1][0] * AtA[2][2] - AtA[1][2] * AtA[2][0]) +
          AtA[0][2] * (AtA[1][0] * AtA[2][1] - AtA[1][1] * AtA[2][0]);
      const double tr = AtA[0][0] + AtA[1][1] + AtA[2][2];
      if (!(det > 1e-20 * tr * tr * tr)) continue;
      double inv[3][3];
      inv[0][0] = (AtA[1][1] * AtA[2][2] - AtA[1][2] * AtA[2][1]) / det;
      inv[0][1] = (AtA[0][2] * AtA[2][1] - AtA[0][1] * AtA[2][2]) / det;
      inv[0][2] = (AtA[0][1] * AtA[1][2] - AtA[0][2] * AtA[1][1]) / det;
      inv[1][0] = (AtA[1][2] * AtA[2][0] - AtA[1][0] * AtA[2][2]) / det;
      inv[1][1] = (AtA[0][0] * AtA[2][2] - AtA[0][2] * AtA[2][0]) / det;
      inv[1][2] = (AtA[0][2] * AtA[1][0] - AtA[0][0] * AtA[1][2]) / det;
      inv[2][0] = (AtA[1][0] * AtA[2][1] - AtA[1][1] * AtA[2][0]) / det;
      inv[2][1] = (AtA[0][1] * AtA[2][0] - AtA[0][0] * AtA[2][1]) / det;
      inv[2][2] = (AtA[0][0] * AtA[1][1] - AtA[0][1] * AtA[1][0]) / det;
      for (size_t j = 0; j < idx.size(); ++j)
        for (int a = 0; a < 3; ++a) {
          double v = 0;
          for (int b = 0; b < 3; ++b) v += inv[a][b] * rows[j * 3 + b];
          w[(mask * n + idx[j]) * 3 + a] = v;
        }
      valid[mask] = 1;
    }
  }
  bool solve(const double* I, double sat, double beta[3]) const {
    size_t mask = 0;
    for (int k = 0; k < steps; ++k)
      if (I[k] >= 0.0 && I[k] < sat) mask |= size_t(1) << k;
    if (!valid[mask]) return false;
    beta[0] = beta[1] = beta[2] = 0;
    for (int k = 0; k < steps; ++k)
      if (mask & (size_t(1) << k))
        for (int a = 0; a < 3; ++a) beta[a] += w[(mask * steps + k) * 3 + a] * I[k];
    return std::isfinite(beta[0]) && std::isfinite(beta[1]) && std::isfinite(beta[2]);
  }
};

// Wrapped phase + modulation of one level at one pixel
// (JSLAB_PhaseShiftDecoder_V2.cpp:69-117). Returns false when the level
// rejects the pixel; *mod receives the normalized modulation.
bool level_phase(const psgc_decode_config& cfg, int level, const double* I,
                 const std::vector<double>& st, const std::vector<double>& ct,
                 const HdrFit* hdr, double* wrap, double* mod) {
  const int inum = cfg.shift_num[level];
  double sin_sum = 0, cos_sum = 0, sum_gray = 0;
  if (cfg.enable_hdr) {
    double beta[3];
    if (!hdr->solve(I, cfg.hdr_saturation_gray, beta) ||
        std::hypot(beta[1], beta[2]) <= cfg.hdr_min_amplitude || beta[0] <= 1e-3) {
      *mod = 0.0;
      return false;
    }
    cos_sum = beta[1] * inum / 2.0;
    sin_sum = beta[2] * inum / 2.0;
    sum_gray = beta[0] * inum;
  } else {
    for (int k = 0; k < inum; ++k) {
      sin_sum -= st[k] * I[k];
      cos_sum += ct[k] * I[k];
      sum_gray += I[k];
    }
  }
  *wrap = psgc::atan2(sin_sum, cos_sum);
  double b = sin_sum * sin_sum + cos_sum * cos_sum;
  const double mean_val = sum_gray / inum;
  b = 2.0 * std::sqrt(b) / inum;
  double norm_mod = 0.0;
  if (mean_val > 1e-3) norm_mod = b / mean_val;
  *mod = norm_mod;
  return !(norm_mod < cfg.modulation_threshold);
}

}  // namespace

extern "C" {

// Full decode, JSLAB_PhaseShiftDecoderV2::setImages + startCalculate
// (JSLAB_PhaseShiftDecoder_V2.cpp:256-295) and the per-pixel body (:54-182).
int orc_decode(const psgc_decode_config* cfg_, const void* const* frames,
               int n_frames, int type, int rows, int cols, int64_t stride,
               double* phase, int32_t* kmap, double* modmap, uint8_t* validmap) {
  const psgc_decode_config cfg = *cfg_;
  if (stride == 0) stride = (int64_t)cols * sample_size(type);
  const size_t npix = (size_t)rows * cols;
  const bool gray = cfg.unwrap_mode == PSGC_UNWRAP_GRAYCODE;
  const int nlev = gray ? 1 : 3;
  int nphase = 0;
  for (int i = 0; i < nlev; ++i) nphase += cfg.shift_num[i];
  const int need = gray ? nphase + 2 * cfg.gray_bits : nphase;
  if (n_frames != need) return PSGC_E_INVALID;

  // sin/cos tables (JSLAB_PhaseShiftDecoder_V2.cpp:203-211).
  std::vector<double> st[3], ct[3];
  for (int i = 0; i < nlev; ++i)
    for (int k = 0; k < cfg.shift_num[i]; ++k) {
      st[i].push_back(std::sin(2.0 * PI * k / cfg.shift_num[i]));
      ct[i].push_back(std::cos(2.0 * PI * k / cfg.shift_num[i]));
    }
  std::vector<HdrFit> hdr;
  if (cfg.enable_hdr)
    for (int i = 0; i < nlev; ++i) hdr.emplace_back(cfg.shift_num[i]);

  // Phase frames: f64 + pre-blur (JSLAB_ImageTypeConvertBodyV2, :18-28);
  // with HDR the raw samples are kept (:262-265).
  std::vector<std::vector<double>> img(nphase);
  const std::vector<double> gk = gauss_kernel_d(3, cfg.preblur_sigma);
  for (int f = 0; f < nphase; ++f) {
    img[f] = to_f64(frames[f], type, stride, rows, cols);
    if (!cfg.enable_hdr)
      blur3_replicate(img[f].data(), img[f].data(), rows, cols, gk[1], gk[0]);
  }

  std::vector<double> ph(npix, 0.0), md(npix, 0.0);
  std::vector<int32_t> kk(npix, cfg.enable_hdr ? -1 : 0);
  std::vector<uint8_t> vd(npix, 0);
  const double f0 = cfg.freq[0], f1 = cfg.freq[1], f2 = cfg.freq[2];

#pragma omp parallel for
  for (int r = 0; r < rows; ++r) {
    for (int x = 0; x < cols; ++x) {
      const size_t p = (size_t)r * cols + x;
      double wrap[3];
      bool ok = true;
      int base = 0;
      for (int i = 0; i < nlev; ++i) {
        double I[32];
        for (int k = 0; k < cfg.shift_num[i]; ++k) I[k] = img[base + k][p];
        double m;
        ok = level_phase(cfg, i, I, st[i], ct[i], cfg.enable_hdr ? &hdr[i] : nullptr,
                         &wrap[i], &m);
        if (i == 0) md[p] = m;
        base += cfg.shift_num[i];
        if (!ok) break;
      }
      if (!ok) continue;

      if (gray) {
        // Gray-code order with complement frames: bit b is set when the
        // pattern frame is brighter than its complement.
        unsigned g = 0;
        for (int b = 0; b < cfg.gray_bits && ok; ++b) {
          const double a = sample_at(frames[nphase + b], type, stride, r, x);
          const double c = sample_at(frames[nphase + cfg.gray_bits + b], type, stride, r, x);
          const double d = a - c;
          if (std::fabs(d) < cfg.gray_min_contrast) ok = false;
          g = (g << 1) | (d > 0 ? 1u : 0u);
        }
        if (!ok) continue;
        unsigned n = g;
        for (unsigned s = 1; s < 32; s <<= 1) n ^= n >> s;  // gray -> binary
        const double w = wrap[0] < 0 ? wrap[0] + 2 * PI : wrap[0];
        ph[p] = w + 2 * PI * (double)n;
        kk[p] = (int32_t)n;
        vd[p] = 255;
        continue;
      }

      // Heterodyne unwrapping (JSLAB_PhaseShiftDecoder_V2.cpp:124-179).
      double phase12, phase23, phase123;
      if (wrap[0] > wrap[1]) phase12 = wrap[0] - wrap[1];
      else phase12 = wrap[0] - wrap[1] + 2 * PI;
      if (f1 > f2) {
        if (wrap[1] > wrap[2]) phase23 = wrap[1] - wrap[2];
        else phase23 = wrap[1] - wrap[2] + 2 * PI;
      } else {
        if (wrap[2] > wrap[1]) phase23 = wrap[2] - wrap[1];
        else phase23 = wrap[2] - wrap[1] + 2 * PI;
      }
      if (phase12 > phase23) phase123 = phase12 - phase23;
      else phase123 = phase12 - phase23 + 2 * PI;
      const int f12 = std::abs(static_cast<int>(f0 - f1));
      const int f23 = std::abs(static_cast<int>(f1 - f2));
      const double uphase12 =
          phase12 + std::round((phase123 * f12 - phase12) / (2 * PI)) * 2 * PI;
      const double uphase1 =
          wrap[0] + std::round((uphase12 * f0 / f12 - wrap[0]) / (2 * PI)) * 2 * PI;
      const double k_val = std::round((uphase12 * f0 / f12 - wrap[0]) / (2 * PI));
      const double uphase23 =
          phase23 + std::round((phase123 * f23 - phase23) / (2 * PI)) * 2 * PI;
      double uphase2 =
          wrap[1] + std::round((uphase23 * f1 / f23 - wrap[1]) / (2 * PI)) * 2 * PI;
      uphase2 = (uphase2 * f0) / f1;
      double uphase3 =
          wrap[2] + std::round((uphase23 * f2 / f23 - wrap[2]) / (2 * PI)) * 2 * PI;
      uphase3 = (uphase3 * f0) / f2;
      ph[p] = cfg.avg_three_phase ? (uphase1 + uphase2 + uphase3) / 3.0 : uphase1;
      kk[p] = (int32_t)k_val;
      vd[p] = 255;
    }
  }

  if (gray) {
    // Period-jump correction: an order that disagrees with the median of
    // the valid absolute phases of its row neighbours (x-2..x+2) by about a
    // full period is moved by that many periods.
    std::vector<double> ph0 = ph;
    std::vector<int32_t> k0 = kk;
#pragma omp parallel for
    for (int r = 0; r < rows; ++r)
      for (int x = 0; x < cols; ++x) {
        const size_t p = (size_t)r * cols + x;
        if (!vd[p]) continue;
        double v[5];
        int n = 0;
        for (int d = -2; d <= 2; ++d) {
          const int xx = x + d;
          if (xx < 0 || xx >= cols || !vd[(size_t)r * cols + xx]) continue;
          v[n++] = ph0[(size_t)r * cols + xx];
        }
        std::sort(v, v + n);
        const double med = v[n / 2];
        const double jump = std::round((med - ph0[p]) / (2 * PI));
        if (jump != 0.0) {
          ph[p] = ph0[p] + jump * 2 * PI;
          kk[p] = k0[p] + (int32_t)jump;
        }
      }
  }

  // Optional post-smooth (JSLAB_PhaseShiftDecoder_V2.cpp:280-294).
  if (cfg.post_smooth) {
    if (cfg.enable_hdr) {
      std::vector<double> wgt(npix), ph2(npix), w2(npix);
      for (size_t p = 0; p < npix; ++p) wgt[p] = vd[p] ? 1.0 : 0.0;  // 255 * (1/255)
      blur3_replicate(ph.data(), ph2.data(), rows, cols, 0.5, 0.25);
      blur3_replicate(wgt.data(), w2.data(), rows, cols, 0.5, 0.25);
      for (size_t p = 0; p < npix; ++p) {
        const double w = std::max(w2[p], 1e-12);
        ph[p] = vd[p] ? ph2[p] / w : 0.0;
      }
    } else {
      blur3_replicate(ph.data(), ph.data(), rows, cols, 0.5, 0.25);
    }
  }

  if (phase) std::memcpy(phase, ph.data(), npix * sizeof(double));
  if (kmap) std::memcpy(kmap, kk.data(), npix * sizeof(int32_t));
  if (modmap) std::memcpy(modmap, md.data(), npix * sizeof(double));
  if (validmap) std::memcpy(validmap, vd.data(), npix);
  return PSGC_OK;
}

// stereoMatch (JSLAB_PhaseEpipolarReconstructor_V2.cpp:933-1006).
int orc_stereo_match(const double* l_phase, const double* r_phase,
                     const int32_t* l_k, const int32_t* r_k, int rows, int cols,
                     int pha_f, double grad_threshold, double* disp,
                     int32_t* match_col) {
  const int max_rank_limit = pha_f;
  const int knum = pha_f;
#pragma omp parallel for
  for (int i = 0; i < rows; ++i) {
    const double* lp = l_phase + (size_t)i * cols;
    const double* rp = r_phase + (size_t)i * cols;
    const int32_t* lk = l_k + (size_t)i * cols;
    const int32_t* rk = r_k + (size_t)i * cols;
    double* dp = disp + (size_t)i * cols;
    int32_t* mc = match_col ? match_col + (size_t)i * cols : nullptr;
    for (int j = 0; j < cols; ++j) {
      dp[j] = 0.0;
      if (mc) mc[j] = -1;
    }
    std::vector<std::vector<int>> table(knum + 1);
    for (int j = 0; j < cols; ++j) {
      const int kv = rk[j];
      if (kv >= 0 && kv <= knum) table[kv].push_back(j);
    }
    for (int j = 0; j < cols; ++j) {
      const int k = lk[j];
      if (k > max_rank_limit || k <= 0) continue;
      if (k > knum) continue;
      if (table[k].empty()) continue;
      double min_cost = 1e9;
      int best = -1;
      const double lv = lp[j];
      for (int c : table[k]) {
        const double diff = std::fabs(rp[c] - lv);
        if (diff < min_cost) {
          min_cost = diff;
          best = c;
        }
      }
      if (min_cost >= grad_threshold) continue;
      if (best < 2 || best >= cols - 1) continue;
      const double c0 = rp[best - 1] - lv;
      const double c1 = rp[best] - lv;
      const double c2 = rp[best + 1] - lv;
      const double num = c1 * (c0 - c2);
      const double den = 2.0 * (c0 - c1) * (c1 - c2);
      double delta = 0.0;
      if (std::fabs(den) > 1e-9) delta = num / den;
      dp[j] = (double)j - (double)best - delta;
      if (mc) mc[j] = best;
    }
  }
  return PSGC_OK;
}

// cv::medianBlur(src, dst, 5) on f32 with BORDER_REPLICATE.
int orc_median5(const float* src, float* dst, int rows, int cols) {
#pragma omp parallel for
  for (int r = 0; r < rows; ++r)
    for (int x = 0; x < cols; ++x) {
      float v[25];
      int n = 0;
      for (int dy = -2; dy <= 2; ++dy) {
        const int yy = std::min(std::max(r + dy, 0), rows - 1);
        for (int dx = -2; dx <= 2; ++dx) {
          const int xx = std::min(std::max(x + dx, 0), cols - 1);
          v[n++] = src[(size_t)yy * cols + xx];
        }
      }
      std::nth_element(v, v + 12, v + 25);
      dst[(size_t)r * cols + x] = v[12];
    }
  return PSGC_OK;
}

// Q reprojection (JSLAB_PhaseEpipolarReconstructor_V2.cpp:414-461).
int orc_reproject(const double* disp, int rows, int cols, const double* Q,
                  const double* R1, double min_depth, double max_depth,
                  int use_rect, float* xyz, uint8_t* P) {
#pragma omp parallel for
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      const size_t p = (size_t)i * cols + j;
      xyz[p * 3 + 0] = xyz[p * 3 + 1] = xyz[p * 3 + 2] = 0.f;
      P[p] = 0;
      const double d = disp[p];
      if (std::fabs(d) < 1e-5) continue;
      const double w = Q[14] * d + Q[15];
      double x = j + Q[3];
      double y = i + Q[7];
      double z = Q[11];
      const double inv_w = 1.0 / w;
      x *= inv_w;
      y *= inv_w;
      z *= inv_w;
      const double rx = R1[0] * x + R1[3] * y + R1[6] * z;
      const double ry = R1[1] * x + R1[4] * y + R1[7] * z;
      const double rz = R1[2] * x + R1[5] * y + R1[8] * z;
      if (use_rect) {
        if (z > min_depth && z < max_depth) {
          xyz[p * 3 + 0] = (float)x;
          xyz[p * 3 + 1] = (float)y;
          xyz[p * 3 + 2] = (float)z;
          P[p] = 1;
        }
      } else if (rz > min_depth && rz < max_depth) {
        xyz[p * 3 + 0] = (float)rx;
        xyz[p * 3 + 1] = (float)ry;
        xyz[p * 3 + 2] = (float)rz;
        P[p] = 1;
      }
    }
  return PSGC_OK;
}

// startCalculate_v2's per-pixel triangulation (JSLAB_PhaseEpipolarReconstructor_V2.cpp:
// 493-545) with evaluateV2CalibPoint (:13-76). cx/cy/cz: rows*cols*4 floats
// (the _rx/_ry/_rz CV_32FC4 maps).
int orc_reproject_poly(const double* disp, int rows, int cols, const float* cxm,
                       const float* cym, const float* czm, const double* R1,
                       double min_depth, double max_depth, int use_rect, float* xyz,
                       uint8_t* P) {
  const float kInvDepth = -1.0f, kFittedRay = -2.0f, kZLinearXY = -3.0f, kRational = -4.0f;
  auto eval = [&](const float* cx, const float* cy, const float* cz, double d, double& x,
                  double& y, double& z) -> bool {
    if (std::abs(cz[0]) <= 1e-9 && std::abs(cz[1]) <= 1e-9 && std::abs(cz[2]) <= 1e-9 &&
        std::abs(cz[3]) <= 1e-9)
      return false;
    if (std::abs(cx[3] - kInvDepth) < 1e-6f) {
      double d2 = d * d;
      double inv_z = cz[0] + cz[1] * d + cz[2] * d2 + cz[3] * d2 * d;
      if (!std::isfinite(inv_z) || std::abs(inv_z) < 1e-12) return false;
      z = 1.0 / inv_z;
      x = cx[0] * z;
      y = cy[0] * z;
      return std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
    }
    if (std::abs(cx[3] - kFittedRay) < 1e-6f) {
      double d2 = d * d;
      if (std::abs(cz[3] - kInvDepth) < 1e-6f) {
        double inv_z = cz[0] + cz[1] * d + cz[2] * d2;
        if (!std::isfinite(inv_z) || std::abs(inv_z) < 1e-12) return false;
        z = 1.0 / inv_z;
      } else {
        z = cz[0] + cz[1] * d + cz[2] * d2 + cz[3] * d2 * d;
      }
      x = cx[0] * z;
      y = cy[0] * z;
      return std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
    }
    if (std::abs(cx[3] - kZLinearXY) < 1e-6f) {
      if (std::abs(cz[3] - kRational) < 1e-6f) {
        double denom = cz[1] * d + 1.0;
        if (!std::isfinite(denom) || std::abs(denom) < 1e-12) return false;
        z = cz[0] / denom;
      } else {
        double d2 = d * d;
        double d3 = d2 * d;
        z = cz[0] + cz[1] * d + cz[2] * d2 + cz[3] * d3;
      }
      x = cx[0] * z + cx[1];
      y = cy[0] * z + cy[1];
      return std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
    }
    if (std::abs(cz[3] - kInvDepth) < 1e-6f) {
      double inv_z = cz[0] + cz[1] * d + cz[2] * d * d;
      if (!std::isfinite(inv_z) || std::abs(inv_z) < 1e-12) return false;
      z = 1.0 / inv_z;
      x = cx[0] * z;
      y = cy[0] * z;
      return std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
    }
    double d2 = d * d;
    double d3 = d2 * d;
    x = cx[0] + cx[1] * d + cx[2] * d2 + cx[3] * d3;
    y = cy[0] + cy[1] * d + cy[2] * d2 + cy[3] * d3;
    z = cz[0] + cz[1] * d + cz[2] * d2 + cz[3] * d3;
    return std::isfinite(x) && std::isfinite(y) && std::isfinite(z);
  };
#pragma omp parallel for
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < cols; ++j) {
      const size_t p = (size_t)i * cols + j;
      xyz[p * 3 + 0] = xyz[p * 3 + 1] = xyz[p * 3 + 2] = 0.f;
      P[p] = 0;
      const double d = disp[p];
      if (std::abs(d) < 1e-5) continue;
      const float* cz = czm + p * 4;
      if (!(std::abs(cz[0]) > 1e-9 || std::abs(cz[1]) > 1e-9 || std::abs(cz[2]) > 1e-9 ||
            std::abs(cz[3]) > 1e-9))
        continue;
      double x = 0.0, y = 0.0, z = 0.0;
      if (!eval(cxm + p * 4, cym + p * 4, cz, d, x, y, z)) continue;
      const double rx = R1[0] * x + R1[3] * y + R1[6] * z;
      const double ry = R1[1] * x + R1[4] * y + R1[7] * z;
      const double rz = R1[2] * x + R1[5] * y + R1[8] * z;
      if (use_rect) {
        if (z > min_depth && z < max_depth) {
          xyz[p * 3 + 0] = (float)x;
          xyz[p * 3 + 1] = (float)y;
          xyz[p * 3 + 2] = (float)z;
          P[p] = 1;
        }
      } else if (rz > min_depth && rz < max_depth) {
        xyz[p * 3 + 0] = (float)rx;
        xyz[p * 3 + 1] = (float)ry;
        xyz[p * 3 + 2] = (float)rz;
        P[p] = 1;
      }
    }
  return PSGC_OK;
}

// applyGaussianFilterToPointCloudV2 (JSLAB_PhaseEpipolarReconstructor_V2.cpp:
// 79-124); semantics GPU = the reference kernel's mask update and division
// (JSLAB_PhaseEpipolarReconstructor_V2_GPU.cu:220-274).
int orc_gaussian(const float* src, const uint8_t* P, float* dst, uint8_t* dstP,
                 int rows, int cols, int win, float sigma, int semantics) {
  const float count_threshold = 0.8f;
  const int half = (win - 1) / 2;
  const std::vector<double> k1 = gauss_kernel_d(win, sigma);
  std::vector<float> kf(win), k2((size_t)win * win);
  for (int i = 0; i < win; ++i) kf[i] = (float)k1[i];
  for (int i = 0; i < win; ++i)
    for (int j = 0; j < win; ++j) k2[(size_t)i * win + j] = kf[i] * kf[j];
  const size_t npix = (size_t)rows * cols;
  std::fill(dst, dst + npix * 3, 0.f);
  if (semantics == PSGC_SEM_CPU) std::memcpy(dstP, P, npix);
  else std::fill(dstP, dstP + npix, (uint8_t)0);
#pragma omp parallel for
  for (int r = half; r < rows - half; ++r)
    for (int x = half; x < cols - half; ++x) {
      const size_t p = (size_t)r * cols + x;
      if (P[p] == 0) continue;
      const float cx = src[p * 3];
      float s0 = 0, s1 = 0, s2 = 0, count = 0;
      for (int ik = -half; ik <= half; ++ik)
        for (int jk = -half; jk <= half; ++jk) {
          const size_t q = (size_t)(r + ik) * cols + (x + jk);
          if (P[q] == 0) continue;
          if (std::fabs(src[q * 3] - cx) < 3.0f) {
            const float w = k2[(size_t)(ik + half) * win + (jk + half)];
            s0 += src[q * 3 + 0] * w;
            s1 += src[q * 3 + 1] * w;
            s2 += src[q * 3 + 2] * w;
            count += w;
          }
        }
      if (count > count_threshold) {
        if (semantics == PSGC_SEM_CPU) {
          const float inv = 1.f / count;
          dst[p * 3 + 0] = s0 * inv;
          dst[p * 3 + 1] = s1 * inv;
          dst[p * 3 + 2] = s2 * inv;
        } else {
          dst[p * 3 + 0] = s0 / count;
          dst[p * 3 + 1] = s1 / count;
          dst[p * 3 + 2] = s2 / count;
          dstP[p] = 1;
        }
      }
    }
  return PSGC_OK;
}

// applyRadiusOutlierFilterToPointCloud
// (JSLAB_PhaseEpipolarReconstructor_V2.cpp:126-195).
int orc_ror(const float* src, const uint8_t* P, float* dst, uint8_t* dstP,
            int rows, int cols, float radius, int min_neighbors, int window) {
  const size_t npix = (size_t)rows * cols;
  std::fill(dst, dst + npix * 3, 0.f);
  std::fill(dstP, dstP + npix, (uint8_t)0);
  const float r2 = radius * radius;
  const int half = window / 2;
#pragma omp parallel for
  for (int r = 0; r < rows; ++r)
    for (int x = 0; x < cols; ++x) {
      const size_t p = (size_t)r * cols + x;
      if (P[p] == 0) continue;
      const float px = src[p * 3], py = src[p * 3 + 1], pz = src[p * 3 + 2];
      int cnt = 0;
      const int r0 = std::max(0, r - half), r1 = std::min(rows, r + half + 1);
      const int c0 = std::max(0, x - half), c1 = std::min(cols, x + half + 1);
      for (int i = r0; i < r1 && cnt < min_neighbors; ++i)
        for (int j = c0; j < c1; ++j) {
          const size_t q = (size_t)i * cols + j;
          if (P[q] == 0 || q == p) continue;
          const float dx = px - src[q * 3];
          const float dy = py - src[q * 3 + 1];
          const float dz = pz - src[q * 3 + 2];
          const float d2 = dx * dx + dy * dy + dz * dz;
          if (d2 <= r2 && ++cnt >= min_neighbors) break;
        }
      if (cnt >= min_neighbors) {
        dst[p * 3] = px;
        dst[p * 3 + 1] = py;
        dst[p * 3 + 2] = pz;
        dstP[p] = 1;
      }
    }
  return PSGC_OK;
}

// Fixed pairwise reduction: slots [0, n) with n a power of two, summed as
// s[i] += s[i + h] for h = n/2, n/4, ..., 1. The HIP kernels reduce in the
// same order so that the statistics are bit-identical.
static double tree_sum(std::vector<double> s) {
  for (size_t h = s.size() / 2; h >= 1; h /= 2)
    for (size_t i = 0; i < h; ++i) s[i] += s[i + h];
  return s.empty() ? 0.0 : s[0];
}

}  // extern "C"

namespace {

// Threshold statistics of SOR over the eligible points (ok): chunks of 1024
// consecutive points tree-summed, the chunk partials tree-summed over the
// next power of two; sample standard deviation. st = {count, mean, sd,
// mean + std_mul sd}.
void sor_stats_of(const double* md, const uint8_t* ok, size_t npix, float std_mul, double st[4]) {
  const size_t chunk = 1024;
  const size_t nchunks = (npix + chunk - 1) / chunk;
  size_t np2 = 1;
  while (np2 < nchunks) np2 *= 2;
  std::vector<double> cs(np2, 0.0), cq(np2, 0.0);
  std::vector<long long> cn(nchunks, 0);
  for (size_t c = 0; c < nchunks; ++c) {
    std::vector<double> s(chunk, 0.0), q(chunk, 0.0);
    for (size_t t = 0; t < chunk; ++t) {
      const size_t p = c * chunk + t;
      if (p < npix && ok[p]) {
        s[t] = md[p];
        q[t] = md[p] * md[p];
        cn[c]++;
      }
    }
    cs[c] = tree_sum(s);
    cq[c] = tree_sum(q);
  }
  long long n = 0;
  for (auto v : cn) n += v;
  const double sum = tree_sum(cs), sq = tree_sum(cq);
  double mean = 0, sd = 0;
  if (n > 0) mean = sum / (double)n;
  if (n > 1) {
    const double var = (sq - sum * sum / (double)n) / ((double)n - 1.0);
    sd = std::sqrt(std::max(var, 0.0));
  }
  st[0] = (double)n;
  st[1] = mean;
  st[2] = sd;
  st[3] = mean + (double)std_mul * sd;
}

}  // namespace

extern "C" {

// Statistical outlier removal (PCL StatisticalOutlierRemoval definition on
// an organized window); see psgc.h for the exact rule. Reduction order:
// chunks of 1024 consecutive pixels are tree-summed, the chunk partials are
// tree-summed over the next power of two.
int orc_sor(const float* src, const uint8_t* P, float* dst, uint8_t* dstP,
            int rows, int cols, int k, float std_mul, int half,
            double* mean_dist, double* stats) {
  const size_t npix = (size_t)rows * cols;
  std::vector<double> md(npix, 0.0);
  std::vector<uint8_t> ok(npix, 0);
#pragma omp parallel for
  for (int r = 0; r < rows; ++r) {
    std::vector<float> d2s;
    for (int x = 0; x < cols; ++x) {
      const size_t p = (size_t)r * cols + x;
      if (P[p] == 0) continue;
      d2s.clear();
      const float px = src[p * 3], py = src[p * 3 + 1], pz = src[p * 3 + 2];
      for (int i = std::max(0, r - half); i < std::min(rows, r + half + 1); ++i)
        for (int j = std::max(0, x - half); j < std::min(cols, x + half + 1); ++j) {
          const size_t q = (size_t)i * cols + j;
          if (P[q] == 0 || q == p) continue;
          const float dx = px - src[q * 3];
          const float dy = py - src[q * 3 + 1];
          const float dz = pz - src[q * 3 + 2];
          d2s.push_back(dx * dx + dy * dy + dz * dz);
        }
      if ((int)d2s.size() < k) continue;
      std::partial_sort(d2s.begin(), d2s.begin() + k, d2s.end());
      double s = 0;
      for (int m = 0; m < k; ++m) s += std::sqrt((double)d2s[m]);
      md[p] = s / k;
      ok[p] = 1;
    }
  }
  double st[4];
  sor_stats_of(md.data(), ok.data(), npix, std_mul, st);
  const double thr = st[3];
  const long long n = (long long)st[0];
  const double mean = st[1], sd = st[2];

  for (size_t p = 0; p < npix; ++p) {
    const bool keep = ok[p] && md[p] <= thr;
    dstP[p] = keep ? 1 : 0;
    for (int c = 0; c < 3; ++c) dst[p * 3 + c] = keep ? src[p * 3 + c] : 0.f;
  }
  if (mean_dist) std::memcpy(mean_dist, md.data(), npix * sizeof(double));
  if (stats) {
    stats[0] = (double)n;
    stats[1] = mean;
    stats[2] = sd;
    stats[3] = thr;
  }
  return PSGC_OK;
}

}  // extern "C"

// ---- MLS (JSLAB_MLS_Filter.cpp:11-415) ------------------------------------
namespace {

void eigen3(const float C[3][3], float ev[3], float V[3][3]) {
  const float p1 = C[0][1] * C[0][1] + C[0][2] * C[0][2] + C[1][2] * C[1][2];
  if (p1 < 1e-20f) {
    ev[0] = C[0][0];
    ev[1] = C[1][1];
    ev[2] = C[2][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.f : 0.f;
    for (int i = 0; i < 2; ++i)
      for (int j = i + 1; j < 3; ++j)
        if (ev[i] > ev[j]) {
          std::swap(ev[i], ev[j]);
          for (int c = 0; c < 3; ++c) std::swap(V[i][c], V[j][c]);
        }
    return;
  }
  const float trace = C[0][0] + C[1][1] + C[2][2];
  const float q = trace / 3.0f;
  float B[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i][j] = C[i][j];
  B[0][0] -= q;
  B[1][1] -= q;
  B[2][2] -= q;
  const float p2 = B[0][0] * B[0][0] + B[1][1] * B[1][1] + B[2][2] * B[2][2] + 2.0f * p1;
  const float p = std::sqrt(p2 / 6.0f);
  const float inv_p = (p > 1e-15f) ? (1.0f / p) : 0.0f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i][j] *= inv_p;
  const float detB = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) -
                     B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                     B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
  float r = detB * 0.5f;
  r = std::max(-1.0f, std::min(1.0f, r));
  const float phi = psgc::acos_f32(r) / 3.0f;
  const float PIf = 3.14159265358979323846f;
  ev[0] = q + 2.0f * p * psgc::cos_f32(phi + 2.0f * PIf / 3.0f);
  ev[2] = q + 2.0f * p * psgc::cos_f32(phi);
  ev[1] = trace - ev[0] - ev[2];
  for (int idx = 0; idx < 3; ++idx) {
    float A[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i][j] = C[i][j];
    A[0][0] -= ev[idx];
    A[1][1] -= ev[idx];
    A[2][2] -= ev[idx];
    float c01[3] = {A[0][1] * A[1][2] - A[0][2] * A[1][1],
                    A[0][2] * A[1][0] - A[0][0] * A[1][2],
                    A[0][0] * A[1][1] - A[0][1] * A[1][0]};
    float c02[3] = {A[0][1] * A[2][2] - A[0][2] * A[2][1],
                    A[0][2] * A[2][0] - A[0][0] * A[2][2],
                    A[0][0] * A[2][1] - A[0][1] * A[2][0]};
    float c12[3] = {A[1][1] * A[2][2] - A[1][2] * A[2][1],
                    A[1][2] * A[2][0] - A[1][0] * A[2][2],
                    A[1][0] * A[2][1] - A[1][1] * A[2][0]};
    const float l01 = c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2];
    const float l02 = c02[0] * c02[0] + c02[1] * c02[1] + c02[2] * c02[2];
    const float l12 = c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2];
    float* best = c12;
    float bl = l12;
    if (l01 > bl) { best = c01; bl = l01; }
    if (l02 > bl) { best = c02; bl = l02; }
    const float inv = (bl > 1e-20f) ? (1.0f / std::sqrt(bl)) : 0.0f;
    V[idx][0] = best[0] * inv;
    V[idx][1] = best[1] * inv;
    V[idx][2] = best[2] * inv;
  }
}

bool chol(const float A[6][6], const float b[6], float x[6], int nc) {
  float L[6][6] = {};
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j <= i; ++j) {
      float s = A[i][j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (s <= 1e-12f) return false;
        L[i][j] = std::sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  float y[6] = {};
  for (int i = 0; i < nc; ++i) {
    float s = b[i];
    for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
    y[i] = s / L[i][i];
  }
  for (int i = nc - 1; i >= 0; --i) {
    float s = y[i];
    for (int k = i + 1; k < nc; ++k) s -= L[k][i] * x[k];
    x[i] = s / L[i][i];
  }
  return true;
}

// One point of the MLS filter. each(visit) calls visit(pk) for every
// candidate neighbour pk (float[3]) in a fixed order; the radius test, the
// three accumulation passes and the fit follow JSLAB_MLS_Filter.cpp:196-389
// (.cu:214-389). C (GPU semantics): the per-neighbour multiply-adds (d^2,
// the centroid, covariance and normal-equation sums, the local coordinates
// u, v, f) are fused, as a CUDA build contracts `s += a * b`; without C each
// product and sum is rounded (the CPU build). The solve and the evaluation
// at the point are uncontracted in both.
// out receives the projected point, or pi where the fit is skipped.
template <bool C, typename Each>
void mls_point_t(const float pi[3], Each&& each, float inv_2s2, float r2cut, int order,
                 int semantics, float out[3]) {
  using psgc::madd;
  const int minnb[3] = {3, 5, 8};
  const float pi0 = pi[0], pi1 = pi[1], pi2 = pi[2];
  out[0] = pi0;
  out[1] = pi1;
  out[2] = pi2;
  float cx = 0, cy = 0, cz = 0, wsum = 0;
  int nv = 0;
  each([&](const float* pk) {
    const float d2 = psgc::sq3<C>(pk[0] - pi0, pk[1] - pi1, pk[2] - pi2);
    if (d2 > r2cut) return;
    const float w = psgc::exp_f32(-d2 * inv_2s2);
    cx = madd<C>(w, pk[0], cx);
    cy = madd<C>(w, pk[1], cy);
    cz = madd<C>(w, pk[2], cz);
    wsum += w;
    nv++;
  });
  if (nv < minnb[0] || wsum < 1e-12f) return;
  const float iw = 1.0f / wsum;
  cx *= iw;
  cy *= iw;
  cz *= iw;
  float cov[3][3] = {};
  each([&](const float* pk) {
    const float d2 = psgc::sq3<C>(pk[0] - pi0, pk[1] - pi1, pk[2] - pi2);
    if (d2 > r2cut) return;
    const float w = psgc::exp_f32(-d2 * inv_2s2);
    const float ex = pk[0] - cx, ey = pk[1] - cy, ez = pk[2] - cz;
    cov[0][0] = madd<C>(w * ex, ex, cov[0][0]);
    cov[0][1] = madd<C>(w * ex, ey, cov[0][1]);
    cov[0][2] = madd<C>(w * ex, ez, cov[0][2]);
    cov[1][1] = madd<C>(w * ey, ey, cov[1][1]);
    cov[1][2] = madd<C>(w * ey, ez, cov[1][2]);
    cov[2][2] = madd<C>(w * ez, ez, cov[2][2]);
  });
  cov[1][0] = cov[0][1];
  cov[2][0] = cov[0][2];
  cov[2][1] = cov[1][2];
  float ev[3], V[3][3];
  eigen3(cov, ev, V);
  const float nx = V[0][0], ny = V[0][1], nz = V[0][2];
  const float t1x = V[1][0], t1y = V[1][1], t1z = V[1][2];
  const float t2x = V[2][0], t2y = V[2][1], t2z = V[2][2];
  int cur;
  if (semantics == PSGC_SEM_CPU) {
    cur = (nv >= minnb[order]) ? order : 0;
    if (cur > 0 && nv < minnb[cur]) cur = (nv >= minnb[1]) ? 1 : 0;
  } else {
    cur = order;
    if (cur > 0 && nv < minnb[cur]) cur = (nv >= minnb[1]) ? 1 : 0;
  }
  if (cur == 0) {
    const float dot = (pi0 - cx) * nx + (pi1 - cy) * ny + (pi2 - cz) * nz;
    out[0] = pi0 - dot * nx;
    out[1] = pi1 - dot * ny;
    out[2] = pi2 - dot * nz;
    return;
  }
  const int nc = (cur == 1) ? 3 : 6;
  float AtWA[6][6] = {}, AtWf[6] = {};
  each([&](const float* pk) {
    const float d2 = psgc::sq3<C>(pk[0] - pi0, pk[1] - pi1, pk[2] - pi2);
    if (d2 > r2cut) return;
    const float w = psgc::exp_f32(-d2 * inv_2s2);
    const float ex = pk[0] - cx, ey = pk[1] - cy, ez = pk[2] - cz;
    const float u = madd<C>(ez, t1z, madd<C>(ey, t1y, ex * t1x));
    const float v = madd<C>(ez, t2z, madd<C>(ey, t2y, ex * t2x));
    const float f = madd<C>(ez, nz, madd<C>(ey, ny, ex * nx));
    float basis[6] = {1.0f, u, v, 0, 0, 0};
    if (nc == 6) {
      basis[3] = u * u;
      basis[4] = u * v;
      basis[5] = v * v;
    }
    for (int a = 0; a < nc; ++a) {
      AtWf[a] = madd<C>(w * basis[a], f, AtWf[a]);
      for (int b = a; b < nc; ++b) AtWA[a][b] = madd<C>(w * basis[a], basis[b], AtWA[a][b]);
    }
  });
  for (int a = 0; a < nc; ++a)
    for (int b = 0; b < a; ++b) AtWA[a][b] = AtWA[b][a];
  float co[6] = {};
  bool solved = chol(AtWA, AtWf, co, nc);
  if (!solved && cur == 2) {
    float A1[6][6] = {}, b1[6] = {};
    for (int a = 0; a < 3; ++a) {
      b1[a] = AtWf[a];
      for (int b = 0; b < 3; ++b) A1[a][b] = AtWA[a][b];
    }
    solved = chol(A1, b1, co, 3);
    if (solved) cur = 1;
  }
  if (!solved) {
    const float dot = (pi0 - cx) * nx + (pi1 - cy) * ny + (pi2 - cz) * nz;
    out[0] = pi0 - dot * nx;
    out[1] = pi1 - dot * ny;
    out[2] = pi2 - dot * nz;
    return;
  }
  const float ex0 = pi0 - cx, ey0 = pi1 - cy, ez0 = pi2 - cz;
  const float u0 = ex0 * t1x + ey0 * t1y + ez0 * t1z;
  const float v0 = ex0 * t2x + ey0 * t2y + ez0 * t2z;
  float ff = co[0] + co[1] * u0 + co[2] * v0;
  if (cur == 2) ff += co[3] * u0 * u0 + co[4] * u0 * v0 + co[5] * v0 * v0;
  out[0] = cx + u0 * t1x + v0 * t2x + ff * nx;
  out[1] = cy + u0 * t1y + v0 * t2y + ff * ny;
  out[2] = cz + u0 * t1z + v0 * t2z + ff * nz;
}

template <typename Each>
void mls_point(const float pi[3], Each&& each, float inv_2s2, float r2cut, int order,
               int semantics, float out[3]) {
  if (semantics == PSGC_SEM_GPU) mls_point_t<true>(pi, each, inv_2s2, r2cut, order, semantics, out);
  else mls_point_t<false>(pi, each, inv_2s2, r2cut, order, semantics, out);
}

// One MLS iteration over an organized cloud: the window of +-hw pixels,
// row by row (JSLAB_MLS_Filter.cpp:176-389).
void mls_pass(const float* src, const uint8_t* P, float* dst, int rows, int cols,
              int hw, float sigma, int order_in, float r2cut, int semantics) {
  std::memcpy(dst, src, (size_t)rows * cols * 3 * sizeof(float));
  const float inv_2s2 = 1.0f / (2.0f * sigma * sigma);
  const int order = std::max(0, std::min(2, order_in));
#pragma omp parallel for
  for (int r = 0; r < rows; ++r)
    for (int x = 0; x < cols; ++x) {
      const size_t p = (size_t)r * cols + x;
      if (P[p] == 0) continue;
      const int r0 = std::max(0, r - hw), r1 = std::min(rows, r + hw + 1);
      const int c0 = std::max(0, x - hw), c1 = std::min(cols, x + hw + 1);
      auto each = [&](auto&& visit) {
        for (int i = r0; i < r1; ++i)
          for (int j = c0; j < c1; ++j) {
            const size_t q = (size_t)i * cols + j;
            if (P[q] != 0) visit(src + q * 3);
          }
      };
      mls_point(src + p * 3, each, inv_2s2, r2cut, order, semantics, dst + p * 3);
    }
}

}  // namespace

extern "C" {

// applyMLSFilterToPointCloud_CPU (JSLAB_MLS_Filter.cpp:396-415).
int orc_mls(const float* src, const uint8_t* P, float* dst, int rows, int cols,
            int window, float sigma, int order, float max_radius, int iterations,
            int semantics) {
  const float r2 = (max_radius > 0.0f) ? max_radius * max_radius : 9.0f * sigma * sigma;
  const int n_iter = std::max(1, iterations);
  const size_t n3 = (size_t)rows * cols * 3;
  std::vector<float> cur(src, src + n3);
  for (int it = 0; it < n_iter; ++it) {
    mls_pass(cur.data(), P, dst, rows, cols, window, sigma, order, r2, semantics);
    if (it + 1 < n_iter) std::memcpy(cur.data(), dst, n3 * sizeof(float));
  }
  return PSGC_OK;
}

// The MLS filter's 3x3 symmetric eigen solver (symmetricEigen3x3,
// JSLAB_MLS_Filter.cpp:11-106) on n row-major matrices: ev[3] ascending as
// the solver orders them, V[3][3] one eigenvector per row.
int orc_eigen3(const float* C, float* ev, float* V, int n) {
  for (int i = 0; i < n; ++i) {
    float A[3][3], W[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) A[r][c] = C[i * 9 + r * 3 + c];
    eigen3(A, ev + i * 3, W);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) V[i * 9 + r * 3 + c] = W[r][c];
  }
  return PSGC_OK;
}

// The MLS filter's Cholesky solve (choleskySolve, JSLAB_MLS_Filter.cpp:113-156)
// of one nc x nc system (nc <= 6, A row-major nc x nc): 1 if solved, 0 when a
// pivot is <= 1e-12.
int orc_chol(const float* A, const float* b, float* x, int nc) {
  if (nc < 1 || nc > 6) return PSGC_E_INVALID;
  float M[6][6] = {}, bb[6] = {}, xx[6] = {};
  for (int r = 0; r < nc; ++r) {
    bb[r] = b[r];
    for (int c = 0; c < nc; ++c) M[r][c] = A[r * nc + c];
  }
  const bool ok = chol(M, bb, xx, nc);
  for (int r = 0; r < nc; ++r) x[r] = xx[r];
  return ok ? 1 : 0;
}

// JSLAB_PhaseEpipolarReconstructor_V2::startCalculate
// (JSLAB_PhaseEpipolarReconstructor_V2.cpp:397-480) on rectified inputs.
static int reconstruct_impl(const psgc_recon_config* cfg, const double* l_phase,
                            const double* r_phase, const int32_t* l_k, const int32_t* r_k,
                            int rows, int cols, int pha_f, const double* Q,
                            const float* const* coef, const double* R1, float* xyz,
                            uint8_t* P, float* disp_out, int32_t* match_col);

int orc_reconstruct(const psgc_recon_config* cfg, const double* l_phase,
                    const double* r_phase, const int32_t* l_k,
                    const int32_t* r_k, int rows, int cols, int pha_f,
                    const double* Q, const double* R1, float* xyz, uint8_t* P,
                    float* disp_out, int32_t* match_col) {
  return reconstruct_impl(cfg, l_phase, r_phase, l_k, r_k, rows, cols, pha_f, Q, nullptr, R1,
                          xyz, P, disp_out, match_col);
}

// startCalculate_v2 (JSLAB_PhaseEpipolarReconstructor_V2.cpp:482-593): the same
// chain with the per-pixel polynomial triangulation in place of Q.
int orc_reconstruct_poly(const psgc_recon_config* cfg, const double* l_phase,
                         const double* r_phase, const int32_t* l_k, const int32_t* r_k,
                         int rows, int cols, int pha_f, const float* cx, const float* cy,
                         const float* cz, const double* R1, float* xyz, uint8_t* P,
                         float* disp_out, int32_t* match_col) {
  const float* coef[3] = {cx, cy, cz};
  return reconstruct_impl(cfg, l_phase, r_phase, l_k, r_k, rows, cols, pha_f, nullptr, coef, R1,
                          xyz, P, disp_out, match_col);
}

static int reconstruct_impl(const psgc_recon_config* cfg, const double* l_phase,
                            const double* r_phase, const int32_t* l_k, const int32_t* r_k,
                            int rows, int cols, int pha_f, const double* Q,
                            const float* const* coef, const double* R1, float* xyz,
                            uint8_t* P, float* disp_out, int32_t* match_col) {
  const size_t npix = (size_t)rows * cols;
  std::vector<double> disp(npix);
  orc_stereo_match(l_phase, r_phase, l_k, r_k, rows, cols, pha_f,
                   cfg->grad_threshold, disp.data(), match_col);
  // The reference GPU path stores the disparity as f32 (…V2_GPU.cu:382);
  // the CPU path converts to f32 around the median (…V2.cpp:405-410).
  std::vector<float> d32(npix), d32m(npix);
  for (size_t p = 0; p < npix; ++p) d32[p] = (float)disp[p];
  const bool f32_disp = cfg->use_disp_median_filter || cfg->semantics == PSGC_SEM_GPU;
  if (cfg->use_disp_median_filter) {
    orc_median5(d32.data(), d32m.data(), rows, cols);
    d32.swap(d32m);
  }
  if (f32_disp)
    for (size_t p = 0; p < npix; ++p) disp[p] = (double)d32[p];
  if (disp_out)
    for (size_t p = 0; p < npix; ++p) disp_out[p] = (float)disp[p];
  std::vector<float> xyz0(npix * 3), xyzg(npix * 3), xyzr(npix * 3);
  std::vector<uint8_t> P0(npix), Pg(npix), Pr(npix);
  if (coef)
    orc_reproject_poly(disp.data(), rows, cols, coef[0], coef[1], coef[2], R1, cfg->min_depth,
                       cfg->max_depth, cfg->use_rectified_coords, xyz0.data(), P0.data());
  else
    orc_reproject(disp.data(), rows, cols, Q, R1, cfg->min_depth, cfg->max_depth,
                  cfg->use_rectified_coords, xyz0.data(), P0.data());
  orc_gaussian(xyz0.data(), P0.data(), xyzg.data(), Pg.data(), rows, cols,
               cfg->gauss_window, cfg->gauss_sigma, cfg->semantics);
  // CPU semantics: the outlier stage reads the pre-Gaussian mask (…V2.cpp:466).
  const uint8_t* Pin = cfg->semantics == PSGC_SEM_CPU ? P0.data() : Pg.data();
  if (cfg->outlier_mode == PSGC_OUTLIER_ROR) {
    orc_ror(xyzg.data(), Pin, xyzr.data(), Pr.data(), rows, cols, cfg->radius_search,
            cfg->radius_min_neighbors, cfg->radius_window_size);
  } else if (cfg->outlier_mode == PSGC_OUTLIER_SOR) {
    orc_sor(xyzg.data(), Pin, xyzr.data(), Pr.data(), rows, cols, cfg->sor_k,
            cfg->sor_std_mul, cfg->sor_window, nullptr, nullptr);
  } else {
    xyzr = xyzg;
    std::memcpy(Pr.data(), Pin, npix);
  }
  std::memcpy(P, Pr.data(), npix);
  if (cfg->use_mls) {
    orc_mls(xyzr.data(), Pr.data(), xyz, rows, cols, cfg->mls_window_size,
            cfg->mls_spatial_sigma, cfg->mls_polynomial_order, cfg->mls_max_radius,
            cfg->mls_iterations, cfg->semantics);
  } else {
    std::memcpy(xyz, xyzr.data(), npix * 3 * sizeof(float));
  }
  return PSGC_OK;
}

// The right search window of prepareRectifiedROI
// (JSLAB_PhaseEpipolarReconstructor_V2.cpp:629-669): fringe orders 1..pha_f
// present under the mask inside left_rect = {x, y, w, h}, then the first and
// last right column carrying one of them in those rows, padded by 32 and
// clipped; no such column: the full rows.
int orc_roi_right_window(const int32_t* l_k, const int32_t* r_k, int rows, int cols,
                         int pha_f, const int32_t* lr, const uint8_t* mask, int32_t* rr) {
  std::vector<uint8_t> needed((size_t)std::max(pha_f + 1, 1), 0);
  for (int y = 0; y < lr[3]; ++y)
    for (int x = 0; x < lr[2]; ++x) {
      if (mask[(size_t)y * lr[2] + x] == 0) continue;
      const int k = l_k[(size_t)(lr[1] + y) * cols + lr[0] + x];
      if (k > 0 && k <= pha_f) needed[k] = 1;
    }
  int mn = cols, mx = -1;
  for (int y = lr[1]; y < lr[1] + lr[3]; ++y) {
    if (y < 0 || y >= rows) continue;
    for (int x = 0; x < cols; ++x) {
      const int k = r_k[(size_t)y * cols + x];
      if (k > 0 && k <= pha_f && needed[k]) {
        mn = std::min(mn, x);
        mx = std::max(mx, x);
      }
    }
  }
  const int pad = 32;
  if (mx >= mn) {
    const int x0 = std::max(0, mn - pad), x1 = std::min(cols - 1, mx + pad);
    rr[0] = x0;
    rr[2] = x1 - x0 + 1;
  } else {
    rr[0] = 0;
    rr[2] = cols;
  }
  rr[1] = lr[1];
  rr[3] = lr[3];
  return PSGC_OK;
}

// startCalculate_v2_roi(xyz, P, RectifiedROI)
// (JSLAB_PhaseEpipolarReconstructor_V2.cpp:693-914): the match of the left
// crop against the right crop (same rows), only under the mask, with the
// edge test in right-crop columns and the disparity from global columns;
// median on the crop; polynomial triangulation with the coefficients at
// global pixels; Gaussian; ROR, kept only if it leaves at least 10 % of the
// points (else the pre-ROR mask and the Gaussian points) — or SOR / none as
// configured; MLS. Outputs are crop-sized (lr[2] x lr[3]).
int orc_reconstruct_poly_roi(const psgc_recon_config* cfg, const double* l_phase,
                             const double* r_phase, const int32_t* l_k, const int32_t* r_k,
                             int rows, int cols, int pha_f, const int32_t* lr,
                             const int32_t* rr, const uint8_t* mask, const float* cxm,
                             const float* cym, const float* czm, const double* R1,
                             float* xyz, uint8_t* P, float* disp_out) {
  (void)rows;
  const int lx0 = lr[0], y0 = lr[1], lw = lr[2], h = lr[3];
  const int rx0 = rr[0], rw = rr[2];
  const int knum = pha_f;
  const size_t n = (size_t)lw * h;
  std::vector<double> disp(n, 0.0);
#pragma omp parallel for
  for (int r = 0; r < h; ++r) {
    const double* lp = l_phase + (size_t)(y0 + r) * cols + lx0;
    const double* rp = r_phase + (size_t)(y0 + r) * cols + rx0;
    const int32_t* lk = l_k + (size_t)(y0 + r) * cols + lx0;
    const int32_t* rk = r_k + (size_t)(y0 + r) * cols + rx0;
    const uint8_t* mk = mask + (size_t)r * lw;
    std::vector<std::vector<int>> table(knum + 1);
    for (int j = 0; j < rw; ++j)
      if (rk[j] >= 0 && rk[j] <= knum) table[rk[j]].push_back(j);
    for (int j = 0; j < lw; ++j) {
      if (mk[j] == 0) continue;
      const int k = lk[j];
      if (k > pha_f || k <= 0 || k > knum) continue;
      if (table[k].empty()) continue;
      double min_cost = 1e9;
      int best = -1;
      const double lv = lp[j];
      for (int c : table[k]) {
        const double diff = std::abs(rp[c] - lv);
        if (diff < min_cost) {
          min_cost = diff;
          best = c;
        }
      }
      if (min_cost >= cfg->grad_threshold) continue;
      if (best < 2 || best >= rw - 1) continue;
      const double c0 = rp[best - 1] - lv, c1 = rp[best] - lv, c2 = rp[best + 1] - lv;
      const double num = c1 * (c0 - c2);
      const double den = 2.0 * (c0 - c1) * (c1 - c2);
      double delta = 0.0;
      if (std::abs(den) > 1e-9) delta = num / den;
      disp[(size_t)r * lw + j] =
          static_cast<double>(lx0 + j) - static_cast<double>(rx0 + best) - delta;
    }
  }
  std::vector<float> d32(n), d32m(n);
  for (size_t p = 0; p < n; ++p) d32[p] = (float)disp[p];
  const bool f32_disp = cfg->use_disp_median_filter || cfg->semantics == PSGC_SEM_GPU;
  if (cfg->use_disp_median_filter) {
    orc_median5(d32.data(), d32m.data(), h, lw);
    d32.swap(d32m);
  }
  if (f32_disp)
    for (size_t p = 0; p < n; ++p) disp[p] = (double)d32[p];
  if (disp_out)
    for (size_t p = 0; p < n; ++p) disp_out[p] = (float)disp[p];
  // triangulation of the masked crop pixels with the coefficients at their
  // global positions: gather the crop's coefficients, then the pixel loop
  std::vector<float> cx(n * 4), cy(n * 4), cz(n * 4);
  for (int r = 0; r < h; ++r)
    for (int j = 0; j < lw; ++j) {
      const size_t g = ((size_t)(y0 + r) * cols + lx0 + j) * 4, c = ((size_t)r * lw + j) * 4;
      for (int a = 0; a < 4; ++a) {
        cx[c + a] = cxm[g + a];
        cy[c + a] = cym[g + a];
        cz[c + a] = czm[g + a];
      }
    }
  std::vector<double> dmask(disp);
  for (size_t p = 0; p < n; ++p)
    if (mask[p] == 0) dmask[p] = 0.0;
  std::vector<float> xyz0(n * 3), xyzg(n * 3), xyzr(n * 3);
  std::vector<uint8_t> P0(n), Pg(n), Pr(n);
  orc_reproject_poly(dmask.data(), h, lw, cx.data(), cy.data(), cz.data(), R1, cfg->min_depth,
                     cfg->max_depth, cfg->use_rectified_coords, xyz0.data(), P0.data());
  orc_gaussian(xyz0.data(), P0.data(), xyzg.data(), Pg.data(), h, lw, cfg->gauss_window,
               cfg->gauss_sigma, cfg->semantics);
  const uint8_t* Pin = cfg->semantics == PSGC_SEM_CPU ? P0.data() : Pg.data();
  if (cfg->outlier_mode == PSGC_OUTLIER_ROR) {
    orc_ror(xyzg.data(), Pin, xyzr.data(), Pr.data(), h, lw, cfg->radius_search,
            cfg->radius_min_neighbors, cfg->radius_window_size);
    int64_t raw = 0, kept = 0;
    for (size_t p = 0; p < n; ++p) {
      raw += Pin[p] != 0;
      kept += Pr[p] != 0;
    }
    if (raw > 0 && (double)kept < (double)raw * 0.1) {
      xyzr = xyzg;
      std::memcpy(Pr.data(), Pin, n);
    }
  } else if (cfg->outlier_mode == PSGC_OUTLIER_SOR) {
    orc_sor(xyzg.data(), Pin, xyzr.data(), Pr.data(), h, lw, cfg->sor_k, cfg->sor_std_mul,
            cfg->sor_window, nullptr, nullptr);
  } else {
    xyzr = xyzg;
    std::memcpy(Pr.data(), Pin, n);
  }
  std::memcpy(P, Pr.data(), n);
  if (cfg->use_mls)
    orc_mls(xyzr.data(), Pr.data(), xyz, h, lw, cfg->mls_window_size, cfg->mls_spatial_sigma,
            cfg->mls_polynomial_order, cfg->mls_max_radius, cfg->mls_iterations,
            cfg->semantics);
  else
    std::memcpy(xyz, xyzr.data(), n * 3 * sizeof(float));
  return PSGC_OK;
}

// ---- setData remap (JSLAB_PhaseEpipolarReconstructor_V2.cpp:368-377) -------
// cv::remap of OpenCV 4.7.0 (the reference's pinned version, PhaseShiftGrayCode3D.vcxproj:86;
// absent here) with CV_32FC1 map pairs and the default BORDER_CONSTANT 0,
// restated from its published imgwarp.cpp: RemapInvoker converts each map
// pair to fixed point, X = cvRound(map_x * 32) (round half to even; a value
// outside int range, NaN or infinity gives INT_MIN), integer tap
// saturate_cast<short>(X >> 5) and fraction X & 31 (likewise for y);
// remapBilinear for 64F images then reads the float table of
// initInterTab2D(INTER_LINEAR): w = {(1-fy)(1-fx), (1-fy)fx, fy(1-fx), fy fx}
// with fx = (X & 31)/32, exact in float, and sums
// ((v00 w0 + v01 w1) + v10 w2) + v11 w3 in double; a tap outside the source
// reads 0. remapNearest takes the tap saturate_cast<short>(cvRound(map)).
namespace {
int cv_round(float v) {
  return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)std::nearbyint(v) : INT32_MIN;
}
int sat_short(int v) { return v < -32768 ? -32768 : (v > 32767 ? 32767 : v); }
}  // namespace

int orc_remap(int kind, const void* src, int src_rows, int src_cols, const float* map_x,
              const float* map_y, int rows, int cols, void* dst) {
  if (kind != 0 && kind != 1) return PSGC_E_INVALID;
  auto inside = [&](int y, int x) { return x >= 0 && x < src_cols && y >= 0 && y < src_rows; };
#pragma omp parallel for schedule(static)
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) {
      const size_t p = (size_t)r * cols + c;
      if (kind == 1) {
        const int x = sat_short(cv_round(map_x[p])), y = sat_short(cv_round(map_y[p]));
        ((int32_t*)dst)[p] = inside(y, x) ? ((const int32_t*)src)[(size_t)y * src_cols + x] : 0;
        continue;
      }
      const int X = cv_round(map_x[p] * 32.0f), Y = cv_round(map_y[p] * 32.0f);
      const int sx = sat_short(X >> 5), sy = sat_short(Y >> 5);
      const float ax = (float)(X & 31) * (1.0f / 32), ay = (float)(Y & 31) * (1.0f / 32);
      const float w[4] = {(1.0f - ay) * (1.0f - ax), (1.0f - ay) * ax, ay * (1.0f - ax), ay * ax};
      const double* S = (const double*)src;
      auto tap = [&](int y, int x) { return inside(y, x) ? S[(size_t)y * src_cols + x] : 0.0; };
      double v = tap(sy, sx) * (double)w[0];
      v = v + tap(sy, sx + 1) * (double)w[1];
      v = v + tap(sy + 1, sx) * (double)w[2];
      v = v + tap(sy + 1, sx + 1) * (double)w[3];
      ((double*)dst)[p] = v;
    }
  return PSGC_OK;
}

int orc_atan2_many(const double* y, const double* x, double* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = psgc::atan2(y[i], x[i]);
  return PSGC_OK;
}

// ---- Unorganized clouds: voxel-grid filters -------------------------------
// Not in the reference (which filters organized clouds only). The grid:
// psgc_voxel.h's frame over the participating points' bounding box, points
// ordered by cell key, ties by index (a stable sort). Candidate neighbours
// of a point: the cells of the (2R+1)^3 block around its cell, in the order
// dz, dy, dx = -R..R (z outermost), each cell's points in grid order. Sums
// run in that order on both sides.

}  // extern "C"

namespace {

struct CloudGrid {
  psgc::VoxelFrame f;
  std::vector<int64_t> order;    // point index by grid position
  std::vector<uint64_t> ukey;    // occupied cells, ascending
  std::vector<int64_t> start, count;
  std::vector<uint64_t> key_of;  // cell key of each point (participating only)
};

bool cloud_ok(const float* xyz, const uint8_t* valid, int64_t i) {
  return (!valid || valid[i] != 0) &&
         psgc::voxel_point_ok(xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]);
}

int build_grid(const float* xyz, const uint8_t* valid, int64_t n, float voxel, CloudGrid* g) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  int64_t np = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (!cloud_ok(xyz, valid, i)) continue;
    ++np;
    for (int c = 0; c < 3; ++c) {
      mn[c] = std::min(mn[c], xyz[i * 3 + c]);
      mx[c] = std::max(mx[c], xyz[i * 3 + c]);
    }
  }
  if (np == 0) {
    mn[0] = mn[1] = mn[2] = mx[0] = mx[1] = mx[2] = 0.0f;
  }
  if (!psgc::voxel_frame(mn, mx, voxel, &g->f)) return PSGC_E_INVALID;
  g->key_of.assign(n, 0);
  std::vector<std::pair<uint64_t, int64_t>> kv;
  kv.reserve(np);
  for (int64_t i = 0; i < n; ++i) {
    if (!cloud_ok(xyz, valid, i)) continue;
    const uint64_t k = psgc::voxel_key_of(g->f, xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]);
    g->key_of[i] = k;
    kv.emplace_back(k, i);
  }
  std::sort(kv.begin(), kv.end());
  g->order.resize(kv.size());
  g->ukey.clear();
  g->start.clear();
  g->count.clear();
  for (size_t j = 0; j < kv.size(); ++j) {
    g->order[j] = kv[j].second;
    if (j == 0 || kv[j].first != kv[j - 1].first) {
      g->ukey.push_back(kv[j].first);
      g->start.push_back((int64_t)j);
      g->count.push_back(0);
    }
    g->count.back()++;
  }
  return PSGC_OK;
}

// visit(index) for every point of the (2R+1)^3 block of cells around the
// cell of p (canonical order; cells outside the frame skipped)
template <typename F>
void grid_block(const CloudGrid& g, const float* p, int R, F&& visit) {
  const int64_t ix = psgc::voxel_coord(p[0], g.f.ox, g.f.inv);
  const int64_t iy = psgc::voxel_coord(p[1], g.f.oy, g.f.inv);
  const int64_t iz = psgc::voxel_coord(p[2], g.f.oz, g.f.inv);
  for (int dz = -R; dz <= R; ++dz)
    for (int dy = -R; dy <= R; ++dy)
      for (int dx = -R; dx <= R; ++dx) {
        const int64_t cx = ix + dx, cy = iy + dy, cz = iz + dz;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= g.f.X || cy >= g.f.Y || cz >= g.f.Z) continue;
        const uint64_t kk = psgc::voxel_key(g.f, cx, cy, cz);
        auto it = std::lower_bound(g.ukey.begin(), g.ukey.end(), kk);
        if (it == g.ukey.end() || *it != kk) continue;
        const size_t c = (size_t)(it - g.ukey.begin());
        for (int64_t j = g.start[c]; j < g.start[c] + g.count[c]; ++j) visit(g.order[j]);
      }
}

inline float d2_of(const float* a, const float* b) {
  const float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
  return dx * dx + dy * dy + dz * dz;
}

}  // namespace

extern "C" {

// SOR on an unorganized cloud: mean distance to the k nearest other points
// (exact: the search widens ring by ring, R = 1..rmax cells, until the k-th
// distance is below R voxels; a point whose k-th nearest lies farther than
// rmax voxels is not eligible and is removed), threshold mean + std_mul sd
// over the eligible points, keep[i] = eligible and md <= threshold.
// md: 0 where not eligible.
int orc_cloud_sor(const float* xyz, const uint8_t* valid, int64_t n, int k, float std_mul,
                  float voxel, int rmax, uint8_t* keep, double* md_out, double* stats) {
  if (n < 0 || k < 1 || rmax < 1) return PSGC_E_INVALID;
  CloudGrid g;
  const int rc = build_grid(xyz, valid, n, voxel, &g);
  if (rc != PSGC_OK) return rc;
  std::vector<double> md(n, 0.0);
  std::vector<uint8_t> ok(n, 0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    if (!cloud_ok(xyz, valid, i)) continue;
    std::vector<float> d2s;
    for (int R = 1; R <= rmax; ++R) {
      d2s.clear();
      grid_block(g, xyz + i * 3, R, [&](int64_t j) {
        if (j != i) d2s.push_back(d2_of(xyz + i * 3, xyz + j * 3));
      });
      if ((int)d2s.size() < k) continue;
      std::nth_element(d2s.begin(), d2s.begin() + (k - 1), d2s.end());
      const float lim = psgc::cloud_knn_limit(voxel, R);
      if (!(d2s[k - 1] <= lim)) continue;
      std::sort(d2s.begin(), d2s.begin() + k);
      double s = 0;
      for (int m = 0; m < k; ++m) s += std::sqrt((double)d2s[m]);
      md[i] = s / k;
      ok[i] = 1;
      break;
    }
  }
  double st[4];
  sor_stats_of(md.data(), ok.data(), (size_t)n, std_mul, st);
  for (int64_t i = 0; i < n; ++i) {
    if (keep) keep[i] = (ok[i] && md[i] <= st[3]) ? 1 : 0;
    if (md_out) md_out[i] = md[i];
  }
  if (stats) std::memcpy(stats, st, sizeof(st));
  return PSGC_OK;
}

// MLS on an unorganized cloud: the organized filter's fit (orders, fallback,
// semantics) with the neighbours within `max_radius` (3 sigma if 0) of the
// point, itself included, found in the (2R+1)^3 cells around it
// (R = cloud_radius_cells: voxel >= radius, or half-radius cells with R = 2).
// Each iteration rebuilds the grid from the moved points.
int orc_cloud_mls(const float* xyz, const uint8_t* valid, int64_t n, float sigma, int order,
                  float max_radius, int iterations, int semantics, float voxel, float* out) {
  const float r = (max_radius > 0.0f) ? max_radius : 3.0f * sigma;
  const int R = psgc::cloud_radius_cells(&voxel, r);
  if (n < 0 || !(sigma > 0.0f) || R == 0) return PSGC_E_INVALID;
  const float r2 = r * r;
  const float inv_2s2 = 1.0f / (2.0f * sigma * sigma);
  const int ord = std::max(0, std::min(2, order));
  std::vector<float> cur(xyz, xyz + n * 3);
  std::memcpy(out, xyz, (size_t)n * 3 * sizeof(float));
  for (int it = 0; it < std::max(1, iterations); ++it) {
    CloudGrid g;
    const int rc = build_grid(cur.data(), valid, n, voxel, &g);
    if (rc != PSGC_OK) return rc;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
      if (!cloud_ok(cur.data(), valid, i)) continue;
      auto each = [&](auto&& visit) {
        grid_block(g, cur.data() + i * 3, R, [&](int64_t j) { visit(cur.data() + j * 3); });
      };
      mls_point(cur.data() + i * 3, each, inv_2s2, r2, ord, semantics, out + i * 3);
    }
    std::memcpy(cur.data(), out, (size_t)n * 3 * sizeof(float));
  }
  return PSGC_OK;
}

// Gaussian smoothing of an unorganized cloud: the weighted mean of the points
// within `radius` (the point itself included), weights exp(-d^2 / 2 sigma^2);
// candidates from the (2R+1)^3 cells as for the MLS.
int orc_cloud_gaussian(const float* xyz, const uint8_t* valid, int64_t n, float sigma,
                       float radius, float voxel, float* out) {
  if (n < 0 || !(sigma > 0.0f) || !(radius > 0.0f)) return PSGC_E_INVALID;
  const int R = psgc::cloud_radius_cells(&voxel, radius);
  if (R == 0) return PSGC_E_INVALID;
  CloudGrid g;
  const int rc = build_grid(xyz, valid, n, voxel, &g);
  if (rc != PSGC_OK) return rc;
  const float r2 = radius * radius;
  const float inv_2s2 = 1.0f / (2.0f * sigma * sigma);
  std::memcpy(out, xyz, (size_t)n * 3 * sizeof(float));
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    if (!cloud_ok(xyz, valid, i)) continue;
    float sx = 0, sy = 0, sz = 0, sw = 0;
    grid_block(g, xyz + i * 3, R, [&](int64_t j) {
      const float* pk = xyz + j * 3;
      const float d2 = d2_of(xyz + i * 3, pk);
      if (d2 > r2) return;
      const float w = psgc::exp_f32(-d2 * inv_2s2);
      sx += w * pk[0];
      sy += w * pk[1];
      sz += w * pk[2];
      sw += w;
    });
    out[i * 3] = sx / sw;
    out[i * 3 + 1] = sy / sw;
    out[i * 3 + 2] = sz / sw;
  }
  return PSGC_OK;
}

// Brute-force pins for small clouds: for every participating point, its k
// smallest squared distances to the other participating points (ascending;
// +inf where fewer), and the count of points within `radius` (itself
// included).
int orc_cloud_brute(const float* xyz, const uint8_t* valid, int64_t n, int k, float radius,
                    float* knn_d2, int32_t* within) {
  const float r2 = radius * radius;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    std::vector<float> d2s;
    int32_t cnt = 0;
    if (cloud_ok(xyz, valid, i))
      for (int64_t j = 0; j < n; ++j) {
        if (!cloud_ok(xyz, valid, j)) continue;
        const float d2 = d2_of(xyz + i * 3, xyz + j * 3);
        if (d2 <= r2) ++cnt;
        if (j != i) d2s.push_back(d2);
      }
    std::sort(d2s.begin(), d2s.end());
    for (int m = 0; m < k; ++m) knn_d2[i * k + m] = m < (int)d2s.size() ? d2s[m] : INFINITY;
    within[i] = cnt;
  }
  return PSGC_OK;
}

// Cell coordinates of the shared voxel frame (psgc_voxel.h) along one axis:
// out[i] = the cell of coordinate v[i] for origin o and edge voxel, as the
// grid builds of the device and of this restatement assign it.
int orc_voxel_cells(const float* v, int64_t n, float o, float voxel, int64_t* out) {
  if (!(voxel > 0.0f)) return PSGC_E_INVALID;
  const double inv = 1.0 / (double)voxel;
  for (int64_t i = 0; i < n; ++i) out[i] = psgc::voxel_coord(v[i], o, inv);
  return PSGC_OK;
}

// The device's per-point bound of its first k-nearest pass
// (psgc::cloud_knn_point_limit) for points xyz[n][3] in the frame of origin
// o[3] and edge voxel.
int orc_knn_point_limit(const float* xyz, int64_t n, const float* o, float voxel, float* out) {
  if (!(voxel > 0.0f)) return PSGC_E_INVALID;
  psgc::VoxelFrame f{};
  f.ox = o[0];
  f.oy = o[1];
  f.oz = o[2];
  f.voxel = voxel;
  f.inv = 1.0 / (double)voxel;
  for (int64_t i = 0; i < n; ++i)
    out[i] = psgc::cloud_knn_point_limit(f, xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]);
  return PSGC_OK;
}

// The fixed sorting network of the device's k-nearest selection
// (psgc::make_sort_net(k), k <= 26) on every input of zeros and ones (2^k of
// them, 64 at a time as bit lanes): the number left unsorted. By the 0-1
// principle a comparator network that sorts these sorts every input.
int orc_check_sort_net(int k, int64_t* unsorted) {
  if (k < 2 || k > 26 || !unsorted) return PSGC_E_INVALID;
  const psgc::SortNet net = psgc::make_sort_net(k);
  const int64_t words = k <= 6 ? 1 : (int64_t)1 << (k - 6);
  int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad)
  for (int64_t w = 0; w < words; ++w) {
    static const uint64_t lane_bits[6] = {0xAAAAAAAAAAAAAAAAull, 0xCCCCCCCCCCCCCCCCull,
                                          0xF0F0F0F0F0F0F0F0ull, 0xFF00FF00FF00FF00ull,
                                          0xFFFF0000FFFF0000ull, 0xFFFFFFFF00000000ull};
    uint64_t v[26];
    for (int i = 0; i < k; ++i)
      v[i] = i < 6 ? lane_bits[i] : (((w >> (i - 6)) & 1) ? ~0ull : 0ull);
    for (int c = 0; c < net.n; ++c) {
      const uint64_t a = v[net.a[c]], b = v[net.b[c]];
      v[net.a[c]] = a & b;  // min
      v[net.b[c]] = a | b;  // max
    }
    uint64_t ok = ~0ull;
    for (int i = 0; i + 1 < k; ++i) ok &= ~(v[i] & ~v[i + 1]);
    // lanes beyond 2^k (k < 6) repeat earlier inputs; count distinct ones only
    const uint64_t live = k >= 6 ? ~0ull : ((1ull << (1 << k)) - 1);
    bad += __builtin_popcountll(~ok & live);
  }
  *unsorted = bad;
  return PSGC_OK;
}

// Exhaustive comparison of psgc::{exp,acos,cos}_f32 (which = 0, 1, 2) with
// the C library over every `stride`-th float in [lo, hi]: the largest ulp
// distance from the float library routine, how many results differ from it,
// and how many differ from the double routine rounded to float (a correctly
// rounded reference in all but rare double-rounding cases).
int orc_check_f32(int which, float lo, float hi, int stride, int64_t* max_ulp,
                  int64_t* n_diff_libf, int64_t* n_diff_cr, int64_t* n_total) {
  auto key = [](float f) -> int64_t {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i < 0 ? -(int64_t)(i & 0x7fffffff) : (int64_t)i;
  };
  auto from_key = [](int64_t k) -> float {
    int32_t i = k < 0 ? (int32_t)((-k) | 0x80000000LL) : (int32_t)k;
    float f;
    std::memcpy(&f, &i, 4);
    return f;
  };
  const int64_t k0 = key(lo), k1 = key(hi);
  int64_t mu = 0, nd = 0, nc = 0, nt = 0;
#pragma omp parallel for reduction(max : mu) reduction(+ : nd, nc, nt)
  for (int64_t k = k0; k <= k1; k += stride) {
    const float x = from_key(k);
    float ours, libf, cr;
    if (which == 0) {
      ours = psgc::exp_f32(x); libf = std::exp(x); cr = (float)std::exp((double)x);
    } else if (which == 1) {
      ours = psgc::acos_f32(x); libf = std::acos(x); cr = (float)std::acos((double)x);
    } else {
      ours = psgc::cos_f32(x); libf = std::cos(x); cr = (float)std::cos((double)x);
    }
    const int64_t d = std::llabs(key(ours) - key(libf));
    if (d > mu) mu = d;
    nd += ours != libf;
    nc += ours != cr;
    ++nt;
  }
  *max_ulp = mu;
  *n_diff_libf = nd;
  *n_diff_cr = nc;
  *n_total = nt;
  return PSGC_OK;
}

// The organized MLS kernel's weight exponential psgc::exp_f32_w (256-entry
// table, one fused multiply-add for the reduction's integer) against
// psgc::exp_f32, which the restatement calls: the number of floats in
// [lo, hi] (every `stride`-th) whose results differ in any bit.
int orc_check_exp_w(float lo, float hi, int stride, int64_t* n_diff, int64_t* n_total) {
  if (!n_diff || !n_total || stride < 1) return -1;
  auto key = [](float f) -> int64_t {
    int32_t i;
    std::memcpy(&i, &f, 4);
    return i < 0 ? -(int64_t)(i & 0x7fffffff) : (int64_t)i;
  };
  auto from_key = [](int64_t k) -> float {
    int32_t i = k < 0 ? (int32_t)((-k) | 0x80000000LL) : (int32_t)k;
    float f;
    std::memcpy(&f, &i, 4);
    return f;
  };
  const int64_t k0 = key(lo), k1 = key(hi);
  int64_t nd = 0, nt = 0;
#pragma omp parallel for reduction(+ : nd, nt) schedule(static, 1 << 16)
  for (int64_t k = k0; k <= k1; k += stride) {
    const float x = from_key(k);
    const float a = psgc::exp_f32_w(x, psgc::kExp2Tab256), b = psgc::exp_f32(x);
    nd += std::memcmp(&a, &b, 4) != 0;
    ++nt;
  }
  *n_diff = nd;
  *n_total = nt;
  return PSGC_OK;
}

// Threads the OpenMP loops above run on (the timed CPU baseline's core count).
int orc_num_threads() { return omp_get_max_threads(); }

}  // extern "C"
