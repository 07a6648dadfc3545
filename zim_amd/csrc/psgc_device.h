// psgc_device.h — per-pixel device routines shared by the plain and the
// tiled kernels. Each routine states the reference lines it follows; the
// operation order is the reference CPU path's, evaluated without
// floating-point contraction.
#pragma once

#include "psgc_internal.h"
#include "psgc_math.h"

namespace psgc {

// Global-memory accesses through pointers taken from work descriptors or
// argument structs. Without the address space the compiler emits flat
// accesses, which also count against the LDS wait counter, so a kernel's LDS
// reads would wait for its outstanding global loads and stores.
template <typename T>
__device__ inline T ld_global(const T* p, size_t i) {
  return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <typename T>
__device__ inline void st_global(T* p, size_t i, T v) {
  ((__attribute__((address_space(1))) T*)p)[i] = v;
}
// float4 (a class with constructors) through the plain 4-wide vector type
__device__ inline float4 ld_global(const float4* p, size_t i) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f v = ((const __attribute__((address_space(1))) v4f*)p)[i];
  return make_float4(v.x, v.y, v.z, v.w);
}

// ---------------------------------------------------------------- decode --

__device__ inline double load_sample(const DecodeArgs& a, int f, int r, int x) {
  const char* row = (const char*)a.frames[f] + (size_t)r * a.stride;
  if (a.sample_type == PSGC_SAMPLE_U16) return (double)ld_global((const uint16_t*)row, x);
  if (a.sample_type == PSGC_SAMPLE_U8) return (double)ld_global((const uint8_t*)row, x);
  return ld_global((const double*)row, x);
}

// 3x3 pre-blur at (r, x), replicated border: row pass c*x + s*(l + r) on rows
// r-1, r, r+1, then the same combination down the column
// (cv::GaussianBlur in JSLAB_ImageTypeConvertBodyV2,
// JSLAB_PhaseShiftDecoder_V2.cpp:22-26).
__device__ inline double blurred_sample(const DecodeArgs& a, int f, int r, int x) {
  const int xl = x > 0 ? x - 1 : 0, xr = x < a.cols - 1 ? x + 1 : a.cols - 1;
  const int ru = r > 0 ? r - 1 : 0, rd = r < a.rows - 1 ? r + 1 : a.rows - 1;
  const double c = a.blur_c, s = a.blur_s;
  const double hu = c * load_sample(a, f, ru, x) +
                    s * (load_sample(a, f, ru, xl) + load_sample(a, f, ru, xr));
  const double hm = c * load_sample(a, f, r, x) +
                    s * (load_sample(a, f, r, xl) + load_sample(a, f, r, xr));
  const double hd = c * load_sample(a, f, rd, x) +
                    s * (load_sample(a, f, rd, xl) + load_sample(a, f, rd, xr));
  return c * hm + s * (hu + hd);
}

// Wrapped phase and normalized modulation of one level from its sums
// S = -sum sin(d_k) I_k, C = sum cos(d_k) I_k, G = sum I_k
// (JSLAB_PhaseShiftDecoder_V2.cpp:99-117). Returns false when the pixel is
// rejected (modulation below the threshold).
__device__ inline bool level_finish(double S, double C, double G, int n, double thr,
                                    double* wrap, double* mod, const double* atab = kAtanTab) {
  *wrap = psgc::atan2(S, C, atab);
  double b = S * S + C * C;
  double mean;
  if ((n & (n - 1)) == 0) {
    // dividing by a power of two only scales (exactly, for these non-negative
    // sums): the same bits as the multiplication by 2^-log2(n), without the
    // FP64 division sequence
    const double inv = ldexp(1.0, -__builtin_ctz((unsigned)n));
    mean = G * inv;
    b = 2.0 * sqrt(b) * inv;
  } else {
    mean = G / n;
    b = 2.0 * sqrt(b) / n;
  }
  double nm = 0.0;
  if (mean > 1e-3) nm = b / mean;
  *mod = nm;
  return !(nm < thr);
}

// HDR least-squares fit of level i at (r, x) (JSLAB_HDRPhaseFit.h:43-53 as
// used at JSLAB_PhaseShiftDecoder_V2.cpp:73-89): the raw samples with
// 0 <= I < saturation select a subset, whose pseudo-inverse columns give
// beta = (A, C, S) of I = A + C cos d - S sin d. Returns false when the
// subset is rank deficient, beta is not finite, the amplitude |(C, S)| is not
// above hdr_min_amp or A is not above 1e-3.
__device__ inline bool hdr_fit(const DecodeArgs& a, int i, int r, int x, double beta[3]) {
  const int n = a.shift_num[i];
  const int base = a.level_base[i];
  double I[12];
  unsigned mask = 0;
  for (int k = 0; k < n; ++k) {
    I[k] = load_sample(a, base + k, r, x);
    if (I[k] >= 0.0 && I[k] < a.hdr_sat) mask |= 1u << k;
  }
  if (!a.hdr_ok[i][mask]) return false;
  const double* w = a.hdr_w[i] + (size_t)mask * n * 3;
  beta[0] = beta[1] = beta[2] = 0.0;
  for (int k = 0; k < n; ++k)
    if (mask >> k & 1u) {
      beta[0] += w[k * 3 + 0] * I[k];
      beta[1] += w[k * 3 + 1] * I[k];
      beta[2] += w[k * 3 + 2] * I[k];
    }
  if (!(isfinite(beta[0]) && isfinite(beta[1]) && isfinite(beta[2]))) return false;
  if (hypot(beta[1], beta[2]) <= a.hdr_min_amp || beta[0] <= 1e-3) return false;
  return true;
}

// Three-frequency heterodyne unwrapping and fringe order
// (JSLAB_PhaseShiftDecoder_V2.cpp:124-179).
__device__ inline void heterodyne(const DecodeArgs& a, const double wrap[3], double* ph,
                                  int32_t* k) {
  const double f0 = a.freq[0], f1 = a.freq[1], f2 = a.freq[2];
  double phase12, phase23, phase123;
  if (wrap[0] > wrap[1]) phase12 = wrap[0] - wrap[1];
  else phase12 = wrap[0] - wrap[1] + 2 * kPi;
  if (f1 > f2) {
    if (wrap[1] > wrap[2]) phase23 = wrap[1] - wrap[2];
    else phase23 = wrap[1] - wrap[2] + 2 * kPi;
  } else {
    if (wrap[2] > wrap[1]) phase23 = wrap[2] - wrap[1];
    else phase23 = wrap[2] - wrap[1] + 2 * kPi;
  }
  if (phase12 > phase23) phase123 = phase12 - phase23;
  else phase123 = phase12 - phase23 + 2 * kPi;
  const double f12 = (double)a.f12, f23 = (double)a.f23;
  const double u12 = phase12 + round_half_away((phase123 * f12 - phase12) / (2 * kPi)) * 2 * kPi;
  const double kv = round_half_away((u12 * f0 / f12 - wrap[0]) / (2 * kPi));
  const double u1 = wrap[0] + kv * 2 * kPi;
  const double u23 = phase23 + round_half_away((phase123 * f23 - phase23) / (2 * kPi)) * 2 * kPi;
  double u2 = wrap[1] + round_half_away((u23 * f1 / f23 - wrap[1]) / (2 * kPi)) * 2 * kPi;
  u2 = (u2 * f0) / f1;
  double u3 = wrap[2] + round_half_away((u23 * f2 / f23 - wrap[2]) / (2 * kPi)) * 2 * kPi;
  u3 = (u3 * f0) / f2;
  *ph = a.avg_three ? (u1 + u2 + u3) / 3.0 : u1;
  *k = (int32_t)kv;
}

// Gray-code order with complement frames (MSB first) combined with the
// wrapped phase moved to [0, 2 pi). Returns false for low contrast.
__device__ inline bool gray_order(const DecodeArgs& a, int r, int x, double wrap0, double* ph,
                                  int32_t* k) {
  const int nphase = a.shift_num[0];
  unsigned g = 0;
  for (int b = 0; b < a.gray_bits; ++b) {
    const double v = load_sample(a, nphase + b, r, x);
    const double c = load_sample(a, nphase + a.gray_bits + b, r, x);
    const double d = v - c;
    if (fabs(d) < a.gray_min_contrast) return false;
    g = (g << 1) | (d > 0 ? 1u : 0u);
  }
  unsigned n = g;
  for (unsigned s = 1; s < 32; s <<= 1) n ^= n >> s;
  const double w = wrap0 < 0 ? wrap0 + 2 * kPi : wrap0;
  *ph = w + 2 * kPi * (double)n;
  *k = (int32_t)n;
  return true;
}

// sqrt((double)f) for a finite f >= 0: the instruction sequence the
// compiler emits for a double square root (reciprocal square-root estimate,
// then fused multiply-add corrections of the root and the half reciprocal),
// without its scaling of arguments below 2^-767 and its fix-up of zero and
// infinity, which no non-zero float needs; zero is returned as is. The same
// bits as sqrt for every such argument, in about two thirds of the
// instructions (the SOR sums of k square roots per point).
__device__ inline double sqrt_of_float(float f) {
  const double x = (double)f;
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = y * 0.5;
  const double e = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, e, g);
  h = __builtin_fma(h, e, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  return f == 0.0f ? x : g;
}

// k-th smallest of v[0..n) (n small), by partial selection.
__device__ inline double select_kth(double* v, int n, int kth) {
  for (int i = 0; i <= kth; ++i) {
    int mi = i;
    for (int t = i + 1; t < n; ++t)
      if (v[t] < v[mi]) mi = t;
    const double tmp = v[i];
    v[i] = v[mi];
    v[mi] = tmp;
  }
  return v[kth];
}

// ----------------------------------------------------------------- match --

// Parabolic sub-pixel root around the integer match `best` of left column j
// (JSLAB_PhaseEpipolarReconstructor_V2.cpp:992-1003). `rrow` is the right
// phase row.
__device__ inline double subpixel_delta(const double* rrow, int best, double lv) {
  const double c0 = rrow[best - 1] - lv;
  const double c1 = rrow[best] - lv;
  const double c2 = rrow[best + 1] - lv;
  const double num = c1 * (c0 - c2);
  const double den = 2.0 * (c0 - c1) * (c1 - c2);
  double delta = 0.0;
  if (fabs(den) > 1e-9) delta = num / den;
  return delta;
}

__device__ inline double subpixel_disparity(const double* rrow, int best, double lv, int j) {
  return (double)j - (double)best - subpixel_delta(rrow, best, lv);
}

// Q reprojection of one pixel (JSLAB_PhaseEpipolarReconstructor_V2.cpp:
// 423-459). Returns the mask value; out receives the f32 point or zeros.
__device__ inline uint8_t reproject_point(const double* q, const double* r1, int i, int j,
                                          double d, double min_depth, double max_depth,
                                          int use_rect, float out[3]) {
  out[0] = out[1] = out[2] = 0.f;
  if (fabs(d) < 1e-5) return 0;
  const double w = q[14] * d + q[15];
  double x = j + q[3];
  double y = i + q[7];
  double z = q[11];
  const double inv_w = 1.0 / w;
  x *= inv_w;
  y *= inv_w;
  z *= inv_w;
  if (use_rect) {
    if (z > min_depth && z < max_depth) {
      out[0] = (float)x;
      out[1] = (float)y;
      out[2] = (float)z;
      return 1;
    }
    return 0;
  }
  const double rx = r1[0] * x + r1[3] * y + r1[6] * z;
  const double ry = r1[1] * x + r1[4] * y + r1[7] * z;
  const double rz = r1[2] * x + r1[5] * y + r1[8] * z;
  if (rz > min_depth && rz < max_depth) {
    out[0] = (float)rx;
    out[1] = (float)ry;
    out[2] = (float)rz;
    return 1;
  }
  return 0;
}

// Per-pixel polynomial triangulation (evaluateV2CalibPoint and the point
// loop of startCalculate_v2, JSLAB_PhaseEpipolarReconstructor_V2.cpp:13-76,
// 493-545). c = (cx, cy, cz): the pixel's three 4-term coefficient vectors;
// markers in cx.w / cz.w select the form: cx.w = -1 inverse-depth cubic,
// -2 fitted ray (cz.w = -1: inverse-depth quadratic, else depth cubic), -3
// z-linear x/y (cz.w = -4: rational depth, else depth cubic); otherwise
// cz.w = -1 inverse-depth quadratic, else cubics in d for x, y and z. The
// expressions keep the reference's operation order (d^2 * d versus d3,
// cz2 * d * d versus cz2 * d2) so the results are the CPU path's bits.
__device__ inline bool poly_marker(float v, float m) { return fabsf(v - m) < 1e-6f; }

__device__ inline uint8_t reproject_poly(const float4 cx, const float4 cy, const float4 cz,
                                         const double* r1, double d, double min_depth,
                                         double max_depth, int use_rect, float out[3]) {
  out[0] = out[1] = out[2] = 0.f;
  if (fabs(d) < 1e-5) return 0;
  if (!(fabsf(cz.x) > 1e-9 || fabsf(cz.y) > 1e-9 || fabsf(cz.z) > 1e-9 || fabsf(cz.w) > 1e-9))
    return 0;
  double x, y, z;
  const double d2 = d * d;
  if (poly_marker(cx.w, -1.0f)) {
    const double inv_z = cz.x + cz.y * d + cz.z * d2 + cz.w * d2 * d;
    if (!isfinite(inv_z) || fabs(inv_z) < 1e-12) return 0;
    z = 1.0 / inv_z;
    x = cx.x * z;
    y = cy.x * z;
  } else if (poly_marker(cx.w, -2.0f)) {
    if (poly_marker(cz.w, -1.0f)) {
      const double inv_z = cz.x + cz.y * d + cz.z * d2;
      if (!isfinite(inv_z) || fabs(inv_z) < 1e-12) return 0;
      z = 1.0 / inv_z;
    } else {
      z = cz.x + cz.y * d + cz.z * d2 + cz.w * d2 * d;
    }
    x = cx.x * z;
    y = cy.x * z;
  } else if (poly_marker(cx.w, -3.0f)) {
    if (poly_marker(cz.w, -4.0f)) {
      const double denom = cz.y * d + 1.0;
      if (!isfinite(denom) || fabs(denom) < 1e-12) return 0;
      z = cz.x / denom;
    } else {
      const double d3 = d2 * d;
      z = cz.x + cz.y * d + cz.z * d2 + cz.w * d3;
    }
    x = cx.x * z + cx.y;
    y = cy.x * z + cy.y;
  } else if (poly_marker(cz.w, -1.0f)) {
    const double inv_z = cz.x + cz.y * d + cz.z * d * d;
    if (!isfinite(inv_z) || fabs(inv_z) < 1e-12) return 0;
    z = 1.0 / inv_z;
    x = cx.x * z;
    y = cy.x * z;
  } else {
    const double d3 = d2 * d;
    x = cx.x + cx.y * d + cx.z * d2 + cx.w * d3;
    y = cy.x + cy.y * d + cy.z * d2 + cy.w * d3;
    z = cz.x + cz.y * d + cz.z * d2 + cz.w * d3;
  }
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) return 0;
  if (use_rect) {
    if (z > min_depth && z < max_depth) {
      out[0] = (float)x;
      out[1] = (float)y;
      out[2] = (float)z;
      return 1;
    }
    return 0;
  }
  const double rx = r1[0] * x + r1[3] * y + r1[6] * z;
  const double ry = r1[1] * x + r1[4] * y + r1[7] * z;
  const double rz = r1[2] * x + r1[5] * y + r1[8] * z;
  if (rz > min_depth && rz < max_depth) {
    out[0] = (float)rx;
    out[1] = (float)ry;
    out[2] = (float)rz;
    return 1;
  }
  return 0;
}

// Squared distance in f32, evaluated as dx*dx + dy*dy + dz*dz
// (JSLAB_PhaseEpipolarReconstructor_V2.cpp:171-174).
template <bool C = false>
__device__ inline float dist2(float px, float py, float pz, const float* q) {
  return sq3<C>(px - q[0], py - q[1], pz - q[2]);
}

// ------------------------------------------------------------------- MLS --

// symmetricEigen3x3 (JSLAB_MLS_Filter.cpp:11-106): ascending eigenvalues,
// eigenvectors as rows of V.
__device__ inline void eigen3(const float C[3][3], float ev[3], float V[3][3]) {
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
          float t = ev[i];
          ev[i] = ev[j];
          ev[j] = t;
          for (int c = 0; c < 3; ++c) {
            t = V[i][c];
            V[i][c] = V[j][c];
            V[j][c] = t;
          }
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
  const float p = sqrtf(p2 / 6.0f);
  const float inv_p = (p > 1e-15f) ? (1.0f / p) : 0.0f;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) B[i][j] *= inv_p;
  const float detB = B[0][0] * (B[1][1] * B[2][2] - B[1][2] * B[2][1]) -
                     B[0][1] * (B[1][0] * B[2][2] - B[1][2] * B[2][0]) +
                     B[0][2] * (B[1][0] * B[2][1] - B[1][1] * B[2][0]);
  float rr = detB * 0.5f;
  rr = fmaxf(-1.0f, fminf(1.0f, rr));
  const float phi = acos_f32(rr) / 3.0f;
  const float PIf = 3.14159265358979323846f;
  ev[0] = q + 2.0f * p * cos_f32(phi + 2.0f * PIf / 3.0f);
  ev[2] = q + 2.0f * p * cos_f32(phi);
  ev[1] = trace - ev[0] - ev[2];
  for (int idx = 0; idx < 3; ++idx) {
    float A[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i][j] = C[i][j];
    A[0][0] -= ev[idx];
    A[1][1] -= ev[idx];
    A[2][2] -= ev[idx];
    const float c01[3] = {A[0][1] * A[1][2] - A[0][2] * A[1][1],
                          A[0][2] * A[1][0] - A[0][0] * A[1][2],
                          A[0][0] * A[1][1] - A[0][1] * A[1][0]};
    const float c02[3] = {A[0][1] * A[2][2] - A[0][2] * A[2][1],
                          A[0][2] * A[2][0] - A[0][0] * A[2][2],
                          A[0][0] * A[2][1] - A[0][1] * A[2][0]};
    const float c12[3] = {A[1][1] * A[2][2] - A[1][2] * A[2][1],
                          A[1][2] * A[2][0] - A[1][0] * A[2][2],
                          A[1][0] * A[2][1] - A[1][1] * A[2][0]};
    const float l01 = c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2];
    const float l02 = c02[0] * c02[0] + c02[1] * c02[1] + c02[2] * c02[2];
    const float l12 = c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2];
    const float* best = c12;
    float bl = l12;
    if (l01 > bl) { best = c01; bl = l01; }
    if (l02 > bl) { best = c02; bl = l02; }
    const float inv = (bl > 1e-20f) ? (1.0f / sqrtf(bl)) : 0.0f;
    V[idx][0] = best[0] * inv;
    V[idx][1] = best[1] * inv;
    V[idx][2] = best[2] * inv;
  }
}

// choleskySolve (JSLAB_MLS_Filter.cpp:113-156).
__device__ inline bool chol_solve(const float A[6][6], const float b[6], float x[6], int nc) {
  float L[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) L[i][j] = 0.f;
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j <= i; ++j) {
      float s = A[i][j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (s <= 1e-12f) return false;
        L[i][j] = sqrtf(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  float y[6] = {0, 0, 0, 0, 0, 0};
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

// Effective polynomial order after the neighbour-count fallback:
// CPU rule JSLAB_MLS_Filter.cpp:277-279, GPU rule JSLAB_MLS_Filter.cu:283-286.
__device__ inline int mls_order(int order, int nv, int semantics) {
  const int minnb0 = 3, minnb1 = 5, minnb2 = 8;
  const int need = order == 2 ? minnb2 : (order == 1 ? minnb1 : minnb0);
  int cur;
  if (semantics == PSGC_SEM_CPU) cur = (nv >= need) ? order : 0;
  else cur = order;
  if (cur > 0 && nv < (cur == 2 ? minnb2 : minnb1)) cur = (nv >= minnb1) ? 1 : 0;
  return cur;
}

// Projection of point pi onto the polynomial fitted in its local frame, from
// centroid, covariance frame and normal equations
// (JSLAB_MLS_Filter.cpp:283-388). Returns the projected point in out.
struct MlsFit {
  float cx, cy, cz;
  float n[3], t1[3], t2[3];
};

__device__ inline void mls_frame(float cov[3][3], MlsFit* f) {
  cov[1][0] = cov[0][1];
  cov[2][0] = cov[0][2];
  cov[2][1] = cov[1][2];
  float ev[3], V[3][3];
  eigen3(cov, ev, V);
  for (int c = 0; c < 3; ++c) {
    f->n[c] = V[0][c];
    f->t1[c] = V[1][c];
    f->t2[c] = V[2][c];
  }
}

__device__ inline void mls_plane(const MlsFit& f, const float pi[3], float out[3]) {
  const float dot = (pi[0] - f.cx) * f.n[0] + (pi[1] - f.cy) * f.n[1] + (pi[2] - f.cz) * f.n[2];
  out[0] = pi[0] - dot * f.n[0];
  out[1] = pi[1] - dot * f.n[1];
  out[2] = pi[2] - dot * f.n[2];
}

// Evaluates the fitted polynomial at the query point and maps it back to
// world coordinates (JSLAB_MLS_Filter.cpp:377-388).
__device__ inline void mls_eval_poly(const MlsFit& f, const float co[6], int cur,
                                     const float pi[3], float out[3]) {
  const float ex0 = pi[0] - f.cx, ey0 = pi[1] - f.cy, ez0 = pi[2] - f.cz;
  const float u0 = ex0 * f.t1[0] + ey0 * f.t1[1] + ez0 * f.t1[2];
  const float v0 = ex0 * f.t2[0] + ey0 * f.t2[1] + ez0 * f.t2[2];
  float ff = co[0] + co[1] * u0 + co[2] * v0;
  if (cur == 2) ff += co[3] * u0 * u0 + co[4] * u0 * v0 + co[5] * v0 * v0;
  out[0] = f.cx + u0 * f.t1[0] + v0 * f.t2[0] + ff * f.n[0];
  out[1] = f.cy + u0 * f.t1[1] + v0 * f.t2[1] + ff * f.n[1];
  out[2] = f.cz + u0 * f.t1[2] + v0 * f.t2[2] + ff * f.n[2];
}

// Solve (with the order-2 -> 1 fallback of JSLAB_MLS_Filter.cpp:350-374) and
// evaluate at the query point.
__device__ inline void mls_solve_eval(const MlsFit& f, float AtWA[6][6], float AtWf[6], int cur,
                                      const float pi[3], float out[3]) {
  const int nc = cur == 1 ? 3 : 6;
  for (int a = 0; a < nc; ++a)
    for (int b = 0; b < a; ++b) AtWA[a][b] = AtWA[b][a];
  float co[6] = {0, 0, 0, 0, 0, 0};
  bool solved = chol_solve(AtWA, AtWf, co, nc);
  if (!solved && cur == 2) {
    float A1[6][6], b1[6];
    for (int a = 0; a < 6; ++a) {
      b1[a] = a < 3 ? AtWf[a] : 0.f;
      for (int b = 0; b < 6; ++b) A1[a][b] = (a < 3 && b < 3) ? AtWA[a][b] : 0.f;
    }
    solved = chol_solve(A1, b1, co, 3);
    if (solved) cur = 1;
  }
  if (!solved) {
    mls_plane(f, pi, out);
    return;
  }
  mls_eval_poly(f, co, cur, pi, out);
}

// Accumulates one neighbour pk (weight w) into the normal equations.
template <bool C>
__device__ inline void mls_accum_poly(const MlsFit& f, const float* pk, float w, int nc,
                                      float AtWA[6][6], float AtWf[6]) {
  const float ex = pk[0] - f.cx, ey = pk[1] - f.cy, ez = pk[2] - f.cz;
  const float u = madd<C>(ez, f.t1[2], madd<C>(ey, f.t1[1], ex * f.t1[0]));
  const float v = madd<C>(ez, f.t2[2], madd<C>(ey, f.t2[1], ex * f.t2[0]));
  const float fv = madd<C>(ez, f.n[2], madd<C>(ey, f.n[1], ex * f.n[0]));
  float basis[6] = {1.0f, u, v, 0.f, 0.f, 0.f};
  if (nc == 6) {
    basis[3] = u * u;
    basis[4] = u * v;
    basis[5] = v * v;
  }
  for (int a = 0; a < nc; ++a) {
    const float wa = w * basis[a];
    AtWf[a] = madd<C>(wa, fv, AtWf[a]);
    for (int b = a; b < nc; ++b) AtWA[a][b] = madd<C>(wa, basis[b], AtWA[a][b]);
  }
}

// ---- fixed-size variants (register resident) ------------------------------
// The same arithmetic as chol_solve / mls_accum_poly / mls_solve_eval with
// the system size a compile-time constant, so every array index is static and
// the normal equations stay in registers instead of private memory. Only the
// upper triangle of A is accumulated; A[j][i] (j <= i) is the value the
// reference reads after mirroring it into the lower triangle.

// Cholesky solve of the leading N x N block of A (N <= NA).
template <int N, int NA>
__device__ inline bool chol_solve_fixed(const float (&A)[NA][NA], const float (&b)[NA],
                                        float x[6]) {
  float L[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      float s = A[j][i];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (s <= 1e-12f) return false;
        L[i][j] = sqrtf(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  float y[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    float s = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
    y[i] = s / L[i][i];
  }
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    float s = y[i];
#pragma unroll
    for (int k = i + 1; k < N; ++k) s -= L[k][i] * x[k];
    x[i] = s / L[i][i];
  }
  return true;
}

// One neighbour into the NC-term normal equations. For NC = 6 the leading
// 3 x 3 block receives exactly the terms of the NC = 3 system, so a point
// whose order fell back to 1 solves that block.
template <int NC, bool C>
__device__ inline void mls_accum_fixed(const MlsFit& f, const float pk[3], float w,
                                       float (&A)[NC][NC], float (&bv)[NC]) {
  const float ex = pk[0] - f.cx, ey = pk[1] - f.cy, ez = pk[2] - f.cz;
  const float u = madd<C>(ez, f.t1[2], madd<C>(ey, f.t1[1], ex * f.t1[0]));
  const float v = madd<C>(ez, f.t2[2], madd<C>(ey, f.t2[1], ex * f.t2[0]));
  const float fv = madd<C>(ez, f.n[2], madd<C>(ey, f.n[1], ex * f.n[0]));
  const float basis[6] = {1.0f, u, v, u * u, u * v, v * v};
#pragma unroll
  for (int a = 0; a < NC; ++a) {
    const float wa = w * basis[a];  // w * basis[a] * x evaluates as (w * basis[a]) * x
    bv[a] = madd<C>(wa, fv, bv[a]);
#pragma unroll
    for (int b = a; b < NC; ++b) A[a][b] = madd<C>(wa, basis[b], A[a][b]);
  }
}

// Centroid and covariance terms of one neighbour (weights w), as the first
// two passes accumulate them.
template <bool C>
__device__ inline void mls_accum_centroid(const float pk[3], float w, MlsFit* f, float* wsum) {
  f->cx = madd<C>(w, pk[0], f->cx);
  f->cy = madd<C>(w, pk[1], f->cy);
  f->cz = madd<C>(w, pk[2], f->cz);
  *wsum += w;
}

template <bool C>
__device__ inline void mls_accum_cov(const MlsFit& f, const float pk[3], float w,
                                     float (&cov)[3][3]) {
  const float ex = pk[0] - f.cx, ey = pk[1] - f.cy, ez = pk[2] - f.cz;
  const float wx = w * ex, wy = w * ey, wz = w * ez;
  cov[0][0] = madd<C>(wx, ex, cov[0][0]);
  cov[0][1] = madd<C>(wx, ey, cov[0][1]);
  cov[0][2] = madd<C>(wx, ez, cov[0][2]);
  cov[1][1] = madd<C>(wy, ey, cov[1][1]);
  cov[1][2] = madd<C>(wy, ez, cov[1][2]);
  cov[2][2] = madd<C>(wz, ez, cov[2][2]);
}

// mls_solve_eval for an NC-term accumulation (NC = 6 serves orders 1 and 2).
template <int NC>
__device__ inline void mls_solve_eval_fixed(const MlsFit& f, const float (&A)[NC][NC],
                                            const float (&bv)[NC], int cur, const float pi[3],
                                            float out[3]) {
  float co[6] = {0, 0, 0, 0, 0, 0};
  bool solved;
  if (NC == 6 && cur == 2) {
    solved = chol_solve_fixed<NC, NC>(A, bv, co);
    if (!solved) {
      // the reference's fallback starts from zeroed coefficients
      co[0] = co[1] = co[2] = co[3] = co[4] = co[5] = 0.f;
      solved = chol_solve_fixed<3, NC>(A, bv, co);
      if (solved) cur = 1;
    }
  } else {
    solved = chol_solve_fixed<3, NC>(A, bv, co);
  }
  if (!solved) {
    mls_plane(f, pi, out);
    return;
  }
  mls_eval_poly(f, co, cur, pi, out);
}

}  // namespace psgc
