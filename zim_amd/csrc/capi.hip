// capi.hip — the extern "C" entry points of libpsgc.so (include/psgc.h):
// argument validation, context-owned buffers, host staging and the
// orchestration of the per-stage kernels.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdarg>
#include <cstring>
#include <vector>

#include "psgc_internal.h"

namespace psgc {

bool g_prof_on = false;

namespace {
thread_local char g_err[1024] = "";

struct ProfRec {
  std::string name;
  hipEvent_t beg, end;
};
std::vector<ProfRec> g_prof;           // closed and open intervals
std::vector<size_t> g_prof_open;       // indices of open intervals

constexpr double PI = 3.14159265358979323846;
}  // namespace

void prof_begin(const char* name, hipStream_t s) {
  ProfRec r;
  r.name = name;
  if (hipEventCreate(&r.beg) != hipSuccess || hipEventCreate(&r.end) != hipSuccess) return;
  hipEventRecord(r.beg, s);
  g_prof.push_back(r);
  g_prof_open.push_back(g_prof.size() - 1);
}

void prof_end(const char* name, hipStream_t s) {
  for (size_t i = g_prof_open.size(); i-- > 0;) {
    ProfRec& r = g_prof[g_prof_open[i]];
    if (r.name == name) {
      hipEventRecord(r.end, s);
      g_prof_open.erase(g_prof_open.begin() + i);
      return;
    }
  }
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int ctx_buf(psgc_ctx* ctx, int slot, size_t bytes, void** out) {
  if ((int)ctx->bufs.size() < S_NUM_SLOTS) ctx->bufs.resize(S_NUM_SLOTS);
  DevBuf& b = ctx->bufs[slot];
  if (bytes == 0) bytes = 1;
  if (b.ptr == nullptr || b.bytes < bytes) {
    if (b.ptr) {
      PSGC_HIP(hipStreamSynchronize(ctx->stream));
      PSGC_HIP(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    const size_t want = bytes + bytes / 8;  // headroom for slightly larger calls
    if (hipMalloc(&b.ptr, want) != hipSuccess) {
      b.ptr = nullptr;
      (void)hipGetLastError();
      PSGC_FAIL(PSGC_E_ALLOC, "hipMalloc of %zu bytes failed", want);
    }
    b.bytes = want;
  }
  *out = b.ptr;
  return PSGC_OK;
}

int ctx_release(psgc_ctx* ctx) {
  if (ctx->stream) PSGC_HIP(hipStreamSynchronize(ctx->stream));
  host_stage_release(ctx);
  for (auto& b : ctx->bufs)
    if (b.ptr) {
      PSGC_HIP(hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
  return PSGC_OK;
}

// getGaussianKernel(win, sigma) in f64 -> f32, outer product in f32.
void gauss_kernel_f32(int win, float sigma, float* k2) {
  std::vector<double> k(win);
  double sg = sigma;
  if (win == 3 && sg <= 0) {
    k = {0.25, 0.5, 0.25};
  } else {
    if (sg <= 0) sg = ((win - 1) * 0.5 - 1) * 0.3 + 0.8;
    const double scale2 = -0.5 / (sg * sg);
    double sum = 0;
    for (int i = 0; i < win; ++i) {
      const double x = i - (win - 1) * 0.5;
      k[i] = std::exp(scale2 * x * x);
      sum += k[i];
    }
    sum = 1.0 / sum;
    for (int i = 0; i < win; ++i) k[i] *= sum;
  }
  std::vector<float> kf(win);
  for (int i = 0; i < win; ++i) kf[i] = (float)k[i];
  for (int i = 0; i < win; ++i)
    for (int j = 0; j < win; ++j) k2[i * win + j] = kf[i] * kf[j];
}

// getGaussianKernel(3, sigma) in f64: (centre, side).
void blur3_weights(double sigma, double* c, double* s) {
  if (sigma <= 0) {
    *c = 0.5;
    *s = 0.25;
    return;
  }
  const double scale2 = -0.5 / (sigma * sigma);
  double k[3], sum = 0;
  for (int i = 0; i < 3; ++i) {
    const double x = i - (3 - 1) * 0.5;
    k[i] = std::exp(scale2 * x * x);
    sum += k[i];
  }
  sum = 1.0 / sum;
  for (int i = 0; i < 3; ++i) k[i] *= sum;
  *c = k[1];
  *s = k[0];
}

namespace {

int sample_bytes(int t) {
  return t == PSGC_SAMPLE_U8 ? 1 : (t == PSGC_SAMPLE_U16 ? 2 : 8);
}

struct DeviceGuard {
  int rc = PSGC_OK;
  explicit DeviceGuard(psgc_ctx* ctx) {
    if (hipSetDevice(ctx->device) != hipSuccess) rc = PSGC_E_HIP;
  }
};

// Device pointer for a caller buffer: in host mode, a staging slot (filled
// from the host when `upload`); in device mode, the pointer itself.
template <typename T>
int stage_in(psgc_ctx* ctx, int mem, int slot, const T* p, size_t n, const T** out) {
  if (mem == PSGC_MEM_DEVICE) {
    *out = p;
    return PSGC_OK;
  }
  void* d;
  PSGC_TRY(ctx_buf(ctx, slot, n * sizeof(T), &d));
  PSGC_TRY(host_to_device(ctx, d, p, n * sizeof(T)));
  *out = (const T*)d;
  return PSGC_OK;
}

template <typename T>
int stage_out(psgc_ctx* ctx, int mem, int slot, T* p, size_t n, T** out) {
  if (p == nullptr) {
    *out = nullptr;
    return PSGC_OK;
  }
  if (mem == PSGC_MEM_DEVICE) {
    *out = p;
    return PSGC_OK;
  }
  void* d;
  PSGC_TRY(ctx_buf(ctx, slot, n * sizeof(T), &d));
  *out = (T*)d;
  return PSGC_OK;
}

template <typename T>
int copy_back(psgc_ctx* ctx, int mem, T* host, const T* dev, size_t n) {
  if (mem == PSGC_MEM_HOST && host) PSGC_TRY(device_to_host(ctx, host, dev, n * sizeof(T)));
  return PSGC_OK;
}

int finish(psgc_ctx* ctx, int mem) {
  if (mem == PSGC_MEM_HOST) PSGC_HIP(hipStreamSynchronize(ctx->stream));
  return PSGC_OK;
}

int check_mem(int mem) {
  if (mem != PSGC_MEM_DEVICE && mem != PSGC_MEM_HOST)
    PSGC_FAIL(PSGC_E_INVALID, "mem must be PSGC_MEM_DEVICE or PSGC_MEM_HOST");
  return PSGC_OK;
}

int check_size(int rows, int cols) {
  if (rows < 1 || cols < 1 || (int64_t)rows * cols > (int64_t)1 << 31)
    PSGC_FAIL(PSGC_E_INVALID, "bad image size %d x %d", rows, cols);
  return PSGC_OK;
}

}  // namespace

// Decode configuration -> kernel arguments (frames filled by the caller).
int build_decode_args(const psgc_decode_config* cfg, int n_frames, int sample_type, int rows,
                     int cols, int64_t stride, DecodeArgs* a) {
  std::memset(a, 0, sizeof(*a));
  if (!cfg) PSGC_FAIL(PSGC_E_INVALID, "null decode config");
  if (sample_type < PSGC_SAMPLE_U8 || sample_type > PSGC_SAMPLE_F64)
    PSGC_FAIL(PSGC_E_INVALID, "bad sample type %d", sample_type);
  const bool gray = cfg->unwrap_mode == PSGC_UNWRAP_GRAYCODE;
  if (!gray && cfg->unwrap_mode != PSGC_UNWRAP_HETERODYNE3)
    PSGC_FAIL(PSGC_E_INVALID, "bad unwrap_mode %d", cfg->unwrap_mode);
  a->nlev = gray ? 1 : 3;
  int total = 0;
  for (int i = 0; i < a->nlev; ++i) {
    const int n = cfg->shift_num[i];
    if (n < 3 || n > 16) PSGC_FAIL(PSGC_E_INVALID, "shift_num[%d]=%d outside [3,16]", i, n);
    a->shift_num[i] = n;
    a->level_base[i] = total;
    for (int k = 0; k < n; ++k) {
      a->sin_tab[total + k] = std::sin(2.0 * PI * k / n);
      a->cos_tab[total + k] = std::cos(2.0 * PI * k / n);
    }
    total += n;
  }
  // JSLAB_PhaseShiftDecoder::create refuses a zero frequency
  // (JSLAB_PhaseShiftDecoder.cpp:11-28).
  if (!gray) {
    for (int i = 0; i < 3; ++i)
      if (cfg->freq[i] == 0) PSGC_FAIL(PSGC_E_INVALID, "freq[%d] == 0", i);
    a->f12 = std::abs(static_cast<int>(cfg->freq[0] - cfg->freq[1]));
    a->f23 = std::abs(static_cast<int>(cfg->freq[1] - cfg->freq[2]));
    if (a->f12 == 0 || a->f23 == 0)
      PSGC_FAIL(PSGC_E_INVALID, "beat frequencies f12=%d f23=%d must be non-zero", a->f12,
                a->f23);
  } else {
    if (cfg->gray_bits < 1 || cfg->gray_bits > 24)
      PSGC_FAIL(PSGC_E_INVALID, "gray_bits=%d outside [1,24]", cfg->gray_bits);
    a->gray_bits = cfg->gray_bits;
    a->gray_min_contrast = cfg->gray_min_contrast;
    total += 2 * cfg->gray_bits;
  }
  if (n_frames != total)
    PSGC_FAIL(PSGC_E_INVALID, "expected %d frames, got %d", total, n_frames);
  if (n_frames > 64) PSGC_FAIL(PSGC_E_INVALID, "at most 64 frames");
  for (int i = 0; i < 3; ++i) a->freq[i] = cfg->freq[i];
  a->n_frames = n_frames;
  a->sample_type = sample_type;
  a->rows = rows;
  a->cols = cols;
  a->stride = stride;
  a->mod_thresh = cfg->modulation_threshold;
  a->avg_three = cfg->avg_three_phase ? 1 : 0;
  blur3_weights(cfg->preblur_sigma, &a->blur_c, &a->blur_s);
  if (cfg->enable_hdr) {
    // JSLAB_HDRPhaseFit asserts 3 <= steps <= 12 (JSLAB_HDRPhaseFit.h:15);
    // saturation is classified on the raw samples, so there is no pre-blur
    // (JSLAB_PhaseShiftDecoder_V2.cpp:262-265).
    for (int i = 0; i < a->nlev; ++i)
      if (a->shift_num[i] > 12)
        PSGC_FAIL(PSGC_E_INVALID, "HDR fit needs shift_num[%d]=%d <= 12", i, a->shift_num[i]);
    a->hdr = 1;
    a->hdr_sat = cfg->hdr_saturation_gray;
    a->hdr_min_amp = cfg->hdr_min_amplitude;
    a->blur_c = 1.0;
    a->blur_s = 0.0;
  }
  return PSGC_OK;
}

// ---- HDR pseudo-inverse tables (JSLAB_HDRPhaseFit.h:13-40) ---------------
// For every subset (bit mask) of the N equally spaced steps with at least 3
// members, the design rows (1, cos d_k, -sin d_k) give the Gram matrix G; the
// least-squares weights of sample k are G^-1 row_k (the pseudo-inverse
// column that the reference takes from cv::invert(DECOMP_SVD)). G^-1 is the
// adjugate over the determinant. Subsets whose Gram determinant is not above
// 1e-20 trace^3 are rank deficient and stay marked unusable (the reference's
// singular-value test; with distinct angles every subset of 3 or more has
// full rank). Layout per level: weights (2^N * N * 3 doubles), then flags
// (2^N bytes), each block 256-byte aligned.

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

static void hdr_level_sizes(int n, size_t* wbytes, size_t* okbytes) {
  *wbytes = align256(((size_t)1 << n) * n * 3 * sizeof(double));
  *okbytes = align256((size_t)1 << n);
}

size_t hdr_table_bytes(const DecodeArgs& a) {
  size_t b = 0;
  for (int i = 0; i < a.nlev; ++i) {
    size_t wb, ob;
    hdr_level_sizes(a.shift_num[i], &wb, &ob);
    b += wb + ob;
  }
  return b;
}

void hdr_table_fill(const DecodeArgs& a, char* host) {
  for (int i = 0; i < a.nlev; ++i) {
    const int n = a.shift_num[i];
    size_t wb, ob;
    hdr_level_sizes(n, &wb, &ob);
    double* W = (double*)host;
    uint8_t* ok = (uint8_t*)(host + wb);
    std::memset(host, 0, wb + ob);
    double cs[16], sn[16];
    for (int k = 0; k < n; ++k) {
      const double d = 2.0 * PI * k / n;
      cs[k] = std::cos(d);
      sn[k] = -std::sin(d);
    }
    for (unsigned mask = 0; mask < (1u << n); ++mask) {
      if (__builtin_popcount(mask) < 3) continue;
      double G[3][3] = {};
      for (int k = 0; k < n; ++k) {
        if (!(mask >> k & 1u)) continue;
        const double row[3] = {1.0, cs[k], sn[k]};
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) G[r][c] += row[r] * row[c];
      }
      const double det = G[0][0] * (G[1][1] * G[2][2] - G[1][2] * G[2][1]) -
                         G[0][1] * (G[1][0] * G[2][2] - G[1][2] * G[2][0]) +
                         G[0][2] * (G[1][0] * G[2][1] - G[1][1] * G[2][0]);
      const double tr = G[0][0] + G[1][1] + G[2][2];
      if (!(det > 1e-20 * tr * tr * tr)) continue;
      const double inv[3][3] = {
          {(G[1][1] * G[2][2] - G[1][2] * G[2][1]) / det,
           (G[0][2] * G[2][1] - G[0][1] * G[2][2]) / det,
           (G[0][1] * G[1][2] - G[0][2] * G[1][1]) / det},
          {(G[1][2] * G[2][0] - G[1][0] * G[2][2]) / det,
           (G[0][0] * G[2][2] - G[0][2] * G[2][0]) / det,
           (G[0][2] * G[1][0] - G[0][0] * G[1][2]) / det},
          {(G[1][0] * G[2][1] - G[1][1] * G[2][0]) / det,
           (G[0][1] * G[2][0] - G[0][0] * G[2][1]) / det,
           (G[0][0] * G[1][1] - G[0][1] * G[1][0]) / det}};
      for (int k = 0; k < n; ++k) {
        if (!(mask >> k & 1u)) continue;
        const double row[3] = {1.0, cs[k], sn[k]};
        for (int r = 0; r < 3; ++r) {
          double v = 0.0;
          for (int c = 0; c < 3; ++c) v += inv[r][c] * row[c];
          W[((size_t)mask * n + k) * 3 + r] = v;
        }
      }
      ok[mask] = 1;
    }
    host += wb + ob;
  }
}

void hdr_table_bind(DecodeArgs* a, const char* dev) {
  for (int i = 0; i < a->nlev; ++i) {
    size_t wb, ob;
    hdr_level_sizes(a->shift_num[i], &wb, &ob);
    a->hdr_w[i] = (const double*)dev;
    a->hdr_ok[i] = (const uint8_t*)(dev + wb);
    dev += wb + ob;
  }
}

int check_mls_weights(float sigma, float max_radius) {
  if (!(sigma > 0.0f) || !std::isfinite(sigma))
    PSGC_FAIL(PSGC_E_INVALID, "MLS spatial sigma must be finite and > 0");
  if (!std::isfinite(1.0f / (2.0f * sigma * sigma)))
    PSGC_FAIL(PSGC_E_INVALID, "MLS spatial sigma %g too small: 1/(2 sigma^2) overflows", sigma);
  const float r = max_radius > 0.0f ? max_radius : 3.0f * sigma;
  if (!std::isfinite(r * r))
    PSGC_FAIL(PSGC_E_INVALID, "MLS radius %g: its square overflows", max_radius);
  return PSGC_OK;
}

}  // namespace psgc

using namespace psgc;

extern "C" {

const char* psgc_last_error(void) { return g_err; }
int psgc_abi_version(void) { return PSGC_ABI_VERSION; }

int psgc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

void psgc_decode_config_default(psgc_decode_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->freq[0] = 121.0;  // driver values, main.cpp:422-423
  c->freq[1] = 110.0;
  c->freq[2] = 120.0;
  c->shift_num[0] = c->shift_num[1] = c->shift_num[2] = 8;
  c->modulation_threshold = 1.0;
  c->preblur_sigma = 3.0;
  c->avg_three_phase = 1;
  c->post_smooth = 1;
  c->enable_hdr = 0;
  c->hdr_saturation_gray = 255.0;
  c->hdr_min_amplitude = 2.0;
  c->unwrap_mode = PSGC_UNWRAP_HETERODYNE3;
}

void psgc_recon_config_default(psgc_recon_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->min_depth = 0;
  c->max_depth = 1000;
  c->min_phase_diff = 0.03;
  c->max_phase_diff = 0.4;
  c->grad_threshold = 1;
  c->use_rectified_coords = 1;
  c->use_disp_median_filter = 1;
  c->use_full_left_rectification = 1;
  c->radius_search = 0.7f;
  c->radius_min_neighbors = 10;
  c->radius_window_size = 10;
  c->use_mls = 0;
  c->mls_window_size = 30;
  c->mls_spatial_sigma = 0.6f;
  c->use_gpu_mls = 1;
  c->mls_polynomial_order = 1;
  c->mls_max_radius = 0.0f;
  c->mls_iterations = 3;
  c->semantics = PSGC_SEM_CPU;
  c->gauss_window = 3;
  c->gauss_sigma = 0.7f;
  c->outlier_mode = PSGC_OUTLIER_ROR;
  c->sor_k = 30;
  c->sor_std_mul = 1.0f;
  c->sor_window = 4;
}

void psgc_recon_config_default_gpu(psgc_recon_config* c) {
  if (!c) return;
  psgc_recon_config_default(c);
  c->radius_search = 0.05f;
  c->radius_min_neighbors = 5;
  c->radius_window_size = 5;
  c->use_mls = 1;
  c->mls_window_size = 25;
  c->mls_spatial_sigma = 0.5f;
  c->mls_iterations = 2;
  c->semantics = PSGC_SEM_GPU;
}

int psgc_warmup(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    PSGC_FAIL(PSGC_E_NODEVICE, "no HIP device");
  }
  PSGC_HIP(hipFree(nullptr));
  return PSGC_OK;
}

int psgc_create(psgc_ctx** out, int device) {
  if (!out) PSGC_FAIL(PSGC_E_INVALID, "null output pointer");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    PSGC_FAIL(PSGC_E_NODEVICE, "no HIP device");
  }
  if (device < 0 || device >= n) PSGC_FAIL(PSGC_E_NODEVICE, "device %d of %d", device, n);
  PSGC_HIP(hipSetDevice(device));
  psgc_ctx* ctx = new psgc_ctx();
  ctx->device = device;
  if (hipStreamCreate(&ctx->own_stream) != hipSuccess) {
    delete ctx;
    PSGC_FAIL(PSGC_E_HIP, "hipStreamCreate failed");
  }
  ctx->stream = ctx->own_stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->cu_count = prop.multiProcessorCount;
  ctx->bufs.resize(S_NUM_SLOTS);
  // set-up of the allocation-free compaction on this device (a failure here
  // is reported by psgc_compact_noalloc, which repeats the lookup)
  int64_t* scratch_dev;
  if (compact_noalloc_scratch(device, &scratch_dev) != PSGC_OK) set_error("%s", "");
  *out = ctx;
  return PSGC_OK;
}

int psgc_destroy(psgc_ctx* ctx) {
  if (!ctx) return PSGC_OK;
  hipSetDevice(ctx->device);
  int rc = ctx_release(ctx);
  if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return rc;
}

int psgc_release(psgc_ctx* ctx) {
  PSGC_CTX_ENTER(ctx);
  return ctx_release(ctx);
}

int psgc_set_kernels(psgc_ctx* ctx, int which) {
  PSGC_CTX_ENTER(ctx);
  if (which != PSGC_KERNELS_TILED && which != PSGC_KERNELS_PLAIN)
    PSGC_FAIL(PSGC_E_INVALID, "kernels=%d", which);
  ctx->kernels = which;
  return PSGC_OK;
}

int psgc_set_stream(psgc_ctx* ctx, void* s) {
  PSGC_CTX_ENTER(ctx);
  PSGC_HIP(hipStreamSynchronize(ctx->stream));
  ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
  return PSGC_OK;
}

int psgc_synchronize(psgc_ctx* ctx) {
  PSGC_CTX_ENTER(ctx);
  PSGC_HIP(hipStreamSynchronize(ctx->stream));
  return PSGC_OK;
}

int psgc_decode(psgc_ctx* ctx, const psgc_decode_config* cfg, const void* const* frames,
                int n_frames, int sample_type, int rows, int cols, int64_t row_stride_bytes,
                double* phase, int32_t* k, double* mod, uint8_t* valid, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!frames || !phase || !k) PSGC_FAIL(PSGC_E_INVALID, "null frames/phase/k");
  DecodeArgs a;
  const int sb = sample_bytes(sample_type);
  if (row_stride_bytes == 0) row_stride_bytes = (int64_t)cols * sb;
  PSGC_TRY(build_decode_args(cfg, n_frames, sample_type, rows, cols, row_stride_bytes, &a));
  if (row_stride_bytes < (int64_t)cols * sb)
    PSGC_FAIL(PSGC_E_INVALID, "row stride %lld below %d", (long long)row_stride_bytes, cols * sb);
  for (int f = 0; f < n_frames; ++f)
    if (!frames[f]) PSGC_FAIL(PSGC_E_INVALID, "frame %d is null", f);
  const size_t npix = (size_t)rows * cols;
  if (mem == PSGC_MEM_HOST) {
    const size_t plane = (size_t)cols * sb * rows;
    void* d;
    PSGC_TRY(ctx_buf(ctx, S_FRAMES, plane * n_frames, &d));
    for (int f = 0; f < n_frames; ++f) {
      // through the pinned chunks; pitched frames are packed row by row on
      // the host (a pitched DMA from pageable memory measured about 50x
      // slower than a flat one)
      PSGC_TRY(host_to_device_2d(ctx, (char*)d + plane * f, frames[f], (size_t)row_stride_bytes,
                                 (size_t)cols * sb, rows));
      a.frames[f] = (char*)d + plane * f;
    }
    a.stride = (int64_t)cols * sb;
  } else {
    for (int f = 0; f < n_frames; ++f) a.frames[f] = frames[f];
  }
  double *d_phase, *d_mod, *d_raw;
  int32_t* d_k;
  uint8_t* d_valid;
  PSGC_TRY(stage_out(ctx, mem, S_PHASE, phase, npix, &d_phase));
  PSGC_TRY(stage_out(ctx, mem, S_K, k, npix, &d_k));
  PSGC_TRY(stage_out(ctx, mem, S_MOD, mod, npix, &d_mod));
  const bool gray = a.nlev == 1;
  if (valid) {
    PSGC_TRY(stage_out(ctx, mem, S_VALID, valid, npix, &d_valid));
  } else if (gray) {
    void* v;
    PSGC_TRY(ctx_buf(ctx, S_VALID, npix, &v));
    d_valid = (uint8_t*)v;
  } else {
    d_valid = nullptr;
  }
  const bool smooth = cfg->post_smooth != 0;
  if (smooth || gray) {
    void* v;
    PSGC_TRY(ctx_buf(ctx, S_PHASE_RAW, npix * sizeof(double), &v));
    d_raw = (double*)v;
  } else {
    d_raw = d_phase;
  }
  if (a.hdr) {
    // the HDR-weighted post-smooth needs the validity map
    if (!d_valid) {
      void* v;
      PSGC_TRY(ctx_buf(ctx, S_VALID, npix, &v));
      d_valid = (uint8_t*)v;
    }
    // tables are rebuilt per call: wait for earlier work that may read them
    const size_t tb = hdr_table_bytes(a);
    void* dt;
    PSGC_TRY(ctx_buf(ctx, S_HDR, tb, &dt));
    std::vector<char> host(tb);
    hdr_table_fill(a, host.data());
    PSGC_HIP(hipStreamSynchronize(ctx->stream));
    PSGC_HIP(hipMemcpy(dt, host.data(), tb, hipMemcpyHostToDevice));
    hdr_table_bind(&a, (const char*)dt);
  }
  // the HDR fit needs a level's raw samples together: one thread per pixel
  if (ctx->kernels == PSGC_KERNELS_PLAIN || a.hdr)
    PSGC_TRY(launch_decode_plain(a, d_raw, d_k, d_mod, d_valid, ctx->stream));
  else
    PSGC_TRY(tiled_decode(ctx, a, d_raw, d_k, d_mod, d_valid, ctx->stream));
  auto post_smooth = [&](const double* in, double* out) -> int {
    return a.hdr ? launch_post_smooth_valid(in, d_valid, out, rows, cols, ctx->stream)
                 : launch_post_smooth(in, out, rows, cols, ctx->stream);
  };
  if (gray) {
    // jump correction reads (raw, k) and writes (phase, k)
    void* kt;
    PSGC_TRY(ctx_buf(ctx, S_TMP0, npix * sizeof(int32_t), &kt));
    PSGC_HIP(hipMemcpyAsync(kt, d_k, npix * sizeof(int32_t), hipMemcpyDeviceToDevice,
                            ctx->stream));
    PSGC_TRY(launch_gray_fix(a, d_raw, d_phase, (const int32_t*)kt, d_k, d_valid, ctx->stream));
    if (smooth) {
      PSGC_HIP(hipMemcpyAsync(d_raw, d_phase, npix * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
      PSGC_TRY(post_smooth(d_raw, d_phase));
    }
  } else if (smooth) {
    PSGC_TRY(post_smooth(d_raw, d_phase));
  }
  PSGC_TRY(copy_back(ctx, mem, phase, d_phase, npix));
  PSGC_TRY(copy_back(ctx, mem, k, d_k, npix));
  PSGC_TRY(copy_back(ctx, mem, mod, d_mod, npix));
  PSGC_TRY(copy_back(ctx, mem, valid, d_valid, npix));
  return finish(ctx, mem);
}

int psgc_stereo_match(psgc_ctx* ctx, const double* l_phase, const double* r_phase,
                      const int32_t* l_k, const int32_t* r_k, int rows, int cols, int pha_f,
                      double grad_threshold, double* disp, int32_t* match_col, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!l_phase || !r_phase || !l_k || !r_k || !disp)
    PSGC_FAIL(PSGC_E_INVALID, "null input/output");
  if (pha_f < 0 || pha_f > 65535) PSGC_FAIL(PSGC_E_INVALID, "pha_f=%d", pha_f);
  const size_t npix = (size_t)rows * cols;
  const double *lp, *rp;
  const int32_t *lk, *rk;
  PSGC_TRY(stage_in(ctx, mem, S_LPH, l_phase, npix, &lp));
  PSGC_TRY(stage_in(ctx, mem, S_RPH, r_phase, npix, &rp));
  PSGC_TRY(stage_in(ctx, mem, S_LK, l_k, npix, &lk));
  PSGC_TRY(stage_in(ctx, mem, S_RK, r_k, npix, &rk));
  double* d_disp;
  int32_t* d_mc;
  PSGC_TRY(stage_out(ctx, mem, S_DISP64, disp, npix, &d_disp));
  PSGC_TRY(stage_out(ctx, mem, S_MATCH, match_col, npix, &d_mc));
  void* span;
  PSGC_TRY(ctx_buf(ctx, S_SPAN, (size_t)rows * (pha_f + 1) * 2 * sizeof(int32_t), &span));
  PSGC_TRY(launch_match_plain(lp, rp, lk, rk, rows, cols, pha_f, grad_threshold,
                              (int32_t*)span, d_disp, nullptr, d_mc, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, disp, d_disp, npix));
  PSGC_TRY(copy_back(ctx, mem, match_col, d_mc, npix));
  return finish(ctx, mem);
}

int psgc_median5(psgc_ctx* ctx, const float* src, float* dst, int rows, int cols, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!src || !dst) PSGC_FAIL(PSGC_E_INVALID, "null src/dst");
  const size_t npix = (size_t)rows * cols;
  const float* s;
  float* d;
  PSGC_TRY(stage_in(ctx, mem, S_DISP, src, npix, &s));
  PSGC_TRY(stage_out(ctx, mem, S_DISP2, dst, npix, &d));
  if (s == d) PSGC_FAIL(PSGC_E_INVALID, "median5 cannot run in place");
  if (ctx->kernels == PSGC_KERNELS_TILED)
    PSGC_TRY(tiled_median(ctx, s, d, rows, cols, ctx->stream));
  else
    PSGC_TRY(launch_median5(s, d, rows, cols, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, dst, d, npix));
  return finish(ctx, mem);
}

int psgc_reproject(psgc_ctx* ctx, const float* disp, int rows, int cols, const double Q[16],
                   const double R1[9], double min_depth, double max_depth,
                   int use_rectified_coords, float* xyz, uint8_t* P, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!disp || !Q || !R1 || !xyz || !P) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  const size_t npix = (size_t)rows * cols;
  double qr[25];
  std::memcpy(qr, Q, 16 * sizeof(double));
  std::memcpy(qr + 16, R1, 9 * sizeof(double));
  const float* dd;
  float* dx;
  uint8_t* dp;
  PSGC_TRY(stage_in(ctx, mem, S_DISP, disp, npix, &dd));
  PSGC_TRY(stage_out(ctx, mem, S_XYZ0, xyz, npix * 3, &dx));
  PSGC_TRY(stage_out(ctx, mem, S_P0, P, npix, &dp));
  PSGC_TRY(launch_reproject(dd, nullptr, rows, cols, qr, min_depth, max_depth,
                            use_rectified_coords, dx, dp, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, xyz, dx, npix * 3));
  PSGC_TRY(copy_back(ctx, mem, P, dp, npix));
  return finish(ctx, mem);
}

int psgc_filter_gaussian(psgc_ctx* ctx, const float* xyz, const uint8_t* P, float* dst_xyz,
                         uint8_t* dst_P, int rows, int cols, int win, float sigma, int semantics,
                         int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!xyz || !P || !dst_xyz || !dst_P) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (win < 1 || win > 7 || (win & 1) == 0) PSGC_FAIL(PSGC_E_INVALID, "win=%d", win);
  const size_t npix = (size_t)rows * cols;
  float k2[49];
  gauss_kernel_f32(win, sigma, k2);
  const float* sx;
  const uint8_t* sp;
  float* dx;
  uint8_t* dp;
  PSGC_TRY(stage_in(ctx, mem, S_XYZ0, xyz, npix * 3, &sx));
  PSGC_TRY(stage_in(ctx, mem, S_P0, P, npix, &sp));
  PSGC_TRY(stage_out(ctx, mem, S_XYZ1, dst_xyz, npix * 3, &dx));
  PSGC_TRY(stage_out(ctx, mem, S_P1, dst_P, npix, &dp));
  PSGC_TRY(launch_gaussian(sx, sp, dx, dp, rows, cols, win, k2, semantics, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, dst_xyz, dx, npix * 3));
  PSGC_TRY(copy_back(ctx, mem, dst_P, dp, npix));
  return finish(ctx, mem);
}

int psgc_cloud_sor(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n, int k,
                   float std_mul, float voxel, int rmax, uint8_t* keep, double* mean_dist,
                   double* stats, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  if (!xyz || !mean_dist || !stats || n < 0) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  const float* sx;
  const uint8_t* sv = nullptr;
  uint8_t* dk;
  double *dmd, *dst;
  PSGC_TRY(stage_in(ctx, mem, S_XYZ0, xyz, (size_t)n * 3, &sx));
  if (valid) PSGC_TRY(stage_in(ctx, mem, S_P0, valid, (size_t)n, &sv));
  PSGC_TRY(stage_out(ctx, mem, S_P1, keep, (size_t)n, &dk));
  PSGC_TRY(stage_out(ctx, mem, S_MEANDIST, mean_dist, (size_t)n, &dmd));
  PSGC_TRY(stage_out(ctx, mem, S_STATS, stats, 4, &dst));
  PSGC_TRY(cloud_sor_run(ctx, sx, sv, n, k, std_mul, voxel, rmax, dk, dmd, dst));
  PSGC_TRY(copy_back(ctx, mem, keep, dk, (size_t)n));
  PSGC_TRY(copy_back(ctx, mem, mean_dist, dmd, (size_t)n));
  PSGC_TRY(copy_back(ctx, mem, stats, dst, 4));
  return finish(ctx, mem);
}

int psgc_cloud_tune_voxel(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                          int k, int rmax, const float* candidates, int n_candidates,
                          float* best_voxel, float* times_ms) {
  PSGC_CTX_ENTER(ctx);
  if (!xyz || n < 0) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  return cloud_tune_voxel(ctx, xyz, valid, n, k, rmax, candidates, n_candidates, best_voxel,
                          times_ms);
}

int psgc_cloud_mls(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                   float spatial_sigma, int polynomial_order, float max_radius, int iterations,
                   int semantics, float voxel, float* out, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  if (!xyz || !out || n < 0) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (iterations < 1 || iterations > 64) PSGC_FAIL(PSGC_E_INVALID, "iterations=%d", iterations);
  const float* sx;
  const uint8_t* sv = nullptr;
  float* dout;
  PSGC_TRY(stage_in(ctx, mem, S_XYZ0, xyz, (size_t)n * 3, &sx));
  if (valid) PSGC_TRY(stage_in(ctx, mem, S_P0, valid, (size_t)n, &sv));
  PSGC_TRY(stage_out(ctx, mem, S_XYZ1, out, (size_t)n * 3, &dout));
  void* tmp = nullptr;
  if (iterations > 1) PSGC_TRY(ctx_buf(ctx, S_XYZ2, (size_t)n * 12 + 16, &tmp));
  PSGC_TRY(cloud_mls_run(ctx, sx, sv, n, spatial_sigma, polynomial_order, max_radius, iterations,
                         semantics, voxel, dout, (float*)tmp));
  PSGC_TRY(copy_back(ctx, mem, out, dout, (size_t)n * 3));
  return finish(ctx, mem);
}

int psgc_cloud_gaussian(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                        float sigma, float radius, float voxel, float* out, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  if (!xyz || !out || n < 0) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  const float* sx;
  const uint8_t* sv = nullptr;
  float* dout;
  PSGC_TRY(stage_in(ctx, mem, S_XYZ0, xyz, (size_t)n * 3, &sx));
  if (valid) PSGC_TRY(stage_in(ctx, mem, S_P0, valid, (size_t)n, &sv));
  PSGC_TRY(stage_out(ctx, mem, S_XYZ1, out, (size_t)n * 3, &dout));
  PSGC_TRY(cloud_gaussian_run(ctx, sx, sv, n, sigma, radius, voxel, dout));
  PSGC_TRY(copy_back(ctx, mem, out, dout, (size_t)n * 3));
  return finish(ctx, mem);
}

int psgc_filter_ror(psgc_ctx* ctx, const float* xyz, const uint8_t* P, float* dst_xyz,
                    uint8_t* dst_P, int rows, int cols, float radius, int min_neighbors,
                    int window_size, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!xyz || !P || !dst_xyz || !dst_P) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (window_size < 0 || window_size > 512) PSGC_FAIL(PSGC_E_INVALID, "window=%d", window_size);
  const size_t npix = (size_t)rows * cols;
  const float* sx;
  const uint8_t* sp;
  float* dx;
  uint8_t* dp;
  PSGC_TRY(stage_in(ctx, mem, S_XYZ0, xyz, npix * 3, &sx));
  PSGC_TRY(stage_in(ctx, mem, S_P0, P, npix, &sp));
  PSGC_TRY(stage_out(ctx, mem, S_XYZ1, dst_xyz, npix * 3, &dx));
  PSGC_TRY(stage_out(ctx, mem, S_P1, dst_P, npix, &dp));
  PSGC_TRY(launch_ror(sx, sp, dx, dp, rows, cols, radius, min_neighbors, window_size,
                      ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, dst_xyz, dx, npix * 3));
  PSGC_TRY(copy_back(ctx, mem, dst_P, dp, npix));
  return finish(ctx, mem);
}

int psgc_filter_sor(psgc_ctx* ctx, const float* xyz, const uint8_t* P, float* dst_xyz,
                    uint8_t* dst_P, int rows, int cols, int k, float std_mul, int half_window,
                    double* mean_dist, double stats[4], int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!xyz || !P || !dst_xyz || !dst_P) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (half_window < 1 || half_window > 64) PSGC_FAIL(PSGC_E_INVALID, "half_window=%d", half_window);
  const size_t npix = (size_t)rows * cols;
  const float* sx;
  const uint8_t* sp;
  float* dx;
  uint8_t* dp;
  PSGC_TRY(stage_in(ctx, mem, S_XYZ0, xyz, npix * 3, &sx));
  PSGC_TRY(stage_in(ctx, mem, S_P0, P, npix, &sp));
  PSGC_TRY(stage_out(ctx, mem, S_XYZ1, dst_xyz, npix * 3, &dx));
  PSGC_TRY(stage_out(ctx, mem, S_P1, dst_P, npix, &dp));
  double* md;
  if (mean_dist && mem == PSGC_MEM_DEVICE) {
    md = mean_dist;
  } else {
    void* v;
    PSGC_TRY(ctx_buf(ctx, S_MEANDIST, npix * sizeof(double), &v));
    md = (double*)v;
  }
  void *part, *st;
  PSGC_TRY(ctx_buf(ctx, S_PARTIAL, sor_partial_doubles(npix) * sizeof(double), &part));
  PSGC_TRY(ctx_buf(ctx, S_STATS, 4 * sizeof(double), &st));
  if (ctx->kernels == PSGC_KERNELS_TILED)
    PSGC_TRY(tiled_sor(ctx, sx, sp, dx, dp, rows, cols, k, std_mul, half_window, md, (double*)part,
                       (double*)st, ctx->stream));
  else
    PSGC_TRY(launch_sor(sx, sp, dx, dp, rows, cols, k, std_mul, half_window, md, (double*)part,
                        (double*)st, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, dst_xyz, dx, npix * 3));
  PSGC_TRY(copy_back(ctx, mem, dst_P, dp, npix));
  if (mean_dist && mem == PSGC_MEM_HOST) PSGC_TRY(copy_back(ctx, mem, mean_dist, md, npix));
  if (stats) {
    // statistics are always returned to host memory
    PSGC_HIP(hipMemcpyAsync(stats, st, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PSGC_HIP(hipStreamSynchronize(ctx->stream));
  }
  return finish(ctx, mem);
}

int psgc_filter_mls(psgc_ctx* ctx, const float* src_xyz, const uint8_t* P, float* dst_xyz,
                    int rows, int cols, int window_size, float spatial_sigma,
                    int polynomial_order, float max_radius, int iterations, int semantics,
                    int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!src_xyz || !P || !dst_xyz) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (window_size < 0 || window_size > 256) PSGC_FAIL(PSGC_E_INVALID, "window=%d", window_size);
  PSGC_TRY(check_mls_weights(spatial_sigma, max_radius));
  const size_t npix = (size_t)rows * cols;
  const float r2 = (max_radius > 0.0f) ? max_radius * max_radius
                                       : 9.0f * spatial_sigma * spatial_sigma;
  const int order = polynomial_order < 0 ? 0 : (polynomial_order > 2 ? 2 : polynomial_order);
  const int n_iter = iterations > 1 ? iterations : 1;
  const float* sx;
  const uint8_t* sp;
  float* dx;
  PSGC_TRY(stage_in(ctx, mem, S_XYZ0, src_xyz, npix * 3, &sx));
  PSGC_TRY(stage_in(ctx, mem, S_P0, P, npix, &sp));
  PSGC_TRY(stage_out(ctx, mem, S_XYZ1, dst_xyz, npix * 3, &dx));
  if (sx == dx) PSGC_FAIL(PSGC_E_INVALID, "MLS cannot run in place");
  // ping-pong: the last pass must land in dx
  void* tmp = nullptr;
  if (n_iter > 1) PSGC_TRY(ctx_buf(ctx, S_XYZ2, npix * 3 * sizeof(float), &tmp));
  if (ctx->kernels == PSGC_KERNELS_TILED) {
    PSGC_TRY(tiled_mls(ctx, sx, sp, dx, (float*)tmp, rows, cols, window_size, spatial_sigma, order,
                       r2, n_iter, semantics, ctx->stream));
    PSGC_TRY(copy_back(ctx, mem, dst_xyz, dx, npix * 3));
    return finish(ctx, mem);
  }
  const float* in = sx;
  for (int it = 0; it < n_iter; ++it) {
    const int remaining = n_iter - 1 - it;
    float* out = (remaining % 2 == 0) ? dx : (float*)tmp;
    PSGC_TRY(launch_mls(in, sp, out, rows, cols, window_size, spatial_sigma, order, r2,
                        semantics, ctx->stream));
    in = out;
  }
  PSGC_TRY(copy_back(ctx, mem, dst_xyz, dx, npix * 3));
  return finish(ctx, mem);
}

}  // extern "C"

namespace psgc {
namespace {

// psgc_reconstruct / psgc_reconstruct_poly: Q reprojection when coef_h is
// NULL, else the polynomial triangulation with the three maps coef_h[0..2].
int reconstruct_impl(psgc_ctx* ctx, const psgc_recon_config* cfg, const double* l_phase,
                     const double* r_phase, const int32_t* l_k, const int32_t* r_k, int rows,
                     int cols, int pha_f, const double* Q, const float* const* coef_h,
                     const double* R1, float* xyz, uint8_t* P, float* disp, int32_t* match_col,
                     int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!cfg || !l_phase || !r_phase || !l_k || !r_k || (!Q && !coef_h) || !R1 || !xyz || !P)
    PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (coef_h && (!coef_h[0] || !coef_h[1] || !coef_h[2]))
    PSGC_FAIL(PSGC_E_INVALID, "null coefficient map");
  if (pha_f < 0 || pha_f > 65535) PSGC_FAIL(PSGC_E_INVALID, "pha_f=%d", pha_f);
  const int gw = cfg->gauss_window;
  if (gw < 1 || gw > 7 || (gw & 1) == 0) PSGC_FAIL(PSGC_E_INVALID, "gauss_window=%d", gw);
  if (cfg->outlier_mode < PSGC_OUTLIER_NONE || cfg->outlier_mode > PSGC_OUTLIER_SOR)
    PSGC_FAIL(PSGC_E_INVALID, "outlier_mode=%d", cfg->outlier_mode);
  if (cfg->use_mls) PSGC_TRY(check_mls_weights(cfg->mls_spatial_sigma, cfg->mls_max_radius));
  const size_t npix = (size_t)rows * cols;
  const double *lp, *rp;
  const int32_t *lk, *rk;
  PSGC_TRY(stage_in(ctx, mem, S_LPH, l_phase, npix, &lp));
  PSGC_TRY(stage_in(ctx, mem, S_RPH, r_phase, npix, &rp));
  PSGC_TRY(stage_in(ctx, mem, S_LK, l_k, npix, &lk));
  PSGC_TRY(stage_in(ctx, mem, S_RK, r_k, npix, &rk));
  const float* coef[3] = {nullptr, nullptr, nullptr};
  if (coef_h)
    for (int a = 0; a < 3; ++a)
      PSGC_TRY(stage_in(ctx, mem, S_COEF0 + a, coef_h[a], npix * 4, &coef[a]));
  const float* const* coef_d = coef_h ? coef : nullptr;
  if (ctx->kernels == PSGC_KERNELS_TILED) {
    float *d_xyz, *d_disp;
    uint8_t* d_P;
    int32_t* d_mc;
    PSGC_TRY(stage_out(ctx, mem, S_TMP0, xyz, npix * 3, &d_xyz));
    PSGC_TRY(stage_out(ctx, mem, S_TMP1, P, npix, &d_P));
    PSGC_TRY(stage_out(ctx, mem, S_DISP2, disp, npix, &d_disp));
    PSGC_TRY(stage_out(ctx, mem, S_MATCH, match_col, npix, &d_mc));
    PSGC_TRY(tiled_reconstruct(ctx, *cfg, lp, rp, lk, rk, rows, cols, pha_f, Q, coef_d, R1,
                               d_xyz, d_P, d_disp, d_mc, ctx->stream));
    PSGC_TRY(copy_back(ctx, mem, xyz, d_xyz, npix * 3));
    PSGC_TRY(copy_back(ctx, mem, P, d_P, npix));
    PSGC_TRY(copy_back(ctx, mem, disp, d_disp, npix));
    PSGC_TRY(copy_back(ctx, mem, match_col, d_mc, npix));
    return finish(ctx, mem);
  }
  void *span, *d64, *d32, *d32m, *x0, *x1, *x2, *p0, *p1, *p2;
  PSGC_TRY(ctx_buf(ctx, S_SPAN, (size_t)rows * (pha_f + 1) * 2 * sizeof(int32_t), &span));
  PSGC_TRY(ctx_buf(ctx, S_DISP64, npix * sizeof(double), &d64));
  PSGC_TRY(ctx_buf(ctx, S_DISP, npix * sizeof(float), &d32));
  PSGC_TRY(ctx_buf(ctx, S_DISP2, npix * sizeof(float), &d32m));
  PSGC_TRY(ctx_buf(ctx, S_XYZ0, npix * 3 * sizeof(float), &x0));
  PSGC_TRY(ctx_buf(ctx, S_XYZ1, npix * 3 * sizeof(float), &x1));
  PSGC_TRY(ctx_buf(ctx, S_XYZ2, npix * 3 * sizeof(float), &x2));
  PSGC_TRY(ctx_buf(ctx, S_P0, npix, &p0));
  PSGC_TRY(ctx_buf(ctx, S_P1, npix, &p1));
  PSGC_TRY(ctx_buf(ctx, S_P2, npix, &p2));
  int32_t* d_mc;
  PSGC_TRY(stage_out(ctx, mem, S_MATCH, match_col, npix, &d_mc));
  hipStream_t s = ctx->stream;
  PSGC_TRY(launch_match_plain(lp, rp, lk, rk, rows, cols, pha_f, cfg->grad_threshold,
                              (int32_t*)span, (double*)d64, (float*)d32, d_mc, s));
  const float* dfinal = (const float*)d32;
  if (cfg->use_disp_median_filter) {
    PSGC_TRY(launch_median5((const float*)d32, (float*)d32m, rows, cols, s));
    dfinal = (const float*)d32m;
  }
  // CPU semantics without the median reprojects the f64 disparity
  // (JSLAB_PhaseEpipolarReconstructor_V2.cpp:401-430).
  const bool use64 = !cfg->use_disp_median_filter && cfg->semantics == PSGC_SEM_CPU;
  if (coef_d) {
    PSGC_TRY(launch_reproject_poly(dfinal, use64 ? (const double*)d64 : nullptr, rows, cols,
                                   coef_d, R1, cfg->min_depth, cfg->max_depth,
                                   cfg->use_rectified_coords, (float*)x0, (uint8_t*)p0, s));
  } else {
    double qr[25];
    std::memcpy(qr, Q, 16 * sizeof(double));
    std::memcpy(qr + 16, R1, 9 * sizeof(double));
    PSGC_TRY(launch_reproject(dfinal, use64 ? (const double*)d64 : nullptr, rows, cols, qr,
                              cfg->min_depth, cfg->max_depth, cfg->use_rectified_coords,
                              (float*)x0, (uint8_t*)p0, s));
  }
  float k2[49];
  gauss_kernel_f32(gw, cfg->gauss_sigma, k2);
  PSGC_TRY(launch_gaussian((const float*)x0, (const uint8_t*)p0, (float*)x1, (uint8_t*)p1, rows,
                           cols, gw, k2, cfg->semantics, s));
  const uint8_t* pin = cfg->semantics == PSGC_SEM_CPU ? (const uint8_t*)p0 : (const uint8_t*)p1;
  float* xo = (float*)x2;
  uint8_t* po = (uint8_t*)p2;
  if (cfg->outlier_mode == PSGC_OUTLIER_ROR) {
    PSGC_TRY(launch_ror((const float*)x1, pin, xo, po, rows, cols, cfg->radius_search,
                        cfg->radius_min_neighbors, cfg->radius_window_size, s));
  } else if (cfg->outlier_mode == PSGC_OUTLIER_SOR) {
    void *md, *part, *st;
    PSGC_TRY(ctx_buf(ctx, S_MEANDIST, npix * sizeof(double), &md));
    PSGC_TRY(ctx_buf(ctx, S_PARTIAL, sor_partial_doubles(npix) * sizeof(double), &part));
    PSGC_TRY(ctx_buf(ctx, S_STATS, 4 * sizeof(double), &st));
    PSGC_TRY(launch_sor((const float*)x1, pin, xo, po, rows, cols, cfg->sor_k, cfg->sor_std_mul,
                        cfg->sor_window, (double*)md, (double*)part, (double*)st, s));
  } else {
    xo = (float*)x1;
    po = (uint8_t*)pin;
  }
  float* d_xyz;
  uint8_t* d_P;
  PSGC_TRY(stage_out(ctx, mem, S_TMP0, xyz, npix * 3, &d_xyz));
  PSGC_TRY(stage_out(ctx, mem, S_TMP1, P, npix, &d_P));
  if (cfg->use_mls) {
    const float r2 = (cfg->mls_max_radius > 0.0f)
                         ? cfg->mls_max_radius * cfg->mls_max_radius
                         : 9.0f * cfg->mls_spatial_sigma * cfg->mls_spatial_sigma;
    const int order = cfg->mls_polynomial_order < 0 ? 0
                      : (cfg->mls_polynomial_order > 2 ? 2 : cfg->mls_polynomial_order);
    const int n_iter = cfg->mls_iterations > 1 ? cfg->mls_iterations : 1;
    // ping-pong between x0 (free now) and the output
    const float* in = xo;
    for (int it = 0; it < n_iter; ++it) {
      const int remaining = n_iter - 1 - it;
      float* out = (remaining % 2 == 0) ? d_xyz : (float*)x0;
      if (out == in) out = (float*)x1;
      PSGC_TRY(launch_mls(in, po, out, rows, cols, cfg->mls_window_size, cfg->mls_spatial_sigma,
                          order, r2, cfg->semantics, s));
      in = out;
    }
    if (in != d_xyz)
      PSGC_HIP(hipMemcpyAsync(d_xyz, in, npix * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
  } else {
    PSGC_HIP(hipMemcpyAsync(d_xyz, xo, npix * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  PSGC_HIP(hipMemcpyAsync(d_P, po, npix, hipMemcpyDeviceToDevice, s));
  if (disp) {
    if (mem == PSGC_MEM_HOST)
      PSGC_HIP(hipMemcpyAsync(disp, dfinal, npix * sizeof(float), hipMemcpyDeviceToHost, s));
    else
      PSGC_HIP(hipMemcpyAsync(disp, dfinal, npix * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  PSGC_TRY(copy_back(ctx, mem, xyz, d_xyz, npix * 3));
  PSGC_TRY(copy_back(ctx, mem, P, d_P, npix));
  PSGC_TRY(copy_back(ctx, mem, match_col, d_mc, npix));
  return finish(ctx, mem);
}

}  // namespace
}  // namespace psgc

extern "C" {

int psgc_reconstruct(psgc_ctx* ctx, const psgc_recon_config* cfg, const double* l_phase,
                     const double* r_phase, const int32_t* l_k, const int32_t* r_k, int rows,
                     int cols, int pha_f, const double Q[16], const double R1[9], float* xyz,
                     uint8_t* P, float* disp, int32_t* match_col, int mem) {
  if (!Q) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  return reconstruct_impl(ctx, cfg, l_phase, r_phase, l_k, r_k, rows, cols, pha_f, Q, nullptr,
                          R1, xyz, P, disp, match_col, mem);
}

int psgc_reconstruct_poly(psgc_ctx* ctx, const psgc_recon_config* cfg, const double* l_phase,
                          const double* r_phase, const int32_t* l_k, const int32_t* r_k,
                          int rows, int cols, int pha_f, const float* coef_x,
                          const float* coef_y, const float* coef_z, const double R1[9],
                          float* xyz, uint8_t* P, float* disp, int32_t* match_col, int mem) {
  const float* coef[3] = {coef_x, coef_y, coef_z};
  return reconstruct_impl(ctx, cfg, l_phase, r_phase, l_k, r_k, rows, cols, pha_f, nullptr,
                          coef, R1, xyz, P, disp, match_col, mem);
}

int psgc_reproject_poly(psgc_ctx* ctx, const double* disp, int rows, int cols,
                        const float* coef_x, const float* coef_y, const float* coef_z,
                        const double R1[9], double min_depth, double max_depth,
                        int use_rectified_coords, float* xyz, uint8_t* P, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!disp || !coef_x || !coef_y || !coef_z || !R1 || !xyz || !P)
    PSGC_FAIL(PSGC_E_INVALID, "null argument");
  const size_t npix = (size_t)rows * cols;
  const double* dd;
  const float* c[3];
  float* dx;
  uint8_t* dp;
  PSGC_TRY(stage_in(ctx, mem, S_DISP64, disp, npix, &dd));
  PSGC_TRY(stage_in(ctx, mem, S_COEF0, coef_x, npix * 4, &c[0]));
  PSGC_TRY(stage_in(ctx, mem, S_COEF1, coef_y, npix * 4, &c[1]));
  PSGC_TRY(stage_in(ctx, mem, S_COEF2, coef_z, npix * 4, &c[2]));
  PSGC_TRY(stage_out(ctx, mem, S_XYZ0, xyz, npix * 3, &dx));
  PSGC_TRY(stage_out(ctx, mem, S_P0, P, npix, &dp));
  PSGC_TRY(launch_reproject_poly(nullptr, dd, rows, cols, c, R1, min_depth, max_depth,
                                 use_rectified_coords, dx, dp, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, xyz, dx, npix * 3));
  PSGC_TRY(copy_back(ctx, mem, P, dp, npix));
  return finish(ctx, mem);
}

}  // extern "C"

namespace psgc {
namespace {

int check_roi(int rows, int cols, const int32_t* lr, const int32_t* rr) {
  if (!lr) PSGC_FAIL(PSGC_E_INVALID, "null left_rect");
  if (lr[2] < 1 || lr[3] < 1 || lr[0] < 0 || lr[1] < 0 || lr[0] + lr[2] > cols ||
      lr[1] + lr[3] > rows)
    PSGC_FAIL(PSGC_E_INVALID, "left_rect {%d,%d,%d,%d} outside %dx%d", lr[0], lr[1], lr[2],
              lr[3], cols, rows);
  if (rr && (rr[1] != lr[1] || rr[3] != lr[3] || rr[2] < 1 || rr[0] < 0 ||
             rr[0] + rr[2] > cols))
    PSGC_FAIL(PSGC_E_INVALID, "right_rect {%d,%d,%d,%d} must share left_rect's rows", rr[0],
              rr[1], rr[2], rr[3]);
  return PSGC_OK;
}

}  // namespace
}  // namespace psgc

extern "C" {

int psgc_roi_right_window(psgc_ctx* ctx, const int32_t* l_k, const int32_t* r_k, int rows,
                          int cols, int pha_f, const int32_t left_rect[4], const uint8_t* mask,
                          int32_t right_rect[4], int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!l_k || !r_k || !mask || !right_rect) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (pha_f < 0 || pha_f > 65535) PSGC_FAIL(PSGC_E_INVALID, "pha_f=%d", pha_f);
  PSGC_TRY(check_roi(rows, cols, left_rect, nullptr));
  const size_t npix = (size_t)rows * cols;
  const size_t nroi = (size_t)left_rect[2] * left_rect[3];
  const int32_t *lk, *rk;
  const uint8_t* mk;
  PSGC_TRY(stage_in(ctx, mem, S_LK, l_k, npix, &lk));
  PSGC_TRY(stage_in(ctx, mem, S_RK, r_k, npix, &rk));
  PSGC_TRY(stage_in(ctx, mem, S_MASK, mask, nroi, &mk));
  void* buf;
  PSGC_TRY(ctx_buf(ctx, S_CNT, ((size_t)pha_f + 3) * sizeof(int32_t), &buf));
  int32_t* mnmx = (int32_t*)buf;
  int32_t* needed = mnmx + 2;
  const int32_t init[2] = {cols, -1};
  PSGC_HIP(hipMemsetAsync(needed, 0, ((size_t)pha_f + 1) * sizeof(int32_t), ctx->stream));
  PSGC_HIP(hipMemcpyAsync(mnmx, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
  PSGC_TRY(launch_roi_right_window(lk, rk, rows, cols, pha_f, left_rect, mk, needed, mnmx,
                                   ctx->stream));
  int32_t got[2];
  PSGC_HIP(hipMemcpyAsync(got, mnmx, sizeof(got), hipMemcpyDeviceToHost, ctx->stream));
  PSGC_HIP(hipStreamSynchronize(ctx->stream));
  const int pad = 32;
  if (got[1] >= got[0]) {
    const int x0 = got[0] - pad > 0 ? got[0] - pad : 0;
    const int x1 = got[1] + pad < cols - 1 ? got[1] + pad : cols - 1;
    right_rect[0] = x0;
    right_rect[2] = x1 - x0 + 1;
  } else {
    right_rect[0] = 0;
    right_rect[2] = cols;
  }
  right_rect[1] = left_rect[1];
  right_rect[3] = left_rect[3];
  return PSGC_OK;
}

int psgc_reconstruct_poly_roi(psgc_ctx* ctx, const psgc_recon_config* cfg, const double* l_phase,
                              const double* r_phase, const int32_t* l_k, const int32_t* r_k,
                              int rows, int cols, int pha_f, const int32_t left_rect[4],
                              const int32_t right_rect[4], const uint8_t* mask,
                              const float* coef_x, const float* coef_y, const float* coef_z,
                              const double R1[9], float* xyz, uint8_t* P, float* disp, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!cfg || !l_phase || !r_phase || !l_k || !r_k || !mask || !coef_x || !coef_y || !coef_z ||
      !R1 || !xyz || !P)
    PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (pha_f < 0 || pha_f > 65535) PSGC_FAIL(PSGC_E_INVALID, "pha_f=%d", pha_f);
  PSGC_TRY(check_roi(rows, cols, left_rect, right_rect));
  if (!right_rect) PSGC_FAIL(PSGC_E_INVALID, "null right_rect");
  const int gw = cfg->gauss_window;
  if (gw < 1 || gw > 7 || (gw & 1) == 0) PSGC_FAIL(PSGC_E_INVALID, "gauss_window=%d", gw);
  if (cfg->outlier_mode < PSGC_OUTLIER_NONE || cfg->outlier_mode > PSGC_OUTLIER_SOR)
    PSGC_FAIL(PSGC_E_INVALID, "outlier_mode=%d", cfg->outlier_mode);
  if (cfg->use_mls) PSGC_TRY(check_mls_weights(cfg->mls_spatial_sigma, cfg->mls_max_radius));
  const size_t npix = (size_t)rows * cols;
  const int lw = left_rect[2], h = left_rect[3];
  const size_t n = (size_t)lw * h;
  hipStream_t s = ctx->stream;
  const double *lp, *rp;
  const int32_t *lk, *rk;
  const uint8_t* mk;
  const float* coef[3];
  PSGC_TRY(stage_in(ctx, mem, S_LPH, l_phase, npix, &lp));
  PSGC_TRY(stage_in(ctx, mem, S_RPH, r_phase, npix, &rp));
  PSGC_TRY(stage_in(ctx, mem, S_LK, l_k, npix, &lk));
  PSGC_TRY(stage_in(ctx, mem, S_RK, r_k, npix, &rk));
  PSGC_TRY(stage_in(ctx, mem, S_MASK, mask, n, &mk));
  PSGC_TRY(stage_in(ctx, mem, S_COEF0, coef_x, npix * 4, &coef[0]));
  PSGC_TRY(stage_in(ctx, mem, S_COEF1, coef_y, npix * 4, &coef[1]));
  PSGC_TRY(stage_in(ctx, mem, S_COEF2, coef_z, npix * 4, &coef[2]));
  void *span, *d64, *d32, *d32m, *x0, *x1, *x2, *x3, *p0, *p1, *p2, *p3;
  PSGC_TRY(ctx_buf(ctx, S_SPAN, (size_t)h * (pha_f + 1) * 2 * sizeof(int32_t), &span));
  PSGC_TRY(ctx_buf(ctx, S_DISP64, n * sizeof(double), &d64));
  PSGC_TRY(ctx_buf(ctx, S_DISP, n * sizeof(float), &d32));
  PSGC_TRY(ctx_buf(ctx, S_DISP2, n * sizeof(float), &d32m));
  PSGC_TRY(ctx_buf(ctx, S_XYZ0, n * 12, &x0));
  PSGC_TRY(ctx_buf(ctx, S_XYZ1, n * 12, &x1));
  PSGC_TRY(ctx_buf(ctx, S_XYZ2, n * 12, &x2));
  PSGC_TRY(ctx_buf(ctx, S_XYZ3, n * 12, &x3));
  PSGC_TRY(ctx_buf(ctx, S_P0, n, &p0));
  PSGC_TRY(ctx_buf(ctx, S_P1, n, &p1));
  PSGC_TRY(ctx_buf(ctx, S_P2, n, &p2));
  PSGC_TRY(ctx_buf(ctx, S_P3, n, &p3));
  PSGC_TRY(launch_roi_match(lp, rp, lk, rk, cols, left_rect, right_rect, mk, pha_f,
                            cfg->grad_threshold, (int32_t*)span, (double*)d64, (float*)d32, s));
  const float* dfinal = (const float*)d32;
  if (cfg->use_disp_median_filter) {
    PSGC_TRY(launch_median5((const float*)d32, (float*)d32m, h, lw, s));
    dfinal = (const float*)d32m;
  }
  const bool use64 = !cfg->use_disp_median_filter && cfg->semantics == PSGC_SEM_CPU;
  PSGC_TRY(launch_roi_poly(use64 ? (const double*)d64 : nullptr, dfinal, mk, cols, left_rect,
                           coef, R1, cfg->min_depth, cfg->max_depth, cfg->use_rectified_coords,
                           (float*)x0, (uint8_t*)p0, s));
  float k2[49];
  gauss_kernel_f32(gw, cfg->gauss_sigma, k2);
  PSGC_TRY(launch_gaussian((const float*)x0, (const uint8_t*)p0, (float*)x1, (uint8_t*)p1, h, lw,
                           gw, k2, cfg->semantics, s));
  const uint8_t* pin = cfg->semantics == PSGC_SEM_CPU ? (const uint8_t*)p0 : (const uint8_t*)p1;
  const float* xo = (const float*)x1;
  const uint8_t* po = pin;
  if (cfg->outlier_mode == PSGC_OUTLIER_ROR) {
    PSGC_TRY(launch_ror((const float*)x1, pin, (float*)x2, (uint8_t*)p2, h, lw,
                        cfg->radius_search, cfg->radius_min_neighbors, cfg->radius_window_size,
                        s));
    void* cnt;
    PSGC_TRY(ctx_buf(ctx, S_CNT, 2 * sizeof(unsigned long long), &cnt));
    PSGC_HIP(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), s));
    PSGC_TRY(launch_roi_ror_select((const float*)x1, pin, (const float*)x2, (const uint8_t*)p2,
                                   n, (unsigned long long*)cnt, (float*)x3, (uint8_t*)p3, s));
    xo = (const float*)x3;
    po = (const uint8_t*)p3;
  } else if (cfg->outlier_mode == PSGC_OUTLIER_SOR) {
    void *md, *part, *st;
    PSGC_TRY(ctx_buf(ctx, S_MEANDIST, n * sizeof(double), &md));
    PSGC_TRY(ctx_buf(ctx, S_PARTIAL, sor_partial_doubles(n) * sizeof(double), &part));
    PSGC_TRY(ctx_buf(ctx, S_STATS, 4 * sizeof(double), &st));
    PSGC_TRY(launch_sor((const float*)x1, pin, (float*)x2, (uint8_t*)p2, h, lw, cfg->sor_k,
                        cfg->sor_std_mul, cfg->sor_window, (double*)md, (double*)part,
                        (double*)st, s));
    xo = (const float*)x2;
    po = (const uint8_t*)p2;
  }
  float* d_xyz;
  uint8_t* d_P;
  PSGC_TRY(stage_out(ctx, mem, S_TMP0, xyz, n * 3, &d_xyz));
  PSGC_TRY(stage_out(ctx, mem, S_TMP1, P, n, &d_P));
  if (cfg->use_mls) {
    const float r2 = (cfg->mls_max_radius > 0.0f)
                         ? cfg->mls_max_radius * cfg->mls_max_radius
                         : 9.0f * cfg->mls_spatial_sigma * cfg->mls_spatial_sigma;
    const int order = cfg->mls_polynomial_order < 0 ? 0
                      : (cfg->mls_polynomial_order > 2 ? 2 : cfg->mls_polynomial_order);
    const int n_iter = cfg->mls_iterations > 1 ? cfg->mls_iterations : 1;
    // ping-pong between x0 (free now) and the output
    const float* in = xo;
    for (int it = 0; it < n_iter; ++it) {
      const int remaining = n_iter - 1 - it;
      float* out = (remaining % 2 == 0) ? d_xyz : (float*)x0;
      PSGC_TRY(launch_mls(in, po, out, h, lw, cfg->mls_window_size, cfg->mls_spatial_sigma,
                          order, r2, cfg->semantics, s));
      in = out;
    }
  } else {
    PSGC_HIP(hipMemcpyAsync(d_xyz, xo, n * 12, hipMemcpyDeviceToDevice, s));
  }
  PSGC_HIP(hipMemcpyAsync(d_P, po, n, hipMemcpyDeviceToDevice, s));
  if (disp) {
    if (mem == PSGC_MEM_HOST) PSGC_TRY(device_to_host(ctx, disp, dfinal, n * sizeof(float)));
    else PSGC_HIP(hipMemcpyAsync(disp, dfinal, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  PSGC_TRY(copy_back(ctx, mem, xyz, d_xyz, n * 3));
  PSGC_TRY(copy_back(ctx, mem, P, d_P, n));
  return finish(ctx, mem);
}

int psgc_load_coef_bin(const char* path, int* rows, int* cols, float* out) {
  if (!path || !rows || !cols) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  FILE* fp = std::fopen(path, "rb");
  if (!fp) PSGC_FAIL(PSGC_E_INVALID, "cannot open %s", path);
  int32_t hdr[3];
  const bool ok_hdr = std::fread(hdr, sizeof(int32_t), 3, fp) == 3;
  if (!ok_hdr || hdr[1] < 1 || hdr[2] < 1 || (int64_t)hdr[1] * hdr[2] > ((int64_t)1 << 31)) {
    std::fclose(fp);
    PSGC_FAIL(PSGC_E_INVALID, "%s: bad header", path);
  }
  constexpr int kCV_32FC4 = 29;  // CV_MAKETYPE(CV_32F, 4)
  if (hdr[0] != kCV_32FC4) {
    std::fclose(fp);
    PSGC_FAIL(PSGC_E_INVALID, "%s: type %d is not CV_32FC4", path, hdr[0]);
  }
  *rows = hdr[1];
  *cols = hdr[2];
  if (out) {
    const size_t n = (size_t)hdr[1] * hdr[2] * 4;
    const bool ok = std::fread(out, sizeof(float), n, fp) == n;
    std::fclose(fp);
    if (!ok) PSGC_FAIL(PSGC_E_INVALID, "%s: truncated data", path);
  } else {
    std::fclose(fp);
  }
  return PSGC_OK;
}

int psgc_rectify_maps(psgc_ctx* ctx, const double K_raw[9], const double R[9],
                      const double K_rect[9], int rows, int cols, float* map_x, float* map_y,
                      int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!K_raw || !R || !K_rect || !map_x || !map_y) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  // K_rect^-1 (upper triangular 3x3 with K[8] = 1 assumed general)
  const double* K = K_rect;
  const double det = K[0] * (K[4] * K[8] - K[5] * K[7]) - K[1] * (K[3] * K[8] - K[5] * K[6]) +
                     K[2] * (K[3] * K[7] - K[4] * K[6]);
  if (det == 0) PSGC_FAIL(PSGC_E_INVALID, "singular K_rect");
  double Ki[9] = {(K[4] * K[8] - K[5] * K[7]) / det, (K[2] * K[7] - K[1] * K[8]) / det,
                  (K[1] * K[5] - K[2] * K[4]) / det, (K[5] * K[6] - K[3] * K[8]) / det,
                  (K[0] * K[8] - K[2] * K[6]) / det, (K[2] * K[3] - K[0] * K[5]) / det,
                  (K[3] * K[7] - K[4] * K[6]) / det, (K[1] * K[6] - K[0] * K[7]) / det,
                  (K[0] * K[4] - K[1] * K[3]) / det};
  double T[9], M[9];
  for (int a = 0; a < 3; ++a)  // T = R^T Ki
    for (int b = 0; b < 3; ++b) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += R[k * 3 + a] * Ki[k * 3 + b];
      T[a * 3 + b] = s;
    }
  for (int a = 0; a < 3; ++a)  // M = K_raw T
    for (int b = 0; b < 3; ++b) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += K_raw[a * 3 + k] * T[k * 3 + b];
      M[a * 3 + b] = s;
    }
  void* md;
  PSGC_TRY(ctx_buf(ctx, S_QR, sizeof(M), &md));
  PSGC_HIP(hipMemcpyAsync(md, M, sizeof(M), hipMemcpyHostToDevice, ctx->stream));
  const size_t np = (size_t)rows * cols;
  float *mx, *my;
  PSGC_TRY(stage_out(ctx, mem, S_MAPX, map_x, np, &mx));
  PSGC_TRY(stage_out(ctx, mem, S_MAPY, map_y, np, &my));
  PSGC_TRY(launch_rect_maps(rows, cols, (const double*)md, mx, my, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, map_x, mx, np));
  PSGC_TRY(copy_back(ctx, mem, map_y, my, np));
  // the small M buffer must stay valid until the kernel ran
  PSGC_HIP(hipStreamSynchronize(ctx->stream));
  return finish(ctx, mem);
}

int psgc_remap(psgc_ctx* ctx, int kind, const void* src, int src_rows, int src_cols,
               const float* map_x, const float* map_y, int rows, int cols, void* dst, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  PSGC_TRY(check_size(src_rows, src_cols));
  if (!src || !map_x || !map_y || !dst) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (kind != 0 && kind != 1) PSGC_FAIL(PSGC_E_INVALID, "kind=%d", kind);
  const size_t es = kind == 0 ? 8 : 4;
  const size_t np = (size_t)rows * cols, ns = (size_t)src_rows * src_cols;
  const char* s;
  const float *mx, *my;
  char* d;
  PSGC_TRY(stage_in(ctx, mem, S_SRC, (const char*)src, ns * es, &s));
  PSGC_TRY(stage_in(ctx, mem, S_MAPX, map_x, np, &mx));
  PSGC_TRY(stage_in(ctx, mem, S_MAPY, map_y, np, &my));
  PSGC_TRY(stage_out(ctx, mem, S_DST, (char*)dst, np * es, &d));
  PSGC_TRY(launch_remap(kind, s, d, mx, my, rows, cols, src_rows, src_cols, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, (char*)dst, d, np * es));
  return finish(ctx, mem);
}

int psgc_compact(psgc_ctx* ctx, const float* xyz, const uint8_t* P, const uint8_t* gray, int rows,
                 int cols, float* out_xyz, uint8_t* out_gray, int64_t* count, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!xyz || !P || !out_xyz || !count) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  const size_t np = (size_t)rows * cols;
  const float* x;
  const uint8_t *p, *g = nullptr;
  PSGC_TRY(stage_in(ctx, mem, S_XYZ0, xyz, np * 3, &x));
  PSGC_TRY(stage_in(ctx, mem, S_P0, P, np, &p));
  if (gray) PSGC_TRY(stage_in(ctx, mem, S_P1, gray, np, &g));
  float* ox;
  uint8_t* og;
  PSGC_TRY(stage_out(ctx, mem, S_XYZ1, out_xyz, np * 3, &ox));
  PSGC_TRY(stage_out(ctx, mem, S_P2, out_gray, np, &og));
  void* scratch;
  PSGC_TRY(ctx_buf(ctx, S_CNT, compact_scratch_int64(np) * sizeof(int64_t), &scratch));
  int64_t* total = (int64_t*)scratch + compact_scratch_int64(np) - 1;
  PSGC_TRY(launch_compact(x, p, g, np, (int64_t*)scratch, ox, og, total, ctx->stream));
  PSGC_HIP(hipMemcpyAsync(count, total, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  PSGC_HIP(hipStreamSynchronize(ctx->stream));
  const size_t n = (size_t)*count;
  PSGC_TRY(copy_back(ctx, mem, out_xyz, ox, n * 3));
  PSGC_TRY(copy_back(ctx, mem, out_gray, og, n));
  return finish(ctx, mem);
}

int psgc_compact_noalloc(psgc_ctx* ctx, const float* xyz, const uint8_t* P, const uint8_t* gray,
                         int rows, int cols, float* out_xyz, uint8_t* out_gray, int64_t* count) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_size(rows, cols));
  if (!xyz || !P || !out_xyz || !count) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  return launch_compact_noalloc(ctx->device, xyz, P, gray, (size_t)rows * cols, out_xyz, out_gray,
                                count, ctx->stream);
}

int psgc_write_ply(const char* path, const float* xyz, const uint8_t* gray, int64_t n) {
  if (!path || (n > 0 && !xyz) || n < 0) PSGC_FAIL(PSGC_E_INVALID, "bad arguments");
  return write_ply(path, xyz, gray, n);
}

int psgc_write_raw_f64(const char* raw_path, const char* info_path, const double* data, int rows,
                       int cols) {
  if (!raw_path || rows < 0 || cols < 0 || ((int64_t)rows * cols > 0 && !data))
    PSGC_FAIL(PSGC_E_INVALID, "bad arguments");
  return write_raw_f64(raw_path, info_path, data, rows, cols);
}

int psgc_adjoint_forward(psgc_ctx* ctx, const double theta[9], const double nominal[5],
                         const double* phL, const double* phR, const int32_t* match_col,
                         int rows, int cols, double* xyz, uint8_t* valid, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!theta || !nominal || !phL || !phR || !match_col || !xyz)
    PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (!(theta[0] > 0) || !(nominal[0] > 0)) PSGC_FAIL(PSGC_E_INVALID, "focal must be > 0");
  const size_t np = (size_t)rows * cols;
  const double *l, *r;
  const int32_t* m;
  PSGC_TRY(stage_in(ctx, mem, S_LPH, phL, np, &l));
  PSGC_TRY(stage_in(ctx, mem, S_RPH, phR, np, &r));
  PSGC_TRY(stage_in(ctx, mem, S_MATCH, match_col, np, &m));
  double* dx;
  uint8_t* dv;
  PSGC_TRY(stage_out(ctx, mem, S_ADJ0, xyz, np * 3, &dx));
  PSGC_TRY(stage_out(ctx, mem, S_P0, valid, np, &dv));
  PSGC_TRY(launch_adj_forward(theta, nominal, l, r, m, rows, cols, dx, dv, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, xyz, dx, np * 3));
  PSGC_TRY(copy_back(ctx, mem, valid, dv, np));
  return finish(ctx, mem);
}

int psgc_adjoint_forward_batch(psgc_ctx* ctx, const psgc_adjoint_view* views, int n) {
  PSGC_CTX_ENTER(ctx);
  if (!views) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  return launch_adj_batch(ctx, views, n, false);
}

int psgc_adjoint_backward_batch(psgc_ctx* ctx, const psgc_adjoint_view* views, int n) {
  PSGC_CTX_ENTER(ctx);
  if (!views) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  return launch_adj_batch(ctx, views, n, true);
}

int psgc_adjoint_backward(psgc_ctx* ctx, const double theta[9], const double nominal[5],
                          const double* phL, const double* phR, const int32_t* match_col,
                          int rows, int cols, const double* grad_xyz, double* grad_theta,
                          double* grad_phL, double* grad_phR, int mem) {
  PSGC_CTX_ENTER(ctx);
  PSGC_TRY(check_mem(mem));
  PSGC_TRY(check_size(rows, cols));
  if (!theta || !nominal || !phL || !phR || !match_col || !grad_xyz || !grad_theta)
    PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (!(theta[0] > 0) || !(nominal[0] > 0)) PSGC_FAIL(PSGC_E_INVALID, "focal must be > 0");
  const size_t np = (size_t)rows * cols;
  const double *l, *r, *g;
  const int32_t* m;
  PSGC_TRY(stage_in(ctx, mem, S_LPH, phL, np, &l));
  PSGC_TRY(stage_in(ctx, mem, S_RPH, phR, np, &r));
  PSGC_TRY(stage_in(ctx, mem, S_MATCH, match_col, np, &m));
  PSGC_TRY(stage_in(ctx, mem, S_ADJ0, grad_xyz, np * 3, &g));
  double *gt, *gl, *gr;
  PSGC_TRY(stage_out(ctx, mem, S_STATS, grad_theta, 9, &gt));
  PSGC_TRY(stage_out(ctx, mem, S_ADJ1, grad_phL, np, &gl));
  PSGC_TRY(stage_out(ctx, mem, S_ADJ2, grad_phR, np, &gr));
  PSGC_TRY(launch_adj_backward(theta, nominal, l, r, m, rows, cols, g, gt, gl, gr, ctx->stream));
  PSGC_TRY(copy_back(ctx, mem, grad_theta, gt, 9));
  PSGC_TRY(copy_back(ctx, mem, grad_phL, gl, np));
  PSGC_TRY(copy_back(ctx, mem, grad_phR, gr, np));
  return finish(ctx, mem);
}

int psgc_profile_enable(int on) {
  for (auto& r : g_prof) {
    hipEventDestroy(r.beg);
    hipEventDestroy(r.end);
  }
  g_prof.clear();
  g_prof_open.clear();
  g_prof_on = on != 0;
  return PSGC_OK;
}

int psgc_profile_read(const char* name, int64_t* launches, double* total_ms) {
  if (!name || !launches || !total_ms) PSGC_FAIL(PSGC_E_INVALID, "null argument");
  PSGC_HIP(hipDeviceSynchronize());
  int64_t n = 0;
  double ms = 0;
  for (auto& r : g_prof) {
    if (r.name != name) continue;
    float t = 0;
    if (hipEventElapsedTime(&t, r.beg, r.end) == hipSuccess) {
      ms += t;
      ++n;
    }
  }
  *launches = n;
  *total_ms = ms;
  return PSGC_OK;
}

int psgc_rectify_ideal(double fx, double fy, double cx1, double cy1, double cx2, double baseline,
                       double Q[16], double R1[9]) {
  if (!Q || !R1) PSGC_FAIL(PSGC_E_INVALID, "null Q/R1");
  if (!(fx > 0) || !(fy > 0) || baseline == 0) PSGC_FAIL(PSGC_E_INVALID, "bad intrinsics");
  (void)fy;
  // Q of cv::stereoRectify (flags 0) for R = I, T = (-B, 0, 0):
  // [1 0 0 -cx1; 0 1 0 -cy1; 0 0 0 f; 0 0 -1/Tx (cx1-cx2)/Tx], Tx = -B.
  const double Tx = -baseline;
  const double q[16] = {1, 0, 0, -cx1, 0, 1, 0, -cy1, 0, 0, 0, fx, 0, 0, -1.0 / Tx,
                        (cx1 - cx2) / Tx};
  std::memcpy(Q, q, sizeof(q));
  const double r[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  std::memcpy(R1, r, sizeof(r));
  return PSGC_OK;
}

}  // extern "C"

namespace {

// Rotation vector of a rotation matrix (cvRodrigues2, matrix to vector): the
// axis from the skew part, the angle from the trace; near 0 and pi the
// symmetric part gives the axis.
void rodrigues_vec(const double R[9], double r[3]) {
  const double rx = R[7] - R[5], ry = R[2] - R[6], rz = R[3] - R[1];
  const double s = std::sqrt((rx * rx + ry * ry + rz * rz) * 0.25);
  double c = (R[0] + R[4] + R[8] - 1) * 0.5;
  c = c > 1. ? 1. : c < -1. ? -1. : c;
  const double theta = std::acos(c);
  if (s < 1e-5) {
    if (c > 0) {
      r[0] = r[1] = r[2] = 0;
      return;
    }
    // theta ~ pi: R ~ 2 n n^T - I, so |n_i| from the diagonal and the signs
    // of n_y, n_z relative to n_x >= 0 from the first row
    const double t0 = std::sqrt(std::max((R[0] + 1) * 0.5, 0.));
    const double t1 = std::sqrt(std::max((R[4] + 1) * 0.5, 0.)) * (R[1] < 0 ? -1. : 1.);
    double t2 = std::sqrt(std::max((R[8] + 1) * 0.5, 0.)) * (R[2] < 0 ? -1. : 1.);
    // with n_x ~ 0 the first row carries no sign; the (1, 2) entry does
    if (std::fabs(t0) < std::fabs(t1) && std::fabs(t0) < std::fabs(t2) && (R[5] > 0) != (t1 * t2 > 0))
      t2 = -t2;
    const double n = std::sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    r[0] = t0 * theta / n;
    r[1] = t1 * theta / n;
    r[2] = t2 * theta / n;
    return;
  }
  const double v = 1. / (2 * s) * theta;
  r[0] = rx * v;
  r[1] = ry * v;
  r[2] = rz * v;
}

// Rotation matrix of a rotation vector (cvRodrigues2, vector to matrix).
void rodrigues_mat(const double r[3], double R[9]) {
  const double theta = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (theta < 2.220446049250313e-16) {
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1. : 0.;
    return;
  }
  const double c = std::cos(theta), s = std::sin(theta), c1 = 1. - c;
  const double x = r[0] / theta, y = r[1] / theta, z = r[2] / theta;
  const double rrt[9] = {x * x, x * y, x * z, x * y, y * y, y * z, x * z, y * z, z * z};
  const double rx[9] = {0, -z, y, z, 0, -x, -y, x, 0};
  for (int i = 0; i < 9; ++i) R[i] = c * ((i % 4 == 0) ? 1. : 0.) + c1 * rrt[i] + s * rx[i];
}

void matmul3(const double A[9], const double B[9], bool bt, double C[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A[i * 3 + k] * (bt ? B[j * 3 + k] : B[k * 3 + j]);
      C[i * 3 + j] = s;
    }
}

}  // namespace

extern "C" {

int psgc_stereo_rectify(const double K1[9], const double K2[9], const double R[9],
                        const double T[3], int width, int height, double R1[9], double R2[9],
                        double P1[12], double P2[12], double Q[16]) {
  if (!K1 || !K2 || !R || !T || !R1 || !R2 || !P1 || !P2 || !Q)
    PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (width < 1 || height < 1) PSGC_FAIL(PSGC_E_INVALID, "image size %dx%d", width, height);
  if (!(K1[0] != 0) || !(K1[4] != 0) || !(K2[0] != 0) || !(K2[4] != 0))
    PSGC_FAIL(PSGC_E_INVALID, "focal lengths");
  // half of the relative rotation for each camera
  double om[3], rr[9], t[3];
  rodrigues_vec(R, om);
  for (double& v : om) v *= -0.5;
  rodrigues_mat(om, rr);
  for (int i = 0; i < 3; ++i) t[i] = rr[i * 3] * T[0] + rr[i * 3 + 1] * T[1] + rr[i * 3 + 2] * T[2];
  const int idx = std::fabs(t[0]) > std::fabs(t[1]) ? 0 : 1;
  const double c = t[idx], nt = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  if (!(nt > 0)) PSGC_FAIL(PSGC_E_INVALID, "zero baseline");
  double uu[3] = {0, 0, 0};
  uu[idx] = c > 0 ? 1 : -1;
  // the rotation taking the baseline onto the axis
  double ww[3] = {t[1] * uu[2] - t[2] * uu[1], t[2] * uu[0] - t[0] * uu[2],
                  t[0] * uu[1] - t[1] * uu[0]};
  const double nw = std::sqrt(ww[0] * ww[0] + ww[1] * ww[1] + ww[2] * ww[2]);
  if (nw > 0)
    for (double& v : ww) v *= std::acos(std::fabs(c) / nt) / nw;
  double wR[9];
  rodrigues_mat(ww, wR);
  matmul3(wR, rr, true, R1);
  matmul3(wR, rr, false, R2);
  for (int i = 0; i < 3; ++i) t[i] = R2[i * 3] * T[0] + R2[i * 3 + 1] * T[1] + R2[i * 3 + 2] * T[2];
  // common focal length; principal points from the projected image corners
  const double nx = width, ny = height;
  const double fc = (idx == 0 ? K1[4] + K2[4] : K1[0] + K2[0]) * 0.5;
  double cc[2][2];
  for (int k = 0; k < 2; ++k) {
    const double* A = k == 0 ? K1 : K2;
    const double* Rk = k == 0 ? R1 : R2;
    double sx = 0, sy = 0;
    for (int i = 0; i < 4; ++i) {
      const float u = (float)((i % 2) * (nx - 1)), v = (float)((i < 2 ? 0 : 1) * (ny - 1));
      // undistortPoints with zero distortion, as float
      const float xn = (float)(((double)u - A[2]) * (1. / A[0]));
      const float yn = (float)(((double)v - A[5]) * (1. / A[4]));
      const double X = Rk[0] * xn + Rk[1] * yn + Rk[2];
      const double Y = Rk[3] * xn + Rk[4] * yn + Rk[5];
      const double Z = Rk[6] * xn + Rk[7] * yn + Rk[8];
      sx += (double)(float)(fc * (X / Z));
      sy += (double)(float)(fc * (Y / Z));
    }
    cc[k][0] = (nx - 1) / 2 - sx / 4;
    cc[k][1] = (ny - 1) / 2 - sy / 4;
  }
  if (idx == 0) cc[0][1] = cc[1][1] = (cc[0][1] + cc[1][1]) * 0.5;
  else cc[0][0] = cc[1][0] = (cc[0][0] + cc[1][0]) * 0.5;
  for (int i = 0; i < 12; ++i) P1[i] = P2[i] = 0;
  P1[0] = P1[5] = P2[0] = P2[5] = fc;
  P1[2] = cc[0][0];
  P1[6] = cc[0][1];
  P2[2] = cc[1][0];
  P2[6] = cc[1][1];
  P1[10] = P2[10] = 1;
  P2[idx * 4 + 3] = t[idx] * fc;
  const double q[16] = {1, 0, 0, -cc[0][0], 0, 1, 0, -cc[0][1], 0, 0, 0, fc, 0, 0,
                        -1. / t[idx],
                        (idx == 0 ? cc[0][0] - cc[1][0] : cc[0][1] - cc[1][1]) / t[idx]};
  std::memcpy(Q, q, sizeof(q));
  return PSGC_OK;
}

int psgc_full_left_rectification(const double K1[9], int width, int height, const double R1[9],
                                 double P1[12], double P2[12], double Q[16], int* out_width,
                                 int* out_height) {
  if (!K1 || !R1 || !P1 || !P2 || !Q || !out_width || !out_height)
    PSGC_FAIL(PSGC_E_INVALID, "null argument");
  if (width < 1 || height < 1) PSGC_FAIL(PSGC_E_INVALID, "image size %dx%d", width, height);
  if (!(K1[0] != 0) || !(K1[4] != 0)) PSGC_FAIL(PSGC_E_INVALID, "K1 focal lengths");
  // undistortPoints with zero distortion: its iterations leave the
  // normalised point unchanged, then RR = P1[:3,:3] R1 maps it
  double RR[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      RR[i][j] = P1[i * 4 + 0] * R1[0 * 3 + j] + P1[i * 4 + 1] * R1[1 * 3 + j] +
                 P1[i * 4 + 2] * R1[2 * 3 + j];
  const double ifx = 1.0 / K1[0], ify = 1.0 / K1[4], cx = K1[2], cy = K1[5];
  const float cu[4] = {0.0f, (float)width - 1.0f, (float)width - 1.0f, 0.0f};
  const float cv[4] = {0.0f, 0.0f, (float)height - 1.0f, (float)height - 1.0f};
  float min_x = 0, max_x = 0, min_y = 0, max_y = 0;
  for (int c = 0; c < 4; ++c) {
    const double x = ((double)cu[c] - cx) * ifx, y = ((double)cv[c] - cy) * ify;
    const double xx = RR[0][0] * x + RR[0][1] * y + RR[0][2];
    const double yy = RR[1][0] * x + RR[1][1] * y + RR[1][2];
    const double ww = 1.0 / (RR[2][0] * x + RR[2][1] * y + RR[2][2]);
    const float px = (float)(xx * ww), py = (float)(yy * ww);
    if (!std::isfinite(px) || !std::isfinite(py))
      PSGC_FAIL(PSGC_E_INVALID, "corner %d maps to infinity", c);
    if (c == 0) {
      min_x = max_x = px;
      min_y = max_y = py;
    } else {
      min_x = std::min(min_x, px);
      max_x = std::max(max_x, px);
      min_y = std::min(min_y, py);
      max_y = std::max(max_y, py);
    }
  }
  const double shift_x = -std::floor(min_x), shift_y = -std::floor(min_y);
  const float wf = std::ceil(max_x - min_x + 1.0f), hf = std::ceil(max_y - min_y + 1.0f);
  if (!(wf <= 65536.0f) || !(hf <= 65536.0f))
    PSGC_FAIL(PSGC_E_INVALID, "rectified canvas %.0fx%.0f", wf, hf);
  P1[2] += shift_x;
  P1[6] += shift_y;
  P2[2] += shift_x;
  P2[6] += shift_y;
  Q[3] -= shift_x;
  Q[7] -= shift_y;
  *out_width = (int)wf;
  *out_height = (int)hf;
  return PSGC_OK;
}

}  // extern "C"
