// compact.hip — device-side rectification maps and remap, valid-point
// compaction, and the binary PLY writer.
//
// remap:    cv::remap semantics of the reference's setData
//           (JSLAB_PhaseEpipolarReconstructor_V2.cpp:368-395): phase and
//           modulation INTER_LINEAR, orders INTER_NEAREST, BORDER_CONSTANT 0,
//           with OpenCV 4.7.0's fixed-point coordinates (1/32 px for
//           bilinear) and float weight table.
// maps:     initUndistortRectifyMap for zero distortion
//           (…V2.cpp:333-338): raw = K_raw R^T K_rect^-1 (u, v, 1).
// compact:  stable row-major compaction of the points with P != 0 (the
//           order of the reference's PLY writers, main.cpp:85-162), by
//           wavefront ballot + per-block prefix sums.
// PLY:      binary little-endian, float x y z [+ uchar red green blue]
//           (main.cpp:132-162).

#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "psgc_internal.h"

namespace psgc {
namespace {

__global__ void rect_maps_kernel(int rows, int cols, const double* __restrict__ Mdev,
                                 float* __restrict__ mx, float* __restrict__ my) {
  // M = K_raw R^T K_rect^-1 (3x3 row-major)
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y * blockDim.y + threadIdx.y;
  if (u >= cols || v >= rows) return;
  const double x = Mdev[0] * u + Mdev[1] * v + Mdev[2];
  const double y = Mdev[3] * u + Mdev[4] * v + Mdev[5];
  const double w = Mdev[6] * u + Mdev[7] * v + Mdev[8];
  const size_t p = (size_t)v * cols + u;
  mx[p] = (float)(x / w);
  my[p] = (float)(y / w);
}

// cv::remap's fixed-point coordinates (OpenCV 4.7.0 RemapInvoker):
// cvRound (half to even; NaN, infinities and values outside int range give
// INT_MIN) and saturate_cast<short>.
__device__ inline int cv_round(float v) {
  return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)rintf(v) : INT32_MIN;
}
__device__ inline int sat_short(int v) { return min(max(v, -32768), 32767); }

// INTER_LINEAR on a f64 map: the coordinate is quantised to 1/32 px, the
// four float weights of OpenCV's bilinear table (exact products of k/32)
// multiply the taps in double, summed in the table's order; taps outside the
// source read 0 (BORDER_CONSTANT).
__global__ void remap_f64_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                 const float* __restrict__ mx, const float* __restrict__ my,
                                 int rows, int cols, int src_rows, int src_cols) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (size_t)rows * cols) return;
  const int X = cv_round(mx[p] * 32.0f), Y = cv_round(my[p] * 32.0f);
  const int sx = sat_short(X >> 5), sy = sat_short(Y >> 5);
  const float ax = (float)(X & 31) * (1.0f / 32), ay = (float)(Y & 31) * (1.0f / 32);
  auto tap = [&](int y, int x) -> double {
    return (x >= 0 && x < src_cols && y >= 0 && y < src_rows) ? src[(size_t)y * src_cols + x]
                                                               : 0.0;
  };
  double v = tap(sy, sx) * (double)((1.0f - ay) * (1.0f - ax));
  v = v + tap(sy, sx + 1) * (double)((1.0f - ay) * ax);
  v = v + tap(sy + 1, sx) * (double)(ay * (1.0f - ax));
  v = v + tap(sy + 1, sx + 1) * (double)(ay * ax);
  dst[p] = v;
}

__global__ void remap_i32_nearest_kernel(const int32_t* __restrict__ src, int32_t* __restrict__ dst,
                                         const float* __restrict__ mx,
                                         const float* __restrict__ my, int rows, int cols,
                                         int src_rows, int src_cols) {
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (size_t)rows * cols) return;
  const int x = sat_short(cv_round(mx[p])), y = sat_short(cv_round(my[p]));
  dst[p] = (x >= 0 && x < src_cols && y >= 0 && y < src_rows) ? src[(size_t)y * src_cols + x] : 0;
}

constexpr int CB = 256;   // threads per block
constexpr int CPT = 4;    // pixels per thread
constexpr int CPB = CB * CPT;

__global__ void __launch_bounds__(CB) compact_count_kernel(const uint8_t* __restrict__ P,
                                                           size_t n, int64_t* __restrict__ counts) {
  __shared__ int wsum[CB / 64];
  const size_t base = (size_t)blockIdx.x * CPB;
  int c = 0;
  for (int q = 0; q < CPT; ++q) {
    const size_t p = base + (size_t)q * CB + threadIdx.x;
    c += (p < n && P[p] != 0) ? 1 : 0;
  }
  // wavefront sum
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t s = 0;
    for (int w = 0; w < CB / 64; ++w) s += wsum[w];
    counts[blockIdx.x] = s;
  }
}

// Exclusive scan of the block counts in one workgroup (sequential chunks).
__global__ void __launch_bounds__(1024) compact_scan_kernel(int64_t* __restrict__ counts,
                                                            size_t nblocks,
                                                            int64_t* __restrict__ total) {
  __shared__ int64_t part[1024];
  const size_t per = (nblocks + 1023) / 1024;
  const size_t b0 = threadIdx.x * per;
  int64_t s = 0;
  for (size_t i = b0; i < b0 + per && i < nblocks; ++i) s += counts[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int t = 0; t < 1024; ++t) {
      const int64_t v = part[t];
      part[t] = acc;
      acc += v;
    }
    *total = acc;
  }
  __syncthreads();
  int64_t acc = part[threadIdx.x];
  for (size_t i = b0; i < b0 + per && i < nblocks; ++i) {
    const int64_t v = counts[i];
    counts[i] = acc;
    acc += v;
  }
}

// Scatter: pixel order inside a block is q-major over the CB threads, so the
// prefix is built per sub-row of CB pixels with wavefront ballots.
__global__ void __launch_bounds__(CB) compact_scatter_kernel(
    const float* __restrict__ xyz, const uint8_t* __restrict__ P, const uint8_t* __restrict__ gray,
    size_t n, const int64_t* __restrict__ offsets, float* __restrict__ out,
    uint8_t* __restrict__ out_gray) {
  __shared__ int wcnt[CB / 64];
  __shared__ int64_t running;
  const size_t base = (size_t)blockIdx.x * CPB;
  if (threadIdx.x == 0) running = offsets[blockIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = 0; q < CPT; ++q) {
    const size_t p = base + (size_t)q * CB + threadIdx.x;
    const bool v = p < n && P[p] != 0;
    const unsigned long long m = __ballot(v);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int64_t off = running;
    for (int w = 0; w < wave; ++w) off += wcnt[w];
    if (v) {
      const int64_t o = off + before;
      out[o * 3 + 0] = xyz[p * 3 + 0];
      out[o * 3 + 1] = xyz[p * 3 + 1];
      out[o * 3 + 2] = xyz[p * 3 + 2];
      if (out_gray) out_gray[o] = gray ? gray[p] : 255;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0;
      for (int w = 0; w < CB / 64; ++w) s += wcnt[w];
      running += s;
    }
    __syncthreads();
  }
}

}  // namespace

int launch_rect_maps(int rows, int cols, const double* M_dev, float* mx, float* my,
                     hipStream_t s) {
  dim3 b(16, 16), g((cols + 15) / 16, (rows + 15) / 16);
  rect_maps_kernel<<<g, b, 0, s>>>(rows, cols, M_dev, mx, my);
  PSGC_LAUNCHED();
  return PSGC_OK;
}

int launch_remap(int kind, const void* src, void* dst, const float* mx, const float* my, int rows,
                 int cols, int src_rows, int src_cols, hipStream_t s) {
  const size_t n = (size_t)rows * cols;
  const unsigned g = (unsigned)((n + 255) / 256);
  PSGC_PROF_BEGIN("remap", s);
  if (kind == 0)
    remap_f64_kernel<<<g, 256, 0, s>>>((const double*)src, (double*)dst, mx, my, rows, cols,
                                        src_rows, src_cols);
  else
    remap_i32_nearest_kernel<<<g, 256, 0, s>>>((const int32_t*)src, (int32_t*)dst, mx, my, rows,
                                               cols, src_rows, src_cols);
  PSGC_LAUNCHED();
  PSGC_PROF_END("remap", s);
  return PSGC_OK;
}

size_t compact_scratch_int64(size_t n) { return (n + CPB - 1) / CPB + 1; }

namespace {

int compact_launch(const float* xyz, const uint8_t* P, const uint8_t* gray, size_t n,
                   int64_t* scratch, float* out, uint8_t* out_gray, int64_t* total,
                   hipStream_t s, bool profile) {
  const size_t nb = (n + CPB - 1) / CPB;
  if (profile) PSGC_PROF_BEGIN("compact", s);
  compact_count_kernel<<<(unsigned)nb, CB, 0, s>>>(P, n, scratch);
  PSGC_LAUNCHED();
  compact_scan_kernel<<<1, 1024, 0, s>>>(scratch, nb, total);
  PSGC_LAUNCHED();
  compact_scatter_kernel<<<(unsigned)nb, CB, 0, s>>>(xyz, P, gray, n, scratch, out, out_gray);
  PSGC_LAUNCHED();
  if (profile) PSGC_PROF_END("compact", s);
  return PSGC_OK;
}

// Scratch of the allocation-free compaction: one int64 count per block of
// CPB pixels (overwritten by every call) and the total in the last entry.
// Each GPU holds its own copy from code-object load to process exit.
static_assert(PSGC_COMPACT_NOALLOC_MAX_POINTS % CPB == 0, "whole blocks");
constexpr size_t NOALLOC_SCRATCH_INT64 = PSGC_COMPACT_NOALLOC_MAX_POINTS / CPB + 1;
__device__ __align__(256) int64_t scratch[NOALLOC_SCRATCH_INT64];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
int64_t* g_scratch_ptr[MAX_DEVICES] = {};

}  // namespace

int launch_compact(const float* xyz, const uint8_t* P, const uint8_t* gray, size_t n,
                   int64_t* scratch, float* out, uint8_t* out_gray, int64_t* total,
                   hipStream_t s) {
  return compact_launch(xyz, P, gray, n, scratch, out, out_gray, total, s, true);
}

// The device address of `scratch` on `device` (which must be current): looked
// up once per device and kept; host code never passes the array's own name.
int compact_noalloc_scratch(int device, int64_t** out) {
  if (device < 0 || device >= MAX_DEVICES) PSGC_FAIL(PSGC_E_NODEVICE, "device %d", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int64_t* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      PSGC_FAIL(PSGC_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                hipGetErrorString(e));
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return PSGC_OK;
}

int launch_compact_noalloc(int device, const float* xyz, const uint8_t* P, const uint8_t* gray,
                           size_t n, float* out, uint8_t* out_gray, int64_t* count,
                           hipStream_t s) {
  if (n > PSGC_COMPACT_NOALLOC_MAX_POINTS)
    PSGC_FAIL(PSGC_E_INVALID, "%zu points exceed the allocation-free limit of %zu", n,
              (size_t)PSGC_COMPACT_NOALLOC_MAX_POINTS);
  int64_t* sc;
  PSGC_TRY(compact_noalloc_scratch(device, &sc));
  return compact_launch(xyz, P, gray, n, sc, out, out_gray, count, s, false);
}

int write_ply(const char* path, const float* xyz, const uint8_t* gray, int64_t n) {
  FILE* fp = std::fopen(path, "wb");
  if (!fp) PSGC_FAIL(PSGC_E_INVALID, "cannot open %s", path);
  std::fprintf(fp, "ply\nformat binary_little_endian 1.0\nelement vertex %lld\n", (long long)n);
  std::fprintf(fp, "property float x\nproperty float y\nproperty float z\n");
  if (gray) std::fprintf(fp, "property uchar red\nproperty uchar green\nproperty uchar blue\n");
  std::fprintf(fp, "end_header\n");
  bool ok = true;
  if (!gray) {
    ok = std::fwrite(xyz, sizeof(float) * 3, (size_t)n, fp) == (size_t)n;
  } else {
    std::vector<unsigned char> rec(15);
    for (int64_t i = 0; i < n && ok; ++i) {
      std::memcpy(rec.data(), xyz + i * 3, 12);
      rec[12] = rec[13] = rec[14] = gray[i];
      ok = std::fwrite(rec.data(), 1, 15, fp) == 15;
    }
  }
  if (std::fclose(fp) != 0) ok = false;
  if (!ok) PSGC_FAIL(PSGC_E_INVALID, "write to %s failed", path);
  return PSGC_OK;
}

// Row-major doubles with no header; the optional text file names the shape
// (the dump format of JSLAB_PhaseDisparityExporter.cpp:115-141).
int write_raw_f64(const char* raw_path, const char* info_path, const double* data, int rows,
                  int cols) {
  FILE* fp = std::fopen(raw_path, "wb");
  if (!fp) PSGC_FAIL(PSGC_E_INVALID, "cannot open %s", raw_path);
  const size_t n = (size_t)rows * cols;
  bool ok = n == 0 || std::fwrite(data, sizeof(double), n, fp) == n;
  if (std::fclose(fp) != 0) ok = false;
  if (!ok) PSGC_FAIL(PSGC_E_INVALID, "write to %s failed", raw_path);
  if (!info_path) return PSGC_OK;
  FILE* fi = std::fopen(info_path, "w");
  if (!fi) PSGC_FAIL(PSGC_E_INVALID, "cannot open %s", info_path);
  ok = std::fprintf(fi, "rows %d\ncols %d\ntype CV_64FC1 (double)\nlayout row-major, one double per pixel\n",
                    rows, cols) > 0;
  if (std::fclose(fi) != 0) ok = false;
  if (!ok) PSGC_FAIL(PSGC_E_INVALID, "write to %s failed", info_path);
  return PSGC_OK;
}

}  // namespace psgc
