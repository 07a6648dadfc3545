// psgc_internal.h — context, error plumbing and kernel entry declarations
// shared by the translation units of libpsgc.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "psgc.h"
#include "psgc_voxel.h"

namespace psgc {

void set_error(const char* fmt, ...);

#define PSGC_FAIL(code, ...)        \
  do {                              \
    ::psgc::set_error(__VA_ARGS__); \
    return (code);                  \
  } while (0)

#define PSGC_HIP(expr)                                                        \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess)                                                     \
      PSGC_FAIL(PSGC_E_HIP, "%s failed: %s (%s:%d)", #expr,                  \
                hipGetErrorString(_e), __FILE__, __LINE__);                   \
  } while (0)

#define PSGC_TRY(expr)           \
  do {                           \
    int _rc = (expr);            \
    if (_rc != PSGC_OK) return _rc; \
  } while (0)

// Launch-error check right after a kernel launch.
#define PSGC_LAUNCHED() PSGC_HIP(hipGetLastError())

// One growable device allocation owned by a context.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Buffer slots of a context (staging for host-pointer calls and the
// intermediates of the multi-stage calls).
enum Slot : int {
  S_FRAMES = 0, S_PHASE, S_PHASE_RAW, S_K, S_MOD, S_VALID,
  S_LPH, S_RPH, S_LK, S_RK, S_DISP64, S_DISP, S_DISP2, S_MATCH,
  S_XYZ0, S_XYZ1, S_XYZ2, S_P0, S_P1, S_P2, S_SPAN, S_MEANDIST, S_PARTIAL,
  S_STATS, S_QR, S_TMP0, S_TMP1, S_ADJ0, S_ADJ1, S_ADJ2, S_MAPX, S_MAPY, S_SRC, S_DST, S_CNT,
  S_HDR, S_COEF0, S_COEF1, S_COEF2, S_XYZ3, S_P3, S_MASK,
  // unorganized-cloud grid: keys and indices (sort in/out), sorted points,
  // occupied cells, sort/scan scratch, small counters, retry list
  S_CL_KEY0, S_CL_KEY1, S_CL_IDX0, S_CL_IDX1, S_CL_PTS, S_CL_UKEY, S_CL_START, S_CL_COUNT,
  S_CL_TEMP, S_CL_MISC, S_CL_RETRY, S_CL_TUNE_MD, S_CL_TUNE_ST, S_ADJV,
  // plane-fit objectives: view table, per-chunk partial sums, per-view sums
  S_PLANEV, S_PLANEP, S_PLANEO,
  S_NUM_SLOTS
};

}  // namespace psgc

struct psgc_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::vector<psgc::DevBuf> bufs;
  int cu_count = 0;
  int kernels = PSGC_KERNELS_TILED;
  // pinned staging chunks of the host-pointer calls, their last DMA's event
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  bool pin_busy[2] = {false, false};
  bool pin_failed = false;
};

namespace psgc {

// Entry of a C-ABI call on a context: a non-null context, its device current.
#define PSGC_CTX_ENTER(ctx)                                              \
  do {                                                                   \
    if (!(ctx)) PSGC_FAIL(PSGC_E_INVALID, "null context");               \
    if (hipSetDevice((ctx)->device) != hipSuccess)                       \
      PSGC_FAIL(PSGC_E_HIP, "hipSetDevice(%d) failed", (ctx)->device);    \
  } while (0)

// Returns a device buffer of at least `bytes` for `slot` (grown with a
// stream synchronisation so that in-flight kernels never see a freed block).
int ctx_buf(psgc_ctx* ctx, int slot, size_t bytes, void** out);
int ctx_release(psgc_ctx* ctx);

// Host-memory transfers through the context's pinned staging chunks
// (hoststage.hip). host_to_device(_2d) returns once the caller's memory has
// been read (the device copy is ordered on the context's stream);
// device_to_host returns when the data is in the caller's memory.
int host_to_device(psgc_ctx* ctx, void* dst, const void* src, size_t bytes);
int host_to_device_2d(psgc_ctx* ctx, void* dst, const void* src, size_t spitch, size_t width,
                      size_t rows);
int device_to_host(psgc_ctx* ctx, void* dst, const void* src, size_t bytes);
void host_stage_release(psgc_ctx* ctx);

// ---- kernel launchers (all asynchronous on `s`, device pointers) --------

struct DecodeArgs {
  const void* frames[64];  // device planes
  int n_frames;
  int sample_type;
  int rows, cols;
  int64_t stride;          // row stride in bytes
  int nlev;                // 3 heterodyne, 1 Gray code
  int shift_num[3];
  int level_base[3];
  double sin_tab[48];      // concatenated per level
  double cos_tab[48];
  double freq[3];
  int f12, f23;
  double mod_thresh;
  int avg_three;
  double blur_c, blur_s;   // pre-blur weights (centre, side)
  int gray_bits;
  double gray_min_contrast;
  // HDR: per level, a least-squares fit of I = A + C cos d - S sin d over the
  // unsaturated raw samples instead of the blurred sums
  int hdr;
  double hdr_sat, hdr_min_amp;
  const double* hdr_w[3];  // [mask][k][3] pseudo-inverse columns (device)
  const uint8_t* hdr_ok[3];  // [mask] 1 if the subset has full rank (device)
};

// HDR pseudo-inverse tables of every level of `a`, one host/device blob:
// hdr_table_bytes sizes it, hdr_table_fill writes it on the host and
// hdr_table_bind points a->hdr_w / a->hdr_ok into its device copy.
size_t hdr_table_bytes(const DecodeArgs& a);
void hdr_table_fill(const DecodeArgs& a, char* host);
void hdr_table_bind(DecodeArgs* a, const char* dev);

int launch_decode_plain(const DecodeArgs& a, double* phase_raw, int32_t* k,
                        double* mod, uint8_t* valid, hipStream_t s);
int launch_gray_fix(const DecodeArgs& a, const double* phase_in, double* phase_out,
                    const int32_t* k_in, int32_t* k_out, const uint8_t* valid,
                    hipStream_t s);
int launch_post_smooth(const double* in, double* out, int rows, int cols,
                       hipStream_t s);
int launch_post_smooth_valid(const double* in, const uint8_t* valid, double* out, int rows,
                             int cols, hipStream_t s);

int launch_match_plain(const double* lp, const double* rp, const int32_t* lk,
                       const int32_t* rk, int rows, int cols, int pha_f,
                       double grad_thr, int32_t* span, double* disp64,
                       float* disp32, int32_t* match_col, hipStream_t s);
int launch_median5(const float* src, float* dst, int rows, int cols,
                   hipStream_t s);
int launch_reproject(const float* disp32, const double* disp64, int rows,
                     int cols, const double* QR, double min_depth,
                     double max_depth, int use_rect, float* xyz, uint8_t* P,
                     hipStream_t s);
// coef: three device maps of rows*cols float4 (x, y, z coefficient vectors).
int launch_reproject_poly(const float* disp32, const double* disp64, int rows, int cols,
                          const float* const* coef, const double* R1, double min_depth,
                          double max_depth, int use_rect, float* xyz, uint8_t* P,
                          hipStream_t s);
int launch_gaussian(const float* xyz, const uint8_t* P, float* dst,
                    uint8_t* dstP, int rows, int cols, int win,
                    const float* k2_host, int semantics, hipStream_t s);
int launch_ror(const float* xyz, const uint8_t* P, float* dst, uint8_t* dstP,
               int rows, int cols, float radius, int min_nb, int window,
               hipStream_t s);
int launch_sor(const float* xyz, const uint8_t* P, float* dst, uint8_t* dstP,
               int rows, int cols, int k, float std_mul, int half,
               double* mean_dist, double* partial, double* stats_dev,
               hipStream_t s);
int launch_sor_finish(const float* xyz, double* md, float* dst, uint8_t* dstP,
                      size_t npix, float std_mul, double* partial, double* stats,
                      hipStream_t s);
size_t sor_partial_doubles(size_t npix);
// SOR statistics of one pair within a batch (chunk_base: its first chunk in
// the concatenation of all pairs' chunks of 1024 points).
struct SorStatJob {
  const double* md;
  size_t npix, nchunks, np2;
  double* part;
  double* stats;
  float std_mul;
  int64_t chunk_base;
};
int launch_sor_stats_batch(const SorStatJob* jobs_dev, int njobs, int64_t total_chunks,
                           hipStream_t s);
int launch_sor_apply(const float* xyz, double* md, float* dst, uint8_t* dstP, size_t npix,
                     const double* stats, hipStream_t s);
size_t sor_chunks(size_t npix);
int launch_sor_stats(const double* md, size_t npix, float std_mul, double* partial,
                     double* stats, hipStream_t s);

// ROI-restricted reconstruction (roi.hip); lr / rr = {x, y, w, h} on the host.
int launch_roi_right_window(const int32_t* lk, const int32_t* rk, int rows, int cols, int pha_f,
                            const int32_t* lr, const uint8_t* mask, int32_t* needed,
                            int32_t* mnmx, hipStream_t s);
int launch_roi_match(const double* lp, const double* rp, const int32_t* lk, const int32_t* rk,
                     int cols, const int32_t* lr, const int32_t* rr, const uint8_t* mask,
                     int pha_f, double grad_thr, int32_t* span, double* disp64, float* disp32,
                     hipStream_t s);
int launch_roi_poly(const double* disp64, const float* disp32, const uint8_t* mask, int cols,
                    const int32_t* lr, const float* const* coef, const double* R1,
                    double min_depth, double max_depth, int use_rect, float* xyz, uint8_t* P,
                    hipStream_t s);
int launch_roi_ror_select(const float* xyz_g, const uint8_t* P_in, const float* xyz_r,
                          const uint8_t* P_r, size_t n, unsigned long long* cnt, float* xyz,
                          uint8_t* P, hipStream_t s);

// Tiled single-image / single-pair paths (pipeline.hip).
int tiled_decode(psgc_ctx* ctx, const DecodeArgs& a, double* phase_raw, int32_t* k,
                 double* mod, uint8_t* valid, hipStream_t s);
// coef: NULL for the Q reprojection, else the three device coefficient maps.
int tiled_reconstruct(psgc_ctx* ctx, const psgc_recon_config& rc, const double* lp,
                      const double* rp, const int32_t* lk, const int32_t* rk, int rows,
                      int cols, int pha_f, const double* Q, const float* const* coef,
                      const double* R1, float* xyz, uint8_t* P, float* disp, int32_t* mc,
                      hipStream_t s);
int launch_mls(const float* src, const uint8_t* P, float* dst, int rows,
               int cols, int hw, float sigma, int order, float r2cut,
               int semantics, hipStream_t s);

// Kernel timing with HIP events on the launch stream (off unless enabled
// through psgc_profile_enable). Every launcher brackets its kernels with
// PSGC_PROF_BEGIN/END under a kernel-family name.
void prof_begin(const char* name, hipStream_t s);
void prof_end(const char* name, hipStream_t s);
extern bool g_prof_on;
#define PSGC_PROF_BEGIN(name, s) \
  do { if (::psgc::g_prof_on) ::psgc::prof_begin(name, s); } while (0)
#define PSGC_PROF_END(name, s) \
  do { if (::psgc::g_prof_on) ::psgc::prof_end(name, s); } while (0)

int launch_adj_forward(const double* theta, const double* nominal, const double* phL,
                       const double* phR, const int32_t* mc, int rows, int cols, double* xyz,
                       uint8_t* valid, hipStream_t s);
int launch_adj_backward(const double* theta, const double* nominal, const double* phL,
                        const double* phR, const int32_t* mc, int rows, int cols,
                        const double* gxyz, double* gth, double* gphL, double* gphR,
                        hipStream_t s);

int launch_rect_maps(int rows, int cols, const double* M_dev, float* mx, float* my,
                     hipStream_t s);
int launch_remap(int kind, const void* src, void* dst, const float* mx, const float* my, int rows,
                 int cols, int src_rows, int src_cols, hipStream_t s);
// MLS weights exp(-d^2 / (2 sigma^2)) need a finite positive sigma whose
// 1/(2 sigma^2) is finite, and a finite squared cutoff radius.
int check_mls_weights(float sigma, float max_radius);
int compact_noalloc_scratch(int device, int64_t** out);
int launch_compact_noalloc(int device, const float* xyz, const uint8_t* P, const uint8_t* gray,
                           size_t n, float* out, uint8_t* out_gray, int64_t* count,
                           hipStream_t s);
size_t compact_scratch_int64(size_t n);
int launch_compact(const float* xyz, const uint8_t* P, const uint8_t* gray, size_t n,
                   int64_t* scratch, float* out, uint8_t* out_gray, int64_t* total,
                   hipStream_t s);
int write_ply(const char* path, const float* xyz, const uint8_t* gray, int64_t n);
int write_raw_f64(const char* raw_path, const char* info_path, const double* data, int rows,
                  int cols);

int tiled_median(psgc_ctx* ctx, const float* src, float* dst, int rows, int cols,
                 hipStream_t s);
int tiled_sor(psgc_ctx* ctx, const float* xyz, const uint8_t* P, float* dst, uint8_t* dstP,
              int rows, int cols, int k, float std_mul, int half, double* md, double* partial,
              double* stats, hipStream_t s);
int tiled_mls(psgc_ctx* ctx, const float* src, const uint8_t* P, float* dst, float* tmp,
              int rows, int cols, int hw, float sigma, int order, float r2, int iterations,
              int semantics, hipStream_t s);

// Validates a decode configuration and fills everything but the frames.
int build_decode_args(const psgc_decode_config* cfg, int n_frames, int sample_type, int rows,
                      int cols, int64_t stride, DecodeArgs* a);

// Host-side helpers.
void gauss_kernel_f32(int win, float sigma, float* k2 /* win*win */);
void blur3_weights(double sigma, double* c, double* s);

// ---- unorganized clouds (cloud.hip) --------------------------------------

// Grid of a cloud's participating points: sorted points (x, y, z, index),
// occupied cells (ascending keys, first position, count).
struct CloudGridDev {
  VoxelFrame f;
  int32_t np = 0, ncells = 0;
  const float4* pts = nullptr;
  const uint64_t* ukey = nullptr;
  const int32_t* start = nullptr;
  const int32_t* count = nullptr;
};
int cloud_grid_build(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                     float voxel, CloudGridDev* g);
int cloud_sor_run(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n, int k,
                  float std_mul, float voxel, int rmax, uint8_t* keep, double* md, double* stats);
int cloud_tune_voxel(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n, int k,
                     int rmax, const float* candidates, int ncand, float* best_voxel,
                     float* times_ms);
// tmp: n x 3 floats of scratch when iterations > 1
int cloud_mls_run(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n, float sigma,
                  int order, float max_radius, int iterations, int semantics, float voxel,
                  float* out, float* tmp);
int cloud_gaussian_run(psgc_ctx* ctx, const float* xyz, const uint8_t* valid, int64_t n,
                       float sigma, float radius, float voxel, float* out);

// Batched adjoint (adjoint.hip): one forward or backward launch over views.
int launch_adj_batch(psgc_ctx* ctx, const psgc_adjoint_view* views, int n, bool backward);

}  // namespace psgc
