// psgc_jslab.cpp — the reference's three cv::Mat GPU launchers, implemented on
// the psgc C ABI. A drop-in build of PhaseShiftGrayCode3D compiles this one
// translation unit in place of JSLAB_PhaseShiftDecoder_V2_GPU.cu,
// JSLAB_PhaseEpipolarReconstructor_V2_GPU.cu and JSLAB_MLS_Filter.cu, and
// links libpsgc.so. It includes the reference's own headers, so the
// declarations (and the default arguments of the MLS launcher) are theirs.
//
// Definitions (PhaseShiftGrayCode3D/src/):
//   JSLAB_DecodePhase_GPU            JSLAB_PhaseShiftDecoder_V2_GPU.h:50-52
//   JSLAB_DecodePhase_GPU_Warmup     JSLAB_PhaseShiftDecoder_V2_GPU.h:56
//   JSLAB_DecodePhase_GPU_Release    JSLAB_PhaseShiftDecoder_V2_GPU.h:60
//   JSLAB_ReconstructPipeline_GPU    JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:131-137
//   applyMLSFilterToPointCloud_GPU   JSLAB_MLS_Filter.h:42-47
//
// Errors: a failing psgc call throws std::runtime_error with
// psgc_last_error(); the reference's launchers checked no CUDA status.
// Threading: one psgc context per calling thread, created on first use.
#ifndef JSLAB_USE_CUDA
#define JSLAB_USE_CUDA 1  // JSLAB_MLS_Filter.h declares the GPU launcher under it
#endif

#include "JSLAB_MLS_Filter.h"
#include "JSLAB_PhaseEpipolarReconstructor_V2_GPU.h"
#include "JSLAB_PhaseShiftDecoder_V2_GPU.h"

#include <stdexcept>
#include <string>
#include <vector>

#include "psgc.h"

namespace {

void check(int rc) {
  if (rc != PSGC_OK) throw std::runtime_error(std::string("psgc: ") + psgc_last_error());
}

struct ThreadCtx {
  psgc_ctx* ctx = nullptr;
  ~ThreadCtx() {
    if (ctx) psgc_destroy(ctx);
  }
};
thread_local ThreadCtx t_ctx;

psgc_ctx* ctx() {
  if (!t_ctx.ctx) check(psgc_create(&t_ctx.ctx, 0));
  return t_ctx.ctx;
}

// A dense copy when the rows of m are padded (the ABI takes dense maps).
cv::Mat dense(const cv::Mat& m) { return m.isContinuous() ? m : m.clone(); }

void require(bool ok, const char* what) {
  if (!ok) throw std::invalid_argument(std::string("psgc: ") + what);
}

}  // namespace

void JSLAB_DecodePhase_GPU_Warmup() { check(psgc_warmup()); }

void JSLAB_DecodePhase_GPU_Release() {
  if (t_ctx.ctx) check(psgc_release(t_ctx.ctx));
}

// The caller (JSLAB_PhaseShiftDecoderV2_GPU::startCalculate,
// JSLAB_PhaseShiftDecoder_V2_GPU.cpp:26-37) post-smooths the phase on the host
// afterwards, so the library's post-smooth stays off here; the reference's
// GPU decoder has no HDR fit, so HDR stays off too. Frames are read at their
// own depth: CV_8U as the reference uploads them, CV_16U and CV_64F exactly
// (the reference saturated every other depth to 8 bit,
// JSLAB_PhaseShiftDecoder_V2_GPU.cu:252-258).
void JSLAB_DecodePhase_GPU(const std::vector<cv::Mat>& images,
                           const JSLAB_PhaseShiftDecoder::Config& cfg, cv::Mat& phase,
                           cv::Mat& k_map, cv::Mat& mod_map) {
  const int total = cfg.shift_num[0] + cfg.shift_num[1] + cfg.shift_num[2];
  if (total <= 0 || images.empty()) return;  // as the reference launcher
  require((int)images.size() >= total, "fewer frames than the shift counts name");
  const int rows = images[0].rows, cols = images[0].cols, type = images[0].type();
  int sample;
  if (type == CV_8UC1) sample = PSGC_SAMPLE_U8;
  else if (type == CV_16UC1) sample = PSGC_SAMPLE_U16;
  else if (type == CV_64FC1) sample = PSGC_SAMPLE_F64;
  else throw std::invalid_argument("psgc: frames must be CV_8UC1, CV_16UC1 or CV_64FC1");
  std::vector<cv::Mat> keep(total);
  std::vector<const void*> frames(total);
  for (int i = 0; i < total; ++i) {
    require(images[i].rows == rows && images[i].cols == cols && images[i].type() == type,
            "frames differ in size or type");
    keep[i] = dense(images[i]);
    frames[i] = keep[i].data;
  }
  psgc_decode_config d;
  psgc_decode_config_default(&d);
  for (int i = 0; i < 3; ++i) {
    d.freq[i] = cfg.freq[i];
    d.shift_num[i] = cfg.shift_num[i];
  }
  d.modulation_threshold = cfg.modulation_threshold;
  d.preblur_sigma = cfg.preblur_sigma;
  d.avg_three_phase = cfg.avg_three_phase ? 1 : 0;
  d.post_smooth = 0;
  d.enable_hdr = 0;
  phase.create(rows, cols, CV_64FC1);
  k_map.create(rows, cols, CV_32SC1);
  mod_map.create(rows, cols, CV_64FC1);
  check(psgc_decode(ctx(), &d, frames.data(), total, sample, rows, cols, 0, phase.ptr<double>(),
                    k_map.ptr<int>(), mod_map.ptr<double>(), nullptr, PSGC_MEM_HOST));
}

// Match -> 5x5 median -> Q reprojection -> Gaussian -> radius outlier filter,
// with the reference GPU kernels' semantics (the Gaussian clears the mask of
// the points it rejects; the outlier filter reads that mask). min_disp and
// max_disp are unused, as in the reference
// (JSLAB_PhaseEpipolarReconstructor_V2_GPU.cu:368-372); MLS is not part of
// this launcher (the caller runs applyMLSFilterToPointCloud afterwards,
// JSLAB_PhaseEpipolarReconstructor_V2_GPU.cpp:119-127).
void JSLAB_ReconstructPipeline_GPU(const cv::Mat& l_phase, const cv::Mat& r_phase,
                                   const cv::Mat& l_k, const cv::Mat& r_k, int pha_f,
                                   double min_disp, double max_disp, const double Q[16],
                                   const double R1[9],
                                   const JSLAB_PhaseEpipolarReconstructor_V2_GPU::Config& cfg,
                                   cv::Mat& xyz, cv::Mat& P, cv::Mat& disp_map) {
  (void)min_disp;
  (void)max_disp;
  const int rows = l_phase.rows, cols = l_phase.cols;
  require(l_phase.type() == CV_64FC1 && r_phase.type() == CV_64FC1, "phase maps must be CV_64FC1");
  require(l_k.type() == CV_32SC1 && r_k.type() == CV_32SC1, "order maps must be CV_32SC1");
  for (const cv::Mat* m : {&r_phase, &l_k, &r_k})
    require(m->rows == rows && m->cols == cols, "maps differ in size");
  psgc_recon_config r;
  psgc_recon_config_default_gpu(&r);
  r.min_depth = cfg.min_depth;
  r.max_depth = cfg.max_depth;
  r.min_phase_diff = cfg.min_phase_diff;
  r.max_phase_diff = cfg.max_phase_diff;
  r.grad_threshold = cfg.grad_threshold;
  r.use_rectified_coords = cfg.use_rectified_coords ? 1 : 0;
  r.use_disp_median_filter = cfg.use_disp_median_filter ? 1 : 0;
  r.radius_search = cfg.radius_search;
  r.radius_min_neighbors = cfg.radius_min_neighbors;
  r.radius_window_size = cfg.radius_window_size;
  r.use_mls = 0;
  r.semantics = PSGC_SEM_GPU;
  const cv::Mat lp = dense(l_phase), rp = dense(r_phase), lk = dense(l_k), rk = dense(r_k);
  xyz.create(rows, cols, CV_32FC3);
  P.create(rows, cols, CV_8UC1);
  disp_map.create(rows, cols, CV_32FC1);
  check(psgc_reconstruct(ctx(), &r, lp.ptr<double>(), rp.ptr<double>(), lk.ptr<int>(),
                         rk.ptr<int>(), rows, cols, pha_f, Q, R1, xyz.ptr<float>(),
                         P.ptr<unsigned char>(), disp_map.ptr<float>(), nullptr, PSGC_MEM_HOST));
}

// The organized MLS filter with the reference GPU kernel's semantics (per-
// neighbour sums fused as nvcc contracts them; fallback 2 -> 1 -> 0).
void applyMLSFilterToPointCloud_GPU(const cv::Mat& src_xyz, const cv::Mat& P, cv::Mat& dst_xyz,
                                    int window_size, float spatial_sigma, int polynomial_order,
                                    float max_radius, int iterations) {
  require(src_xyz.type() == CV_32FC3 && P.type() == CV_8UC1, "xyz must be CV_32FC3, P CV_8UC1");
  require(P.rows == src_xyz.rows && P.cols == src_xyz.cols, "xyz and P differ in size");
  const cv::Mat s = dense(src_xyz), m = dense(P);
  // dst may alias src (the reference's callers pass distinct Mats; a fresh
  // buffer keeps an aliasing caller correct as well)
  cv::Mat out(s.rows, s.cols, CV_32FC3);
  check(psgc_filter_mls(ctx(), s.ptr<float>(), m.ptr<unsigned char>(), out.ptr<float>(), s.rows,
                        s.cols, window_size, spatial_sigma, polynomial_order, max_radius,
                        iterations, PSGC_SEM_GPU, PSGC_MEM_HOST));
  dst_xyz = out;
}
