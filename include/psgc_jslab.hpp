// psgc_jslab.hpp — the reference's cv::Mat GPU launchers implemented on the
// psgc C ABI, for a drop-in build of PhaseShiftGrayCode3D: replace
// JSLAB_PhaseShiftDecoder_V2_GPU.cu, JSLAB_PhaseEpipolarReconstructor_V2_GPU.cu
// and JSLAB_MLS_Filter.cu with one translation unit that includes this
// header (after the reference headers) and link libpsgc.so.
//
// Signatures replaced (PhaseShiftGrayCode3D/src/):
//   JSLAB_DecodePhase_GPU / _Warmup / _Release   JSLAB_PhaseShiftDecoder_V2_GPU.h:50-60
//   JSLAB_ReconstructPipeline_GPU               JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:131-137
//   applyMLSFilterToPointCloud_GPU              JSLAB_MLS_Filter.h:42-47
// Compiled only where OpenCV is available.
#pragma once

#if defined(__has_include)
#if __has_include(<opencv2/core.hpp>)
#define PSGC_HAVE_OPENCV 1
#endif
#endif

#include <stdexcept>
#include <string>
#include <vector>

#include "psgc.h"

namespace psgc_jslab {

inline void check(int rc) {
  if (rc != PSGC_OK) throw std::runtime_error(std::string("psgc: ") + psgc_last_error());
}

// One context per thread (the reference used process-global buffers).
inline psgc_ctx* thread_ctx() {
  thread_local struct Holder {
    psgc_ctx* ctx = nullptr;
    Holder() { check(psgc_create(&ctx, 0)); }
    ~Holder() { psgc_destroy(ctx); }
  } h;
  return h.ctx;
}

}  // namespace psgc_jslab

#ifdef PSGC_HAVE_OPENCV
#include <opencv2/core.hpp>

inline void JSLAB_DecodePhase_GPU_Warmup() { psgc_jslab::check(psgc_warmup()); }
inline void JSLAB_DecodePhase_GPU_Release() { psgc_release(psgc_jslab::thread_ctx()); }

// JSLAB_PhaseShiftDecoder::Config -> psgc_decode_config (same fields).
template <typename Config>
inline psgc_decode_config psgc_decode_config_from(const Config& c) {
  psgc_decode_config d;
  psgc_decode_config_default(&d);
  for (int i = 0; i < 3; ++i) {
    d.freq[i] = c.freq[i];
    d.shift_num[i] = c.shift_num[i];
  }
  d.modulation_threshold = c.modulation_threshold;
  d.preblur_sigma = c.preblur_sigma;
  d.avg_three_phase = c.avg_three_phase;
  d.post_smooth = c.post_smooth;
  d.enable_hdr = c.enable_hdr;
  d.hdr_saturation_gray = c.hdr_saturation_gray;
  d.hdr_min_amplitude = c.hdr_min_amplitude;
  return d;
}

// Same contract as the reference launcher; the post-smooth the reference
// runs on the host afterwards (JSLAB_PhaseShiftDecoder_V2_GPU.cpp:35-36) is
// applied by the library when cfg.post_smooth is set, so callers drop theirs.
template <typename Config>
inline void JSLAB_DecodePhase_GPU(const std::vector<cv::Mat>& images, const Config& cfg,
                                  cv::Mat& phase, cv::Mat& k_map, cv::Mat& mod_map) {
  const psgc_decode_config d = psgc_decode_config_from(cfg);
  const int rows = images[0].rows, cols = images[0].cols;
  const int type = images[0].depth() == CV_8U ? PSGC_SAMPLE_U8
                   : images[0].depth() == CV_16U ? PSGC_SAMPLE_U16 : PSGC_SAMPLE_F64;
  std::vector<const void*> frames;
  for (const cv::Mat& m : images) frames.push_back(m.data);
  phase.create(rows, cols, CV_64FC1);
  k_map.create(rows, cols, CV_32SC1);
  mod_map.create(rows, cols, CV_64FC1);
  psgc_jslab::check(psgc_decode(psgc_jslab::thread_ctx(), &d, frames.data(), (int)frames.size(),
                                type, rows, cols, (int64_t)images[0].step[0],
                                phase.ptr<double>(), k_map.ptr<int>(), mod_map.ptr<double>(),
                                nullptr, PSGC_MEM_HOST));
}

template <typename Config>
inline psgc_recon_config psgc_recon_config_from(const Config& c, int semantics) {
  psgc_recon_config r;
  psgc_recon_config_default(&r);
  r.min_depth = c.min_depth;
  r.max_depth = c.max_depth;
  r.min_phase_diff = c.min_phase_diff;
  r.max_phase_diff = c.max_phase_diff;
  r.grad_threshold = c.grad_threshold;
  r.use_rectified_coords = c.use_rectified_coords;
  r.use_disp_median_filter = c.use_disp_median_filter;
  r.radius_search = c.radius_search;
  r.radius_min_neighbors = c.radius_min_neighbors;
  r.radius_window_size = c.radius_window_size;
  r.use_mls = 0;  // the reference applies MLS after the launcher
  r.semantics = semantics;
  return r;
}

template <typename Config>
inline void JSLAB_ReconstructPipeline_GPU(const cv::Mat& l_phase, const cv::Mat& r_phase,
                                          const cv::Mat& l_k, const cv::Mat& r_k, int pha_f,
                                          double /*min_disp*/, double /*max_disp*/,
                                          const double Q[16], const double R1[9],
                                          const Config& cfg, cv::Mat& xyz, cv::Mat& P,
                                          cv::Mat& disp_map) {
  const psgc_recon_config r = psgc_recon_config_from(cfg, PSGC_SEM_GPU);
  const int rows = l_phase.rows, cols = l_phase.cols;
  cv::Mat lp = l_phase.isContinuous() ? l_phase : l_phase.clone();
  cv::Mat rp = r_phase.isContinuous() ? r_phase : r_phase.clone();
  cv::Mat lk = l_k.isContinuous() ? l_k : l_k.clone();
  cv::Mat rk = r_k.isContinuous() ? r_k : r_k.clone();
  xyz.create(rows, cols, CV_32FC3);
  P.create(rows, cols, CV_8UC1);
  disp_map.create(rows, cols, CV_32FC1);
  psgc_jslab::check(psgc_reconstruct(psgc_jslab::thread_ctx(), &r, lp.ptr<double>(),
                                     rp.ptr<double>(), lk.ptr<int>(), rk.ptr<int>(), rows, cols,
                                     pha_f, Q, R1, xyz.ptr<float>(), P.ptr<uchar>(),
                                     disp_map.ptr<float>(), nullptr, PSGC_MEM_HOST));
}

inline void applyMLSFilterToPointCloud_GPU(const cv::Mat& src_xyz, const cv::Mat& P,
                                           cv::Mat& dst_xyz, int window_size, float spatial_sigma,
                                           int polynomial_order = 0, float max_radius = 0.0f,
                                           int iterations = 1) {
  cv::Mat s = src_xyz.isContinuous() ? src_xyz : src_xyz.clone();
  cv::Mat m = P.isContinuous() ? P : P.clone();
  dst_xyz.create(src_xyz.rows, src_xyz.cols, CV_32FC3);
  psgc_jslab::check(psgc_filter_mls(psgc_jslab::thread_ctx(), s.ptr<float>(), m.ptr<uchar>(),
                                    dst_xyz.ptr<float>(), s.rows, s.cols, window_size,
                                    spatial_sigma, polynomial_order, max_radius, iterations,
                                    PSGC_SEM_GPU, PSGC_MEM_HOST));
}
#endif  // PSGC_HAVE_OPENCV
