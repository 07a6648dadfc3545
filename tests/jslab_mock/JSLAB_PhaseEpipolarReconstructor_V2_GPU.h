// Test mock: the GPU reconstructor's Config (fields, order and defaults of
// PhaseShiftGrayCode3D/src/JSLAB_PhaseEpipolarReconstructor_V2_GPU.h:27-61)
// and the launcher prototype (:131-137).
#pragma once
#include <opencv.hpp>
#include "JSLAB_StereoCalibParams_V1.h"
#include "dllexport.h"

class DLL_EXPORT JSLAB_PhaseEpipolarReconstructor_V2_GPU {
 public:
  struct Config {
    double min_depth = 0;
    double max_depth = 1000;
    double min_phase_diff = 0.03;
    double max_phase_diff = 0.4;
    double grad_threshold = 1;
    bool use_rectified_coords = true;
    bool use_disp_median_filter = true;
    float radius_search = 0.05f;
    int radius_min_neighbors = 5;
    int radius_window_size = 5;
    bool use_mls = true;
    int mls_window_size = 25;
    float mls_spatial_sigma = 0.5f;
    bool use_gpu_mls = true;
    int mls_polynomial_order = 1;
    float mls_max_radius = 0.0f;
    int mls_iterations = 2;
  };
};

void JSLAB_ReconstructPipeline_GPU(
    const cv::Mat& l_phase, const cv::Mat& r_phase,
    const cv::Mat& l_k, const cv::Mat& r_k,
    int pha_f, double min_disp, double max_disp,
    const double Q[16], const double R1[9],
    const JSLAB_PhaseEpipolarReconstructor_V2_GPU::Config& cfg,
    cv::Mat& xyz, cv::Mat& P, cv::Mat& disp_map);
