// Test mock: the MLS prototypes of PhaseShiftGrayCode3D/src/JSLAB_MLS_Filter.h:30-57,
// default arguments included.
#pragma once
#include "JSLAB_Config.h"
#include <opencv.hpp>
#include "dllexport.h"

DLL_EXPORT void applyMLSFilterToPointCloud_CPU(
    const cv::Mat& src_xyz, const cv::Mat& P, cv::Mat& dst_xyz, int window_size,
    float spatial_sigma, int polynomial_order = 0, float max_radius = 0.0f, int iterations = 1);

#ifdef JSLAB_USE_CUDA
DLL_EXPORT void applyMLSFilterToPointCloud_GPU(
    const cv::Mat& src_xyz, const cv::Mat& P, cv::Mat& dst_xyz, int window_size,
    float spatial_sigma, int polynomial_order = 0, float max_radius = 0.0f, int iterations = 1);
#endif

DLL_EXPORT void applyMLSFilterToPointCloud(
    const cv::Mat& src_xyz, const cv::Mat& P, cv::Mat& dst_xyz, int window_size,
    float spatial_sigma, bool use_gpu, int polynomial_order = 0, float max_radius = 0.0f,
    int iterations = 1);
