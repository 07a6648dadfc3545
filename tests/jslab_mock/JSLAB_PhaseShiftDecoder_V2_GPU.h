// Test mock: the launcher prototypes of
// PhaseShiftGrayCode3D/src/JSLAB_PhaseShiftDecoder_V2_GPU.h:50-60.
#pragma once
#include "JSLAB_PhaseShiftDecoder.h"
#include <opencv2/core.hpp>
#include <vector>

void JSLAB_DecodePhase_GPU(const std::vector<cv::Mat>& images,
                           const JSLAB_PhaseShiftDecoder::Config& cfg,
                           cv::Mat& phase, cv::Mat& k_map, cv::Mat& mod_map);
DLL_EXPORT void JSLAB_DecodePhase_GPU_Warmup();
void JSLAB_DecodePhase_GPU_Release();
