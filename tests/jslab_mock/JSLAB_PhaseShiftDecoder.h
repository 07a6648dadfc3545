// Test mock of the decoder interface: the Config the GPU launcher takes
// (fields, order and defaults of PhaseShiftGrayCode3D/src/JSLAB_PhaseShiftDecoder.h:22-54).
#pragma once
#include "dllexport.h"
#include <opencv2/core.hpp>

class DLL_EXPORT JSLAB_PhaseShiftDecoder {
 public:
  struct Config {
    double freq[3];
    int shift_num[3];
    double modulation_threshold = 1.0;
    double preblur_sigma = 3.0;
    bool avg_three_phase = true;
    bool post_smooth = true;
    bool enable_hdr = false;
    double hdr_saturation_gray = 255.0;
    double hdr_min_amplitude = 2.0;
  };
  virtual ~JSLAB_PhaseShiftDecoder() = default;
};
