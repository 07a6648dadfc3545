// A caller translation unit that uses the three launchers only through the
// reference's declarations (here: the mock headers beside this file), as
// JSLAB_PhaseShiftDecoderV2_GPU::startCalculate and
// JSLAB_PhaseEpipolarReconstructor_V2_GPU::startCalculate do. Linked with
// zim_amd/jslab/psgc_jslab.cpp and libpsgc.so by tests/test_jslab_dropin.py.
//
//   caller probe        warm up; prints "ok" or "error: <message>"
//   caller run <dir>    decodes dir/frames_{L,R}.u8 and reconstructs, then
//                       runs MLS twice; writes the outputs into <dir>
#define JSLAB_USE_CUDA 1
#include "JSLAB_MLS_Filter.h"
#include "JSLAB_PhaseEpipolarReconstructor_V2_GPU.h"
#include "JSLAB_PhaseShiftDecoder_V2_GPU.h"

#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <string>
#include <vector>

static std::vector<char> slurp(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

static void dump(const std::string& path, const cv::Mat& m) {
  std::ofstream f(path, std::ios::binary);
  for (int i = 0; i < m.rows; ++i)
    f.write(reinterpret_cast<const char*>(m.ptr<uchar>(i)), (std::streamsize)(m.cols * m.elemSize()));
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "probe";
  try {
    JSLAB_DecodePhase_GPU_Warmup();
    if (mode == "probe") {
      std::printf("ok\n");
      return 0;
    }
    const std::string dir = argv[2];
    int rows, cols, n, pha_f, hw_full, order_full, iters_full, hw_default;
    double thr;
    float r_search, sigma, radius_full;
    {
      std::ifstream m(dir + "/meta.txt");
      m >> rows >> cols >> n >> pha_f >> thr >> r_search >> hw_default >> hw_full >> sigma >>
          order_full >> radius_full >> iters_full;
    }
    JSLAB_PhaseShiftDecoder::Config dc;
    dc.freq[0] = 70;
    dc.freq[1] = 64;
    dc.freq[2] = 59;
    dc.shift_num[0] = dc.shift_num[1] = dc.shift_num[2] = n / 3;
    dc.modulation_threshold = thr;  // post_smooth stays true: the launcher must not blur
    cv::Mat phase[2], k[2], mod[2];
    for (int cam = 0; cam < 2; ++cam) {
      std::vector<char> raw = slurp(dir + (cam ? "/frames_R.u8" : "/frames_L.u8"));
      // left frames as padded views (row pitch cols + 16), right frames dense
      const size_t pitch = cam ? (size_t)cols : (size_t)cols + 16;
      std::vector<uchar> store((size_t)n * rows * pitch);
      std::vector<cv::Mat> frames;
      for (int f = 0; f < n; ++f) {
        uchar* base = store.data() + (size_t)f * rows * pitch;
        for (int i = 0; i < rows; ++i)
          std::memcpy(base + i * pitch, raw.data() + ((size_t)f * rows + i) * cols, cols);
        frames.emplace_back(rows, cols, CV_8UC1, base, pitch);
      }
      JSLAB_DecodePhase_GPU(frames, dc, phase[cam], k[cam], mod[cam]);
      const std::string s = cam ? "R" : "L";
      dump(dir + "/phase_" + s + ".f64", phase[cam]);
      dump(dir + "/k_" + s + ".i32", k[cam]);
      dump(dir + "/mod_" + s + ".f64", mod[cam]);
    }
    std::vector<char> qr = slurp(dir + "/QR.f64");
    const double* Q = reinterpret_cast<const double*>(qr.data());
    JSLAB_PhaseEpipolarReconstructor_V2_GPU::Config rc;
    rc.radius_search = r_search;
    cv::Mat xyz, P, disp;
    JSLAB_ReconstructPipeline_GPU(phase[0], phase[1], k[0], k[1], pha_f, 0.0, 0.0, Q, Q + 16, rc,
                                  xyz, P, disp);
    dump(dir + "/xyz.f32", xyz);
    dump(dir + "/P.u8", P);
    dump(dir + "/disp.f32", disp);
    cv::Mat a, b;
    applyMLSFilterToPointCloud_GPU(xyz, P, a, hw_default, sigma);  // the header's defaults
    applyMLSFilterToPointCloud_GPU(xyz, P, b, hw_full, sigma, order_full, radius_full, iters_full);
    dump(dir + "/mls_default.f32", a);
    dump(dir + "/mls_full.f32", b);
    JSLAB_DecodePhase_GPU_Release();
    std::printf("ok\n");
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 3;
  }
  return 0;
}
