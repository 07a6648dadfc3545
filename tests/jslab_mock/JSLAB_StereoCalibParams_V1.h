// Test mock: only named by the reconstructor class declaration.
#pragma once
struct JSLAB_StereoCalibParams_V1 {};
