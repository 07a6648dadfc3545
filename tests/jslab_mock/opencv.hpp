// Test mock of OpenCV's umbrella header.
#pragma once
#include "opencv2/core.hpp"
