// Test mock: the few cv::Mat members that the drop-in launchers and the test
// caller use (no OpenCV in this image). Row-major, reference-counted storage.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>

#define CV_8U 0
#define CV_16U 2
#define CV_32S 4
#define CV_32F 5
#define CV_64F 6
#define CV_MAKETYPE(depth, cn) ((depth) + (((cn) - 1) << 3))
#define CV_8UC1 CV_MAKETYPE(CV_8U, 1)
#define CV_16UC1 CV_MAKETYPE(CV_16U, 1)
#define CV_32SC1 CV_MAKETYPE(CV_32S, 1)
#define CV_32FC1 CV_MAKETYPE(CV_32F, 1)
#define CV_32FC3 CV_MAKETYPE(CV_32F, 3)
#define CV_64FC1 CV_MAKETYPE(CV_64F, 1)

typedef unsigned char uchar;

namespace cv {

class Mat {
 public:
  int rows = 0, cols = 0;
  uchar* data = nullptr;
  struct Step {
    size_t s[2] = {0, 0};
    size_t operator[](int i) const { return s[i]; }
  } step;

  Mat() = default;
  Mat(int r, int c, int type) { create(r, c, type); }
  // a view of caller memory with a row pitch of `pitch` bytes
  Mat(int r, int c, int type, void* ext, size_t pitch) : rows(r), cols(c), type_(type) {
    data = static_cast<uchar*>(ext);
    step.s[0] = pitch;
    step.s[1] = elemSize();
  }

  int type() const { return type_; }
  int depth() const { return type_ & 7; }
  int channels() const { return (type_ >> 3) + 1; }
  size_t elemSize() const {
    static const size_t sz[8] = {1, 1, 2, 2, 4, 4, 8, 8};
    return sz[depth()] * channels();
  }
  bool empty() const { return data == nullptr || rows == 0 || cols == 0; }
  bool isContinuous() const { return step.s[0] == (size_t)cols * elemSize(); }

  void create(int r, int c, int type) {
    if (buf_ && r == rows && c == cols && type == type_) return;
    rows = r;
    cols = c;
    type_ = type;
    buf_.reset(new uchar[(size_t)r * c * elemSize()], std::default_delete<uchar[]>());
    data = buf_.get();
    step.s[0] = (size_t)c * elemSize();
    step.s[1] = elemSize();
  }
  Mat clone() const {
    Mat m(rows, cols, type_);
    for (int i = 0; i < rows; ++i)
      std::memcpy(m.data + i * m.step.s[0], data + i * step.s[0], (size_t)cols * elemSize());
    return m;
  }
  template <typename T> T* ptr(int row = 0) { return reinterpret_cast<T*>(data + row * step.s[0]); }
  template <typename T> const T* ptr(int row = 0) const {
    return reinterpret_cast<const T*>(data + row * step.s[0]);
  }

 private:
  int type_ = 0;
  std::shared_ptr<uchar> buf_;
};

}  // namespace cv
