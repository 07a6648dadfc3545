// Compile-time index tables for the k-nearest selection of the outlier
// filter: the window's ring order and a sorting network. Plain C++17
// (constexpr), usable on the host and on the device.
#pragma once
#include <cstdint>

namespace psgc {

// Window entries in ring order: Chebyshev distance 1, 2, ..., each ring as
// four runs (top row left to right, right column downwards, bottom row right
// to left, left column upwards). Entry i is at offset (dy[i], dx[i]); rings
// 1..r hold the first (2r + 1)^2 - 1 entries.
constexpr int kRingMax = 8;
constexpr int kRingEntries = (2 * kRingMax + 1) * (2 * kRingMax + 1) - 1;
struct RingOrder {
  int8_t dy[kRingEntries], dx[kRingEntries];
};
constexpr RingOrder make_ring_order() {
  RingOrder o{};
  int i = 0;
  for (int rad = 1; rad <= kRingMax; ++rad) {
    for (int t = -rad; t < rad; ++t, ++i) { o.dy[i] = (int8_t)-rad; o.dx[i] = (int8_t)t; }
    for (int t = -rad; t < rad; ++t, ++i) { o.dy[i] = (int8_t)t; o.dx[i] = (int8_t)rad; }
    for (int t = rad; t > -rad; --t, ++i) { o.dy[i] = (int8_t)rad; o.dx[i] = (int8_t)t; }
    for (int t = rad; t > -rad; --t, ++i) { o.dy[i] = (int8_t)t; o.dx[i] = (int8_t)-rad; }
  }
  return o;
}
constexpr RingOrder kRing = make_ring_order();

// Batcher's odd-even merge sort network on 32 inputs, restricted to the
// first k: a comparator touching an index >= k is dropped (such inputs are
// +inf, the largest value, already in its final place, and a comparator
// would leave it there). Comparator c orders (a[c], b[c]), a[c] < b[c].
struct SortNet {
  uint8_t a[192], b[192];
  int n;
};
constexpr SortNet make_sort_net(int k) {
  SortNet s{};
  s.n = 0;
  constexpr int n = 32;
  for (int p = 1; p < n; p += p)
    for (int q = p; q > 0; q /= 2)
      for (int j = q % p; j + q < n; j += q + q)
        for (int i = 0; i < q && i + j + q < n; ++i)
          if ((i + j) / (p + p) == (i + j + q) / (p + p) && i + j + q < k) {
            s.a[s.n] = (uint8_t)(i + j);
            s.b[s.n] = (uint8_t)(i + j + q);
            ++s.n;
          }
  return s;
}

}  // namespace psgc
