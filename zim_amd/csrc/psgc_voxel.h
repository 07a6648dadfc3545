// Voxel frame of an unorganized point cloud: the cell of a point and the
// linear cell key, shared by the device kernels and the CPU restatement so
// that both put every point in the same cell (the same float operations).
#pragma once
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define PSGC_VX_HD __host__ __device__
#else
#define PSGC_VX_HD
#endif

namespace psgc {

// Cells of edge `voxel` from the origin (the bounding box's minimum less two
// cells; two empty cells past the maximum), so every occupied cell has a full
// ring of neighbours inside the frame, and a block cell two away that falls
// outside has a linear key (key + offset) of an empty margin cell or of no
// cell at all; key = ix + X (iy + Y iz).
struct VoxelFrame {
  float ox, oy, oz;
  float inv;     // 1 / voxel
  float voxel;
  int64_t X, Y, Z;
};

PSGC_VX_HD inline int64_t voxel_coord(float v, float o, float inv) {
  return (int64_t)floorf((v - o) * inv);
}

PSGC_VX_HD inline uint64_t voxel_key(const VoxelFrame& f, int64_t ix, int64_t iy, int64_t iz) {
  return (uint64_t)(ix + f.X * (iy + f.Y * iz));
}

PSGC_VX_HD inline uint64_t voxel_key_of(const VoxelFrame& f, float x, float y, float z) {
  return voxel_key(f, voxel_coord(x, f.ox, f.inv), voxel_coord(y, f.oy, f.inv),
                   voxel_coord(z, f.oz, f.inv));
}

// Frame for a bounding box [mn, mx] (host side). Returns false when the
// cell count would exceed 2^40 (voxel too small for the extent).
inline bool voxel_frame(const float mn[3], const float mx[3], float voxel, VoxelFrame* f) {
  if (!(voxel > 0.0f)) return false;
  f->voxel = voxel;
  f->inv = 1.0f / voxel;
  f->ox = mn[0] - 2.0f * voxel;
  f->oy = mn[1] - 2.0f * voxel;
  f->oz = mn[2] - 2.0f * voxel;
  f->X = voxel_coord(mx[0], f->ox, f->inv) + 3;
  f->Y = voxel_coord(mx[1], f->oy, f->inv) + 3;
  f->Z = voxel_coord(mx[2], f->oz, f->inv) + 3;
  const double cells = (double)f->X * (double)f->Y * (double)f->Z;
  return cells < 1099511627776.0;
}

// Number of key bits a radix sort of the cell keys needs.
inline int voxel_key_bits(const VoxelFrame& f) {
  const uint64_t top = (uint64_t)(f.X * f.Y * f.Z);
  int b = 1;
  while (b < 64 && (top >> b) != 0) ++b;
  return b;
}

// A point takes part when it is finite (and, with a mask, marked valid).
PSGC_VX_HD inline bool voxel_point_ok(float x, float y, float z) {
  return x == x && y == y && z == z && x - x == 0.0f && y - y == 0.0f && z - z == 0.0f;
}

// k-nearest search over the (2R+1)^3 cells around a point: every point
// outside that block lies at least R voxels away, so a k-th squared distance
// at or below this limit (R voxels less a margin far above float rounding)
// makes the k found the exact k nearest.
PSGC_VX_HD inline float cloud_knn_limit(float voxel, int R) {
  const float rs = voxel * (float)R;
  return rs * rs * 0.99999f;
}

// Cells of edge voxel around a point's own cell, per axis, that hold every
// neighbour within radius r for the radius filters (MLS, Gaussian): 1 when
// voxel >= r (a 3x3x3 block), 2 when 2 voxel >= r (5x5x5: half-radius cells
// cover 15.6 r^3 instead of 27 r^3); 0 = voxel too small.
PSGC_VX_HD inline int cloud_radius_cells(float voxel, float r) {
  return voxel >= r ? 1 : (2.0f * voxel >= r ? 2 : 0);
}

}  // namespace psgc
