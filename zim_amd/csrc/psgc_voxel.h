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
// Cell coordinates are computed in double: the difference of two floats is
// exact there (for any two coordinates within 2^29 of each other's
// magnitude), and the roundings of 1 / voxel and of the product move a
// coordinate by at most 2^-52 of its value in cells: under 1e-9 cells up to
// 4.5 M cells along an axis, and under 2^-22 cells for the 2^30 cells per axis
// that voxel_frame allows. A point therefore lies within 2^-22 cells of the
// cell it is assigned to, which the margins below exceed; in float the error
// grows with the extent (about 1e-4 cells across 1000 cells) and would
// exceed them.
struct VoxelFrame {
  float ox, oy, oz;
  double inv;    // 1 / voxel
  float voxel;
  int64_t X, Y, Z;
};

PSGC_VX_HD inline int64_t voxel_coord(float v, float o, double inv) {
  return (int64_t)floor(((double)v - (double)o) * inv);
}

PSGC_VX_HD inline uint64_t voxel_key(const VoxelFrame& f, int64_t ix, int64_t iy, int64_t iz) {
  return (uint64_t)(ix + f.X * (iy + f.Y * iz));
}

PSGC_VX_HD inline uint64_t voxel_key_of(const VoxelFrame& f, float x, float y, float z) {
  return voxel_key(f, voxel_coord(x, f.ox, f.inv), voxel_coord(y, f.oy, f.inv),
                   voxel_coord(z, f.oz, f.inv));
}

// Frame for a bounding box [mn, mx] (host side). Returns false when the
// cell count would exceed 2^40, or 2^30 along one axis (voxel too small for
// the extent).
inline bool voxel_frame(const float mn[3], const float mx[3], float voxel, VoxelFrame* f) {
  if (!(voxel > 0.0f)) return false;
  f->voxel = voxel;
  f->inv = 1.0 / (double)voxel;
  f->ox = mn[0] - 2.0f * voxel;
  f->oy = mn[1] - 2.0f * voxel;
  f->oz = mn[2] - 2.0f * voxel;
  f->X = voxel_coord(mx[0], f->ox, f->inv) + 3;
  f->Y = voxel_coord(mx[1], f->oy, f->inv) + 3;
  f->Z = voxel_coord(mx[2], f->oz, f->inv) + 3;
  const double cells = (double)f->X * (double)f->Y * (double)f->Z;
  const int64_t axis = (int64_t)1 << 30;
  return cells < 1099511627776.0 && f->X < axis && f->Y < axis && f->Z < axis;
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
// outside that block lies at least R voxels away (less 2^-21 voxels of cell
// assignment error), so a k-th squared distance at or below this limit (R
// voxels less a relative margin of 1e-5, above that error and the float
// rounding of the squared distances, about 4e-7) makes the k found the exact
// k nearest.
PSGC_VX_HD inline float cloud_knn_limit(float voxel, int R) {
  const float rs = voxel * (float)R;
  return rs * rs * 0.99999f;
}

// The same bound for one point of the 3x3x3 search: a point outside the
// block lies at least voxel * (1 + m) away, m being the point's distance in
// cells to the nearest face of its own cell (0 to 1/2), so a k-th squared
// distance at or below (voxel (1 + m))^2 less a relative margin of 4e-5
// (above the error of m, at most 2^-52 of the point's coordinate in cells,
// and the float rounding of the distances) settles the point as well.
PSGC_VX_HD inline float cloud_knn_point_limit(const VoxelFrame& f, float x, float y, float z) {
  const double c[3] = {((double)x - (double)f.ox) * f.inv, ((double)y - (double)f.oy) * f.inv,
                       ((double)z - (double)f.oz) * f.inv};
  double m = 0.5;
  for (int a = 0; a < 3; ++a) {
    const double t = c[a] - floor(c[a]);
    m = fmin(m, fmin(t, 1.0 - t));
  }
  const double b = (double)f.voxel * (1.0 + m);
  return (float)(b * b * (1.0 - 4e-5));
}

// Cells around a point's own cell, per axis, that hold every neighbour
// within radius r for the radius filters (MLS, Gaussian): 1 when voxel >= r
// (a 3x3x3 block), 2 when 2 voxel >= r (5x5x5: half-radius cells cover
// 15.6 r^3 instead of 27 r^3); 0 = voxel too small. A neighbour passes the
// float test d^2 <= r^2 up to r (1 + 2^-22) away and sits within 2^-22 cells
// of its cell, so the cell edge used is at least r (1 + 2^-20) (or half of
// it): *voxel is raised to that when below it (by under a millionth; the
// neighbour sets do not depend on the cell edge).
PSGC_VX_HD inline int cloud_radius_cells(float* voxel, float r) {
  const float need = r + r * 0x1p-20f;
  if (*voxel >= r) {
    *voxel = *voxel >= need ? *voxel : need;
    return 1;
  }
  if (2.0f * *voxel >= r) {
    *voxel = 2.0f * *voxel >= need ? *voxel : 0.5f * need;
    return 2;
  }
  return 0;
}

}  // namespace psgc
