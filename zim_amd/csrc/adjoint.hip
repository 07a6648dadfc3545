// adjoint.hip — differentiable rectification + sub-pixel root + Q
// triangulation, and its hand-written adjoint.
//
// Model (no reference counterpart; it differentiates the reference's
// sub-pixel root, JSLAB_PhaseEpipolarReconstructor_V2.cpp:992-1003, and Q
// reprojection, :430-438, through a calibration-dependent rectification):
//   theta = (f, cx1, cy1, cx2, cy2, B, wx, wy, wz). The raw phase maps were
//   taken by cameras with intrinsics (f0, c0L) and (f0, c0R) (`nominal`).
//   Left rectified pixel (j, i) samples the raw left map at
//     u = f0/f (j - cx1) + c0Lx,  v = f0/f (i - cy1) + c0Ly;
//   right rectified pixel (x, i) samples the raw right map at the
//   projection of R(w) K_R^-1 (x, i, 1), R(w) = Rz(wz) Ry(wy) Rx(wx),
//   K_R = [f 0 cx2; 0 f cy2; 0 0 1], with the raw intrinsics (f0, c0R).
//   Catmull-Rom bicubic sampling (C1 in theta); taps outside the map
//   invalidate the point.
//   With the integer match b fixed, c_m = phiR(i, b-1+m) - phiL(i, j),
//   delta = c1 (c0 - c2) / (2 (c0 - c1)(c1 - c2)), d = j - b - delta and
//   X = (j - cx1) B / D, Y = (i - cy1) B / D, Z = f B / D, D = d - (cx1 - cx2)
//   (Q with R1 = I). Everything is FP64.
//   valid = 1 marks a triangulated point whose samples all satisfy
//   Itoh's condition (adjacent unwrapped phase taps differ by less than pi:
//   no fringe-order error, surface edge or region boundary inside a
//   sample), 2 a triangulated point with a sample across such a jump (its
//   coordinates are not a smooth function of theta there), 0 no point.
// The adjoint accumulates dL/dtheta (per-workgroup reduction, then f64
// atomics) and scatters dL/dphi into the raw maps with f64 atomics.

#include <vector>

#include "psgc_internal.h"

namespace psgc {
namespace {

struct Theta {
  double f, cx1, cy1, cx2, cy2, B, wx, wy, wz;
};
struct Nominal {
  double f0, cLx, cLy, cRx, cRy;
};

__device__ __attribute__((always_inline)) inline void rot(const Theta& t, double R[3][3],
                                                         double dR[3][3][3]) {
  const double cx = cos(t.wx), sx = sin(t.wx), cy = cos(t.wy), sy = sin(t.wy), cz = cos(t.wz),
               sz = sin(t.wz);
  const double Rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
  const double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
  const double Rz[3][3] = {{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}};
  const double dRx[3][3] = {{0, 0, 0}, {0, -sx, -cx}, {0, cx, -sx}};
  const double dRy[3][3] = {{-sy, 0, cy}, {0, 0, 0}, {-cy, 0, -sy}};
  const double dRz[3][3] = {{-sz, -cz, 0}, {cz, -sz, 0}, {0, 0, 0}};
  auto mul3 = [](const double A[3][3], const double Bm[3][3], const double Cm[3][3],
                 double O[3][3]) {
    double T[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += Bm[i][k] * Cm[k][j];
        T[i][j] = s;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += A[i][k] * T[k][j];
        O[i][j] = s;
      }
  };
  mul3(Rz, Ry, Rx, R);
  mul3(Rz, Ry, dRx, dR[0]);
  mul3(Rz, dRy, Rx, dR[1]);
  mul3(dRz, Ry, Rx, dR[2]);
}

// Catmull-Rom cubic weights of taps -1, 0, 1, 2 at fraction t, and their
// derivatives: the interpolant is C1 across pixel boundaries, so the
// triangulated points are continuously differentiable in theta (a bilinear
// sample has a kink at every pixel boundary, which central differences of
// the calibration objective straddle).
__device__ inline void catmull_rom(double t, double w[4], double dw[4]) {
  const double t2 = t * t, t3 = t2 * t;
  w[0] = 0.5 * (-t3 + 2.0 * t2 - t);
  w[1] = 0.5 * (3.0 * t3 - 5.0 * t2 + 2.0);
  w[2] = 0.5 * (-3.0 * t3 + 4.0 * t2 + t);
  w[3] = 0.5 * (t3 - t2);
  dw[0] = 0.5 * (-3.0 * t2 + 4.0 * t - 1.0);
  dw[1] = 0.5 * (9.0 * t2 - 10.0 * t);
  dw[2] = 0.5 * (-9.0 * t2 + 8.0 * t + 1.0);
  dw[3] = 0.5 * (3.0 * t2 - 2.0 * t);
}

constexpr double kPi = 3.14159265358979323846;

// Bicubic sample of map m at (u, v) over the 4 x 4 taps around it; false if
// a tap would lie outside the map (u in [1, cols - 2], v in [1, rows - 2]).
struct Samp {
  double val, du, dv;  // value and its partial derivatives in u, v
  double a, c;         // fractions within the tap cell
  int x0, y0;          // tap (x0 - 1 .. x0 + 2, y0 - 1 .. y0 + 2)
  bool smooth;         // Itoh's condition: adjacent taps within pi
};

__device__ inline bool bicubic(const double* m, int rows, int cols, double u, double v, Samp* s) {
  if (rows < 4 || cols < 4) return false;
  if (!(u >= 1.0 && u <= (double)(cols - 2) && v >= 1.0 && v <= (double)(rows - 2))) return false;
  int x0 = (int)floor(u), y0 = (int)floor(v);
  if (x0 > cols - 3) x0 = cols - 3;
  if (y0 > rows - 3) y0 = rows - 3;
  s->x0 = x0;
  s->y0 = y0;
  s->a = u - (double)x0;
  s->c = v - (double)y0;
  double wx[4], dwx[4], wy[4], dwy[4];
  catmull_rom(s->a, wx, dwx);
  catmull_rom(s->c, wy, dwy);
  double val = 0.0, du = 0.0, dv = 0.0;
  // Itoh's condition as one running maximum of the adjacent-tap differences
  // (with the ROCm 7.2 compiler, a chain of && over the sixteen tap
  // comparisons folded to a constant for gfx950 at -O2 and above)
  double jump = 0.0, nan_chk = 0.0;
  double prev[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = 0; j < 4; ++j) {
    const double* row = m + (size_t)(y0 - 1 + j) * cols + (x0 - 1);
    const double p[4] = {row[0], row[1], row[2], row[3]};
    double rv = 0.0, rd = 0.0;
    for (int i = 0; i < 4; ++i) {
      val += wy[j] * wx[i] * p[i];
      rv += wx[i] * p[i];
      rd += dwx[i] * p[i];
      if (i > 0) {
        const double d = p[i] - p[i - 1];
        jump = fmax(jump, fabs(d));
        nan_chk += d - d;  // NaN unless d is finite
      }
      if (j > 0) {
        const double d = p[i] - prev[i];
        jump = fmax(jump, fabs(d));
        nan_chk += d - d;
      }
      prev[i] = p[i];
    }
    du += wy[j] * rd;
    dv += dwy[j] * rv;
  }
  // fmax keeps the other operand of a NaN, so non-finite differences are
  // caught by nan_chk (every comparison with a NaN difference fails)
  const bool smooth = jump < kPi && nan_chk == 0.0;
  s->val = val;
  s->du = du;
  s->dv = dv;
  s->smooth = smooth;
  return true;
}

// Adds g times each tap's weight of sample s into the gradient map gm.
__device__ inline void bicubic_scatter(const Samp& s, int cols, double g, double* gm) {
  double wx[4], dwx[4], wy[4], dwy[4];
  catmull_rom(s.a, wx, dwx);
  catmull_rom(s.c, wy, dwy);
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i)
      atomicAdd(&gm[(size_t)(s.y0 - 1 + j) * cols + (s.x0 - 1 + i)], g * (wy[j] * wx[i]));
}

// Right remap of rectified (x, i): raw coordinates (right_uv) and their
// derivatives with respect to theta (indices 0..8; right_map, the same
// coordinates and their derivatives).
__device__ inline void right_uv(const Theta& t, const Nominal& n, const double (&R)[3][3],
                                double x, double i, double r[3], double q[3], double* u,
                                double* v) {
  r[0] = (x - t.cx2) / t.f;
  r[1] = (i - t.cy2) / t.f;
  r[2] = 1.0;
  for (int a = 0; a < 3; ++a) q[a] = R[a][0] * r[0] + R[a][1] * r[1] + R[a][2] * r[2];
  *u = n.f0 * q[0] / q[2] + n.cRx;
  *v = n.f0 * q[1] / q[2] + n.cRy;
}

__device__ inline void right_map(const Theta& t, const Nominal& n, const double (&R)[3][3],
                                 const double (&dR)[3][3][3], double x, double i, double* u,
                                 double* v, double du[9], double dv[9]) {
  double r[3], q[3];
  right_uv(t, n, R, x, i, r, q, u, v);
  // d(u, v)/dq
  const double iz = 1.0 / q[2];
  const double duq[3] = {n.f0 * iz, 0.0, -n.f0 * q[0] * iz * iz};
  const double dvq[3] = {0.0, n.f0 * iz, -n.f0 * q[1] * iz * iz};
  for (int k = 0; k < 9; ++k) du[k] = dv[k] = 0.0;
  // dr/df = -(r0, r1, 0)/f, dr/dcx2 = (-1/f, 0, 0), dr/dcy2 = (0, -1/f, 0)
  const double drs[3][3] = {{-r[0] / t.f, -r[1] / t.f, 0.0}, {-1.0 / t.f, 0, 0}, {0, -1.0 / t.f, 0}};
  const int idx[3] = {0, 3, 4};
  for (int s = 0; s < 3; ++s) {
    double dq[3];
    for (int a = 0; a < 3; ++a) dq[a] = R[a][0] * drs[s][0] + R[a][1] * drs[s][1] + R[a][2] * drs[s][2];
    du[idx[s]] = duq[0] * dq[0] + duq[1] * dq[1] + duq[2] * dq[2];
    dv[idx[s]] = dvq[0] * dq[0] + dvq[1] * dq[1] + dvq[2] * dq[2];
  }
  for (int s = 0; s < 3; ++s) {
    double dq[3];
    for (int a = 0; a < 3; ++a)
      dq[a] = dR[s][a][0] * r[0] + dR[s][a][1] * r[1] + dR[s][a][2] * r[2];
    du[6 + s] = duq[0] * dq[0] + duq[1] * dq[1] + duq[2] * dq[2];
    dv[6 + s] = dvq[0] * dq[0] + dvq[1] * dq[1] + dvq[2] * dq[2];
  }
}

// A pixel's triangulated point and what its adjoint needs of it (the
// derivatives of the right samples' coordinates are recomputed where they
// are consumed, one sample at a time, so they are never all live together).
struct PointEval {
  bool ok;
  bool smooth;  // every sample's taps satisfy Itoh's condition
  double xyz[3];
  // for the adjoint (GRAD)
  double dd_dc[3];  // d d / d c_m
  double dxyz_dd[3];
  double iD, iD2;   // 1 / D, 1 / D^2
  Samp bl, br[3];
};

template <bool GRAD>
__device__ inline void eval_point(const Theta& t, const Nominal& n, const double (&R)[3][3],
                                  const double* phL, const double* phR, int rows, int cols, int i,
                                  int j, int b, PointEval* e) {
  e->ok = false;
  if (b < 1 || b > cols - 2) return;
  const double ul = n.f0 / t.f * (j - t.cx1) + n.cLx;
  const double vl = n.f0 / t.f * (i - t.cy1) + n.cLy;
  if (!bicubic(phL, rows, cols, ul, vl, &e->bl)) return;
  double c[3];
  for (int m = 0; m < 3; ++m) {
    double r[3], q[3], u, v;
    right_uv(t, n, R, (double)(b - 1 + m), (double)i, r, q, &u, &v);
    if (!bicubic(phR, rows, cols, u, v, &e->br[m])) return;
    c[m] = e->br[m].val - e->bl.val;
  }
  const double N = c[1] * (c[0] - c[2]);
  const double Dn = 2.0 * (c[0] - c[1]) * (c[1] - c[2]);
  if (!(fabs(Dn) > 1e-9)) return;
  const double delta = N / Dn;
  if (GRAD) {
    const double dN[3] = {c[1], c[0] - c[2], -c[1]};
    const double dD[3] = {2.0 * (c[1] - c[2]), 2.0 * (c[0] - 2.0 * c[1] + c[2]),
                          -2.0 * (c[0] - c[1])};
    for (int m = 0; m < 3; ++m) e->dd_dc[m] = -(dN[m] * Dn - N * dD[m]) / (Dn * Dn);
  }
  const double d = (double)j - (double)b - delta;
  const double D = d - (t.cx1 - t.cx2);
  if (!(D > 0.0)) return;
  const double X = (j - t.cx1) * t.B / D, Y = (i - t.cy1) * t.B / D, Z = t.f * t.B / D;
  e->xyz[0] = X;
  e->xyz[1] = Y;
  e->xyz[2] = Z;
  if (GRAD) {
    const double iD = 1.0 / D, iD2 = iD * iD;
    e->iD = iD;
    e->iD2 = iD2;
    e->dxyz_dd[0] = -(j - t.cx1) * t.B * iD2;
    e->dxyz_dd[1] = -(i - t.cy1) * t.B * iD2;
    e->dxyz_dd[2] = -t.f * t.B * iD2;
  }
  e->ok = true;
  e->smooth = e->bl.smooth && e->br[0].smooth && e->br[1].smooth && e->br[2].smooth;
}

__device__ inline void adj_forward_px(const Theta& t, const Nominal& n, const double (&R)[3][3],
                                      const double* phL, const double* phR, const int32_t* mc,
                                      int rows, int cols, size_t p, double* xyz, uint8_t* valid) {
  const int i = (int)(p / cols), j = (int)(p % cols);
  PointEval e;
  eval_point<false>(t, n, R, phL, phR, rows, cols, i, j, mc[p], &e);
  xyz[p * 3 + 0] = e.ok ? e.xyz[0] : 0.0;
  xyz[p * 3 + 1] = e.ok ? e.xyz[1] : 0.0;
  xyz[p * 3 + 2] = e.ok ? e.xyz[2] : 0.0;
  if (valid) valid[p] = e.ok ? (e.smooth ? 1 : 2) : 0;
}

// One pixel's contribution to dL/dtheta (acc) and its scatter into the raw
// phase gradients.
__device__ inline void adj_backward_px(const Theta& t, const Nominal& n,
                                       const double (&R)[3][3], const double (&dR)[3][3][3],
                                       const double* phL, const double* phR, const int32_t* mc,
                                       int rows, int cols, size_t p, const double* gxyz,
                                       double* gphL, double* gphR, double (&acc)[9]) {
  const int i = (int)(p / cols), j = (int)(p % cols);
  const int b = mc[p];
  PointEval e;
  eval_point<true>(t, n, R, phL, phR, rows, cols, i, j, b, &e);
  if (!e.ok) return;
  const double g[3] = {gxyz[p * 3], gxyz[p * 3 + 1], gxyz[p * 3 + 2]};
  const double gd = g[0] * e.dxyz_dd[0] + g[1] * e.dxyz_dd[1] + g[2] * e.dxyz_dd[2];
  {
    // direct dependence of the point on theta through Q
    const double iD = e.iD, iD2 = e.iD2;
    double dth[3][9];
    for (int a = 0; a < 3; ++a)
      for (int k = 0; k < 9; ++k) dth[a][k] = 0.0;
    dth[2][0] = t.B * iD;                                // f
    dth[0][1] = -t.B * iD + (j - t.cx1) * t.B * iD2;     // cx1: D depends on -cx1
    dth[1][1] = (i - t.cy1) * t.B * iD2;
    dth[2][1] = t.f * t.B * iD2;
    dth[1][2] = -t.B * iD;                               // cy1
    dth[0][3] = -(j - t.cx1) * t.B * iD2;                // cx2
    dth[1][3] = -(i - t.cy1) * t.B * iD2;
    dth[2][3] = -t.f * t.B * iD2;
    dth[0][5] = (j - t.cx1) * iD;                        // B
    dth[1][5] = (i - t.cy1) * iD;
    dth[2][5] = t.f * iD;
    for (int k = 0; k < 9; ++k) acc[k] = g[0] * dth[0][k] + g[1] * dth[1][k] + g[2] * dth[2][k];
  }
  double gL = 0.0;
  for (int m = 0; m < 3; ++m) {
    const double gc = gd * e.dd_dc[m];  // dL/dc_m
    gL -= gc;
    // through the right sample
    double u, v, dur[9], dvr[9];
    right_map(t, n, R, dR, (double)(b - 1 + m), (double)i, &u, &v, dur, dvr);
    for (int k = 0; k < 9; ++k) acc[k] += gc * (e.br[m].du * dur[k] + e.br[m].dv * dvr[k]);
    if (gphR) bicubic_scatter(e.br[m], cols, gc, gphR);
  }
  // the left sample's coordinates depend on f, cx1, cy1
  double dul[9], dvl[9];
  for (int k = 0; k < 9; ++k) dul[k] = dvl[k] = 0.0;
  dul[0] = -n.f0 / (t.f * t.f) * (j - t.cx1);
  dvl[0] = -n.f0 / (t.f * t.f) * (i - t.cy1);
  dul[1] = -n.f0 / t.f;
  dvl[2] = -n.f0 / t.f;
  for (int k = 0; k < 9; ++k) acc[k] += gL * (e.bl.du * dul[k] + e.bl.dv * dvl[k]);
  if (gphL) bicubic_scatter(e.bl, cols, gL, gphL);
}

// Workgroup sum of acc, added to gth (f64 atomics).
__device__ inline void reduce_theta(double (&acc)[9], double* gth) {
  __shared__ double red[9][256];
  for (int k = 0; k < 9; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int h = 128; h >= 1; h /= 2) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 9; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < 9) atomicAdd(&gth[threadIdx.x], red[threadIdx.x][0]);
}

// The rotation of theta's angles and its derivatives, computed once per
// workgroup (one thread) into LDS: uniform for every pixel of a view.
struct Rot {
  double R[3][3];
  double dR[3][3][3];
};

__device__ __attribute__((always_inline)) inline void rot_shared(const Theta& t, Rot* r) {
  if (threadIdx.x == 0) rot(t, r->R, r->dR);
  __syncthreads();
}

__global__ void __launch_bounds__(256) adj_forward_kernel(Theta t, Nominal n,
                                                          const double* __restrict__ phL,
                                                          const double* __restrict__ phR,
                                                          const int32_t* __restrict__ mc, int rows,
                                                          int cols, double* __restrict__ xyz,
                                                          uint8_t* __restrict__ valid) {
  __shared__ Rot rs;
  rot_shared(t, &rs);
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (size_t)rows * cols) return;
  adj_forward_px(t, n, rs.R, phL, phR, mc, rows, cols, p, xyz, valid);
}

__global__ void __launch_bounds__(256) adj_backward_kernel(
    Theta t, Nominal n, const double* __restrict__ phL, const double* __restrict__ phR,
    const int32_t* __restrict__ mc, int rows, int cols, const double* __restrict__ gxyz,
    double* __restrict__ gth, double* __restrict__ gphL, double* __restrict__ gphR) {
  __shared__ Rot rs;
  rot_shared(t, &rs);
  const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (p < (size_t)rows * cols)
    adj_backward_px(t, n, rs.R, rs.dR, phL, phR, mc, rows, cols, p, gxyz, gphL, gphR, acc);
  reduce_theta(acc, gth);
}

// ---- many views in one launch ---------------------------------------------
// Each view owns a run of whole 256-thread workgroups (block_base = the
// first), so a workgroup's view, and its rotation, are uniform.

struct AdjView {
  Theta t;
  Nominal n;
  const double* phL;
  const double* phR;
  const int32_t* mc;
  int rows, cols;
  double* xyz;
  uint8_t* valid;
  const double* gxyz;
  double* gth;
  double* gphL;
  double* gphR;
  int64_t block_base;
};

__device__ inline int view_of_block(const AdjView* v, int nv, int64_t b) {
  int lo = 0, hi = nv - 1;  // last view with block_base <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (v[mid].block_base <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(256) adj_forward_batch_kernel(const AdjView* __restrict__ views,
                                                                int nv) {
  __shared__ Rot rs;
  const AdjView& V = views[view_of_block(views, nv, blockIdx.x)];
  rot_shared(V.t, &rs);
  const size_t p = (size_t)(blockIdx.x - V.block_base) * 256 + threadIdx.x;
  if (p >= (size_t)V.rows * V.cols) return;
  adj_forward_px(V.t, V.n, rs.R, V.phL, V.phR, V.mc, V.rows, V.cols, p, V.xyz, V.valid);
}

__global__ void __launch_bounds__(256) adj_backward_batch_kernel(
    const AdjView* __restrict__ views, int nv) {
  __shared__ Rot rs;
  const AdjView& V = views[view_of_block(views, nv, blockIdx.x)];
  rot_shared(V.t, &rs);
  const size_t p = (size_t)(blockIdx.x - V.block_base) * 256 + threadIdx.x;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (p < (size_t)V.rows * V.cols)
    adj_backward_px(V.t, V.n, rs.R, rs.dR, V.phL, V.phR, V.mc, V.rows, V.cols, p, V.gxyz, V.gphL,
                    V.gphR, acc);
  reduce_theta(acc, V.gth);
}

__global__ void zero_f64_kernel(double* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

}  // namespace

int launch_adj_forward(const double* theta, const double* nominal, const double* phL,
                       const double* phR, const int32_t* mc, int rows, int cols, double* xyz,
                       uint8_t* valid, hipStream_t s) {
  Theta t{theta[0], theta[1], theta[2], theta[3], theta[4], theta[5], theta[6], theta[7], theta[8]};
  Nominal n{nominal[0], nominal[1], nominal[2], nominal[3], nominal[4]};
  const size_t np = (size_t)rows * cols;
  PSGC_PROF_BEGIN("adj_forward", s);
  adj_forward_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(t, n, phL, phR, mc, rows, cols,
                                                                  xyz, valid);
  PSGC_LAUNCHED();
  PSGC_PROF_END("adj_forward", s);
  return PSGC_OK;
}

int launch_adj_backward(const double* theta, const double* nominal, const double* phL,
                        const double* phR, const int32_t* mc, int rows, int cols,
                        const double* gxyz, double* gth, double* gphL, double* gphR,
                        hipStream_t s) {
  Theta t{theta[0], theta[1], theta[2], theta[3], theta[4], theta[5], theta[6], theta[7], theta[8]};
  Nominal n{nominal[0], nominal[1], nominal[2], nominal[3], nominal[4]};
  const size_t np = (size_t)rows * cols;
  PSGC_PROF_BEGIN("adj_backward", s);
  zero_f64_kernel<<<1, 64, 0, s>>>(gth, 9);
  PSGC_LAUNCHED();
  if (gphL) {
    zero_f64_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(gphL, np);
    PSGC_LAUNCHED();
  }
  if (gphR) {
    zero_f64_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(gphR, np);
    PSGC_LAUNCHED();
  }
  adj_backward_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(t, n, phL, phR, mc, rows, cols,
                                                                   gxyz, gth, gphL, gphR);
  PSGC_LAUNCHED();
  PSGC_PROF_END("adj_backward", s);
  return PSGC_OK;
}

// Batched forward/backward over views (device pointers); zeroes the
// gradient outputs first (backward).
int launch_adj_batch(psgc_ctx* ctx, const psgc_adjoint_view* views, int n, bool backward) {
  if (n < 1) PSGC_FAIL(PSGC_E_INVALID, "no views");
  hipStream_t s = ctx->stream;
  std::vector<AdjView> hv(n);
  int64_t blocks = 0;
  for (int v = 0; v < n; ++v) {
    const psgc_adjoint_view& w = views[v];
    if (w.rows < 1 || w.cols < 1 || (int64_t)w.rows * w.cols > ((int64_t)1 << 31))
      PSGC_FAIL(PSGC_E_INVALID, "view %d: size %d x %d", v, w.rows, w.cols);
    if (!w.phL || !w.phR || !w.match_col || !(w.theta[0] > 0) || !(w.nominal[0] > 0))
      PSGC_FAIL(PSGC_E_INVALID, "view %d: null map or non-positive focal", v);
    if (!backward && !w.xyz) PSGC_FAIL(PSGC_E_INVALID, "view %d: null xyz", v);
    if (backward && (!w.grad_xyz || !w.grad_theta))
      PSGC_FAIL(PSGC_E_INVALID, "view %d: null grad_xyz or grad_theta", v);
    AdjView& a = hv[v];
    a.t = Theta{w.theta[0], w.theta[1], w.theta[2], w.theta[3], w.theta[4],
                w.theta[5], w.theta[6], w.theta[7], w.theta[8]};
    a.n = Nominal{w.nominal[0], w.nominal[1], w.nominal[2], w.nominal[3], w.nominal[4]};
    a.phL = w.phL;
    a.phR = w.phR;
    a.mc = w.match_col;
    a.rows = w.rows;
    a.cols = w.cols;
    a.xyz = w.xyz;
    a.valid = w.valid;
    a.gxyz = w.grad_xyz;
    a.gth = w.grad_theta;
    a.gphL = w.grad_phL;
    a.gphR = w.grad_phR;
    a.block_base = blocks;
    blocks += ((int64_t)w.rows * w.cols + 255) / 256;
  }
  void* d;
  PSGC_TRY(ctx_buf(ctx, S_ADJV, sizeof(AdjView) * n, &d));
  PSGC_HIP(hipMemcpyAsync(d, hv.data(), sizeof(AdjView) * n, hipMemcpyHostToDevice, s));
  if (backward) {
    for (int v = 0; v < n; ++v) {
      const size_t np = (size_t)hv[v].rows * hv[v].cols;
      zero_f64_kernel<<<1, 64, 0, s>>>(hv[v].gth, 9);
      PSGC_LAUNCHED();
      if (hv[v].gphL) {
        zero_f64_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(hv[v].gphL, np);
        PSGC_LAUNCHED();
      }
      if (hv[v].gphR) {
        zero_f64_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(hv[v].gphR, np);
        PSGC_LAUNCHED();
      }
    }
    PSGC_PROF_BEGIN("adj_backward_batch", s);
    adj_backward_batch_kernel<<<(unsigned)blocks, 256, 0, s>>>((const AdjView*)d, n);
    PSGC_LAUNCHED();
    PSGC_PROF_END("adj_backward_batch", s);
  } else {
    PSGC_PROF_BEGIN("adj_forward_batch", s);
    adj_forward_batch_kernel<<<(unsigned)blocks, 256, 0, s>>>((const AdjView*)d, n);
    PSGC_LAUNCHED();
    PSGC_PROF_END("adj_forward_batch", s);
  }
  // hv is pageable host memory that ends with this call: wait for the copy
  // (and the launch) before returning
  PSGC_HIP(hipStreamSynchronize(s));
  return PSGC_OK;
}

}  // namespace psgc
