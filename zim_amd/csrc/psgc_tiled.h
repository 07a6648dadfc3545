// psgc_tiled.h — work descriptors of the LDS-tiled, batch-capable kernels.
//
// Every tiled kernel walks a flat tile space that concatenates the tiles of
// all jobs of a batch (one job = one image of one stereo pair). Descriptors
// live in device memory; a launch covers the whole batch, so one captured
// graph of a handful of launches processes any number of pairs of uneven
// size. The grid is persistent (a few workgroups per CU) and XCD-aware:
// workgroups that the hardware places on the same XCD (blockIdx % 8) own one
// contiguous range of tiles, so the halos that neighbouring tiles share stay
// in that XCD's L2.
#pragma once

#include "psgc_internal.h"

namespace psgc {

constexpr int kNumXcd = 8;

// Decode of one camera image: the frames and configuration are in `a`.
struct DecodeJob {
  DecodeArgs a;
  double* phase;    // raw (pre post-smooth) absolute phase
  int32_t* k;
  double* mod;      // may be null
  uint8_t* valid;   // may be null
  int tiles_x, tiles_y;
  int64_t tile_base;
};

// Order-constrained match of one pair, one workgroup per row.
struct MatchJob {
  const double* lp;  // left phase (raw when smooth != 0)
  const double* rp;
  const int32_t* lk;
  const int32_t* rk;
  int rows, cols, pha_f, smooth;
  double grad_thr;
  float* disp;       // f32 disparity (0 = none)
  double* disp64;    // may be null
  int32_t* match_col;  // may be null
  int64_t tile_base; // first row index in the flat space
};

// Point-cloud stages of one pair on 16x16 tiles.
struct CloudJob {
  int rows, cols;
  int tiles_x, tiles_y;
  int64_t tile_base;
  // median
  const float* disp_in;
  float* disp_med;
  // reproject + Gaussian
  const float* disp_src;     // disparity to reproject (f32)
  const double* disp_src64;  // if set, reproject this f64 disparity instead
  double q[16], r1[9];
  const float4* coef[3];     // if set: per-pixel polynomial triangulation instead of q
  double min_depth, max_depth;
  int use_rect, semantics, gwin;
  float gk[49];
  float* xyz0;               // reprojected points (may be null)
  uint8_t* P0;               // reprojection mask (may be null)
  float* xyzg;               // Gaussian output
  uint8_t* Pg;               // Gaussian output mask (GPU semantics) or copy
  // SOR
  const float* sor_xyz;
  const uint8_t* sor_P;
  int sor_k, sor_half;
  double* md;                // mean distance, -1 where not eligible
  // MLS
  const float* mls_src;
  const uint8_t* mls_P;      // base mask
  const double* mls_md;      // if set: mask &= (md >= 0 && md <= *mls_thr)
  const double* mls_thr;
  float* mls_dst;
  uint8_t* mls_Pout;         // if set, receives the effective mask
  int mls_hw, mls_order, mls_sem;
  float mls_inv2s2, mls_r2;
};

// Flat tile -> job by forward linear search from `from`, the job of this
// workgroup's previous tile: a workgroup visits its tiles in increasing
// order, so over a whole launch it walks the job table once.
template <typename J>
__device__ inline int find_job(const J* jobs, int njobs, int64_t t, int from = 0) {
  int j = from;
  while (j + 1 < njobs && jobs[j + 1].tile_base <= t) ++j;
  return j;
}

// Tile range handled by this workgroup's XCD, and the first tile and
// stride of this workgroup inside it.
struct XcdRange {
  int64_t begin, end, first, stride;
};

__device__ inline XcdRange xcd_range(int64_t total) {
  const int nblk = gridDim.x;
  const int xcd = blockIdx.x % kNumXcd;
  const int per = (nblk + kNumXcd - 1) / kNumXcd;        // blocks per XCD
  const int nx = min(kNumXcd, nblk);
  const int64_t chunk = (total + nx - 1) / nx;
  XcdRange r;
  r.begin = min((int64_t)xcd * chunk, total);
  r.end = min(r.begin + chunk, total);
  r.first = r.begin + blockIdx.x / kNumXcd;
  r.stride = per;
  if (xcd >= nx) r.first = r.end;                         // idle (tiny grids)
  return r;
}

// Launchers (asynchronous on s; jobs in device memory).
// Decode tiles are kDecodeTileCols x tile_rows (tile_rows 2 or 4: output rows
// per thread); the jobs' tiles_x / tiles_y / tile_base must be laid out for it.
constexpr int kDecodeTileCols = 256;
int launch_decode_tiled(const DecodeJob* jobs_dev, int njobs, int64_t total_tiles, int grid,
                        int tile_rows, hipStream_t s);
// threads: the row tile, 128, 256 or 512 threads per row.
int launch_match_tiled(const MatchJob* jobs_dev, int njobs, int64_t total_rows, int grid,
                       int max_cols, int max_knum, int threads, hipStream_t s);
int launch_median_tiled(const CloudJob* jobs, int njobs, int64_t total, int grid, hipStream_t s);
int launch_repro_gauss_tiled(const CloudJob* jobs, int njobs, int64_t total, int grid,
                             int max_half, hipStream_t s);
int launch_sor_knn_tiled(const CloudJob* jobs, int njobs, int64_t total, int grid, int max_half,
                         int min_k, int max_k, hipStream_t s);
// MLS kernel variants: kMlsPlain evaluates each weight in all three passes;
// kMlsCached keeps the first in-radius weights of each point in LDS between
// the passes (the rest are evaluated again; classes above half window 10
// have no cache). Results are identical. Jobs are launched by window class:
// half windows up to 6, 8, 10, 12 and 14 (the class fixes the LDS tile, so
// small windows do not pay for the largest halo).
enum { kMlsPlain = 0, kMlsCached = 1 };
// Largest half windows the tiled SOR and MLS kernels stage in LDS.
constexpr int kSorTiledMaxHalf = 8;
constexpr int kMlsTiledMaxHalf = 14;
constexpr int kMlsClasses = 5;
constexpr int kMlsClassHalf[kMlsClasses] = {6, 8, 10, 12, 14};
inline int mls_class(int hw) {
  int c = 0;
  while (c < kMlsClasses - 1 && hw > kMlsClassHalf[c]) ++c;
  return c;
}
// Semantics present among the MLS jobs of a launch (a bit each): CPU
// semantics accumulate with rounded products and sums, GPU semantics with
// fused multiply-adds.
enum { kSemBitCpu = 1, kSemBitGpu = 2 };
inline int mls_sem_bit(int semantics) { return semantics == PSGC_SEM_GPU ? kSemBitGpu : kSemBitCpu; }
// One launch over jobs of one window class (every job's half window at most
// kMlsClassHalf[cls]).
int launch_mls_tiled(const CloudJob* jobs, int njobs, int64_t total, int grid, int cls,
                     int variant, int sems, hipStream_t s);

// Shared-memory bytes of the row-match kernel for a given width and order count.
size_t match_tiled_lds(int cols, int knum);
int match_tiled_max_lds();

}  // namespace psgc
